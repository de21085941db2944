// lbm_oracle.hpp — CPU restatement of the reference's D2Q9 algorithm.
//
// TEST INFRASTRUCTURE ONLY. This is the correctness oracle. Only tests/, the smoke check
// in __graft_entry__.py and the cpu_baseline leg of bench.py may load or call it, and
// only as the checker. The product (liblbm_hip.so) never links or calls this code.
//
// It follows the reference's unfused stage order (src/main.cu:96-114):
//   stream -> swap -> boundaries -> uncorrected macroscopics -> reset forces -> IBM
//   -> Guo velocity correction (+ adapter averages) -> equilibrium -> collision,
// on separate buffers exactly as `LBM<2>` does (src/core/lbm.cuh:36-43), templated on
// the arithmetic type so the same code is the FP32 oracle and the FP64 "host run".
// Every function cites the reference lines it restates. Reference quirks are opt-in
// (lbm_config.quirks); the default is the corrected behaviour.
//
// Parity pinning: every function is checked by tests/test_oracle_pinning.py or
// tests/test_adjoint_cpu.py against something it does not restate (moment identities,
// closed forms, fixed points, operator equivalences, adjointness, finite differences).
// Parity unpinned: the exact FP32 rounding sequence of the reference's pow/double-literal
// equilibrium (equilibrium.cu:12-37), the float-atomic summation orders of
// macroscopics.cuh:116-118 and IBM_impl.cu:150-151 (nondeterministic in the reference),
// and device libm (sqrtf). All three act at FP32 rounding level.
#pragma once

#include "../include/lbm_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace oracle {

constexpr int Q = 9;
// Velocity set, order 0,E,N,W,S,NE,NW,SW,SE: lbm_constants.cuh:19-29
constexpr int CX[Q] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int CY[Q] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
// Opposite directions: lbm_constants.cuh:31
constexpr int OPP[Q] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
// Lallemand-Luo moment basis rows [rho, e, eps, jx, qx, jy, qy, pxx, pxy]: lbm_constants.cuh:33-43
constexpr int MM[Q * Q] = {
     1,  1,  1,  1,  1,  1,  1,  1,  1,
    -4, -1, -1, -1, -1,  2,  2,  2,  2,
     4, -2, -2, -2, -2,  1,  1,  1,  1,
     0,  1,  0, -1,  0,  1, -1, -1,  1,
     0, -2,  0,  2,  0,  1, -1, -1,  1,
     0,  0,  1,  0, -1,  1,  1, -1, -1,
     0,  0, -2,  0,  2,  1,  1, -1, -1,
     0,  1, -1,  1, -1,  0,  0,  0,  0,
     0,  0,  0,  0,  0,  1, -1,  1, -1};
// M^-1 as (numerator, denominator): lbm_constants.cuh:45-55
constexpr int MINV_NUM[Q * Q] = {
    1, -1,  1,  0,  0,  0,  0,  0,  0,
    1, -1, -1,  1, -1,  0,  0,  1,  0,
    1, -1, -1,  0,  0,  1, -1, -1,  0,
    1, -1, -1, -1,  1,  0,  0,  1,  0,
    1, -1, -1,  0,  0, -1,  1, -1,  0,
    1,  1,  1,  1,  1,  1,  1,  0,  1,
    1,  1,  1, -1, -1,  1,  1,  0, -1,
    1,  1,  1, -1, -1, -1, -1,  0,  1,
    1,  1,  1,  1,  1, -1, -1,  0, -1};
constexpr int MINV_DEN[Q * Q] = {
    9,  9,  9, 1,  1,  1,  1, 1, 1,
    9, 36, 18, 6,  6,  1,  1, 4, 1,
    9, 36, 18, 1,  1,  6,  6, 4, 1,
    9, 36, 18, 6,  6,  1,  1, 4, 1,
    9, 36, 18, 1,  1,  6,  6, 4, 1,
    9, 18, 36, 6, 12,  6, 12, 1, 4,
    9, 18, 36, 6, 12,  6, 12, 1, 4,
    9, 18, 36, 6, 12,  6, 12, 1, 4,
    9, 18, 36, 6, 12,  6, 12, 1, 4};

template <class R> inline R weight(int q) {  // lbm_constants.cuh:13-17
    if (q == 0) return R(4) / R(9);
    return q < 5 ? R(1) / R(9) : R(1) / R(36);
}
template <class R> inline R minv(int q, int k) {
    return R(MINV_NUM[q * Q + k]) / R(MINV_DEN[q * Q + k]);
}

// ---- reduced-precision storage (round to nearest even) ---------------------------------
inline uint16_t f32_to_f16(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    uint32_t sign = (u >> 16) & 0x8000u;
    uint32_t ex = (u >> 23) & 0xffu;
    uint32_t mant = u & 0x7fffffu;
    if (ex == 0xffu) return uint16_t(sign | 0x7c00u | (mant ? 0x200u : 0u));
    int e = int(ex) - 127 + 15;
    if (e >= 31) return uint16_t(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return uint16_t(sign);
        mant |= 0x800000u;
        int shift = 14 - e;
        uint32_t hm = mant >> shift;
        uint32_t rem = mant & ((1u << shift) - 1u);
        uint32_t halfway = 1u << (shift - 1);
        if (rem > halfway || (rem == halfway && (hm & 1u))) hm++;
        return uint16_t(sign | hm);
    }
    uint32_t hm = mant >> 13;
    uint32_t rem = mant & 0x1fffu;
    uint32_t h = sign | (uint32_t(e) << 10) | hm;
    if (rem > 0x1000u || (rem == 0x1000u && (hm & 1u))) h++;
    return uint16_t(h);
}
inline float f16_to_f32(uint16_t h) {
    uint32_t sign = uint32_t(h & 0x8000u) << 16;
    uint32_t ex = (h >> 10) & 0x1fu;
    uint32_t mant = h & 0x3ffu;
    uint32_t u;
    if (ex == 0) {
        if (mant == 0) {
            u = sign;
        } else {  // subnormal: normalise
            int e = -1;
            do { mant <<= 1; e++; } while (!(mant & 0x400u));
            u = sign | (uint32_t(127 - 15 - e) << 23) | ((mant & 0x3ffu) << 13);
        }
    } else if (ex == 31) {
        u = sign | 0x7f800000u | (mant << 13);
    } else {
        u = sign | ((ex - 15 + 127) << 23) | (mant << 13);
    }
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
inline uint16_t f32_to_bf16(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40u);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
inline float bf16_to_f32(uint16_t h) {
    uint32_t u = uint32_t(h) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// ---- node-local physics -----------------------------------------------------------------

// Second-order equilibrium. Default: w rho (1 + 3cu + 4.5cu^2 - 1.5u^2) in R.
// literal == true restates equilibrium.cu:5-39 as written: cs from 1/sqrt(3.0f) in float,
// pow() and double literals promote the bracket to double (quirk Q15).
template <class R>
inline void equilibrium(R rho, R ux, R uy, R* feq, bool literal) {
    if constexpr (std::is_floating_point<R>::value) {
      if (literal) {
        float cs = 1.0f / std::sqrt(3.0f);
        float cs2 = cs * cs;
        float cs4 = cs2 * cs2;
        float u_dot_u = float(ux) * float(ux) + float(uy) * float(uy);
        for (int q = 0; q < Q; q++) {
            float cu = float(CX[q]) * float(ux) + float(CY[q]) * float(uy);
            double br = 1 + 0.5 * std::pow(double(cu), 2) / cs4 - 0.5 * u_dot_u / cs2 + 1.0 * cu / cs2;
            feq[q] = R(float(weight<float>(q) * float(rho) * br));
        }
        return;
      }
    }
    R usq = ux * ux + uy * uy;
    for (int q = 0; q < Q; q++) {
        R cu = R(CX[q]) * ux + R(CY[q]) * uy;
        feq[q] = weight<R>(q) * rho * (R(1) + R(3) * cu + R(4.5) * cu * cu - R(1.5) * usq);
    }
}

// BGK<2>::apply with Guo forcing (BGK.cuh:13-51), node-local.
template <class R>
inline void collide_bgk(R* f, const R* feq, R ux, R uy, R fx, R fy, R omega) {
    const R cs2 = R(1) / R(3);
    for (int q = 0; q < Q; q++) {
        R cx = R(CX[q]), cy = R(CY[q]);
        R force_term = weight<R>(q) * ((R(1) - R(0.5) * omega) * ((cx - ux) / cs2 + (cx * ux + cy * uy) * cx / (cs2 * cs2)) * fx +
                                       (R(1) - R(0.5) * omega) * ((cy - uy) / cs2 + (cx * ux + cy * uy) * cy / (cs2 * cs2)) * fy);
        f[q] = f[q] - omega * (f[q] - feq[q]) + force_term;
    }
}

// MRT<2>::compute_forcing_term (MRT.cu:4-24). Default: the moments of Guo's source in
// the M basis; quirk Q10 keeps the reference's rows 4/5 (F[4]=Fy, F[5]=-Fx).
template <class R>
inline void mrt_forcing(R* F, R ux, R uy, R fx, R fy, bool literal_rows) {
    F[0] = R(0);
    F[1] = R(6) * (fx * ux + fy * uy);
    F[2] = R(-6) * (fx * ux + fy * uy);
    F[3] = fx;
    if (literal_rows) {
        F[4] = fy;
        F[5] = -fx;
    } else {
        F[4] = -fx;
        F[5] = fy;
    }
    F[6] = -fy;
    F[7] = R(2) * (fx * ux - fy * uy);
    F[8] = fx * uy + fy * ux;
}

// MRT<2>::apply (MRT.cu:26-76): m = M f, m_eq = M f_eq, relax, f = M^-1 m*.
template <class R>
inline void collide_mrt(R* f, const R* feq, R ux, R uy, R fx, R fy, const R* S, bool literal_rows) {
    R F[Q];
    mrt_forcing<R>(F, ux, uy, fx, fy, literal_rows);
    R m_post[Q];
    for (int k = 0; k < Q; k++) {
        R m = 0, meq = 0;
        for (int i = 0; i < Q; i++) {
            m += R(MM[k * Q + i]) * f[i];
            meq += R(MM[k * Q + i]) * feq[i];
        }
        R source = (R(1) - S[k] / R(2)) * F[k];
        m_post[k] = m - S[k] * (m - meq) + source;
    }
    for (int k = 0; k < Q; k++) {
        R acc = 0;
        for (int i = 0; i < Q; i++) acc += minv<R>(k, i) * m_post[i];
        f[k] = acc;
    }
}

// CM<2,A>::cm_matrix_inverse (CM.cuh:140-250): closed-form T^-1(u), row-major 9x9.
template <class R>
inline void cm_matrix_inverse(R* Mi, R ux, R uy) {
    R ux2 = ux * ux, uy2 = uy * uy, uxuy = ux * uy;
    R ux2uy = uy * (ux * ux), uxuy2 = ux * (uy * uy), ux2uy2 = (ux * ux) * (uy * uy);
    R x3 = -uxuy2, x5 = -uy;
    const R h = R(0.5), qr = R(0.25), e = R(0.125);
    Mi[0] = -ux2 + ux2uy2 - uy2 + R(1);
    Mi[1] = R(-2) * ux + R(2) * uxuy2;
    Mi[2] = R(2) * ux2uy + R(2) * x5;
    Mi[3] = h * ux2 + h * uy2 - R(1);
    Mi[4] = -h * ux2 + h * uy2;
    Mi[5] = R(4) * uxuy;
    Mi[6] = R(2) * uy;
    Mi[7] = R(2) * ux;
    Mi[8] = R(1);

    Mi[9] = h * ux + h * ux2 - h * ux2uy2 + h * x3;
    Mi[10] = ux - h * uy2 + x3 + h;
    Mi[11] = -ux2uy - uxuy;
    Mi[12] = -qr * ux - qr * ux2 - qr * uy2 + qr;
    Mi[13] = qr * ux + qr * ux2 - qr * uy2 + qr;
    Mi[14] = R(-2) * uxuy + x5;
    Mi[15] = x5;
    Mi[16] = -ux - h;
    Mi[17] = -h;

    Mi[18] = -h * ux2uy - h * ux2uy2 + h * uy + h * uy2;
    Mi[19] = -uxuy + x3;
    Mi[20] = -h * ux2 - ux2uy + uy + h;
    Mi[21] = -qr * ux2 - qr * uy2 + qr * x5 + qr;
    Mi[22] = qr * ux2 - qr * uy2 + qr * x5 - qr;
    Mi[23] = -ux - R(2) * uxuy;
    Mi[24] = x5 - h;
    Mi[25] = -ux;
    Mi[26] = -h;

    Mi[27] = -h * ux + h * ux2 - h * ux2uy2 + h * uxuy2;
    Mi[28] = ux + h * uy2 + x3 - h;
    Mi[29] = -ux2uy + uxuy;
    Mi[30] = qr * ux - qr * ux2 - qr * uy2 + qr;
    Mi[31] = -qr * ux + qr * ux2 - qr * uy2 + qr;
    Mi[32] = R(-2) * uxuy + uy;
    Mi[33] = x5;
    Mi[34] = h - ux;
    Mi[35] = -h;

    Mi[36] = h * ux2uy - h * ux2uy2 + h * uy2 + h * x5;
    Mi[37] = uxuy + x3;
    Mi[38] = h * ux2 - ux2uy + uy - h;
    Mi[39] = -qr * ux2 + qr * uy - qr * uy2 + qr;
    Mi[40] = qr * ux2 + qr * uy - qr * uy2 - qr;
    Mi[41] = ux - R(2) * uxuy;
    Mi[42] = x5 + h;
    Mi[43] = -ux;
    Mi[44] = -h;

    Mi[45] = qr * ux2uy + qr * ux2uy2 + qr * uxuy + qr * uxuy2;
    Mi[46] = h * uxuy + h * uxuy2 + qr * uy + qr * uy2;
    Mi[47] = qr * ux + qr * ux2 + h * ux2uy + h * uxuy;
    Mi[48] = e * ux + e * ux2 + e * uy + e * uy2;
    Mi[49] = -e * ux - e * ux2 + e * uy + e * uy2;
    Mi[50] = h * ux + uxuy + h * uy + qr;
    Mi[51] = h * uy + qr;
    Mi[52] = h * ux + qr;
    Mi[53] = qr;

    Mi[54] = qr * ux2uy + qr * ux2uy2 - qr * uxuy + qr * x3;
    Mi[55] = h * uxuy + h * uxuy2 - qr * uy2 + qr * x5;
    Mi[56] = -qr * ux + qr * ux2 + h * ux2uy - h * uxuy;
    Mi[57] = -e * ux + e * ux2 + e * uy + e * uy2;
    Mi[58] = e * ux - e * ux2 + e * uy + e * uy2;
    Mi[59] = h * ux + uxuy + h * x5 - qr;
    Mi[60] = h * uy + qr;
    Mi[61] = h * ux - qr;
    Mi[62] = qr;

    Mi[63] = -qr * ux2uy + qr * ux2uy2 + qr * uxuy + qr * x3;
    Mi[64] = -h * uxuy + h * uxuy2 + qr * uy - qr * uy2;
    Mi[65] = qr * ux - qr * ux2 + h * ux2uy - h * uxuy;
    Mi[66] = -e * ux + e * ux2 + e * uy2 + e * x5;
    Mi[67] = e * ux - e * ux2 + e * uy2 + e * x5;
    Mi[68] = -h * ux + uxuy + h * x5 + qr;
    Mi[69] = h * uy - qr;
    Mi[70] = h * ux - qr;
    Mi[71] = qr;

    Mi[72] = -qr * ux2uy + qr * ux2uy2 - qr * uxuy + qr * uxuy2;
    Mi[73] = -h * uxuy + h * uxuy2 + qr * uy2 + qr * x5;
    Mi[74] = -qr * ux - qr * ux2 + h * ux2uy + h * uxuy;
    Mi[75] = e * ux + e * ux2 + e * uy2 + e * x5;
    Mi[76] = -e * ux - e * ux2 + e * uy2 + e * x5;
    Mi[77] = -h * ux + uxuy + h * uy - qr;
    Mi[78] = h * uy - qr;
    Mi[79] = h * ux + qr;
    Mi[80] = qr;
}

// OptimalAdapter::compute_higher_order_relaxation (adapters.cuh:48-111), without the
// diagnostic printf (quirk Q14).
template <class R>
inline R optimal_adapter_rate(R rho, R j_mag, R pi_mag, const R* avg) {
    R s_p[4] = {rho / avg[0], j_mag / avg[1], pi_mag / avg[2], R(1)};
    const R theta[4] = {R(0.0003), R(-0.00775), R(0.00016), R(0.0087)};
    R tau_star = 0;
    for (int i = 0; i < 4; i++) tau_star += theta[i] * s_p[i];
    tau_star = tau_star > R(0) ? tau_star : R(0.005);
    tau_star = std::min(tau_star, R(1.5));
    return R(1) / (R(3) * tau_star + R(0.5));
}

// CM<2,A>::apply (CM.cuh:27-138). avg == nullptr selects NoAdapter.
template <class R>
inline void collide_cm(R* f, R ux, R uy, R Fx, R Fy, const R* S, const R* avg) {
    R rho = 0;
    for (int i = 0; i < Q; i++) rho += f[i];
    R k[Q] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    R pi[3] = {0, 0, 0};
    k[0] = rho;
    for (int i = 0; i < Q; i++) {
        R c_cx = R(CX[i]) - ux;
        R c_cy = R(CY[i]) - uy;
        R c_cx2 = c_cx * c_cx;
        R c_cy2 = c_cy * c_cy;
        R f_i = f[i];
        k[1] += f_i * (c_cx);
        k[2] += f_i * (c_cy);
        k[3] += f_i * (c_cx2 + c_cy2);
        k[4] += f_i * (c_cx2 - c_cy2);
        k[5] += f_i * c_cx * c_cy;
        k[6] += f_i * c_cx2 * c_cy;
        k[7] += f_i * c_cx * c_cy2;
        k[8] += f_i * c_cx2 * c_cy2;
        pi[0] += f_i * R(CX[i]) * R(CX[i]);
        pi[1] += f_i * R(CX[i]) * R(CY[i]);
        pi[2] += f_i * R(CY[i]) * R(CY[i]);
    }
    R pi_mag = std::sqrt(pi[0] * pi[0] + R(2) * pi[1] * pi[1] + pi[2] * pi[2]);
    const R cs2 = R(1) / R(3);
    R k_eq[Q] = {rho, 0, 0, R(2) * rho * cs2, 0, 0, 0, 0, rho * cs2 * cs2};
    R F[Q] = {0, Fx, Fy, 0, 0, 0, Fy * cs2, Fx * cs2, 0};
    R j_mag = std::sqrt(ux * ux + uy * uy) * rho;
    R hor = 0;
    if (avg) hor = optimal_adapter_rate<R>(rho, j_mag, pi_mag, avg);
    R k_post[Q];
    for (int i = 0; i < Q; i++) {
        R s = (avg && i > 5) ? hor : S[i];  // AdapterBase::i_star = 5 (adapters.cuh:7-14)
        k_post[i] = k[i] - s * (k[i] - k_eq[i]) + (R(1) - R(0.5) * s) * F[i];
    }
    R Ti[Q * Q];
    cm_matrix_inverse<R>(Ti, ux, uy);
    for (int i = 0; i < Q; i++) {
        R acc = 0;
        for (int j = 0; j < Q; j++) acc += Ti[i * Q + j] * k_post[j];
        f[i] = acc;
    }
}

// delta4 (IBMUtils.cuh:23-33): Peskin 4-point kernel.
template <class R>
inline R delta4(R r) {
    R rabs = std::fabs(r);
    if (rabs < R(1)) return R(0.125) * (R(3) - R(2) * rabs + std::sqrt(R(1) + R(4) * rabs - R(4) * r * r));
    if (rabs < R(2)) return R(0.125) * (R(5) - R(2) * rabs - std::sqrt(R(-7) + R(12) * rabs - R(4) * r * r));
    return R(0);
}

// ---- the solver -------------------------------------------------------------------------

template <class R>
struct Solver {
    lbm_config cfg;
    int nx, ny;
    size_t N;
    bool eq_literal = false;       // quirk Q15 (oracle only)
    std::vector<R> f, f_back;      // SoA [q][node] (lbm_constants.cuh:320-323)
    std::vector<R> f_eq;
    std::vector<R> rho, u, force;  // u, force SoA [c][node]
    std::vector<R> pi_mag;
    std::vector<uint8_t> flags;
    std::vector<R> force_field;    // AoS [node][2]; empty => uniform cfg.force
    R avg[3] = {0, 0, 0};          // MomentInfo (lbm.cuh:25-31)
    int64_t timestep = 0;
    // IBM (IBMManager.cuh:31-51)
    std::vector<R> lag_xy, lag_ds, lag_u, lag_rho, lag_force, lag_force_sum;
    int ibm_iters = 3;
    std::vector<R> eul_u_prev, eul_force_iter;

    Solver(const lbm_config& c) : cfg(c), nx(c.nx), ny(c.ny), N(size_t(c.nx) * size_t(c.ny)) {
        if (nx < 3 || ny < 3) throw std::runtime_error("grid must be at least 3x3");
        f.assign(Q * N, R(0));
        f_back.assign(Q * N, R(0));
        f_eq.assign(Q * N, R(0));
        rho.assign(N, R(1));
        u.assign(2 * N, R(0));
        force.assign(2 * N, R(0));
        pi_mag.assign(N, R(0));
        flags.assign(N, 0);
    }

    R S(int k) const { return R(cfg.S[k]); }
    bool quirk(uint32_t b) const { return (cfg.quirks & b) != 0; }

    // Storage rounding of the shifted deviation f - w rho0 (FP32 arithmetic, as the GPU).
    R store_round(int q, R v) const {
        if (cfg.storage == LBM_F32) return v;
        float w0 = weight<float>(q) * cfg.rho0;
        float d = float(v) - w0;
        float back = cfg.storage == LBM_F16_SHIFTED ? f16_to_f32(f32_to_f16(d)) : bf16_to_f32(f32_to_bf16(d));
        return R(back + w0);
    }
    void quantize(std::vector<R>& buf) {
        if (cfg.storage == LBM_F32) return;
        for (int q = 0; q < Q; q++)
            for (size_t n = 0; n < N; n++) buf[q * N + n] = store_round(q, buf[q * N + n]);
    }

    // init_kernel + init_node (init.cuh:10-43): f = f_back = f_eq(rho, u).
    void init_from_macros(const R* rho_in, const R* u_aos) {
        for (size_t n = 0; n < N; n++) {
            rho[n] = rho_in[n];
            u[n] = u_aos[2 * n];
            u[N + n] = u_aos[2 * n + 1];
            R feq[Q];
            equilibrium<R>(rho[n], u[n], u[N + n], feq, eq_literal);
            for (int q = 0; q < Q; q++) f[q * N + n] = f_back[q * N + n] = f_eq[q * N + n] = feq[q];
        }
        quantize(f);
        quantize(f_back);
        timestep = 0;
    }

    // stream_node (streaming.cu:3-33): push f[node][i] to the neighbour's f_back slot,
    // wrapping on periodic axes; out-of-domain targets are skipped (slot stays stale, Q2).
    // The rest population is copied unless quirk Q1 (loop from i = 1).
    void stream() {
        int i0 = quirk(LBM_QUIRK_STALE_REST) ? 1 : 0;
#pragma omp parallel for schedule(static)
        for (int y = 0; y < ny; y++) {
            for (int x = 0; x < nx; x++) {
                size_t node = size_t(y) * nx + x;
                for (int i = i0; i < Q; i++) {
                    int xn = x + CX[i], yn = y + CY[i];
                    if (cfg.periodic_x) xn = (xn + nx) % nx;
                    if (cfg.periodic_y) yn = (yn + ny) % ny;
                    if (xn < 0 || xn >= nx || yn < 0 || yn >= ny) continue;
                    f_back[i * N + size_t(yn) * nx + xn] = f[i * N + node];
                }
            }
        }
    }

    // ---- boundary functors, applied to the post-stream node values g[9]. Functors that
    // read another node use `snap`, the post-stream field before any boundary ran.

    // BounceBack<2>::apply (bbDomainBoundary.cuh:15-51), reflection enabled on the
    // non-periodic axes. wall != nullptr gives the appended moving-wall form.
    void bc_bounce_back(R* g, int x, int y, const float* wall) const {
        for (int i = 1; i < Q; i++) {
            int xn = x + CX[i], yn = y + CY[i];
            bool bx = !cfg.periodic_x && (xn < 0 || xn >= nx);
            bool by = !cfg.periodic_y && (yn < 0 || yn >= ny);
            if (bx || by) {
                int o = OPP[i];
                R v = g[i];
                if (wall) {
                    R cu = R(CX[i]) * R(wall[0]) + R(CY[i]) * R(wall[1]);
                    v = v - R(6) * weight<R>(i) * R(cfg.rho0) * cu;
                }
                g[o] = v;
            }
        }
    }

    // RegularizedInlet::apply_top (regularizedInlet.cuh:15-69).
    void bc_regularized_inlet_top(R* g) const {
        const R ux = R(cfg.u_lid), uy = R(0);
        R r = (g[0] + g[1] + g[3] + R(2) * (g[2] + g[5] + g[6])) / (R(1) + uy);
        R feq[Q];
        const R cs2 = R(1) / R(3);
        for (int q = 0; q < Q; q++) {
            R cu = R(CX[q]) * ux + R(CY[q]) * uy;
            R uu = ux * ux + uy * uy;
            feq[q] = weight<R>(q) * r * (R(1) + cu / cs2 + cu * cu / (R(2) * cs2 * cs2) - uu / (R(2) * cs2));
        }
        for (int q = 0; q < Q; q++)
            if (CY[q] < 0) g[q] = feq[q] + (g[OPP[q]] - feq[OPP[q]]);
        R pxx = 0, pyy = 0, pxy = 0;
        for (int q = 0; q < Q; q++) {
            R cx = R(CX[q]), cy = R(CY[q]);
            pxx += cx * cx * g[q];
            pyy += cy * cy * g[q];
            pxy += cx * cy * g[q];
        }
        pxx -= cs2 * r + r * ux * ux;
        pyy -= cs2 * r + r * uy * uy;
        pxy -= r * ux * uy;
        for (int q = 0; q < Q; q++) {
            R cx = R(CX[q]), cy = R(CY[q]);
            R Qxx = cx * cx - cs2, Qyy = cy * cy - cs2, Qxy = cx * cy;
            R fneq = (weight<R>(q) / (R(2) * cs2 * cs2)) * (Qxx * pxx + Qyy * pyy + R(2) * Qxy * pxy);
            g[q] = feq[q] + fneq;
        }
    }

    // RegularizedBounceBack::apply (regularizedBounceBack.cuh:13-100).
    void bc_regularized_bounce_back(R* g, int x, int y) const {
        bool unk[Q] = {false, false, false, false, false, false, false, false, false};
        for (int i = 1; i < Q; i++) {
            if (x == 0) unk[i] = CX[i] > 0;
            else if (x == nx - 1) unk[i] = CX[i] < 0;
            else if (y == 0) unk[i] = CY[i] > 0;
            else if (y == ny - 1) unk[i] = CY[i] < 0;
        }
        R r = 0;
        for (int i = 0; i < Q; i++)
            if (!unk[i]) r += R(unk[OPP[i]] ? 2 : 1) * g[i];
        const R cs2 = R(1) / R(3);
        R feq[Q];
        for (int q = 0; q < Q; q++) feq[q] = weight<R>(q) * r * (R(1) + R(0) / cs2 + R(0) / (R(2) * cs2 * cs2) - R(0) / (R(2) * cs2));
        for (int q = 0; q < Q; q++)
            if (unk[q]) g[q] = feq[q] + (g[OPP[q]] - feq[OPP[q]]);
        R pxx = 0, pyy = 0, pxy = 0;
        for (int q = 0; q < Q; q++) {
            R cx = R(CX[q]), cy = R(CY[q]);
            pxx += cx * cx * g[q];
            pyy += cy * cy * g[q];
            pxy += cx * cy * g[q];
        }
        pxx -= cs2 * r;
        pyy -= cs2 * r;
        for (int q = 0; q < Q; q++) {
            R cx = R(CX[q]), cy = R(CY[q]);
            R Qxx = cx * cx - cs2, Qyy = cy * cy - cs2, Qxy = cx * cy;
            R fneq = (weight<R>(q) / (R(2) * cs2 * cs2)) * (Qxx * pxx + Qyy * pyy + R(2) * Qxy * pxy);
            g[q] = feq[q] + fneq;
        }
    }

    // RegularizedCornerBounceBack::apply (regularizedBounceBack.cuh:107-222): density
    // from the diagonal neighbour's post-stream populations (clamped to the interior).
    void bc_regularized_corner(R* g, int x, int y, const std::vector<R>& snap) const {
        bool left = x == 0, right = x == nx - 1, bottom = y == 0, top = y == ny - 1;
        if (!((left || right) && (bottom || top))) return;
        bool unk[Q] = {false, false, false, false, false, false, false, false, false};
        for (int i = 1; i < Q; i++)
            unk[i] = (left && CX[i] > 0) || (right && CX[i] < 0) || (bottom && CY[i] > 0) || (top && CY[i] < 0);
        int dx = left ? x + 1 : x - 1;
        int dy = bottom ? y + 1 : y - 1;
        dx = std::max(1, std::min(dx, nx - 2));
        dy = std::max(1, std::min(dy, ny - 2));
        size_t dn = size_t(dy) * nx + dx;
        R r = 0;
        for (int i = 0; i < Q; i++) r += snap[i * N + dn];
        const R cs2 = R(1) / R(3);
        auto feq = [&](int i) { return weight<R>(i) * r * (R(1) + R(0) / cs2 + R(0) / (R(2) * cs2 * cs2) - R(0) / (R(2) * cs2)); };
        for (int i = 0; i < Q; i++)
            if (unk[i]) g[i] = feq(i) - (g[OPP[i]] - feq(OPP[i]));
        R pxx = 0, pyy = 0, pxy = 0;
        for (int q = 0; q < Q; q++) {
            R fneq = g[q] - feq(q);
            R cx = R(CX[q]), cy = R(CY[q]);
            pxx += cx * cx * fneq;
            pyy += cy * cy * fneq;
            pxy += cx * cy * fneq;
        }
        for (int q = 0; q < Q; q++) {
            R cx = R(CX[q]), cy = R(CY[q]);
            R Qxx = cx * cx - cs2, Qyy = cy * cy - cs2, Qxy = cx * cy;
            R fneq = (weight<R>(q) / (R(2) * cs2 * cs2)) * (Qxx * pxx + Qyy * pyy + R(2) * Qxy * pxy);
            g[q] = feq(q) + fneq;
        }
    }

    // ZG_OutflowBoundary<2>::apply (zeroGradientOutflow.cuh:9-59).
    void bc_zg_outflow(R* g, int x, int y, const std::vector<R>& snap) const {
        int n0 = 0, n1 = 0, ix = x, iy = y;
        if (x == 0) { n0 = 1; ix = 1; }
        else if (x == nx - 1) { n0 = -1; ix = nx - 2; }
        else if (y == 0) { n1 = 1; iy = 1; }
        else if (y == ny - 1) { n1 = -1; iy = ny - 2; }
        else return;
        size_t in = size_t(iy) * nx + ix;
        for (int i = 0; i < Q; i++)
            if (CX[i] * n0 + CY[i] * n1 > 0) g[i] = snap[i * N + in];
    }

    // ZouHe::apply_left / apply_top (zouHeInflow.cuh:9-68) with SoA indexing (quirk Q8
    // fixed); quirk Q9 keeps the left inlet's f2-for-f3 density typo.
    void bc_zou_he_left(R* g) const {
        const R ux = R(cfg.u_lid);
        R r = quirk(LBM_QUIRK_ZOUHE_RHO) ? (g[0] + g[2] + g[4] + R(2) * (g[2] + g[6] + g[7])) / (R(1) - ux)
                                         : (g[0] + g[2] + g[4] + R(2) * (g[3] + g[6] + g[7])) / (R(1) - ux);
        g[1] = g[3] + (R(2) / R(3)) * r * ux;
        g[5] = g[7] - R(0.5) * (g[2] - g[4]) + (R(1) / R(6)) * r * ux;
        g[8] = g[6] + R(0.5) * (g[2] - g[4]) + (R(1) / R(6)) * r * ux;
    }
    void bc_zou_he_top(R* g) const {
        const R ux = R(cfg.u_lid), uy = R(0);
        R r = (g[0] + g[1] + g[3] + R(2) * (g[2] + g[5] + g[6])) / (R(1) + uy);
        g[4] = g[2] - (R(2) / R(3)) * r * uy;
        R d13 = g[1] - g[3];
        g[7] = g[5] + R(0.5) * d13 - (R(1) / R(2)) * r * ux - (R(1) / R(6)) * r * uy;
        g[8] = g[6] - R(0.5) * d13 + (R(1) / R(2)) * r * ux - (R(1) / R(6)) * r * uy;
    }

    // CylinderBoundary::apply (cylinderBoundary.cuh:22-34) with SoA indexing (quirk Q8
    // fixed): every post-stream population is sent back along its opposite direction.
    void bc_cylinder(R* g) const {
        R saved[Q];
        for (int j = 0; j < Q; j++) saved[j] = g[j];
        for (int i = 0; i < Q; i++) g[OPP[i]] = saved[i];
    }

    // PressureOutlet::apply (pressureOutlet.cuh:7-42) with SoA indexing (quirk Q8 fixed):
    // the velocity of the node to the left (x - 1, periodic wrap in x) from its post-stream
    // populations, then f = f_eq(rho = 1, u) written as in that functor.
    void bc_pressure_outlet(R* g, int x, int y, const std::vector<R>& snap) const {
        int ix = x - 1;
        if (ix < 0) ix += nx;
        size_t in = size_t(y) * nx + ix;
        R rho_i = 0;
        for (int i = 0; i < Q; i++) rho_i += snap[i * N + in];
        R ux = 0, uy = 0;
        for (int i = 0; i < Q; i++) {
            ux += snap[i * N + in] * R(CX[i]);
            uy += snap[i * N + in] * R(CY[i]);
        }
        ux /= rho_i;
        uy /= rho_i;
        const R target_rho = R(1);
        for (int i = 0; i < Q; i++) {
            R cu = R(CX[i]) * ux + R(CY[i]) * uy;
            R uu = ux * ux + uy * uy;
            R cs2 = R(1) / R(3);
            g[i] = weight<R>(i) * target_rho * (R(1) + cu / cs2 + (cu * cu) / (R(2) * cs2 * cs2) - uu / (R(2) * cs2));
        }
    }

    // boundaries_kernel_2D (boundaries.cuh:10-85): switch on the node's flag.
    void apply_boundaries() {
        std::vector<R> snap = f;
#pragma omp parallel for schedule(static)
        for (int y = 0; y < ny; y++) {
            for (int x = 0; x < nx; x++) {
                size_t node = size_t(y) * nx + x;
                uint8_t fl = flags[node];
                if (fl == LBM_FLUID) continue;
                R g[Q];
                for (int q = 0; q < Q; q++) g[q] = f[q * N + node];
                switch (fl) {
                    case LBM_BOUNCE_BACK: bc_bounce_back(g, x, y, nullptr); break;
                    case LBM_MOVING_WALL_BOUNCE_BACK: bc_bounce_back(g, x, y, cfg.wall_u); break;
                    case LBM_ZOU_HE_TOP: bc_zou_he_top(g); break;
                    case LBM_ZOU_HE_LEFT: bc_zou_he_left(g); break;
                    case LBM_ZG_OUTFLOW: bc_zg_outflow(g, x, y, snap); break;
                    case LBM_REGULARIZED_INLET_TOP: bc_regularized_inlet_top(g); break;
                    case LBM_REGULARIZED_BOUNCE_BACK: bc_regularized_bounce_back(g, x, y); break;
                    case LBM_REGULARIZED_BOUNCE_BACK_CORNER: bc_regularized_corner(g, x, y, snap); break;
                    case LBM_CYLINDER: bc_cylinder(g); break;
                    case LBM_PRESSURE_OUTLET: bc_pressure_outlet(g, x, y, snap); break;
                    default: break;
                }
                for (int q = 0; q < Q; q++) f[q * N + node] = g[q];
            }
        }
    }

    // uncorrected_macroscopics_kernel<2> (macroscopics.cu:5-38).
    void uncorrected_macroscopics() {
#pragma omp parallel for schedule(static)
        for (long long n = 0; n < (long long)N; n++) {
            R r = 0, jx = 0, jy = 0, p0 = 0, p1 = 0, p2 = 0;
            for (int i = 0; i < Q; i++) {
                R fi = f[i * N + n];
                r += fi;
                jx += fi * R(CX[i]);
                jy += fi * R(CY[i]);
                p0 += fi * R(CX[i]) * R(CX[i]);
                p1 += fi * R(CX[i]) * R(CY[i]);
                p2 += fi * R(CY[i]) * R(CY[i]);
            }
            rho[n] = r;
            u[n] = jx * (R(1) / r);
            u[N + n] = jy * (R(1) / r);
            pi_mag[n] = std::sqrt(p0 * p0 + R(2) * p1 * p1 + p2 * p2);
        }
    }

    // reset_forces_kernel + Init::apply_forces (macroscopics.cuh:13-48).
    void reset_forces() {
        for (size_t n = 0; n < N; n++) {
            if (force_field.empty()) {
                force[n] = R(cfg.force[0]);
                force[N + n] = R(cfg.force[1]);
            } else {
                force[n] = force_field[2 * n];
                force[N + n] = force_field[2 * n + 1];
            }
        }
    }

    // correct_macroscopics (macroscopics.cu:99-110, macroscopics.cuh:141-183): Guo
    // half-force velocity correction; then the MomentInfo domain averages, accumulated
    // in FP64 in row order (the reference's float atomics are order-dependent, Q17).
    void correct_macroscopics() {
        for (size_t n = 0; n < N; n++) {
            u[n] += R(0.5) * force[n] / rho[n];
            u[N + n] += R(0.5) * force[N + n] / rho[n];
        }
        if (cfg.collision == LBM_CM && cfg.adapter == LBM_OPTIMAL_ADAPTER) {
            double s0 = 0, s1 = 0, s2 = 0;
            for (int y = 0; y < ny; y++) {
                double r0 = 0, r1 = 0, r2 = 0;
                for (int x = 0; x < nx; x++) {
                    size_t n = size_t(y) * nx + x;
                    R ux = u[n], uy = u[N + n];
                    r0 += double(rho[n]);
                    r1 += double(rho[n] * std::sqrt(ux * ux + uy * uy));
                    r2 += double(pi_mag[n]);
                }
                s0 += r0;
                s1 += r1;
                s2 += r2;
            }
            avg[0] = R(s0 / double(N));
            avg[1] = R(s1 / double(N));
            avg[2] = R(s2 / double(N));
        }
    }

    // compute_equilibrium + collide<CollisionOp> (equilibrium.cuh:19-71, collision.cuh:16-65):
    // every node collides, boundary nodes included (Q18).
    void collide() {
#pragma omp parallel for schedule(static)
        for (long long n = 0; n < (long long)N; n++) {
            R g[Q], feq[Q];
            for (int q = 0; q < Q; q++) g[q] = f[q * N + n];
            R ux = u[n], uy = u[N + n], fx = force[n], fy = force[N + n];
            if (cfg.collision != LBM_CM) {
                equilibrium<R>(rho[n], ux, uy, feq, eq_literal);
                for (int q = 0; q < Q; q++) f_eq[q * N + n] = feq[q];
            }
            R Sv[Q];
            for (int k = 0; k < Q; k++) Sv[k] = S(k);
            if (cfg.collision == LBM_BGK) collide_bgk<R>(g, feq, ux, uy, fx, fy, R(cfg.omega));
            else if (cfg.collision == LBM_MRT) collide_mrt<R>(g, feq, ux, uy, fx, fy, Sv, quirk(LBM_QUIRK_MRT_FORCE_ROWS));
            else collide_cm<R>(g, ux, uy, fx, fy, Sv, cfg.adapter == LBM_OPTIMAL_ADAPTER ? avg : nullptr);
            for (int q = 0; q < Q; q++) f[q * N + n] = g[q];
        }
    }

    // ---- immersed boundary (IBMManager.cuh:222-252, IBM_impl.cu/.cuh) ----------------

    int ibm_w() const { return quirk(LBM_QUIRK_IBM_2X2) ? 2 : 4; }
    int ibm_off() const { return quirk(LBM_QUIRK_IBM_2X2) ? 0 : -1; }

    void set_markers(const R* xy, const R* ds, int n, int iters) {
        lag_xy.assign(xy, xy + 2 * n);
        lag_ds.assign(size_t(n), R(1));
        if (ds) for (int i = 0; i < n; i++) lag_ds[i] = ds[i];
        lag_u.assign(2 * size_t(n), R(0));
        lag_rho.assign(size_t(n), R(0));
        lag_force.assign(2 * size_t(n), R(0));
        lag_force_sum.assign(2 * size_t(n), R(0));
        ibm_iters = iters;
        eul_u_prev.assign(2 * N, R(0));
        eul_force_iter.assign(2 * N, R(0));
    }

    // interpolate_velocities_kernel<2> (IBM_impl.cu:5-51), stencil per Q11 switch.
    void ibm_interpolate() {
        int np = int(lag_rho.size());
        int W = ibm_w(), off = ibm_off();
        for (int l = 0; l < np; l++) {
            R px = lag_xy[2 * l], py = lag_xy[2 * l + 1];
            R gx = std::floor(px), gy = std::floor(py);
            R ux = 0, uy = 0, r = 0;
            for (int i = 0; i < W; i++) {
                for (int j = 0; j < W; j++) {
                    int nxx = int(gx) + off + i, nyy = int(gy) + off + j;
                    if (nxx >= nx || nxx < 0 || nyy >= ny || nyy < 0) continue;
                    R k = delta4<R>(px - R(nxx)) * delta4<R>(py - R(nyy));
                    size_t id = size_t(nyy) * nx + nxx;
                    r += k * rho[id];
                    ux += k * eul_u_prev[id];
                    uy += k * eul_u_prev[N + id];
                }
            }
            lag_u[l] = ux;
            lag_u[np + l] = uy;
            lag_rho[l] = r;
        }
    }
    // compute_lagrangian_kernel (IBM_impl.cuh:7-26): F = 2 rho_L (0 - u_L); Q12 clamp.
    void ibm_lagrangian() {
        int np = int(lag_rho.size());
        for (int l = 0; l < np; l++)
            for (int d = 0; d < 2; d++) {
                R fo = R(2) * lag_rho[l] * (R(0) - lag_u[d * np + l]);
                if (quirk(LBM_QUIRK_IBM_CLAMP)) fo = fo > R(1e-8) ? fo : R(0);
                lag_force[d * np + l] = fo;
            }
    }
    // spread_forces_kernel<2> (IBM_impl.cu:120-154), in marker order; Q13 switch for ds.
    void ibm_spread() {
        int np = int(lag_rho.size());
        int W = ibm_w(), off = ibm_off();
        for (int l = 0; l < np; l++) {
            R px = lag_xy[2 * l], py = lag_xy[2 * l + 1];
            R gx = std::floor(px), gy = std::floor(py);
            R ds = quirk(LBM_QUIRK_IBM_NO_DS) ? R(1) : lag_ds[l];
            for (int i = 0; i < W; i++) {
                for (int j = 0; j < W; j++) {
                    int nxx = int(gx) + off + i, nyy = int(gy) + off + j;
                    if (nxx >= nx || nxx < 0 || nyy >= ny || nyy < 0) continue;
                    R k = delta4<R>(px - R(nxx)) * delta4<R>(py - R(nyy));
                    R w = k * ds;
                    size_t id = size_t(nyy) * nx + nxx;
                    eul_force_iter[id] += w * lag_force[l];
                    eul_force_iter[N + id] += w * lag_force[np + l];
                }
            }
        }
    }
    void ibm_step() {
        int np = int(lag_rho.size());
        if (np == 0) return;
        std::copy(u.begin(), u.end(), eul_u_prev.begin());
        std::fill(lag_force_sum.begin(), lag_force_sum.end(), R(0));
        for (int it = 0; it < ibm_iters; it++) {
            std::fill(eul_force_iter.begin(), eul_force_iter.end(), R(0));
            ibm_interpolate();
            ibm_lagrangian();
            ibm_spread();
            // correct_velocities_kernel (IBM_impl.cuh:28-46); Q12 clamp.
            for (size_t n = 0; n < N; n++)
                for (int d = 0; d < 2; d++) {
                    R cu = eul_u_prev[d * N + n] + eul_force_iter[d * N + n] / (R(2) * rho[n]);
                    if (quirk(LBM_QUIRK_IBM_CLAMP)) cu = cu > R(1e-8) ? cu : R(0);
                    eul_u_prev[d * N + n] = cu;
                }
            // accumulate_forces_kernel (IBM_impl.cuh:48-68)
            for (size_t n = 0; n < 2 * N; n++) force[n] += eul_force_iter[n];
            for (int l = 0; l < 2 * np; l++) lag_force_sum[l] += lag_force[l];
        }
    }

    // One iteration of the reference's time loop (main.cu:96-114).
    void step() {
        timestep++;
        stream();
        std::swap(f, f_back);
        apply_boundaries();
        uncorrected_macroscopics();
        reset_forces();
        ibm_step();
        correct_macroscopics();
        collide();
        quantize(f);
    }
};

}  // namespace oracle
