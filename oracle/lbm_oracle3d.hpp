// lbm_oracle3d.hpp — CPU restatement of the reference's D3Q27 BGK, MRT and CM path.
//
// TEST INFRASTRUCTURE ONLY. This is the correctness oracle for the 3-D step. Only tests/,
// the smoke check in __graft_entry__.py and the cpu_baseline leg of bench.py may load or call
// it, and only as the checker. The product (liblbm_hip.so) never links or calls this code.
//
// It follows the reference's unfused stage order (src/main.cu:96-114) on separate buffers, as
// `LBM<3>` does (src/core/lbm.cuh:36-43): push stream -> swap -> boundaries -> uncorrected
// macroscopics -> reset forces -> Guo velocity correction -> equilibrium -> BGK<3>, MRT<3> or
// CM<3> collision, templated on the arithmetic type (FP32 oracle, FP64 host run). MRT<3> and CM<3>
// are dense: m = M f, m* per moment, f = M^-1 m* (CM: the central-moment matrix T(u) per node).
//
// Parity pinning (tests/test_oracle3d_pinning.py), each against something this file does not
// restate: the equilibrium's moments up to second order (exact identities), the literal
// equilibrium's first moment (17/27) rho u (closed form), the Guo source's moments,
// periodic streaming against numpy.roll, BGK's conserved moments, the rest state as a fixed
// point of a closed bounce-back box, viscous decay of a small
// Taylor-Green vortex against exp(-6 nu k^2 t), and the curvature -F/nu of a force-driven
// channel; for MRT<3>: the basis against its inverse, the collision at unit rates against the
// product-form equilibrium rho*prod_a phi(c_a, u_a), the relaxed second moments, the momentum
// gain per step and the viscous decay rate; for CM<3>: unit omega against the product-form
// equilibrium, the relaxed central second moments and trace, momentum (+F), the forcing central
// moments and the viscous decay rate. Parity unpinned: the FP32 rounding sequence of the
// reference's De Rosis expression (equilibrium.cu:70-96), which acts at FP32 rounding level.
#pragma once

#include "../include/lbm3d_hip.h"
#include "lbm_oracle.hpp"

#include <cmath>
#include <stdexcept>
#include <vector>

namespace oracle3d {

constexpr int Q = 27;
// Velocity set in the reference's order: rest, 6 faces, 12 edges, 8 corners
// (lbm_constants.cuh:76-104).
constexpr int CX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
constexpr int CY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, -1};
constexpr int CZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1};
// Opposite directions (lbm_constants.cuh:106-114).
constexpr int OPP[Q] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15, 26, 25, 24, 23, 22, 21, 20, 19};

// Weights 8/27, 2/27, 1/54, 1/216 by |c|^2 (lbm_constants.cuh:64-71).
template <class R> inline R weight(int q) {
    if (q == 0) return R(8) / R(27);
    if (q < 7) return R(2) / R(27);
    if (q < 19) return R(1) / R(54);
    return R(1) / R(216);
}

// LBM<3>::equilibrium_node (equilibrium.cu:41-96). Default: w rho (1 + 3cu + 4.5cu^2 - 1.5u^2)
// in R. literal == true restates the De Rosis (2017) form as written: the rest and face
// directions are the second-order equilibrium, the edge and corner directions compute
// cs2*s*(s + 2cs4) + 2cs2cs4 - cs4*u^2 scaled by rho/(2 cs2 cs4) w, i.e. a linear
// coefficient 1 instead of 3 (quirk LBM_QUIRK_EQ3D_LINEAR).
template <class R>
inline void equilibrium(R rho, R ux, R uy, R uz, R* feq, bool literal) {
    const R uu = ux * ux + uy * uy + uz * uz;
    if (!literal) {
        for (int q = 0; q < Q; q++) {
            R cu = R(CX[q]) * ux + R(CY[q]) * uy + R(CZ[q]) * uz;
            feq[q] = weight<R>(q) * rho * (R(1) + R(3) * cu + R(4.5) * cu * cu - R(1.5) * uu);
        }
        return;
    }
    const R cs2 = R(1) / R(3);
    const R cs4 = cs2 * cs2;
    const R two_cs4 = R(2) * cs4;
    const R cs2_cs4 = cs2 * cs4;
    const R inv_cs2_cs4 = R(1) / cs2_cs4;
    const R two_cs2_cs4 = R(2) * cs2 * cs4;
    const R half_rho = R(0.5) * rho;
    const R ux2 = ux * ux, uy2 = uy * uy, uz2 = uz * uz;
    feq[0] = half_rho * weight<R>(0) * (R(2) * cs2 - ux2 - uy2 - uz2) / cs2;
    for (int q = 1; q < Q; q++) {
        R s = R(CX[q]) * ux + R(CY[q]) * uy + R(CZ[q]) * uz;
        R bracket;
        if (q < 7) bracket = cs2 * s * s + two_cs2_cs4 + two_cs4 * s - cs4 * (ux2 + uy2 + uz2);
        else bracket = cs2 * s * (s + two_cs4) + two_cs2_cs4 - cs4 * (ux2 + uy2 + uz2);
        feq[q] = inv_cs2_cs4 * half_rho * weight<R>(q) * bracket;
    }
}

// Guo source term of BGK<3>::apply (BGK.cuh:55-96):
// w (1 - omega/2) [ ((c - u)/cs2 + (c.u) c / cs4) . F ].
template <class R>
inline R guo_source(int q, R ux, R uy, R uz, R fx, R fy, R fz, R omega) {
    const R cs2 = R(1) / R(3);
    const R cx = R(CX[q]), cy = R(CY[q]), cz = R(CZ[q]);
    const R cu = cx * ux + cy * uy + cz * uz;
    const R k = R(1) - R(0.5) * omega;
    return weight<R>(q) * (k * ((cx - ux) / cs2 + cu * cx / (cs2 * cs2)) * fx +
                           k * ((cy - uy) / cs2 + cu * cy / (cs2 * cs2)) * fy +
                           k * ((cz - uz) / cs2 + cu * cz / (cs2 * cs2)) * fz);
}

// The De Rosis (2019) raw-moment basis of MRT<3> (lbm_constants.cuh:218-246): row k of M is the
// polynomial P_k(c) evaluated on the velocity set, in the reference's row order.
template <class R>
inline R basis_poly(int k, int cx, int cy, int cz) {
    const R x = R(cx), y = R(cy), z = R(cz), x2 = x * x, y2 = y * y, z2 = z * z;
    switch (k) {
        case 0: return R(1);
        case 1: return x;
        case 2: return y;
        case 3: return z;
        case 4: return x * y;
        case 5: return x * z;
        case 6: return y * z;
        case 7: return x2 - y2;
        case 8: return x2 - z2;
        case 9: return x2 + y2 + z2;
        case 10: return x * y2 + x * z2;
        case 11: return x2 * y + y * z2;
        case 12: return x2 * z + y2 * z;
        case 13: return x * y2 - x * z2;
        case 14: return x2 * y - y * z2;
        case 15: return x2 * z - y2 * z;
        case 16: return x * y * z;
        case 17: return x2 * y2 + x2 * z2 + y2 * z2;
        case 18: return x2 * y2 + x2 * z2 - y2 * z2;
        case 19: return x2 * y2 - x2 * z2;
        case 20: return x2 * y * z;
        case 21: return x * y2 * z;
        case 22: return x * y * z2;
        case 23: return x * y2 * z2;
        case 24: return x2 * y * z2;
        case 25: return x2 * y2 * z;
        default: return x2 * y2 * z2;
    }
}

// M and M^-1 (Gauss-Jordan with partial pivoting in double), row-major 27x27.
struct MrtBasis {
    double M[Q * Q], Minv[Q * Q];
    MrtBasis() {
        double A[Q][2 * Q];
        for (int k = 0; k < Q; k++)
            for (int i = 0; i < Q; i++) {
                M[k * Q + i] = basis_poly<double>(k, CX[i], CY[i], CZ[i]);
                A[k][i] = M[k * Q + i];
                A[k][Q + i] = k == i ? 1.0 : 0.0;
            }
        for (int c = 0; c < Q; c++) {
            int piv = c;
            for (int r = c + 1; r < Q; r++)
                if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
            if (std::fabs(A[piv][c]) < 1e-12) throw std::runtime_error("MRT<3> basis is singular");
            for (int j = 0; j < 2 * Q; j++) std::swap(A[c][j], A[piv][j]);
            const double d = A[c][c];
            for (int j = 0; j < 2 * Q; j++) A[c][j] /= d;
            for (int r = 0; r < Q; r++) {
                if (r == c) continue;
                const double m = A[r][c];
                if (m != 0.0)
                    for (int j = 0; j < 2 * Q; j++) A[r][j] -= m * A[c][j];
            }
        }
        for (int k = 0; k < Q; k++)
            for (int i = 0; i < Q; i++) Minv[k * Q + i] = A[k][Q + i];
    }
};
inline const MrtBasis& mrt_basis() {
    static const MrtBasis b;
    return b;
}

// MRT<3>::compute_forcing_term / compute_forcing_term_component (MRT.cu:79-153): the forcing
// moments F_k in the basis order.
template <class R>
inline void mrt3_forcing(R* F, R ux, R uy, R uz, R fx, R fy, R fz) {
    const R t = R(1) / R(3), n = R(1) / R(9);
    F[0] = 0;
    F[1] = fx;
    F[2] = fy;
    F[3] = fz;
    F[4] = fx * uy + fy * ux;
    F[5] = fx * uz + fz * ux;
    F[6] = fy * uz + fz * uy;
    F[7] = R(2) * fx * ux - R(2) * fy * uy;
    F[8] = R(2) * fx * ux - R(2) * fz * uz;
    F[9] = R(2) * fx * ux + R(2) * fy * uy + R(2) * fz * uz;
    F[10] = R(2) * t * fx;
    F[11] = R(2) * t * fy;
    F[12] = R(2) * t * fz;
    F[13] = F[14] = F[15] = F[16] = 0;
    F[17] = R(4) * t * (fx * ux + fy * uy + fz * uz);
    F[18] = R(4) * t * fx * ux;
    F[19] = R(2) * t * (fy * uy - fz * uz);
    F[20] = t * (fy * uz + fz * uy);
    F[21] = t * (fx * uz + fz * ux);
    F[22] = t * (fx * uy + fy * ux);
    F[23] = n * fx;
    F[24] = n * fy;
    F[25] = n * fz;
    F[26] = R(2) * n * (fx * ux + fy * uy + fz * uz);
}

// MRT<3>::apply (MRT.cu:397-456) with compute_m_post_component (MRT.cu:155-201): moments
// 0-3 conserved (rho, rho u, plus F/2 unless literal), 4-8 relaxed with S[4..8] plus
// (1 - S/2) F, 9-26 set to their equilibria plus F/2; f = M^-1 m*. literal keeps the
// reference's (2 ux^2 + 1) in the equilibrium of moment 19 (MRT.cu:188).
template <class R>
inline void collide_mrt3(R* f, R ux, R uy, R uz, R fx, R fy, R fz, const float* S, bool literal) {
    const auto& B = mrt_basis();
    R rho = 0, m4 = 0, m5 = 0, m6 = 0, m7 = 0, m8 = 0;
    for (int i = 0; i < Q; i++) {
        const R fi = f[i], cx = R(CX[i]), cy = R(CY[i]), cz = R(CZ[i]);
        rho += fi;
        m4 += fi * cx * cy;
        m5 += fi * cx * cz;
        m6 += fi * cy * cz;
        m7 += fi * (cx * cx - cy * cy);
        m8 += fi * (cx * cx - cz * cz);
    }
    R F[Q];
    mrt3_forcing<R>(F, ux, uy, uz, fx, fy, fz);
    const R ux2 = ux * ux, uy2 = uy * uy, uz2 = uz * uz, cs2 = R(1) / R(3), h = R(0.5);
    const R mom = literal ? R(0) : h;
    auto relax = [&](int k, R m, R meq) {
        const R s = R(S[k]);
        return (R(1) - s) * m + s * meq + (R(1) - h * s) * F[k];
    };
    R mp[Q];
    mp[0] = rho;
    mp[1] = rho * ux + mom * F[1];
    mp[2] = rho * uy + mom * F[2];
    mp[3] = rho * uz + mom * F[3];
    mp[4] = relax(4, m4, rho * ux * uy);
    mp[5] = relax(5, m5, rho * ux * uz);
    mp[6] = relax(6, m6, rho * uy * uz);
    mp[7] = relax(7, m7, rho * (ux2 - uy2));
    mp[8] = relax(8, m8, rho * (ux2 - uz2));
    mp[9] = rho * (ux2 + uy2 + uz2 + R(1)) + h * F[9];
    mp[10] = rho * cs2 * ux * (R(3) * uy2 + R(3) * uz2 + R(2)) + h * F[10];
    mp[11] = rho * cs2 * uy * (R(3) * ux2 + R(3) * uz2 + R(2)) + h * F[11];
    mp[12] = rho * cs2 * uz * (R(3) * ux2 + R(3) * uy2 + R(2)) + h * F[12];
    mp[13] = rho * ux * (uy2 - uz2) + h * F[13];
    mp[14] = rho * uy * (ux2 - uz2) + h * F[14];
    mp[15] = rho * uz * (ux2 - uy2) + h * F[15];
    mp[16] = rho * ux * uy * uz + h * F[16];
    mp[17] = rho * cs2 * (R(3) * (ux2 * uy2 + ux2 * uz2 + uy2 * uz2) + R(2) * (ux2 + uy2 + uz2) + R(1)) + h * F[17];
    mp[18] = rho * cs2 * cs2 * (R(9) * (ux2 * uy2 + ux2 * uz2 - uy2 * uz2) + R(6) * ux2 + R(1)) + h * F[18];
    mp[19] = rho * cs2 * (uy2 - uz2) * ((literal ? R(2) : R(3)) * ux2 + R(1)) + h * F[19];
    mp[20] = rho * cs2 * uy * uz * (R(3) * ux2 + R(1)) + h * F[20];
    mp[21] = rho * cs2 * ux * uz * (R(3) * uy2 + R(1)) + h * F[21];
    mp[22] = rho * cs2 * ux * uy * (R(3) * uz2 + R(1)) + h * F[22];
    mp[23] = rho * cs2 * cs2 * ux * (R(3) * uy2 + R(1)) * (R(3) * uz2 + R(1)) + h * F[23];
    mp[24] = rho * cs2 * cs2 * uy * (R(3) * ux2 + R(1)) * (R(3) * uz2 + R(1)) + h * F[24];
    mp[25] = rho * cs2 * cs2 * uz * (R(3) * ux2 + R(1)) * (R(3) * uy2 + R(1)) + h * F[25];
    mp[26] = rho * cs2 * cs2 * cs2 * (R(3) * ux2 + R(1)) * (R(3) * uy2 + R(1)) * (R(3) * uz2 + R(1)) + h * F[26];
    for (int k = 0; k < Q; k++) f[k] = 0;
    for (int i = 0; i < Q; i++)
        for (int k = 0; k < Q; k++) f[k] += R(B.Minv[k * Q + i]) * mp[i];
}

// CM<3, NoAdapter>::apply (CM_D3Q27.cuh:1-105, the De Rosis (2019) universal central-moment
// operator; its sympy-expanded expressions are restated here by their central moments). With
// K_abc = sum f (c_x - u_x)^a (c_y - u_y)^b (c_z - u_z)^c about the corrected u and e = (1, 0, cs2)
// per axis, the post-collision central moments are
//   K*_abc = rho e_a e_b e_c + (F_x/2) [a=1] e_b e_c + (F_y/2) e_a [b=1] e_c + (F_z/2) e_a e_b [c=1]
// except the five deviatoric second moments k4 = K110, k5 = K101, k6 = K011, k7 = K200 - K020,
// k8 = K200 - K002, which become (1 - omega) k (the trace K200 + K020 + K002 is set to rho).
// f* = T(u)^-1 K*, with T(u) the 27x27 central-moment matrix inverted per node (Gauss-Jordan).
template <class R>
inline void collide_cm3(R* f, R ux, R uy, R uz, R fx, R fy, R fz, R omega) {
    const R u[3] = {ux, uy, uz}, F[3] = {fx, fy, fz};
    R rho = 0, k4 = 0, k5 = 0, k6 = 0, k7 = 0, k8 = 0;
    for (int i = 0; i < Q; i++) {
        const R cx = R(CX[i]) - ux, cy = R(CY[i]) - uy, cz = R(CZ[i]) - uz;
        rho += f[i];
        k4 += f[i] * cx * cy;
        k5 += f[i] * cx * cz;
        k6 += f[i] * cy * cz;
        k7 += f[i] * (cx * cx - cy * cy);
        k8 += f[i] * (cx * cx - cz * cz);
    }
    const R e[3] = {R(1), R(0), R(1) / R(3)};
    R K[3][3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            for (int c = 0; c < 3; c++) {
                R v = rho * e[a] * e[b] * e[c];
                if (a == 1) v += R(0.5) * F[0] * e[b] * e[c];
                if (b == 1) v += R(0.5) * F[1] * e[a] * e[c];
                if (c == 1) v += R(0.5) * F[2] * e[a] * e[b];
                K[a][b][c] = v;
            }
    const R w = R(1) - omega;
    K[1][1][0] = w * k4;
    K[1][0][1] = w * k5;
    K[0][1][1] = w * k6;
    K[2][0][0] = (rho + w * (k7 + k8)) / R(3);
    K[0][2][0] = K[2][0][0] - w * k7;
    K[0][0][2] = K[2][0][0] - w * k8;
    // T(u): row (a,b,c), column i: (c_x-u_x)^a (c_y-u_y)^b (c_z-u_z)^c; solve T f = K
    double A[Q][Q + 1];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            for (int c = 0; c < 3; c++) {
                const int r = (a * 3 + b) * 3 + c;
                for (int i = 0; i < Q; i++) {
                    const double dx = double(CX[i]) - double(u[0]), dy = double(CY[i]) - double(u[1]),
                                 dz = double(CZ[i]) - double(u[2]);
                    A[r][i] = std::pow(dx, a) * std::pow(dy, b) * std::pow(dz, c);
                }
                A[r][Q] = double(K[a][b][c]);
            }
    for (int col = 0; col < Q; col++) {
        int piv = col;
        for (int r = col + 1; r < Q; r++)
            if (std::fabs(A[r][col]) > std::fabs(A[piv][col])) piv = r;
        for (int j = 0; j <= Q; j++) std::swap(A[col][j], A[piv][j]);
        const double d = A[col][col];
        for (int j = 0; j <= Q; j++) A[col][j] /= d;
        for (int r = 0; r < Q; r++) {
            if (r == col) continue;
            const double m = A[r][col];
            if (m != 0.0)
                for (int j = 0; j <= Q; j++) A[r][j] -= m * A[col][j];
        }
    }
    for (int i = 0; i < Q; i++) f[i] = R(A[i][Q]);
}

template <class R>
struct Solver3 {
    lbm3d_config cfg;
    int nx, ny, nz;
    size_t N;
    bool eq_literal = false;
    std::vector<R> f, f_back;      // SoA [q][node] (lbm_constants.cuh:320-328)
    std::vector<R> rho, u, force;  // u, force SoA [c][node]
    std::vector<uint8_t> flags;
    std::vector<R> force_field;    // AoS [node][3]; empty => uniform cfg.force
    int64_t timestep = 0;

    explicit Solver3(const lbm3d_config& c)
        : cfg(c), nx(c.nx), ny(c.ny), nz(c.nz), N(size_t(c.nx) * size_t(c.ny) * size_t(c.nz)) {
        if (nx < 3 || ny < 3 || nz < 3) throw std::runtime_error("grid must be at least 3x3x3");
        if (c.collision != LBM_BGK && c.collision != LBM_MRT && c.collision != LBM_CM)
            throw std::runtime_error("3-D operator not restated");
        mrt_basis();
        eq_literal = (c.quirks & LBM_QUIRK_EQ3D_LINEAR) != 0;
        f.assign(Q * N, R(0));
        f_back.assign(Q * N, R(0));
        rho.assign(N, R(1));
        u.assign(3 * N, R(0));
        force.assign(3 * N, R(0));
        flags.assign(N, 0);
    }

    size_t node(int x, int y, int z) const { return (size_t(z) * ny + y) * nx + x; }  // lbm_constants.cuh:351-358

    // Storage rounding of the shifted deviation f - w rho0 (FP32 arithmetic, as the GPU).
    R store_round(int q, R v) const {
        if (cfg.storage == LBM_F32) return v;
        float w0 = weight<float>(q) * cfg.rho0;
        float d = float(v) - w0;
        float back = cfg.storage == LBM_F16_SHIFTED ? oracle::f16_to_f32(oracle::f32_to_f16(d))
                                                    : oracle::bf16_to_f32(oracle::f32_to_bf16(d));
        return R(back + w0);
    }
    void quantize(std::vector<R>& buf) {
        if (cfg.storage == LBM_F32) return;
        for (int q = 0; q < Q; q++)
            for (size_t n = 0; n < N; n++) buf[q * N + n] = store_round(q, buf[q * N + n]);
    }

    // init_kernel<3> + init_node (init.cuh:10-43): f = f_back = f_eq(rho, u).
    void init_from_macros(const R* rho_in, const R* u_aos) {
        for (size_t n = 0; n < N; n++) {
            rho[n] = rho_in[n];
            for (int c = 0; c < 3; c++) u[c * N + n] = u_aos[3 * n + c];
            R feq[Q];
            equilibrium<R>(rho[n], u[n], u[N + n], u[2 * N + n], feq, eq_literal);
            for (int q = 0; q < Q; q++) f[q * N + n] = f_back[q * N + n] = feq[q];
        }
        quantize(f);
        quantize(f_back);
        timestep = 0;
    }

    // LBM<3>::stream_node (streaming.cu:35-65): push f[node][i] to the neighbour's f_back slot,
    // wrapping on periodic axes; out-of-domain targets are skipped (the slot stays stale).
    // The rest population is copied unless LBM_QUIRK_STALE_REST (loop from i = 1).
    void stream() {
        const int i0 = (cfg.quirks & LBM_QUIRK_STALE_REST) ? 1 : 0;
#pragma omp parallel for collapse(2) schedule(static)
        for (int z = 0; z < nz; z++) {
            for (int y = 0; y < ny; y++) {
                for (int x = 0; x < nx; x++) {
                    const size_t n = node(x, y, z);
                    for (int i = i0; i < Q; i++) {
                        int xn = x + CX[i], yn = y + CY[i], zn = z + CZ[i];
                        if (cfg.periodic_x) xn = (xn + nx) % nx;
                        if (cfg.periodic_y) yn = (yn + ny) % ny;
                        if (cfg.periodic_z) zn = (zn + nz) % nz;
                        if (xn < 0 || xn >= nx || yn < 0 || yn >= ny || zn < 0 || zn >= nz) continue;
                        f_back[i * N + node(xn, yn, zn)] = f[i * N + n];
                    }
                }
            }
        }
    }

    // BounceBack<3>::apply (bbDomainBoundary.cuh:52-90): for i = 1..26 in order, a direction
    // whose neighbour lies outside the domain on a reflecting axis copies f_i into f_opp(i).
    // The reference enables all three axes (boundaries.cuh:98); here the reflecting axes are
    // the non-periodic ones, which is the same whenever no axis is periodic.
    void bc_bounce_back(size_t n, int x, int y, int z) {
        for (int i = 1; i < Q; i++) {
            const int xn = x + CX[i], yn = y + CY[i], zn = z + CZ[i];
            const bool bx = !cfg.periodic_x && (xn < 0 || xn >= nx);
            const bool by = !cfg.periodic_y && (yn < 0 || yn >= ny);
            const bool bz = !cfg.periodic_z && (zn < 0 || zn >= nz);
            if (bx || by || bz) f[OPP[i] * N + n] = f[i * N + n];
        }
    }

    // EdgeCornerBounceBack::apply (edgeCornerBounceBack.cuh:10-49), loop as written: the node's
    // inward normals are counted (x, y, z at 0 or N-1) and the functor returns below two; then
    // for i = 0..26, i_out = (NX*cx > 0) || (NY*cy > 0) || (NZ*cz > 0) and o_out the same for
    // opp(i), and i_out && !o_out copies f_i into f_opp(i).
    void bc_corner_edge_bounce_back(size_t n, int x, int y, int z) {
        int normals[3] = {0, 0, 0};
        if (x == 0) normals[0] = 1;
        else if (x == nx - 1) normals[0] = -1;
        if (y == 0) normals[1] = 1;
        else if (y == ny - 1) normals[1] = -1;
        if (z == 0) normals[2] = 1;
        else if (z == nz - 1) normals[2] = -1;
        int num = 0;
        for (int d = 0; d < 3; d++)
            if (normals[d] != 0) num++;
        if (num < 2) return;
        for (int i = 0; i < Q; i++) {
            const int o = OPP[i];
            const bool i_out = (nx * CX[i] > 0) || (ny * CY[i] > 0) || (nz * CZ[i] > 0);
            const bool o_out = (nx * CX[o] > 0) || (ny * CY[o] > 0) || (nz * CZ[o] > 0);
            if (i_out && !o_out) f[o * N + n] = f[i * N + n];
        }
    }

    // RegularizedBoundary::apply (regularizedBoundary3D.cuh:10-137), loops as written, with the
    // imposed velocity (ux, uy, uz): the face is the last match of the x, y, z tests; rho from
    // the known populations (c[d] * s <= 0: parallel once, incoming twice) over (1 - s u_d); its
    // private de_rosis_eq is the solver's equilibrium in its literal form (eq_literal selects it,
    // as for every other equilibrium here); the unknown populations take f_eq + f_neq of their
    // opposite; then Pi(1) from the updated populations and f = f_eq + w / (2 cs^4) Q : Pi.
    void bc_regularized(size_t n, int x, int y, int z, R ux, R uy, R uz) {
        int nd = -1, ns = 0;
        if (x == 0) { nd = 0; ns = +1; }
        if (x == nx - 1) { nd = 0; ns = -1; }
        if (y == 0) { nd = 1; ns = +1; }
        if (y == ny - 1) { nd = 1; ns = -1; }
        if (z == 0) { nd = 2; ns = +1; }
        if (z == nz - 1) { nd = 2; ns = -1; }
        if (nd == -1) return;
        const int* Cd = nd == 0 ? CX : (nd == 1 ? CY : CZ);
        R rho_in = 0, rho_par = 0;
        for (int i = 0; i < Q; i++) {
            if (Cd[i] * ns <= 0) {
                const R fi = f[i * N + n];
                if (Cd[i] * ns == 0) rho_par += fi;
                else rho_in += fi;
            }
        }
        R rho = rho_par + R(2) * rho_in;
        const R un = nd == 0 ? ux : (nd == 1 ? uy : uz);
        rho /= (R(1) - R(ns) * un);
        R feq[Q];
        equilibrium<R>(rho, ux, uy, uz, feq, eq_literal);
        for (int q = 0; q < Q; q++)
            if (Cd[q] * ns > 0) f[q * N + n] = feq[q] + (f[OPP[q] * N + n] - feq[OPP[q]]);
        const R cs2 = R(1) / R(3);
        R Pxx = 0, Pyy = 0, Pzz = 0, Pxy = 0, Pxz = 0, Pyz = 0;
        for (int q = 0; q < Q; q++) {
            const R fi = f[q * N + n];
            Pxx += R(CX[q] * CX[q]) * fi;
            Pyy += R(CY[q] * CY[q]) * fi;
            Pzz += R(CZ[q] * CZ[q]) * fi;
            Pxy += R(CX[q] * CY[q]) * fi;
            Pxz += R(CX[q] * CZ[q]) * fi;
            Pyz += R(CY[q] * CZ[q]) * fi;
        }
        Pxx -= cs2 * rho + rho * ux * ux;
        Pyy -= cs2 * rho + rho * uy * uy;
        Pzz -= cs2 * rho + rho * uz * uz;
        Pxy -= rho * ux * uy;
        Pxz -= rho * ux * uz;
        Pyz -= rho * uy * uz;
        for (int q = 0; q < Q; q++) {
            const R cx = R(CX[q]), cy = R(CY[q]), cz = R(CZ[q]);
            const R fneq = weight<R>(q) / (R(2) * cs2 * cs2) *
                           ((cx * cx - cs2) * Pxx + (cy * cy - cs2) * Pyy + (cz * cz - cs2) * Pzz +
                            R(2) * (cx * cy * Pxy + cx * cz * Pxz + cy * cz * Pyz));
            f[q * N + n] = feq[q] + fneq;
        }
    }

    // ZG_OutflowBoundary<3>::apply (zeroGradientOutflow.cuh:8-59): the first matching face in the
    // order x = 0, x = NX-1, y = 0, y = NY-1, z = 0, z = NZ-1 gives the inward normal n and the
    // interior node one step along it (an interior node returns); for i = 0..26, c_i . n > 0
    // takes the interior node's value. The interior node is read from `snap`, the post-stream
    // state before any functor (boundaries_kernel_3D runs all functors concurrently).
    void bc_zg(size_t n, int x, int y, int z, const std::vector<R>& snap) {
        int nrm[3] = {0, 0, 0};
        int in[3] = {x, y, z};
        if (x == 0) { nrm[0] = 1; in[0] = 1; }
        else if (x == nx - 1) { nrm[0] = -1; in[0] = nx - 2; }
        else if (y == 0) { nrm[1] = 1; in[1] = 1; }
        else if (y == ny - 1) { nrm[1] = -1; in[1] = ny - 2; }
        else if (z == 0) { nrm[2] = 1; in[2] = 1; }
        else if (z == nz - 1) { nrm[2] = -1; in[2] = nz - 2; }
        else return;
        const size_t ni = node(in[0], in[1], in[2]);
        for (int i = 0; i < Q; i++) {
            const int cn = CX[i] * nrm[0] + CY[i] * nrm[1] + CZ[i] * nrm[2];
            if (cn > 0) f[i * N + n] = snap[i * N + ni];
        }
    }

    // RegularizedOutlet::apply (regularizedOutlet.cuh:9-95) at x = NX-1: rho_n, u_n of the
    // x - 1 neighbour's post-stream populations (read from `snap`, as for the Guo functors);
    // f_eq(rho_n, u_n) in the solver's equilibrium (its private de_rosis_eq is the literal
    // form, selected with the rest by eq_literal); the unknown populations (c_x < 0) take
    // f_eq + (f_opp - f_eq_opp); then Pi(1) of the updated node minus that of the equilibrium,
    // and every population f_eq + w / (2 cs^4) Q : Pi(1).
    void bc_regularized_outlet(size_t n, int x, int y, int z, const std::vector<R>& snap) {
        const size_t nb = node(x - 1, y, z);
        R rho = 0, ux = 0, uy = 0, uz = 0;
        for (int q = 0; q < Q; q++) {
            const R v = snap[q * N + nb];
            rho += v;
            ux += v * R(CX[q]);
            uy += v * R(CY[q]);
            uz += v * R(CZ[q]);
        }
        ux /= rho;
        uy /= rho;
        uz /= rho;
        R feq[Q];
        equilibrium<R>(rho, ux, uy, uz, feq, eq_literal);
        for (int q = 0; q < Q; q++)
            if (CX[q] < 0) f[q * N + n] = feq[q] + (f[OPP[q] * N + n] - feq[OPP[q]]);
        const R cs2 = R(1) / R(3);
        R Pxx = 0, Pyy = 0, Pzz = 0, Pxy = 0, Pxz = 0, Pyz = 0;
        for (int q = 0; q < Q; q++) {
            const R fi = f[q * N + n];
            Pxx += R(CX[q] * CX[q]) * fi;
            Pyy += R(CY[q] * CY[q]) * fi;
            Pzz += R(CZ[q] * CZ[q]) * fi;
            Pxy += R(CX[q] * CY[q]) * fi;
            Pxz += R(CX[q] * CZ[q]) * fi;
            Pyz += R(CY[q] * CZ[q]) * fi;
        }
        Pxx -= cs2 * rho + rho * ux * ux;
        Pyy -= cs2 * rho + rho * uy * uy;
        Pzz -= cs2 * rho + rho * uz * uz;
        Pxy -= rho * ux * uy;
        Pxz -= rho * ux * uz;
        Pyz -= rho * uy * uz;
        for (int q = 0; q < Q; q++) {
            const R cx = R(CX[q]), cy = R(CY[q]), cz = R(CZ[q]);
            const R fneq = weight<R>(q) / (R(2) * cs2 * cs2) *
                           ((cx * cx - cs2) * Pxx + (cy * cy - cs2) * Pyy + (cz * cz - cs2) * Pzz +
                            R(2) * (cx * cy * Pxy + cx * cz * Pxz + cy * cz * Pyz));
            f[q * N + n] = feq[q] + fneq;
        }
    }

    // ExtrapolatedCornerEdgeBoundary<3>::apply (cornerBoundaries.cuh:6-36): every population of
    // the node becomes the second-order equilibrium w rho (1 + c.u/cs2 + (c.u)^2/(2 cs4) -
    // u.u/(2 cs2)) of the node one step inward on each axis where the node sits on a domain face
    // (its post-stream state, from `snap`). LBM_QUIRK_CORNER3D_UZ sums u_z with the flat
    // velocity table read at 6 + i, as :23 does.
    void bc_extrapolated(size_t n, int x, int y, int z, const std::vector<R>& snap) {
        const int ix = x == 0 ? 1 : (x == nx - 1 ? nx - 2 : x);
        const int iy = y == 0 ? 1 : (y == ny - 1 ? ny - 2 : y);
        const int iz = z == 0 ? 1 : (z == nz - 1 ? nz - 2 : z);
        const size_t ni = node(ix, iy, iz);
        const bool quirk = (cfg.quirks & LBM_QUIRK_CORNER3D_UZ) != 0;
        R rho = 0, ux = 0, uy = 0, uz = 0;
        for (int i = 0; i < Q; i++) {
            const R v = snap[i * N + ni];
            rho += v;
            ux += v * R(CX[i]);
            uy += v * R(CY[i]);
            const int k = 6 + i;  // C[3*2+i] of the flat [27][3] table
            uz += v * R(quirk ? (k % 3 == 0 ? CX[k / 3] : (k % 3 == 1 ? CY[k / 3] : CZ[k / 3])) : CZ[i]);
        }
        ux /= rho;
        uy /= rho;
        uz /= rho;
        const R cs2 = R(1) / R(3);
        const R uu = ux * ux + uy * uy + uz * uz;
        for (int i = 0; i < Q; i++) {
            const R cu = R(CX[i]) * ux + R(CY[i]) * uy + R(CZ[i]) * uz;
            f[i * N + n] = weight<R>(i) * rho * (R(1) + cu / cs2 + (cu * cu) / (R(2) * cs2 * cs2) - uu / (R(2) * cs2));
        }
    }

    // Flags restated here: FLUID, BOUNCE_BACK, CORNER_EDGE_BOUNCE_BACK, ZG_OUTFLOW, GUO_VELOCITY_INLET (x = 0),
    // GUO_PRESSURE_OUTLET and REGULARIZED_OUTLET (x = NX-1), EXTRAPOLATED_CORNER_EDGE, the
    // regularized wall and inlet; the placements of the reference's 3-D scenarios
    // (windTestScenario.cuh:13-28, flowPastSquareCylinderScenario.cuh:10-31).
    bool has_guo = false;
    void set_flags(const uint8_t* fl) {
        has_guo = false;
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) {
                    const uint8_t v = fl[node(x, y, z)];
                    if (v == LBM_FLUID || v == LBM_BOUNCE_BACK || v == LBM_CORNER_EDGE_BOUNCE_BACK) continue;
                    if (v == LBM_ZG_OUTFLOW) {
                        has_guo = true;  // reads a neighbour: needs the snapshot
                        continue;
                    }
                    if (v == LBM_REGULARIZED_BOUNCE_BACK || v == LBM_REGULARIZED_INLET_LEFT) continue;
                    if (v == LBM_EXTRAPOLATED_CORNER_EDGE) {
                        has_guo = true;
                        continue;
                    }
                    if ((v == LBM_GUO_VELOCITY_INLET && x == 0) ||
                        ((v == LBM_GUO_PRESSURE_OUTLET || v == LBM_REGULARIZED_OUTLET) && x == nx - 1)) {
                        if (cfg.periodic_x) throw std::runtime_error("Guo inlet/outlet need a non-periodic x axis");
                        has_guo = true;
                        continue;
                    }
                    throw std::runtime_error("3-D flag not restated");
                }
        flags.assign(fl, fl + N);
    }

    // GuoVelocityInlet3D::apply (guoInlet3D.cuh:8-60) and GuoPressureOutlet3D::apply
    // (guoOutlet3D.cuh:6-66): the boundary state (rho_b, u_b) from the known populations, then
    // the unknown ones (c_x > 0 at the inlet, < 0 at the outlet) = f_eq(rho_b, u_b) + the
    // neighbour's non-equilibrium part f(x -+ 1) - f_eq(rho_f, u_f). The neighbour is read from
    // `snap`, the post-stream field before any boundary functor (the reference's functors run
    // concurrently, and one may read a neighbour another is rewriting). rho_f, u_f are the
    // neighbour's, or the node's own under LBM_QUIRK_GUO3D_OWN_MOMENTS.
    void bc_guo(size_t n, int x, int y, int z, bool inlet, const std::vector<R>& snap) {
        R g[Q], gn[Q];
        const size_t nb = node(inlet ? x + 1 : x - 1, y, z);
        for (int q = 0; q < Q; q++) {
            g[q] = f[q * N + n];
            gn[q] = snap[q * N + nb];
        }
        R sum = 0;
        for (int q = 0; q < Q; q++) {
            if (inlet ? CX[q] < 0 : CX[q] > 0) sum += R(2) * g[q];
            else if (CX[q] == 0) sum += g[q];
        }
        R rb, uxb, uyb = 0, uzb = 0;
        if (inlet) {
            uxb = R(cfg.u_inlet);
            rb = sum / (R(1) - uxb);
        } else {
            rb = R(1);
            uxb = sum / rb - R(1);
            R jy = 0, jz = 0;
            for (int q = 0; q < Q; q++)
                if (CX[q] >= 0) {
                    jy += g[q] * R(CY[q]);
                    jz += g[q] * R(CZ[q]);
                }
            uyb = jy / rb;
            uzb = jz / rb;
        }
        R feqb[Q], feqf[Q];
        equilibrium<R>(rb, uxb, uyb, uzb, feqb, eq_literal);
        const R* mom = (cfg.quirks & LBM_QUIRK_GUO3D_OWN_MOMENTS) ? g : gn;
        R rf = 0, jxf = 0, jyf = 0, jzf = 0;
        for (int q = 0; q < Q; q++) {
            rf += mom[q];
            jxf += mom[q] * R(CX[q]);
            jyf += mom[q] * R(CY[q]);
            jzf += mom[q] * R(CZ[q]);
        }
        equilibrium<R>(rf, jxf / rf, jyf / rf, jzf / rf, feqf, eq_literal);
        for (int q = 0; q < Q; q++)
            if (inlet ? CX[q] > 0 : CX[q] < 0) f[q * N + n] = feqb[q] + (gn[q] - feqf[q]);
    }

    // boundaries_kernel_3D (boundaries.cuh:89-138) for the restated flags.
    void apply_boundaries() {
        std::vector<R> snap;
        if (has_guo) snap = f;
#pragma omp parallel for collapse(2) schedule(static)
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) {
                    const size_t n = node(x, y, z);
                    if (flags[n] == LBM_BOUNCE_BACK) bc_bounce_back(n, x, y, z);
                    else if (flags[n] == LBM_CORNER_EDGE_BOUNCE_BACK) bc_corner_edge_bounce_back(n, x, y, z);
                    else if (flags[n] == LBM_ZG_OUTFLOW) bc_zg(n, x, y, z, snap);
                    else if (flags[n] == LBM_REGULARIZED_BOUNCE_BACK) bc_regularized(n, x, y, z, R(0), R(0), R(0));
                    else if (flags[n] == LBM_REGULARIZED_INLET_LEFT) bc_regularized(n, x, y, z, R(cfg.u_inlet), R(0), R(0));
                    else if (flags[n] == LBM_GUO_VELOCITY_INLET) bc_guo(n, x, y, z, true, snap);
                    else if (flags[n] == LBM_GUO_PRESSURE_OUTLET) bc_guo(n, x, y, z, false, snap);
                    else if (flags[n] == LBM_REGULARIZED_OUTLET) bc_regularized_outlet(n, x, y, z, snap);
                    else if (flags[n] == LBM_EXTRAPOLATED_CORNER_EDGE) bc_extrapolated(n, x, y, z, snap);
                }
    }

    // uncorrected_macroscopics_kernel<3> (macroscopics.cu:39-100): sums over i = 0..26 in
    // order, u scaled by 1/rho.
    void uncorrected_macroscopics() {
#pragma omp parallel for schedule(static)
        for (long long ln = 0; ln < (long long)N; ln++) {
            const size_t n = size_t(ln);
            R r = 0, jx = 0, jy = 0, jz = 0;
            for (int i = 0; i < Q; i++) {
                const R fi = f[i * N + n];
                r += fi;
                jx += fi * R(CX[i]);
                jy += fi * R(CY[i]);
                jz += fi * R(CZ[i]);
            }
            rho[n] = r;
            u[n] = jx * (R(1) / r);
            u[N + n] = jy * (R(1) / r);
            u[2 * N + n] = jz * (R(1) / r);
        }
    }

    // reset_forces + Init::apply_forces (macroscopics.cuh:13-48): the uniform force or the
    // per-node field.
    void reset_forces() {
        for (size_t n = 0; n < N; n++)
            for (int c = 0; c < 3; c++)
                force[c * N + n] = force_field.empty() ? R(cfg.force[c]) : force_field[3 * n + c];
    }

    // correct_macroscopics_kernel<3> (macroscopics.cu:112-128): u += F / (2 rho).
    void correct_macroscopics() {
        for (size_t n = 0; n < N; n++)
            for (int c = 0; c < 3; c++) u[c * N + n] += R(0.5) * force[c * N + n] / rho[n];
    }

    // equilibrium + BGK<3>::apply (BGK.cuh:55-96): f - omega (f - f_eq) + Guo source;
    // or MRT<3>::apply.
    void collide() {
        if (cfg.collision == LBM_CM) {
#pragma omp parallel for schedule(static)
            for (long long ln = 0; ln < (long long)N; ln++) {
                const size_t n = size_t(ln);
                R g[Q];
                for (int q = 0; q < Q; q++) g[q] = f[q * N + n];
                collide_cm3<R>(g, u[n], u[N + n], u[2 * N + n], force[n], force[N + n], force[2 * N + n], R(cfg.omega));
                for (int q = 0; q < Q; q++) f[q * N + n] = g[q];
            }
            return;
        }
        if (cfg.collision == LBM_MRT) {
            const bool half = (cfg.quirks & LBM_QUIRK_MRT3_LITERAL) != 0;
#pragma omp parallel for schedule(static)
            for (long long ln = 0; ln < (long long)N; ln++) {
                const size_t n = size_t(ln);
                R g[Q];
                for (int q = 0; q < Q; q++) g[q] = f[q * N + n];
                collide_mrt3<R>(g, u[n], u[N + n], u[2 * N + n], force[n], force[N + n], force[2 * N + n], cfg.S, half);
                for (int q = 0; q < Q; q++) f[q * N + n] = g[q];
            }
            return;
        }
        const R om = R(cfg.omega);
#pragma omp parallel for schedule(static)
        for (long long ln = 0; ln < (long long)N; ln++) {
            const size_t n = size_t(ln);
            const R ux = u[n], uy = u[N + n], uz = u[2 * N + n];
            R feq[Q];
            equilibrium<R>(rho[n], ux, uy, uz, feq, eq_literal);
            for (int q = 0; q < Q; q++) {
                const R ft = guo_source<R>(q, ux, uy, uz, force[n], force[N + n], force[2 * N + n], om);
                R& fi = f[q * N + n];
                fi = fi - om * (fi - feq[q]) + ft;
            }
        }
    }

    // ---- immersed boundary, 3-D (IBMManager<3>::multi_direct, IBMManager.cuh:222-252) ----
    std::vector<R> lag_xyz, lag_ds;   // markers AoS [n][3], area per marker
    std::vector<R> lag_force_sum;     // Lagrangian force summed over the iterations, SoA [3][n]
    int ibm_iters = 3;

    void ibm_set_markers(const R* xyz, const R* ds, int n, int iters) {
        lag_xyz.assign(xyz, xyz + 3 * size_t(n));
        lag_ds.assign(size_t(n), R(1));
        if (ds)
            for (int l = 0; l < n; l++) lag_ds[size_t(l)] = ds[l];
        lag_force_sum.assign(3 * size_t(n), R(0));
        ibm_iters = iters;
    }

    // delta4 (IBMUtils.cuh:23-33) in R.
    static R delta4(R r) {
        const R a = std::abs(r);
        if (a < R(1)) return R(0.125) * (R(3) - R(2) * a + std::sqrt(R(1) + R(4) * a - R(4) * r * r));
        if (a < R(2)) return R(0.125) * (R(5) - R(2) * a - std::sqrt(R(-7) + R(12) * a - R(4) * r * r));
        return R(0);
    }

    // The 4x4x4 stencil of marker l (KERNEL_W = 4, OFFSET = -1; IBMUtils.cuh:12-13) in the
    // reference's loop order (x outer, then y, then z), nodes outside the domain skipped:
    // IBM_impl.cu:53-118 (interpolation) and :156-196 (spreading) visit the same nodes.
    template <class F>
    void stencil(int l, F&& fn) const {
        const R px = lag_xyz[3 * size_t(l)], py = lag_xyz[3 * size_t(l) + 1], pz = lag_xyz[3 * size_t(l) + 2];
        if (!(std::abs(px) < R(1e9) && std::abs(py) < R(1e9) && std::abs(pz) < R(1e9))) return;
        const int gx = int(std::floor(px)), gy = int(std::floor(py)), gz = int(std::floor(pz));
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++)
                for (int k = 0; k < 4; k++) {
                    const int x = gx + i - 1, y = gy + j - 1, z = gz + k - 1;
                    if (x >= nx || x < 0 || y >= ny || y < 0 || z >= nz || z < 0) continue;
                    fn(node(x, y, z), delta4(px - R(x)) * delta4(py - R(y)) * delta4(pz - R(z)));
                }
    }

    // multi_direct: u_prev <- u (uncorrected); ITER times: F_iter <- 0; interpolate rho, u_prev
    // to the markers; F_L = 2 rho_L (0 - u_L) (> 1e-8 kept under LBM_QUIRK_IBM_CLAMP,
    // IBM_impl.cuh:7-26); spread F_L phi (times the marker area unless LBM_QUIRK_IBM_NO_DS: the
    // reference spreads without it); u_prev += F_iter / (2 rho) (clamped > 1e-8 under the clamp
    // quirk, :28-47); force += F_iter (:49-68).
    void multi_direct() {
        const int n = int(lag_ds.size());
        if (n == 0) return;
        const bool clamp = (cfg.quirks & LBM_QUIRK_IBM_CLAMP) != 0, no_ds = (cfg.quirks & LBM_QUIRK_IBM_NO_DS) != 0;
        std::vector<R> u_prev(u), f_iter(3 * N, R(0)), lag_f(3 * size_t(n), R(0));
        std::fill(lag_force_sum.begin(), lag_force_sum.end(), R(0));
        for (int it = 0; it < ibm_iters; it++) {
            std::fill(f_iter.begin(), f_iter.end(), R(0));
            for (int l = 0; l < n; l++) {
                R r = 0, ux = 0, uy = 0, uz = 0;
                stencil(l, [&](size_t m, R k) {
                    r += k * rho[m];
                    ux += k * u_prev[m];
                    uy += k * u_prev[N + m];
                    uz += k * u_prev[2 * N + m];
                });
                const R ul[3] = {ux, uy, uz};
                for (int d = 0; d < 3; d++) {
                    R fl = R(2) * r * (R(0) - ul[d]);
                    if (clamp) fl = fl > R(1e-8) ? fl : R(0);
                    lag_f[size_t(d) * n + l] = fl;
                    lag_force_sum[size_t(d) * n + l] += fl;
                }
            }
            for (int l = 0; l < n; l++) {
                const R a = no_ds ? R(1) : lag_ds[size_t(l)];
                stencil(l, [&](size_t m, R k) {
                    for (int d = 0; d < 3; d++) f_iter[d * N + m] += k * lag_f[size_t(d) * n + l] * a;
                });
            }
            for (size_t m = 0; m < N; m++)
                for (int d = 0; d < 3; d++) {
                    R c = u_prev[d * N + m] + f_iter[d * N + m] / (R(2) * rho[m]);
                    if (clamp) c = c > R(1e-8) ? c : R(0);
                    u_prev[d * N + m] = c;
                    force[d * N + m] += f_iter[d * N + m];
                }
        }
    }

    void step() {
        timestep++;
        stream();
        std::swap(f, f_back);
        apply_boundaries();
        uncorrected_macroscopics();
        reset_forces();
        multi_direct();
        correct_macroscopics();
        collide();
        quantize(f);
    }
};

}  // namespace oracle3d
