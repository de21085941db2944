// lbm3d_node.hpp — node-local D3Q27 physics of the fused 3-D step, FP32 in registers.
//
// The 27 populations of one node live in registers. Every loop over directions is unrolled
// against the constexpr velocity table, so c_i u becomes a signed sum of u components and
// the c = 0 terms disappear at compile time. The functions are __host__ __device__: the CPU
// emulation runs the same code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#include "../../include/lbm3d_hip.h"

#define LBM3_HD __host__ __device__ __forceinline__

namespace lbm3 {

constexpr int Q = 27;
// Velocity set in the reference's order: rest, 6 faces, 12 edges, 8 corners
// (lbm_constants.cuh:76-104), and the opposite directions (lbm_constants.cuh:106-114).
constexpr int CX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
constexpr int CY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, -1};
constexpr int CZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1};
constexpr int OPP[Q] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15, 26, 25, 24, 23, 22, 21, 20, 19};

// Weights by shell (lbm_constants.cuh:64-71).
LBM3_HD constexpr float w3(int q) {
    return q == 0 ? 8.0f / 27.0f : (q < 7 ? 2.0f / 27.0f : (q < 19 ? 1.0f / 54.0f : 1.0f / 216.0f));
}

// sum_k c_k v_k for a compile-time direction, as signed adds (no multiply by 0 or +-1).
template <int q>
LBM3_HD float cdot(float vx, float vy, float vz) {
    float s = 0.0f;
    bool first = true;
    auto add = [&](int c, float v) {
        if (c == 0) return;
        if (first) {
            s = c > 0 ? v : -v;
            first = false;
        } else {
            s = c > 0 ? s + v : s - v;
        }
    };
    add(CX[q], vx);
    add(CY[q], vy);
    add(CZ[q], vz);
    return s;
}

struct Macro3 {
    float rho, ux, uy, uz;
};

// uncorrected_macroscopics_kernel<3> (macroscopics.cu:39-100) followed by the Guo half-force
// correction (macroscopics.cu:112-128): sums over i = 0..26 in the reference's order (the
// c = 0 terms add an exact zero and are left out), one reciprocal for the three divisions.
template <int q>
LBM3_HD void mom_acc(const float* g, float& r, float& jx, float& jy, float& jz) {
    r += g[q];
    if (CX[q] > 0) jx += g[q];
    if (CX[q] < 0) jx -= g[q];
    if (CY[q] > 0) jy += g[q];
    if (CY[q] < 0) jy -= g[q];
    if (CZ[q] > 0) jz += g[q];
    if (CZ[q] < 0) jz -= g[q];
}
template <int... qs>
LBM3_HD void mom_all(const float* g, float& r, float& jx, float& jy, float& jz, std::integer_sequence<int, qs...>) {
    (mom_acc<qs>(g, r, jx, jy, jz), ...);
}
LBM3_HD Macro3 macroscopics3(const float* g, float fx, float fy, float fz) {
    float r = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
    mom_all(g, r, jx, jy, jz, std::make_integer_sequence<int, Q>{});
    const float inv = 1.0f / r;
    Macro3 m;
    m.rho = r;
    m.ux = (jx + 0.5f * fx) * inv;
    m.uy = (jy + 0.5f * fy) * inv;
    m.uz = (jz + 0.5f * fz) * inv;
    return m;
}

// Equilibrium (equilibrium.cu:41-96) and BGK<3>::apply with Guo forcing (BGK.cuh:55-96):
//   f_i <- (1 - omega) f_i + omega w_i rho (1 + lin_i c.u + 4.5 (c.u)^2 - 1.5 u^2)
//          + w_i (1 - omega/2) (3 (c - u).F + 9 (c.u)(c.F)),
// lin_i = 3, or `lin_ec` on the edge and corner directions (1 reproduces the reference's
// De Rosis expression, LBM_QUIRK_EQ3D_LINEAR).
struct BgkK {
    float rho, ux, uy, uz, fx, fy, fz;
    float om, om1, k3, k9, uF3, base, lin_ec;
};
template <int q>
LBM3_HD void bgk_dir(float* g, const BgkK& K) {
    const float cu = cdot<q>(K.ux, K.uy, K.uz);
    const float cF = cdot<q>(K.fx, K.fy, K.fz);
    const float lin = q < 7 ? 3.0f : K.lin_ec;
    const float w = w3(q);
    const float feq = w * (K.base + K.rho * cu * (lin + 4.5f * cu));
    const float src = w * (K.k3 * cF - K.uF3 + K.k9 * cu * cF);
    g[q] = K.om1 * g[q] + K.om * feq + src;
}
template <int... qs>
LBM3_HD void bgk_all(float* g, const BgkK& K, std::integer_sequence<int, qs...>) {
    (bgk_dir<qs>(g, K), ...);
}
LBM3_HD void collide_bgk3(float* g, const Macro3& m, float fx, float fy, float fz, float omega, float lin_ec) {
    BgkK K;
    K.rho = m.rho;
    K.ux = m.ux;
    K.uy = m.uy;
    K.uz = m.uz;
    K.fx = fx;
    K.fy = fy;
    K.fz = fz;
    K.om = omega;
    K.om1 = 1.0f - omega;
    const float k = 1.0f - 0.5f * omega;
    K.k3 = 3.0f * k;
    K.k9 = 9.0f * k;
    K.uF3 = K.k3 * (m.ux * fx + m.uy * fy + m.uz * fz);
    K.base = m.rho * (1.0f - 1.5f * (m.ux * m.ux + m.uy * m.uy + m.uz * m.uz));
    K.lin_ec = lin_ec;
    bgk_all(g, K, std::make_integer_sequence<int, Q>{});
}

// MRT<3>::apply (MRT.cu:397-456) with compute_m_post_component (MRT.cu:155-201), without the
// 27x27 matrices. Every post-collision moment the reference sets from (rho, u) is a raw moment
// of the product-form equilibrium f^p = rho * phi_x(c_x) phi_y(c_y) phi_z(c_z), phi(0) = 2/3 -
// u_a^2, phi(+-1) = (1/3 + u_a^2 +- u_a)/2 (1-D moments 1, u_a, 1/3 + u_a^2). So
//   f* = f^p + M^-1 d,
// where d, the post-collision moments minus those of f^p, is non-zero only in the relaxed
// deviatoric second moments 4-8 ((1 - S)(m - m^eq) + (1 - S/2) F_k) and the forcing terms
// (F_k / 2, F_k of MRT.cu:119-153). M^-1 d is taken in two stages: the basis combinations are
// undone into tensor moments T_abc = sum f c_x^a c_y^b c_z^c, which are then inverted one axis
// at a time (h(0) = T_0 - T_2, h(+-1) = (T_2 +- T_1)/2). `literal` keeps the reference's
// momentum without F/2 and its (2 ux^2 + 1) in moment 19 (LBM_QUIRK_MRT3_LITERAL).
struct Mrt3Rates {
    float s[5];  // S[4..8]
    int literal;
};
constexpr int q_of(int cx, int cy, int cz) {
    for (int q = 0; q < Q; q++)
        if (CX[q] == cx && CY[q] == cy && CZ[q] == cz) return q;
    return -1;
}
template <int q>
LBM3_HD void m2_acc(const float* g, float& m4, float& m5, float& m6, float& m7, float& m8) {
    constexpr int xy = CX[q] * CY[q], xz = CX[q] * CZ[q], yz = CY[q] * CZ[q];
    constexpr int d7 = CX[q] * CX[q] - CY[q] * CY[q], d8 = CX[q] * CX[q] - CZ[q] * CZ[q];
    if (xy > 0) m4 += g[q];
    if (xy < 0) m4 -= g[q];
    if (xz > 0) m5 += g[q];
    if (xz < 0) m5 -= g[q];
    if (yz > 0) m6 += g[q];
    if (yz < 0) m6 -= g[q];
    if (d7 > 0) m7 += g[q];
    if (d7 < 0) m7 -= g[q];
    if (d8 > 0) m8 += g[q];
    if (d8 < 0) m8 -= g[q];
}
template <int... qs>
LBM3_HD void m2_all(const float* g, float& m4, float& m5, float& m6, float& m7, float& m8, std::integer_sequence<int, qs...>) {
    (m2_acc<qs>(g, m4, m5, m6, m7, m8), ...);
}
LBM3_HD void collide_mrt3(float* g, const Macro3& m, float fx, float fy, float fz, const Mrt3Rates& R) {
    const float rho = m.rho, ux = m.ux, uy = m.uy, uz = m.uz;
    float m4 = 0.0f, m5 = 0.0f, m6 = 0.0f, m7 = 0.0f, m8 = 0.0f;
    m2_all(g, m4, m5, m6, m7, m8, std::make_integer_sequence<int, Q>{});
    const float ux2 = ux * ux, uy2 = uy * uy, uz2 = uz * uz;
    const float Fu = fx * ux + fy * uy + fz * uz;
    auto dev = [&](int k, float mk, float meq, float Fk) {
        const float sk = R.s[k - 4];
        return (1.0f - sk) * (mk - meq) + (1.0f - 0.5f * sk) * Fk;
    };
    const float half = R.literal ? 0.0f : 0.5f;
    const float d4 = dev(4, m4, rho * ux * uy, fx * uy + fy * ux);
    const float d5 = dev(5, m5, rho * ux * uz, fx * uz + fz * ux);
    const float d6 = dev(6, m6, rho * uy * uz, fy * uz + fz * uy);
    const float d7 = dev(7, m7, rho * (ux2 - uy2), 2.0f * (fx * ux - fy * uy));
    const float d8 = dev(8, m8, rho * (ux2 - uz2), 2.0f * (fx * ux - fz * uz));
    const float d9 = Fu;
    const float d17 = (2.0f / 3.0f) * Fu, d18 = (2.0f / 3.0f) * fx * ux;
    float d19 = (1.0f / 3.0f) * (fy * uy - fz * uz);
    if (R.literal) d19 -= (1.0f / 3.0f) * rho * ux2 * (uy2 - uz2);
    // tensor moments T[a][b][c] of M^-1 d
    float T[3][3][3];
    T[0][0][0] = 0.0f;
    T[1][0][0] = half * fx;
    T[0][1][0] = half * fy;
    T[0][0][1] = half * fz;
    T[1][1][0] = d4;
    T[1][0][1] = d5;
    T[0][1][1] = d6;
    T[2][0][0] = (d7 + d8 + d9) * (1.0f / 3.0f);
    T[0][2][0] = T[2][0][0] - d7;
    T[0][0][2] = T[2][0][0] - d8;
    T[1][2][0] = T[1][0][2] = fx * (1.0f / 6.0f);  // (d10 +- d13)/2, d10 = fx/3, d13 = 0
    T[2][1][0] = T[0][1][2] = fy * (1.0f / 6.0f);
    T[2][0][1] = T[0][2][1] = fz * (1.0f / 6.0f);
    T[1][1][1] = 0.0f;
    const float sxy = 0.5f * (d17 + d18);
    T[0][2][2] = 0.5f * (d17 - d18);
    T[2][2][0] = 0.5f * (sxy + d19);
    T[2][0][2] = 0.5f * (sxy - d19);
    T[2][1][1] = (1.0f / 6.0f) * (fy * uz + fz * uy);
    T[1][2][1] = (1.0f / 6.0f) * (fx * uz + fz * ux);
    T[1][1][2] = (1.0f / 6.0f) * (fx * uy + fy * ux);
    T[1][2][2] = fx * (1.0f / 18.0f);
    T[2][1][2] = fy * (1.0f / 18.0f);
    T[2][2][1] = fz * (1.0f / 18.0f);
    T[2][2][2] = (1.0f / 9.0f) * Fu;
    // invert along z, y, x: index 0 -> c = -1, 1 -> c = 0, 2 -> c = +1
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const float t0 = T[a][b][0], t1 = T[a][b][1], t2 = T[a][b][2];
            T[a][b][0] = 0.5f * (t2 - t1);
            T[a][b][1] = t0 - t2;
            T[a][b][2] = 0.5f * (t2 + t1);
        }
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float t0 = T[a][0][c], t1 = T[a][1][c], t2 = T[a][2][c];
            T[a][0][c] = 0.5f * (t2 - t1);
            T[a][1][c] = t0 - t2;
            T[a][2][c] = 0.5f * (t2 + t1);
        }
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float t0 = T[0][b][c], t1 = T[1][b][c], t2 = T[2][b][c];
            T[0][b][c] = 0.5f * (t2 - t1);
            T[1][b][c] = t0 - t2;
            T[2][b][c] = 0.5f * (t2 + t1);
        }
    // product-form equilibrium
    const float third = 1.0f / 3.0f;
    const float px[3] = {0.5f * (third + ux2 - ux), 2.0f * third - ux2, 0.5f * (third + ux2 + ux)};
    const float py[3] = {0.5f * (third + uy2 - uy), 2.0f * third - uy2, 0.5f * (third + uy2 + uy)};
    const float pz[3] = {0.5f * (third + uz2 - uz), 2.0f * third - uz2, 0.5f * (third + uz2 + uz)};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float ra = rho * px[a];
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const float rab = ra * py[b];
#pragma unroll
            for (int c = 0; c < 3; c++) g[q_of(a - 1, b - 1, c - 1)] = rab * pz[c] + T[a][b][c];
        }
    }
}

// CM<3, NoAdapter>::apply (CM_D3Q27.cuh:1-105; De Rosis 2019 universal central moments), in
// factorised form. Central moments K_abc = sum f (c-u)_x^a (c-u)_y^b (c-u)_z^c about the
// corrected u. After collision, with e = (1, 0, 1/3) per axis:
//   K*_abc = rho e_a e_b e_c + (F_x/2)[a=1] e_b e_c + (F_y/2) e_a [b=1] e_c + (F_z/2) e_a e_b [c=1],
// except the deviatoric second moments K110, K101, K011, K200-K020, K200-K002, which become
// (1 - omega) times their pre-collision values (trace K200 + K020 + K002 = rho). Those five come
// from the raw second moments shifted by u (j = rho u - F/2). f* is then rebuilt one axis at a
// time from central moments (M0, M1, M2) about u_a:
//   h(0) = M0 (1 - u^2) - 2u M1 - M2,   h(+-1) = (M0 (u^2 +- u) + M1 (2u +- 1) + M2) / 2.
LBM3_HD void cm_shift_inv(float& m0, float& m1, float& m2, float u) {
    const float u2 = u * u;
    const float hm = 0.5f * (m0 * (u2 - u) + m1 * (2.0f * u - 1.0f) + m2);
    const float h0 = m0 * (1.0f - u2) - 2.0f * u * m1 - m2;
    const float hp = 0.5f * (m0 * (u2 + u) + m1 * (2.0f * u + 1.0f) + m2);
    m0 = hm;
    m1 = h0;
    m2 = hp;
}
LBM3_HD void collide_cm3(float* g, const Macro3& m, float fx, float fy, float fz, float omega) {
    const float rho = m.rho, ux = m.ux, uy = m.uy, uz = m.uz;
    float m4 = 0.0f, m5 = 0.0f, m6 = 0.0f, m7 = 0.0f, m8 = 0.0f;
    m2_all(g, m4, m5, m6, m7, m8, std::make_integer_sequence<int, Q>{});
    const float jx = rho * ux - 0.5f * fx, jy = rho * uy - 0.5f * fy, jz = rho * uz - 0.5f * fz;
    const float w = 1.0f - omega;
    const float k4 = m4 - ux * jy - uy * jx + rho * ux * uy;
    const float k5 = m5 - ux * jz - uz * jx + rho * ux * uz;
    const float k6 = m6 - uy * jz - uz * jy + rho * uy * uz;
    const float k7 = m7 - 2.0f * (ux * jx - uy * jy) + rho * (ux * ux - uy * uy);
    const float k8 = m8 - 2.0f * (ux * jx - uz * jz) + rho * (ux * ux - uz * uz);
    const float e[3] = {1.0f, 0.0f, 1.0f / 3.0f};
    const float hf[3] = {0.5f * fx, 0.5f * fy, 0.5f * fz};
    float K[3][3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                float v = rho * e[a] * e[b] * e[c];
                if (a == 1) v += hf[0] * e[b] * e[c];
                if (b == 1) v += hf[1] * e[a] * e[c];
                if (c == 1) v += hf[2] * e[a] * e[b];
                K[a][b][c] = v;
            }
    K[1][1][0] = w * k4;
    K[1][0][1] = w * k5;
    K[0][1][1] = w * k6;
    K[2][0][0] = (1.0f / 3.0f) * (rho + w * (k7 + k8));
    K[0][2][0] = K[2][0][0] - w * k7;
    K[0][0][2] = K[2][0][0] - w * k8;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) cm_shift_inv(K[a][b][0], K[a][b][1], K[a][b][2], uz);
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) cm_shift_inv(K[a][0][c], K[a][1][c], K[a][2][c], uy);
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
        for (int c = 0; c < 3; c++) cm_shift_inv(K[0][b][c], K[1][b][c], K[2][b][c], ux);
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int c = 0; c < 3; c++) g[q_of(a - 1, b - 1, c - 1)] = K[a][b][c];
}

// Equilibrium of direction q (equilibrium.cu:41-96), lin_ec as in collide_bgk3; base is
// rho (1 - 1.5 u^2).
template <int q>
LBM3_HD float feq3(float rho, float ux, float uy, float uz, float base, float lin_ec) {
    const float cu = cdot<q>(ux, uy, uz);
    const float lin = q < 7 ? 3.0f : lin_ec;
    return w3(q) * (base + rho * cu * (lin + 4.5f * cu));
}

// The unknown populations of GuoVelocityInlet3D (c_x > 0 at x = 0, guoInlet3D.cuh:8-60) and
// GuoPressureOutlet3D (c_x < 0 at x = NX-1, guoOutlet3D.cuh:6-66): f_eq(rho_b, u_b) plus the
// neighbour's non-equilibrium part gn_q - f_eq_q(rho_f, u_f). The boundary state comes from the
// node's own known populations: inlet rho_b = (sum_{c_x=0} g + 2 sum_{c_x<0} g) / (1 - u_in),
// u_b = (u_in, 0, 0); outlet rho_b = 1, u_x,b = sum_{c_x=0} g + 2 sum_{c_x>0} g - 1 and
// u_y,b, u_z,b the momentum of the c_x >= 0 populations.
struct GuoBc {
    bool inlet;
    float u_in, lin_ec;
    float rf, uxf, uyf, uzf;  // neighbour (or own) density and velocity
};
template <bool INLET, int q>
LBM3_HD void guo_known(const float* g, float& sum, float& jy, float& jz) {
    if (INLET ? CX[q] < 0 : CX[q] > 0) sum += 2.0f * g[q];
    else if (CX[q] == 0) sum += g[q];
    if (!INLET && CX[q] >= 0) {
        if (CY[q] > 0) jy += g[q];
        if (CY[q] < 0) jy -= g[q];
        if (CZ[q] > 0) jz += g[q];
        if (CZ[q] < 0) jz -= g[q];
    }
}
template <bool INLET, int... qs>
LBM3_HD void guo_known_all(const float* g, float& sum, float& jy, float& jz, std::integer_sequence<int, qs...>) {
    (guo_known<INLET, qs>(g, sum, jy, jz), ...);
}
// gn_unknown[k]: the neighbour's post-stream value of the k-th unknown direction (increasing q).
template <bool INLET>
LBM3_HD void bc_guo3(float* g, const float* gn_unknown, const GuoBc& B) {
    float sum = 0.0f, jy = 0.0f, jz = 0.0f;
    guo_known_all<INLET>(g, sum, jy, jz, std::make_integer_sequence<int, Q>{});
    float rb, uxb, uyb = 0.0f, uzb = 0.0f;
    if (INLET) {
        uxb = B.u_in;
        rb = sum / (1.0f - uxb);
    } else {
        rb = 1.0f;
        uxb = sum - 1.0f;
        uyb = jy;
        uzb = jz;
    }
    const float base_b = rb * (1.0f - 1.5f * (uxb * uxb + uyb * uyb + uzb * uzb));
    const float base_f = B.rf * (1.0f - 1.5f * (B.uxf * B.uxf + B.uyf * B.uyf + B.uzf * B.uzf));
    int k = 0;
#define LBM3_GUO_DIR(q)                                                                                \
    if (INLET ? CX[q] > 0 : CX[q] < 0) {                                                               \
        g[q] = feq3<q>(rb, uxb, uyb, uzb, base_b, B.lin_ec) +                                          \
               (gn_unknown[k] - feq3<q>(B.rf, B.uxf, B.uyf, B.uzf, base_f, B.lin_ec));                 \
        k++;                                                                                           \
    }
    LBM3_GUO_DIR(0) LBM3_GUO_DIR(1) LBM3_GUO_DIR(2) LBM3_GUO_DIR(3) LBM3_GUO_DIR(4) LBM3_GUO_DIR(5)
    LBM3_GUO_DIR(6) LBM3_GUO_DIR(7) LBM3_GUO_DIR(8) LBM3_GUO_DIR(9) LBM3_GUO_DIR(10) LBM3_GUO_DIR(11)
    LBM3_GUO_DIR(12) LBM3_GUO_DIR(13) LBM3_GUO_DIR(14) LBM3_GUO_DIR(15) LBM3_GUO_DIR(16) LBM3_GUO_DIR(17)
    LBM3_GUO_DIR(18) LBM3_GUO_DIR(19) LBM3_GUO_DIR(20) LBM3_GUO_DIR(21) LBM3_GUO_DIR(22) LBM3_GUO_DIR(23)
    LBM3_GUO_DIR(24) LBM3_GUO_DIR(25) LBM3_GUO_DIR(26)
#undef LBM3_GUO_DIR
}

// BounceBack<3>::apply (bbDomainBoundary.cuh:52-90): for i = 1..26 in order, a direction whose
// neighbour lies outside the domain on a non-periodic axis copies g_i into g_opp(i).
// Sequential as written, so edge and corner nodes match the reference's loop order.
struct Box3 {
    int nx, ny, nz, px, py, pz;
};
template <int i>
LBM3_HD void bb_dir(float* g, int x, int y, int z, const Box3& B) {
    bool out = false;
    if (CX[i] != 0 && !B.px) out = out || (x + CX[i] < 0 || x + CX[i] >= B.nx);
    if (CY[i] != 0 && !B.py) out = out || (y + CY[i] < 0 || y + CY[i] >= B.ny);
    if (CZ[i] != 0 && !B.pz) out = out || (z + CZ[i] < 0 || z + CZ[i] >= B.nz);
    if (out) g[OPP[i]] = g[i];
}
template <int... is>
LBM3_HD void bb_all(float* g, int x, int y, int z, const Box3& B, std::integer_sequence<int, is...>) {
    (bb_dir<is + 1>(g, x, y, z, B), ...);
}
LBM3_HD void bc_bounce_back3(float* g, int x, int y, int z, const Box3& B) {
    bb_all(g, x, y, z, B, std::make_integer_sequence<int, Q - 1>{});
}

// RegularizedBoundary::apply (regularizedBoundary3D.cuh:10-137) with the imposed velocity u:
//  - the face: the last match in the order x = 0, x = NX-1, y = 0, y = NY-1, z = 0, z = NZ-1
//    gives the axis d and the inward sign s (an interior node returns);
//  - rho = (sum_{c_d s = 0} g + 2 sum_{c_d s < 0} g) / (1 - s u_d);
//  - the unknown populations (c_d s > 0) become f_eq + (g_opp - f_eq_opp);
//  - Pi = sum c c g - rho/3 I - rho u u, then every population becomes f_eq + 4.5 w Q : Pi with
//    Q = c c - I/3 (the w / (2 cs^4) projection).
// The functor's own equilibrium (de_rosis_eq) is the solver's (lin_ec 1 = the reference's form).
template <int D, int S, int q>
constexpr int reg_cn() {
    return (D == 0 ? CX[q] : (D == 1 ? CY[q] : CZ[q])) * S;
}
template <int D, int S, int q>
LBM3_HD void reg_known(const float* g, float& par, float& inc) {
    if constexpr (reg_cn<D, S, q>() == 0) par += g[q];
    else if constexpr (reg_cn<D, S, q>() < 0) inc += g[q];
}
template <int D, int S, int q>
LBM3_HD void reg_unknown(float* g, float rho, float ux, float uy, float uz, float base, float lin_ec) {
    if constexpr (reg_cn<D, S, q>() > 0)
        g[q] = feq3<q>(rho, ux, uy, uz, base, lin_ec) + (g[OPP[q]] - feq3<OPP[q]>(rho, ux, uy, uz, base, lin_ec));
}
template <int q>
LBM3_HD void reg_pi(const float* g, float* P) {
    if constexpr (CX[q] != 0) P[0] += g[q];
    if constexpr (CY[q] != 0) P[1] += g[q];
    if constexpr (CZ[q] != 0) P[2] += g[q];
    if constexpr (CX[q] * CY[q] != 0) P[3] += float(CX[q] * CY[q]) * g[q];
    if constexpr (CX[q] * CZ[q] != 0) P[4] += float(CX[q] * CZ[q]) * g[q];
    if constexpr (CY[q] * CZ[q] != 0) P[5] += float(CY[q] * CZ[q]) * g[q];
}
template <int q>
LBM3_HD void reg_rebuild(float* g, const float* P, float rho, float ux, float uy, float uz, float base, float lin_ec) {
    constexpr float t = 1.0f / 3.0f;
    const float qp = (float(CX[q] * CX[q]) - t) * P[0] + (float(CY[q] * CY[q]) - t) * P[1] +
                     (float(CZ[q] * CZ[q]) - t) * P[2] +
                     2.0f * (float(CX[q] * CY[q]) * P[3] + float(CX[q] * CZ[q]) * P[4] + float(CY[q] * CZ[q]) * P[5]);
    g[q] = feq3<q>(rho, ux, uy, uz, base, lin_ec) + 4.5f * w3(q) * qp;
}
template <int D, int S, int... qs>
LBM3_HD void bc_regularized3_face(float* g, float ux, float uy, float uz, float lin_ec, std::integer_sequence<int, qs...>) {
    float par = 0.0f, inc = 0.0f;
    (reg_known<D, S, qs>(g, par, inc), ...);
    const float un = D == 0 ? ux : (D == 1 ? uy : uz);
    const float rho = (par + 2.0f * inc) / (1.0f - float(S) * un);
    const float base = rho * (1.0f - 1.5f * (ux * ux + uy * uy + uz * uz));
    (reg_unknown<D, S, qs>(g, rho, ux, uy, uz, base, lin_ec), ...);
    float P[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    (reg_pi<qs>(g, P), ...);
    P[0] -= rho / 3.0f + rho * ux * ux;
    P[1] -= rho / 3.0f + rho * uy * uy;
    P[2] -= rho / 3.0f + rho * uz * uz;
    P[3] -= rho * ux * uy;
    P[4] -= rho * ux * uz;
    P[5] -= rho * uy * uz;
    (reg_rebuild<qs>(g, P, rho, ux, uy, uz, base, lin_ec), ...);
}
// One instantiation per face (axis and inward sign known at compile time).
LBM3_HD void bc_regularized3(float* g, int x, int y, int z, const Box3& B, float ux, float uy, float uz, float lin_ec) {
    int face = -1;
    if (x == 0) face = 0;
    if (x == B.nx - 1) face = 1;
    if (y == 0) face = 2;
    if (y == B.ny - 1) face = 3;
    if (z == 0) face = 4;
    if (z == B.nz - 1) face = 5;
    constexpr auto qs = std::make_integer_sequence<int, Q>{};
    switch (face) {
        case 0: bc_regularized3_face<0, 1>(g, ux, uy, uz, lin_ec, qs); break;
        case 1: bc_regularized3_face<0, -1>(g, ux, uy, uz, lin_ec, qs); break;
        case 2: bc_regularized3_face<1, 1>(g, ux, uy, uz, lin_ec, qs); break;
        case 3: bc_regularized3_face<1, -1>(g, ux, uy, uz, lin_ec, qs); break;
        case 4: bc_regularized3_face<2, 1>(g, ux, uy, uz, lin_ec, qs); break;
        case 5: bc_regularized3_face<2, -1>(g, ux, uy, uz, lin_ec, qs); break;
        default: break;
    }
}

// RegularizedOutlet::apply (regularizedOutlet.cuh:9-95) at x = NX-1, given the density and
// velocity (rho, u) of the x - 1 neighbour's post-stream populations: the unknown populations
// (c_x < 0) become f_eq + (g_opp - f_eq_opp) with f_eq(rho, u) (the functor's own de_rosis_eq
// is the solver's equilibrium: lin_ec as everywhere else); then Pi = sum c c g - rho/3 I -
// rho u u of the completed node, and every population f_eq + 4.5 w Q : Pi.
template <int... qs>
LBM3_HD void bc_regularized_outlet3_all(float* g, float rho, float ux, float uy, float uz, float lin_ec,
                                        std::integer_sequence<int, qs...>) {
    const float base = rho * (1.0f - 1.5f * (ux * ux + uy * uy + uz * uz));
    (reg_unknown<0, -1, qs>(g, rho, ux, uy, uz, base, lin_ec), ...);
    float P[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    (reg_pi<qs>(g, P), ...);
    P[0] -= rho / 3.0f + rho * ux * ux;
    P[1] -= rho / 3.0f + rho * uy * uy;
    P[2] -= rho / 3.0f + rho * uz * uz;
    P[3] -= rho * ux * uy;
    P[4] -= rho * ux * uz;
    P[5] -= rho * uy * uz;
    (reg_rebuild<qs>(g, P, rho, ux, uy, uz, base, lin_ec), ...);
}
LBM3_HD void bc_regularized_outlet3(float* g, float rho, float ux, float uy, float uz, float lin_ec) {
    bc_regularized_outlet3_all(g, rho, ux, uy, uz, lin_ec, std::make_integer_sequence<int, Q>{});
}

// ExtrapolatedCornerEdgeBoundary<3>::apply (cornerBoundaries.cuh:6-36): every population of the
// node becomes w rho (1 + c.u/cs2 + (c.u)^2/(2 cs4) - u.u/(2 cs2)) of the inward node's
// post-stream density and velocity: the second-order equilibrium whatever lin_ec (the functor
// writes it out rather than calling the solver's equilibrium).
template <int... qs>
LBM3_HD void bc_extrapolated3_all(float* g, float rho, float ux, float uy, float uz, std::integer_sequence<int, qs...>) {
    const float base = rho * (1.0f - 1.5f * (ux * ux + uy * uy + uz * uz));
    ((g[qs] = feq3<qs>(rho, ux, uy, uz, base, 3.0f)), ...);
}
LBM3_HD void bc_extrapolated3(float* g, float rho, float ux, float uy, float uz) {
    bc_extrapolated3_all(g, rho, ux, uy, uz, std::make_integer_sequence<int, Q>{});
}

// Component k % 3 of direction k / 3 of the velocity table read as a flat [27][3] array: the
// value cornerBoundaries.cuh:23 multiplies f_i by for u_z (k = 6 + i) under LBM_QUIRK_CORNER3D_UZ.
LBM3_HD constexpr int flat_c(int k) { return k % 3 == 0 ? CX[k / 3] : (k % 3 == 1 ? CY[k / 3] : CZ[k / 3]); }

// EdgeCornerBounceBack::apply (edgeCornerBounceBack.cuh:10-49): on a node with at least two
// domain faces (x, y, z at 0 or N-1, periodic or not), every direction with no negative and at
// least one positive component copies g_i into g_opp(i). The functor computes the inward
// normals only to count them; which populations it reflects does not depend on them. The 7
// source directions and their 7 opposites are disjoint, so the loop order does not matter.
template <int i>
LBM3_HD void ceb_dir(float* g) {
    if constexpr (CX[i] >= 0 && CY[i] >= 0 && CZ[i] >= 0 && (CX[i] + CY[i] + CZ[i]) > 0) g[OPP[i]] = g[i];
}
template <int... is>
LBM3_HD void ceb_all(float* g, std::integer_sequence<int, is...>) {
    (ceb_dir<is>(g), ...);
}
LBM3_HD void bc_corner_edge_bounce_back3(float* g, int x, int y, int z, const Box3& B) {
    const int faces = (x == 0) + (x == B.nx - 1) + (y == 0) + (y == B.ny - 1) + (z == 0) + (z == B.nz - 1);
    if (faces >= 2) ceb_all(g, std::make_integer_sequence<int, Q>{});
}

}  // namespace lbm3
