// lbm_node.hpp — node-local D2Q9 physics of the fused step, FP32 in registers.
//
// Everything here acts on one node's nine populations held in registers. The forms are
// chosen for gfx950 VALU cost (sparse moment transforms, delta updates, no divisions by
// constants); they are mathematically the reference operators cited on each function and
// are checked against the oracle's literal restatement.
//
// Functions are __host__ __device__ so the same step body also runs in the host emulation
// used by the CPU tests (bounds and parity checks before any GPU launch).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/lbm_hip.h"

#define LBM_HD __host__ __device__ __forceinline__

namespace lbm {

constexpr int Q = 9;
// D2Q9 velocity set in the reference's order 0,E,N,W,S,NE,NW,SW,SE (lbm_constants.cuh:19-31).
constexpr int CX[Q] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int CY[Q] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
constexpr int OPP[Q] = {0, 3, 4, 1, 2, 7, 8, 5, 6};

LBM_HD constexpr float wgt(int q) { return q == 0 ? 4.0f / 9.0f : (q < 5 ? 1.0f / 9.0f : 1.0f / 36.0f); }

// 1/x. The host divides (correctly rounded). The GPU takes the 1-ulp hardware reciprocal
// (v_rcp_f32) and one Newton step, r + r (1 - x r) with the residual in an FMA: within half
// an ulp but for rare ties, and unbiased; 3 instructions against ~10 for the IEEE division
// (8192^2 strips +2.7 % FP32 BGK, +3.7 % FP16 CM). The raw reciprocal alone is biased: over
// 1000 Taylor-Green steps it moved the validation error 2 % away from the FP64 run.
LBM_HD float recip(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const float r = __builtin_amdgcn_rcpf(x);
    return __builtin_fmaf(r, __builtin_fmaf(-x, r, 1.0f), r);
#else
    return 1.0f / x;
#endif
}

// ---- macroscopics: uncorrected_macroscopics_kernel<2> (macroscopics.cu:5-38) and the
// Guo half-force correction (macroscopics.cu:99-110), in the reference's summation order.
struct Macro {
    float rho, ux, uy;
};
LBM_HD Macro macroscopics(const float* g, float fx, float fy) {
    // The reference's sums over i = 0..8 of g_i, g_i c_ix, g_i c_iy, in that order; the
    // terms with c = 0 add an exact zero and are left out, those with c = -1 subtract.
    const float r = (((((((g[0] + g[1]) + g[2]) + g[3]) + g[4]) + g[5]) + g[6]) + g[7]) + g[8];
    const float jx = ((((g[1] - g[3]) + g[5]) - g[6]) - g[7]) + g[8];
    const float jy = ((((g[2] - g[4]) + g[5]) + g[6]) - g[7]) - g[8];
    // one reciprocal per node; the reference divides three times (macroscopics.cu:29-30,
    // 108-109)
    float inv = recip(r);
    Macro m;
    m.rho = r;
    m.ux = (jx + 0.5f * fx) * inv;
    m.uy = (jy + 0.5f * fy) * inv;
    return m;
}

// Raw second-moment magnitude |Pi| = sqrt(Pxx^2 + 2Pxy^2 + Pyy^2) (macroscopics.cu:24-32).
LBM_HD float pi_magnitude(const float* g) {
    float d = g[5] + g[6] + g[7] + g[8];
    float pxx = g[1] + g[3] + d;
    float pyy = g[2] + g[4] + d;
    float pxy = g[5] - g[6] + g[7] - g[8];
    return sqrtf(pxx * pxx + 2.0f * pxy * pxy + pyy * pyy);
}

// Second-order equilibrium (equilibrium.cu:5-39), FP32 polynomial.
LBM_HD float feq(int q, float rho, float ux, float uy, float usq15) {
    float cu = float(CX[q]) * ux + float(CY[q]) * uy;
    return wgt(q) * rho * (1.0f + 3.0f * cu + 4.5f * cu * cu - usq15);
}

// BGK<2>::apply with Guo forcing (BGK.cuh:13-51). Directions come in opposite pairs
// (q, opp q): the equilibrium and the Guo term split into a part even in c (shared by the
// pair) and a part odd in c (added to one, subtracted from the other), so each pair costs
// one evaluation.
LBM_HD void collide_bgk(float* g, const Macro& m, float fx, float fy, float omega, bool forced) {
    const float ux = m.ux, uy = m.uy, rho = m.rho;
    const float base = 1.0f - 1.5f * (ux * ux + uy * uy);
    const float om1 = 1.0f - omega;
    const float gf = 1.0f - 0.5f * omega;
    const float uf3 = 3.0f * (ux * fx + uy * fy);
    {  // rest population
        float fe = wgt(0) * rho * base;
        float v = om1 * g[0] + omega * fe;
        if (forced) v += wgt(0) * gf * (-uf3);
        g[0] = v;
    }
    constexpr int PQ[4] = {1, 2, 5, 6};  // pairs (1,3), (2,4), (5,7), (6,8)
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int q = PQ[k], o = OPP[q];
        const float cx = float(CX[q]), cy = float(CY[q]);
        const float cu = cx * ux + cy * uy;
        const float wr = wgt(q) * rho;
        const float sym = wr * (base + 4.5f * cu * cu);
        const float asym = 3.0f * wr * cu;
        float vq = om1 * g[q] + omega * (sym + asym);
        float vo = om1 * g[o] + omega * (sym - asym);
        if (forced) {
            const float cf = cx * fx + cy * fy;
            const float fs = wgt(q) * gf * (9.0f * cu * cf - uf3);
            const float fa = wgt(q) * gf * 3.0f * cf;
            vq += fs + fa;
            vo += fs - fa;
        }
        g[q] = vq;
        g[o] = vo;
    }
}

// MRT<2>::apply (MRT.cu:26-76) in delta form f* = f + M^-1 dm, with the Lallemand-Luo
// basis (lbm_constants.cuh:33-55) written out sparsely and m_eq in closed form.
// Forcing moments per MRT.cu:4-24; `literal_rows` keeps its rows 4/5 swap (quirk Q10).
// `forced` (uniform) = 0 skips the forcing moments, which are then all zero.
// CONS: S[0] = S[3] = S[5] = 0 (density and momentum conserved, as DEFAULT_MRT_S_MATRIX and
// every MRT scenario set them): those three moment changes are their forcing terms alone
// (-0 (m - m_eq) + F = F exactly), so their m and m_eq are not formed.
template <bool CONS>
LBM_HD void collide_mrt_t(float* g, const Macro& m, float fx, float fy, const float* S, bool literal_rows, bool forced) {
    const float rho = m.rho, ux = m.ux, uy = m.uy;
    float a = g[1] + g[2] + g[3] + g[4];
    float d = g[5] + g[6] + g[7] + g[8];
    float px = g[1] - g[3], dxs = g[5] - g[6] - g[7] + g[8];
    float py = g[2] - g[4], dys = g[5] + g[6] - g[7] - g[8];
    float mm[Q];
    mm[1] = -4.0f * g[0] - a + 2.0f * d;
    mm[2] = 4.0f * g[0] - 2.0f * a + d;
    mm[4] = -2.0f * px + dxs;
    mm[6] = -2.0f * py + dys;
    mm[7] = g[1] - g[2] + g[3] - g[4];
    mm[8] = g[5] - g[6] + g[7] - g[8];
    if (!CONS) {
        mm[0] = g[0] + a + d;
        mm[3] = px + dxs;
        mm[5] = py + dys;
    }
    float usq = ux * ux + uy * uy;
    float meq[Q] = {rho,
                    rho * (-2.0f + 3.0f * usq),
                    rho * (1.0f - 3.0f * usq),
                    rho * ux,
                    -rho * ux,
                    rho * uy,
                    -rho * uy,
                    rho * (ux * ux - uy * uy),
                    rho * ux * uy};
    float dm[Q];
#pragma unroll
    for (int k = 0; k < Q; k++) {
        if (CONS && (k == 0 || k == 3 || k == 5)) dm[k] = 0.0f;
        else dm[k] = -S[k] * (mm[k] - meq[k]);
    }
    if (forced) {
        float fu = fx * ux + fy * uy;
        float F[Q] = {0.0f, 6.0f * fu, -6.0f * fu, fx, -fx, fy, -fy, 2.0f * (fx * ux - fy * uy), fx * uy + fy * ux};
        if (literal_rows) {
            F[4] = fy;
            F[5] = -fx;
        }
#pragma unroll
        for (int k = 0; k < Q; k++) {
            if (CONS && (k == 0 || k == 3 || k == 5)) dm[k] = F[k];
            else dm[k] += (1.0f - 0.5f * S[k]) * F[k];
        }
    }
    // M^-1 rows (lbm_constants.cuh:45-55)
    const float r9 = 1.0f / 9.0f, r18 = 1.0f / 18.0f, r36 = 1.0f / 36.0f, r6 = 1.0f / 6.0f, r12 = 1.0f / 12.0f;
    float c0 = CONS ? 0.0f : dm[0] * r9;
    float b1 = c0 - dm[1] * r36 - dm[2] * r18;
    float b2 = c0 + dm[1] * r18 + dm[2] * r36;
    float X = (dm[3] - dm[4]) * r6;
    float Y = (dm[5] - dm[6]) * r6;
    float P = 0.25f * dm[7];
    float XD = dm[3] * r6 + dm[4] * r12;
    float YD = dm[5] * r6 + dm[6] * r12;
    float Rr = 0.25f * dm[8];
    g[0] += c0 - dm[1] * r9 + dm[2] * r9;
    g[1] += b1 + X + P;
    g[2] += b1 + Y - P;
    g[3] += b1 - X + P;
    g[4] += b1 - Y - P;
    g[5] += b2 + XD + YD + Rr;
    g[6] += b2 - XD + YD - Rr;
    g[7] += b2 - XD - YD + Rr;
    g[8] += b2 + XD - YD - Rr;
}
LBM_HD void collide_mrt(float* g, const Macro& m, float fx, float fy, const float* S, bool literal_rows, bool forced) {
    if (S[0] == 0.0f && S[3] == 0.0f && S[5] == 0.0f) collide_mrt_t<true>(g, m, fx, fy, S, literal_rows, forced);
    else collide_mrt_t<false>(g, m, fx, fy, S, literal_rows, forced);
}

// OptimalAdapter::compute_higher_order_relaxation (adapters.cuh:48-111), no printf.
LBM_HD float optimal_adapter_rate(float rho, float j_mag, float pi_mag, const float* avg) {
    float s0 = rho / avg[0], s1 = j_mag / avg[1], s2 = pi_mag / avg[2];
    float t = 0.0f;
    t += 0.0003f * s0;
    t += -0.00775f * s1;
    t += 0.00016f * s2;
    t += 0.0087f * 1.0f;
    t = t > 0.0f ? t : 0.005f;
    t = fminf(t, 1.5f);
    return 1.0f / (3.0f * t + 0.5f);
}

// CM<2,A>::apply (CM.cuh:27-138): central moments k about the corrected u, relaxed with
// the De Rosis forcing vector, mapped back with T^-1(u). Here both transforms are the
// factorised (x then y) binomial shifts of the 3x3 velocity grid, applied to the moment
// change dk, so f* = f + T^-1(u) dk. `avg` != nullptr selects OptimalAdapter for k6..k8.
LBM_HD void collide_cm(float* g, const Macro& m, float Fx, float Fy, const float* S, const float* avg) {
    const float ux = m.ux, uy = m.uy;
    // populations grouped by c_x, each ordered c_y = -1, 0, +1
    float fm[3] = {g[7], g[3], g[6]};  // c_x = -1 : c_y = -1, 0, +1
    float f0[3] = {g[4], g[0], g[2]};  // c_x =  0
    float fp[3] = {g[8], g[1], g[5]};  // c_x = +1
    float A0[3], A1[3], A2[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        float s = fm[j] + f0[j] + fp[j];
        float dd = fp[j] - fm[j];
        float t = fp[j] + fm[j];
        A0[j] = s;
        A1[j] = dd - ux * s;
        A2[j] = t - 2.0f * ux * dd + ux * ux * s;
    }
    float kap[3][3];  // kap[a][b] = kappa_ab
    const float* Aa[3] = {A0, A1, A2};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float am = Aa[a][0], a0 = Aa[a][1], ap = Aa[a][2];
        float s = am + a0 + ap;
        float dd = ap - am;
        kap[a][0] = s;
        kap[a][1] = dd - uy * s;
        kap[a][2] = (ap + am) - 2.0f * uy * dd + uy * uy * s;
    }
    const float rho = kap[0][0];
    float k[Q] = {rho, kap[1][0], kap[0][1], kap[2][0] + kap[0][2], kap[2][0] - kap[0][2],
                  kap[1][1], kap[2][1], kap[1][2], kap[2][2]};
    const float cs2 = 1.0f / 3.0f;
    float keq[Q] = {rho, 0.f, 0.f, 2.0f * rho * cs2, 0.f, 0.f, 0.f, 0.f, rho * cs2 * cs2};
    float F[Q] = {0.f, Fx, Fy, 0.f, 0.f, 0.f, Fy * cs2, Fx * cs2, 0.f};
    float hor = 0.0f;
    if (avg) {
        float j_mag = sqrtf(ux * ux + uy * uy) * rho;
        hor = optimal_adapter_rate(rho, j_mag, pi_magnitude(g), avg);
    }
    float dk[Q];
#pragma unroll
    for (int i = 0; i < Q; i++) {
        float s = (avg && i > 5) ? hor : S[i];
        dk[i] = -s * (k[i] - keq[i]) + (1.0f - 0.5f * s) * F[i];
    }
    // back to kappa_ab
    float dkap[3][3];
    dkap[0][0] = dk[0];
    dkap[1][0] = dk[1];
    dkap[0][1] = dk[2];
    dkap[2][0] = 0.5f * (dk[3] + dk[4]);
    dkap[0][2] = 0.5f * (dk[3] - dk[4]);
    dkap[1][1] = dk[5];
    dkap[2][1] = dk[6];
    dkap[1][2] = dk[7];
    dkap[2][2] = dk[8];
    // inverse y shift: central -> raw y-moments -> rows
    float B[3][3];  // B[a][j], j = c_y + 1
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float M0 = dkap[a][0], M1 = dkap[a][1], M2 = dkap[a][2];
        float R1 = M1 + uy * M0;
        float R2 = M2 + 2.0f * uy * M1 + uy * uy * M0;
        B[a][2] = 0.5f * (R2 + R1);
        B[a][0] = 0.5f * (R2 - R1);
        B[a][1] = M0 - R2;
    }
    // inverse x shift per row
    float dfm[3], df0[3], dfp[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        float M0 = B[0][j], M1 = B[1][j], M2 = B[2][j];
        float R1 = M1 + ux * M0;
        float R2 = M2 + 2.0f * ux * M1 + ux * ux * M0;
        dfp[j] = 0.5f * (R2 + R1);
        dfm[j] = 0.5f * (R2 - R1);
        df0[j] = M0 - R2;
    }
    g[7] += dfm[0];
    g[3] += dfm[1];
    g[6] += dfm[2];
    g[4] += df0[0];
    g[0] += df0[1];
    g[2] += df0[2];
    g[8] += dfp[0];
    g[1] += dfp[1];
    g[5] += dfp[2];
}

// ---- boundary functors on the node's post-stream populations g[9] -------------------

struct Geom {
    int nx, ny, px, py;
};

// BounceBack<2>::apply (bbDomainBoundary.cuh:15-51) with reflection on the non-periodic
// axes, sequential as written (so corner results match the reference's loop order).
// `moving` adds Ladd's wall-momentum term -6 w_i rho0 (c_i . u_w) for the appended
// MOVING_WALL_BOUNCE_BACK flag.
LBM_HD void bc_bounce_back(float* g, int x, int y, const Geom& G, bool moving, float wux, float wuy, float rho0) {
#pragma unroll
    for (int i = 1; i < Q; i++) {
        int xn = x + CX[i], yn = y + CY[i];
        bool bx = !G.px && (xn < 0 || xn >= G.nx);
        bool by = !G.py && (yn < 0 || yn >= G.ny);
        if (bx || by) {
            float v = g[i];
            if (moving) v = v - 6.0f * wgt(i) * rho0 * (float(CX[i]) * wux + float(CY[i]) * wuy);
            g[OPP[i]] = v;
        }
    }
}

// CylinderBoundary::apply (cylinderBoundary.cuh:22-34): full-way bounce-back, every
// post-stream population sent back along its opposite direction.
LBM_HD void bc_cylinder(float* g) {
    float s[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) s[q] = g[q];
#pragma unroll
    for (int q = 0; q < Q; q++) g[OPP[q]] = s[q];
}

// PressureOutlet::apply (pressureOutlet.cuh:7-42): f = f_eq(rho = 1, u_i), u_i the velocity
// of the interior node's post-stream populations gi (the node to the left).
LBM_HD void bc_pressure_outlet(float* g, const float* gi) {
    const float r = (((((((gi[0] + gi[1]) + gi[2]) + gi[3]) + gi[4]) + gi[5]) + gi[6]) + gi[7]) + gi[8];
    const float jx = ((((gi[1] - gi[3]) + gi[5]) - gi[6]) - gi[7]) + gi[8];
    const float jy = ((((gi[2] - gi[4]) + gi[5]) + gi[6]) - gi[7]) - gi[8];
    const float ux = jx / r, uy = jy / r;
    const float usq15 = 1.5f * (ux * ux + uy * uy);
#pragma unroll
    for (int q = 0; q < Q; q++) g[q] = feq(q, 1.0f, ux, uy, usq15);
}

// Regularisation of the non-equilibrium part from Pi^(1) (Latt et al. 2008) shared by the
// regularized functors: g_q = feq_q + w_q/(2cs^4) Q_q:Pi.
LBM_HD void regularize(float* g, const float* fe, float pxx, float pyy, float pxy) {
    const float cs2 = 1.0f / 3.0f;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        float cx = float(CX[q]), cy = float(CY[q]);
        float qxx = cx * cx - cs2, qyy = cy * cy - cs2, qxy = cx * cy;
        g[q] = fe[q] + 4.5f * wgt(q) * (qxx * pxx + qyy * pyy + 2.0f * qxy * pxy);
    }
}

// RegularizedInlet::apply_top (regularizedInlet.cuh:15-69), lid velocity (u_lid, 0).
LBM_HD void bc_regularized_inlet_top(float* g, float u_lid) {
    const float ux = u_lid;
    float r = g[0] + g[1] + g[3] + 2.0f * (g[2] + g[5] + g[6]);
    float fe[Q];
    float usq15 = 1.5f * ux * ux;
#pragma unroll
    for (int q = 0; q < Q; q++) fe[q] = feq(q, r, ux, 0.0f, usq15);
    g[4] = fe[4] + (g[2] - fe[2]);
    g[7] = fe[7] + (g[5] - fe[5]);
    g[8] = fe[8] + (g[6] - fe[6]);
    float d = g[5] + g[6] + g[7] + g[8];
    float pxx = g[1] + g[3] + d - (r / 3.0f + r * ux * ux);
    float pyy = g[2] + g[4] + d - r / 3.0f;
    float pxy = g[5] - g[6] + g[7] - g[8];
    regularize(g, fe, pxx, pyy, pxy);
}

// RegularizedBounceBack::apply (regularizedBounceBack.cuh:13-100): no-slip wall, the
// unknowns chosen by the node's position (x edges first).
LBM_HD void bc_regularized_bounce_back(float* g, int x, int y, const Geom& G) {
    bool unk[Q];
#pragma unroll
    for (int i = 0; i < Q; i++) {
        bool u = false;
        if (i > 0) {
            if (x == 0) u = CX[i] > 0;
            else if (x == G.nx - 1) u = CX[i] < 0;
            else if (y == 0) u = CY[i] > 0;
            else if (y == G.ny - 1) u = CY[i] < 0;
        }
        unk[i] = u;
    }
    float r = 0.0f;
#pragma unroll
    for (int i = 0; i < Q; i++)
        if (!unk[i]) r += (unk[OPP[i]] ? 2.0f : 1.0f) * g[i];
    float fe[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) fe[q] = wgt(q) * r;
#pragma unroll
    for (int q = 0; q < Q; q++)
        if (unk[q]) g[q] = fe[q] + (g[OPP[q]] - fe[OPP[q]]);
    float d = g[5] + g[6] + g[7] + g[8];
    float pxx = g[1] + g[3] + d - r / 3.0f;
    float pyy = g[2] + g[4] + d - r / 3.0f;
    float pxy = g[5] - g[6] + g[7] - g[8];
    regularize(g, fe, pxx, pyy, pxy);
}

// RegularizedCornerBounceBack::apply (regularizedBounceBack.cuh:107-222). `rho_diag` is the
// density of the diagonal interior neighbour's post-stream populations.
LBM_HD void bc_regularized_corner(float* g, int x, int y, const Geom& G, float rho_diag) {
    bool left = x == 0, right = x == G.nx - 1, bottom = y == 0, top = y == G.ny - 1;
    if (!((left || right) && (bottom || top))) return;
    const float r = rho_diag;
    float fe[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) fe[q] = wgt(q) * r;
#pragma unroll
    for (int i = 1; i < Q; i++) {
        bool u = (left && CX[i] > 0) || (right && CX[i] < 0) || (bottom && CY[i] > 0) || (top && CY[i] < 0);
        if (u) g[i] = fe[i] - (g[OPP[i]] - fe[OPP[i]]);
    }
    float n[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) n[q] = g[q] - fe[q];
    float d = n[5] + n[6] + n[7] + n[8];
    float pxx = n[1] + n[3] + d;
    float pyy = n[2] + n[4] + d;
    float pxy = n[5] - n[6] + n[7] - n[8];
    regularize(g, fe, pxx, pyy, pxy);
}

// ZouHe::apply_left (zouHeInflow.cuh:9-32) with per-node population indexing (Q8 fixed);
// `literal_rho` keeps the f2-for-f3 density typo (Q9).
LBM_HD void bc_zou_he_left(float* g, float u_in, bool literal_rho) {
    const float ux = u_in;
    float s = literal_rho ? g[2] : g[3];
    float r = (g[0] + g[2] + g[4] + 2.0f * (s + g[6] + g[7])) / (1.0f - ux);
    g[1] = g[3] + (2.0f / 3.0f) * r * ux;
    g[5] = g[7] - 0.5f * (g[2] - g[4]) + (1.0f / 6.0f) * r * ux;
    g[8] = g[6] + 0.5f * (g[2] - g[4]) + (1.0f / 6.0f) * r * ux;
}

// ZouHe::apply_top (zouHeInflow.cuh:35-68), lid velocity (u_lid, 0).
LBM_HD void bc_zou_he_top(float* g, float u_lid) {
    const float ux = u_lid;
    float r = g[0] + g[1] + g[3] + 2.0f * (g[2] + g[5] + g[6]);
    g[4] = g[2];
    float d13 = g[1] - g[3];
    g[7] = g[5] + 0.5f * d13 - 0.5f * r * ux;
    g[8] = g[6] - 0.5f * d13 + 0.5f * r * ux;
}

}  // namespace lbm
