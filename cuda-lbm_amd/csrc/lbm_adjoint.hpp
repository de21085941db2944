// lbm_adjoint.hpp — transposed Jacobians of the fused step's node update (no counterpart
// in the reference).
//
// Forward node update (per node, pre-collision gathered populations g, body force F):
//   rho = sum g,  u = (sum g c + F/2) / rho,
//   f*  = C(g; F)  (BGK or MRT with Guo forcing, lbm_node.hpp).
// Given lambda = dL/df* at the node, these return mu = (df*/dg)^T lambda and the node's
// contribution (df*/dF)^T lambda. With du/dg_p = (c_p - u)/rho and du/dF = I/(2 rho), every
// term reduces to a scalar A and a vector B per node:
//   mu_p = diag_p + A + B . (c_p - u) / rho.
#pragma once

#include "lbm_node.hpp"

namespace lbm {

struct AdjOut {
    float mu[Q];
    float gfx, gfy;  // (df*/dF)^T lambda
};

// d feq_q / du  = w rho (3c + 9 (c.u) c - 3u);  d Sguo_q / du = w (-3F + 9 (c.F) c);
// d Sguo_q / dF = w (3 (c - u) + 9 (c.u) c), where Sguo is Guo's source without (1 - s/2).

// BGK: f*_q = (1-w) g_q + w feq_q + (1 - w/2) Sguo_q.
LBM_HD AdjOut adjoint_bgk(const float* g, const float* lam, float fx, float fy, float omega) {
    float r = 0.f, jx = 0.f, jy = 0.f;
#pragma unroll
    for (int i = 0; i < Q; i++) {
        r += g[i];
        jx += g[i] * float(CX[i]);
        jy += g[i] * float(CY[i]);
    }
    const float inv = 1.0f / r;
    const float ux = (jx + 0.5f * fx) * inv, uy = (jy + 0.5f * fy) * inv;
    const float usq15 = 1.5f * (ux * ux + uy * uy);
    const float gf = 1.0f - 0.5f * omega;
    float A = 0.f, Bx = 0.f, By = 0.f, Gx = 0.f, Gy = 0.f;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const float cx = float(CX[q]), cy = float(CY[q]), w = wgt(q), l = lam[q];
        const float cu = cx * ux + cy * uy, cf = cx * fx + cy * fy;
        A += l * w * (1.0f + 3.0f * cu + 4.5f * cu * cu - usq15);
        // omega dfeq/du + (1 - omega/2) dSguo/du
        Bx += l * w * (omega * r * (3.0f * cx + 9.0f * cu * cx - 3.0f * ux) + gf * (-3.0f * fx + 9.0f * cf * cx));
        By += l * w * (omega * r * (3.0f * cy + 9.0f * cu * cy - 3.0f * uy) + gf * (-3.0f * fy + 9.0f * cf * cy));
        Gx += l * w * gf * (3.0f * (cx - ux) + 9.0f * cu * cx);
        Gy += l * w * gf * (3.0f * (cy - uy) + 9.0f * cu * cy);
    }
    AdjOut o;
    const float om1 = 1.0f - omega;
#pragma unroll
    for (int p = 0; p < Q; p++)
        o.mu[p] = om1 * lam[p] + omega * A + (Bx * (float(CX[p]) - ux) + By * (float(CY[p]) - uy)) * inv;
    o.gfx = 0.5f * Bx * inv + Gx;
    o.gfy = 0.5f * By * inv + Gy;
    return o;
}

// M^T v and M^-T v for the Lallemand-Luo basis (lbm_constants.cuh:33-55), written out.
LBM_HD void mt_apply(const float* v, float* o) {  // o = M^T v
    o[0] = v[0] - 4.f * v[1] + 4.f * v[2];
    o[1] = v[0] - v[1] - 2.f * v[2] + v[3] - 2.f * v[4] + v[7];
    o[2] = v[0] - v[1] - 2.f * v[2] + v[5] - 2.f * v[6] - v[7];
    o[3] = v[0] - v[1] - 2.f * v[2] - v[3] + 2.f * v[4] + v[7];
    o[4] = v[0] - v[1] - 2.f * v[2] - v[5] + 2.f * v[6] - v[7];
    o[5] = v[0] + 2.f * v[1] + v[2] + v[3] + v[4] + v[5] + v[6] + v[8];
    o[6] = v[0] + 2.f * v[1] + v[2] - v[3] - v[4] + v[5] + v[6] - v[8];
    o[7] = v[0] + 2.f * v[1] + v[2] - v[3] - v[4] - v[5] - v[6] + v[8];
    o[8] = v[0] + 2.f * v[1] + v[2] + v[3] + v[4] - v[5] - v[6] - v[8];
}
LBM_HD void minvt_apply(const float* l, float* o) {  // o = M^-T l (columns of M^-1)
    const float r9 = 1.0f / 9.0f, r18 = 1.0f / 18.0f, r36 = 1.0f / 36.0f, r6 = 1.0f / 6.0f, r12 = 1.0f / 12.0f;
    float s14 = l[1] + l[2] + l[3] + l[4], s58 = l[5] + l[6] + l[7] + l[8];
    o[0] = (l[0] + s14 + s58) * r9;
    o[1] = -l[0] * r9 - s14 * r36 + s58 * r18;
    o[2] = l[0] * r9 - s14 * r18 + s58 * r36;
    o[3] = (l[1] - l[3] + l[5] - l[6] - l[7] + l[8]) * r6;
    o[4] = (-l[1] + l[3]) * r6 + (l[5] - l[6] - l[7] + l[8]) * r12;
    o[5] = (l[2] - l[4] + l[5] + l[6] - l[7] - l[8]) * r6;
    o[6] = (-l[2] + l[4]) * r6 + (l[5] + l[6] - l[7] - l[8]) * r12;
    o[7] = (l[1] - l[2] + l[3] - l[4]) * 0.25f;
    o[8] = (l[5] - l[6] + l[7] - l[8]) * 0.25f;
}

// MRT: f* = g - A1 g + A1 feq(rho,u) + A2 Sguo(u,F), A1 = M^-1 S M, A2 = M^-1 (I - S/2) M.
LBM_HD AdjOut adjoint_mrt(const float* g, const float* lam, float fx, float fy, const float* S) {
    float r = 0.f, jx = 0.f, jy = 0.f;
#pragma unroll
    for (int i = 0; i < Q; i++) {
        r += g[i];
        jx += g[i] * float(CX[i]);
        jy += g[i] * float(CY[i]);
    }
    const float inv = 1.0f / r;
    const float ux = (jx + 0.5f * fx) * inv, uy = (jy + 0.5f * fy) * inv;
    const float usq15 = 1.5f * (ux * ux + uy * uy);
    float lm[Q], a[Q], b[Q], psi[Q], chi[Q];
    minvt_apply(lam, lm);
#pragma unroll
    for (int k = 0; k < Q; k++) {
        a[k] = S[k] * lm[k];
        b[k] = (1.0f - 0.5f * S[k]) * lm[k];
    }
    mt_apply(a, psi);
    mt_apply(b, chi);
    float A = 0.f, Bx = 0.f, By = 0.f, Gx = 0.f, Gy = 0.f;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const float cx = float(CX[q]), cy = float(CY[q]), w = wgt(q);
        const float cu = cx * ux + cy * uy, cf = cx * fx + cy * fy;
        A += psi[q] * w * (1.0f + 3.0f * cu + 4.5f * cu * cu - usq15);
        Bx += w * (psi[q] * r * (3.0f * cx + 9.0f * cu * cx - 3.0f * ux) + chi[q] * (-3.0f * fx + 9.0f * cf * cx));
        By += w * (psi[q] * r * (3.0f * cy + 9.0f * cu * cy - 3.0f * uy) + chi[q] * (-3.0f * fy + 9.0f * cf * cy));
        Gx += chi[q] * w * (3.0f * (cx - ux) + 9.0f * cu * cx);
        Gy += chi[q] * w * (3.0f * (cy - uy) + 9.0f * cu * cy);
    }
    AdjOut o;
#pragma unroll
    for (int p = 0; p < Q; p++)
        o.mu[p] = lam[p] - psi[p] + A + (Bx * (float(CX[p]) - ux) + By * (float(CY[p]) - uy)) * inv;
    o.gfx = 0.5f * Bx * inv + Gx;
    o.gfy = 0.5f * By * inv + Gy;
    return o;
}

// Loss seed at the final snapshot: u = (sum g c + F/2)/rho, given dL/du.
// mu_p = dL/du . (c_p - u)/rho, dL/dF = dL/du / (2 rho).
LBM_HD AdjOut adjoint_snapshot(const float* g, float fx, float fy, float dux, float duy) {
    float r = 0.f, jx = 0.f, jy = 0.f;
#pragma unroll
    for (int i = 0; i < Q; i++) {
        r += g[i];
        jx += g[i] * float(CX[i]);
        jy += g[i] * float(CY[i]);
    }
    const float inv = 1.0f / r;
    const float ux = (jx + 0.5f * fx) * inv, uy = (jy + 0.5f * fy) * inv;
    AdjOut o;
#pragma unroll
    for (int p = 0; p < Q; p++) o.mu[p] = (dux * (float(CX[p]) - ux) + duy * (float(CY[p]) - uy)) * inv;
    o.gfx = 0.5f * dux * inv;
    o.gfy = 0.5f * duy * inv;
    return o;
}

// Initial state f*_0 = feq(rho0, u0): (dfeq/du0)^T lambda.
LBM_HD void adjoint_init_u(const float* lam, float rho0, float ux, float uy, float& gx, float& gy) {
    float ax = 0.f, ay = 0.f;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const float cx = float(CX[q]), cy = float(CY[q]), cu = cx * ux + cy * uy;
        ax += lam[q] * wgt(q) * rho0 * (3.0f * cx + 9.0f * cu * cx - 3.0f * ux);
        ay += lam[q] * wgt(q) * rho0 * (3.0f * cy + 9.0f * cu * cy - 3.0f * uy);
    }
    gx = ax;
    gy = ay;
}

}  // namespace lbm
