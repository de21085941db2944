// lbm3d_ibm.hpp — immersed-boundary multi-direct forcing for D3Q27 (IBMManager<3>::multi_direct,
// IBMManager.cuh:222-252) on the band of the markers' 4x4x4 stencils, as per-thread bodies shared
// by the kernels of lbm3d.hip and the host emulator.
//
// The reference runs its five per-iteration kernels over the whole grid (memset, interpolate,
// Lagrangian force, atomic spread, correct, accumulate: ~100 B per node per iteration in 3-D)
// for a force that is non-zero on the band alone. Here one pass per step is
//   gather  (one thread per band node): post-stream populations -> rho, uncorrected u, the
//           external force;
//   ITER x  interpolate (one wavefront per marker, its 64 stencil entries summed by a fixed
//           butterfly), then spread as a gather (one thread per band node) over a (marker,
//           weight) list in marker order — deterministic, the reference's scatter order without
//           atomics;
// with the band's fields in three small SoA arrays that stay in L2. The last spread writes the
// total force of each band node into the step's force array, whose other nodes keep the
// external force.
#pragma once

#include "lbm3d_step.hpp"

namespace lbm3 {

struct Ibm3Params {
    Step3Params sp;         // src, dst (stale slots), flags, grid and uniform force of the step
    int n_markers, n_band;
    int clamp;              // LBM_QUIRK_IBM_CLAMP
    int first, last;        // first / last iteration of the pass
    const int* band_node;   // [n_band] global node of each band entry (ascending)
    const int* mk_ptr;      // [n_markers + 1] stencil entries of marker l: mk_ptr[l] .. mk_ptr[l+1]
    const int* mk_band;     // band index of each stencil entry (x outer, y, z inner)
    const float* mk_w;      // phi(dx) phi(dy) phi(dz)
    const int* bn_ptr;      // [n_band + 1] spread entries of band node b, in marker order
    const int* bn_mk;       // marker of each spread entry
    const float* bn_w;      // phi times the marker area (times 1 under LBM_QUIRK_IBM_NO_DS)
    const float* ext;       // external force field SoA [3][plane], or null (sp.fx, sp.fy, sp.fz)
    float* brho;            // band rho [n_band]
    float* bu;              // band velocity SoA [3][n_band]: uncorrected, then corrected per iteration
    float* bf;              // band force SoA [3][n_band]: external + the iterations' spread forces
    float* lag;             // Lagrangian force of the current iteration SoA [3][n_markers]
    float* lag_sum;         // summed over the pass's iterations SoA [3][n_markers]
    float* force_total;     // the step's force array SoA [3][plane]
};

// uncorrected_macroscopics_kernel<3> (macroscopics.cu:40-96) on band node b, and the force reset
// to the external force (reset_forces, macroscopics.cuh:13-48).
template <int STORE, bool BUF, int GUO>
LBM3_HD void ibm3_gather(const Ibm3Params& p, int b) {
    const Step3Params& s = p.sp;
    const int node = p.band_node[b];
    const int nxy = s.nx * s.ny;
    const int z = node / nxy, r0 = node - z * nxy, y = r0 / s.nx, x = r0 - y * s.nx;
    const lbm::PopBuf<typename Store3<STORE>::T, BUF> src(s.src);
    float g[Q];
    post_stream3<STORE, BUF, GUO>(s, src, x, y, z, uint32_t(node), g);
    float r = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        r += g[q];
        jx += g[q] * float(CX[q]);
        jy += g[q] * float(CY[q]);
        jz += g[q] * float(CZ[q]);
    }
    const float inv = 1.0f / r;
    const int nb = p.n_band;
    p.brho[b] = r;
    p.bu[b] = jx * inv;
    p.bu[nb + b] = jy * inv;
    p.bu[2 * nb + b] = jz * inv;
    if (p.ext) {
        for (int c = 0; c < 3; c++) p.bf[c * nb + b] = p.ext[c * s.plane + node];
    } else {
        p.bf[b] = s.fx;
        p.bf[nb + b] = s.fy;
        p.bf[2 * nb + b] = s.fz;
    }
}

// interpolate_velocities_kernel<3> (IBM_impl.cu:53-118) and compute_lagrangian_kernel
// (IBM_impl.cuh:7-26) for marker l: rho_L, u_L over the stencil, F_L = 2 rho_L (0 - u_L).
// A wavefront per marker: lane `lane` of `nlanes` sums the stencil entries lane, lane + nlanes,
// ... into s[4] (rho, u_x, u_y, u_z); the lanes' sums are then combined by the butterfly
// ibm3_combine (the same fixed order on the GPU and in the host emulation) and lane 0
// finishes the marker.
LBM3_HD void ibm3_interp_partial(const Ibm3Params& p, int l, int lane, int nlanes, float* s) {
    const int nb = p.n_band;
    s[0] = s[1] = s[2] = s[3] = 0.0f;
    for (int e = p.mk_ptr[l] + lane; e < p.mk_ptr[l + 1]; e += nlanes) {
        const float k = p.mk_w[e];
        const int b = p.mk_band[e];
        s[0] += k * p.brho[b];
        s[1] += k * p.bu[b];
        s[2] += k * p.bu[nb + b];
        s[3] += k * p.bu[2 * nb + b];
    }
}
// One butterfly level: every lane adds the sums of lane ^ offset (offsets 32, 16, ..., 1).
LBM3_HD void ibm3_combine(float* s, const float* partner) {
    for (int c = 0; c < 4; c++) s[c] += partner[c];
}
LBM3_HD void ibm3_interp_finish(const Ibm3Params& p, int l, const float* s) {
    const int n = p.n_markers;
    for (int d = 0; d < 3; d++) {
        float fl = 2.0f * s[0] * (0.0f - s[1 + d]);
        if (p.clamp) fl = fl > 1e-8f ? fl : 0.0f;
        p.lag[d * n + l] = fl;
        p.lag_sum[d * n + l] = p.first ? fl : p.lag_sum[d * n + l] + fl;
    }
}

// spread_forces_kernel<3> (IBM_impl.cu:156-196) as a gather over band node b's entries, then
// correct_velocities_kernel and accumulate_forces_kernel (IBM_impl.cuh:28-68); the last iteration
// writes the node's total force into the step's force array.
LBM3_HD void ibm3_spread(const Ibm3Params& p, int b) {
    const int nb = p.n_band, n = p.n_markers;
    float f[3] = {0.0f, 0.0f, 0.0f};
    for (int e = p.bn_ptr[b]; e < p.bn_ptr[b + 1]; e++) {
        const float w = p.bn_w[e];
        const int l = p.bn_mk[e];
        for (int d = 0; d < 3; d++) f[d] += w * p.lag[d * n + l];
    }
    const float r2 = 2.0f * p.brho[b];
    const long long node = p.band_node[b];
    for (int d = 0; d < 3; d++) {
        float c = p.bu[d * nb + b] + f[d] / r2;
        if (p.clamp) c = c > 1e-8f ? c : 0.0f;
        p.bu[d * nb + b] = c;
        const float t = p.bf[d * nb + b] + f[d];
        p.bf[d * nb + b] = t;
        if (p.last) p.force_total[d * p.sp.plane + node] = t;
    }
}

}  // namespace lbm3
