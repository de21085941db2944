// lbm3d_step.hpp — per-thread body of the fused D3Q27 pull-scheme step.
//
// One thread owns one node; a wavefront owns 64 consecutive nodes of one x-row, so each of
// the 27 pulls is one wavefront-wide load of 64 consecutive elements of a population plane,
// shifted by the direction's upstream offset, and each of the 27 stores is 64 aligned
// consecutive elements. One read and one write of every population per step replace the
// reference's stream, boundaries, macroscopics, correction, equilibrium and collide kernels
// (main.cu:98-114) and its f_eq buffer.
//
// Streaming rule (streaming.cu:35-65, pulled): population q at node x comes from x - c_q,
// wrapped on periodic axes; when x - c_q lies outside a non-periodic domain the push kernel
// never writes the slot, so it keeps what the destination buffer already holds (the state of
// two steps before), and the pull reads exactly that slot. LBM_QUIRK_STALE_REST does the same
// for the rest population (streaming.cu:43 starts at i = 1).
//
// The body is __host__ __device__: lbm3d.hip launches it on the GPU, and the CPU tests run it
// over the same grid in a host emulation (bounds and parity checks before any GPU launch).
#pragma once

#include "lbm3d_node.hpp"
#include "lbm_step.hpp"

namespace lbm3 {

// Kernel arguments (passed by value).
struct Step3Params {
    const void* src;       // post-collision populations of step t-1, [27][plane]
    void* dst;             // receives step t; holds step t-2 (stale slots)
    const uint8_t* flags;  // per-node flag (FLUID / BOUNCE_BACK / CORNER_EDGE_BOUNCE_BACK / Guo), or null (all FLUID)
    const float* force;    // SoA [3][plane] force field, or null (uniform fx, fy, fz)
    float* out_rho;        // macroscopic outputs, rho [plane] and u SoA [3][plane], or null
    float* out_u;
    int nx, ny, nz;
    int px, py, pz;
    long long plane;       // elements between population planes (N rounded up to 64)
    float omega, rho0, fx, fy, fz;
    float lin_ec;          // linear equilibrium coefficient of edge/corner directions (3 or 1)
    Mrt3Rates mrt;         // MRT<3> rates S[4..8] and the literal flag
    float u_inlet;         // GUO_VELOCITY_INLET and REGULARIZED_INLET_LEFT velocity
    int guo_own;           // LBM_QUIRK_GUO3D_OWN_MOMENTS
    int guo;               // Guo inlet/outlet or x-face ZG outflow flags present
    int guo_ring_outlet;   // (NX-1) % 64 == 0: the outlet's ring nodes run in a launch of their
                           // own before the main launch, which skips them
    int stale_rest;        // LBM_QUIRK_STALE_REST
    int nt;                // store with the non-temporal cache policy (buffers beyond the
                           // Infinity Cache)
    // Workgroup shape: 256 threads = wx wavefronts along x (64*wx consecutive nodes of a row)
    // times 4/wx rows; nbx, nby workgroups along x and y, nz along z.
    int wx;
    int nbx, nby;
};

// Workgroup of 256 threads: wx = min(4, ceil(nx/64)) wavefronts along x, so a workgroup holds
// whole rows up to nx = 256 and the partial cache lines of the x-shifted pulls are shared
// between wavefronts of one workgroup.
LBM3_HD int wave_cols(int nx) { return nx > 192 ? 4 : (nx > 64 ? 2 : 1); }

// (x, y, z) of thread t of logical workgroup L.
LBM3_HD void thread_node(const Step3Params& p, long long L, int t, int& x, int& y, int& z) {
    const int bx = int(L % p.nbx);
    const long long r = L / p.nbx;
    const int by = int(r % p.nby);
    z = int(r / p.nby);
    const int cols = 64 * p.wx;
    x = bx * cols + t % cols;
    y = by * (4 / p.wx) + t / cols;
}

// Population storage: FP32, or FP16/BF16 of the shifted deviation f - w rho0.
template <int STORE>
struct Store3 {
    using T = uint16_t;
    LBM3_HD static float dec(T raw, int q, float rho0) {
        const float d = STORE == LBM_F16_SHIFTED ? lbm::half_bits_to_float(raw) : lbm::bf16_bits_to_float(raw);
        return d + w3(q) * rho0;
    }
    LBM3_HD static T enc(float v, int q, float rho0) {
        const float d = v - w3(q) * rho0;
        return STORE == LBM_F16_SHIFTED ? lbm::float_to_half_bits(d) : lbm::float_to_bf16_bits(d);
    }
};
template <>
struct Store3<LBM_F32> {
    using T = float;
    LBM3_HD static float dec(T raw, int, float) { return raw; }
    LBM3_HD static T enc(float v, int, float) { return v; }
};

// Offsets (in elements) of the upstream node along one axis, for c = -1, 0, +1 (index c + 1);
// ok[c + 1] is false when the upstream node lies outside a non-periodic domain.
LBM3_HD void axis_pull(int x, int n, int periodic, int stride, int* off, bool* ok) {
    off[1] = 0;
    ok[1] = true;
    // c = +1: upstream x - 1
    ok[2] = x > 0 || periodic;
    off[2] = x > 0 ? -stride : (periodic ? (n - 1) * stride : 0);
    // c = -1: upstream x + 1
    ok[0] = x < n - 1 || periodic;
    off[0] = x < n - 1 ? stride : (periodic ? -(n - 1) * stride : 0);
}

// True when `v` holds on every active lane of the wavefront (the host emulation runs one lane).
LBM3_HD bool wave_all(bool v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_ballot_w64(!v) == 0;
#else
    return v;
#endif
}

// Pull of population q. FAST: every upstream node of the wavefront is inside the domain or
// wrapped, so every lane loads from the source buffer. Otherwise each lane loads from the
// source at an in-domain index, and lanes whose upstream node lies outside a non-periodic
// domain then read their own stale slot of the destination buffer through a plain pointer
// (a per-lane choice between two buffer resources would make the compiler loop over the
// distinct resources of the wavefront).
template <int STORE, bool BUF, bool FAST, int q>
LBM3_HD void pull_dir(float* g, const Step3Params& p, const lbm::PopBuf<typename Store3<STORE>::T, BUF>& src,
                      uint32_t node, const int* ox, const bool* vx, const int* oy, const bool* vy, const int* oz,
                      const bool* vz) {
    using T = typename Store3<STORE>::T;
    constexpr int a = CX[q] + 1, b = CY[q] + 1, c = CZ[q] + 1;
    const long long off = (long long)q * p.plane;
    const uint32_t up = uint32_t(int(node) + ox[a] + oy[b] + oz[c]);
    const T* stale = static_cast<const T*>(p.dst) + off + node;
    T raw;
    if (q == 0 && p.stale_rest) {
        raw = *stale;
    } else if (FAST) {
        raw = src.ld(up, off);
    } else {
        const bool ok = vx[a] && vy[b] && vz[c];
        raw = src.ld(ok ? up : node, off);
        if (!ok) raw = *stale;
    }
    g[q] = Store3<STORE>::dec(raw, q, p.rho0);
}
template <int STORE, bool BUF, bool FAST, int... qs>
LBM3_HD void pull_all(float* g, const Step3Params& p, const lbm::PopBuf<typename Store3<STORE>::T, BUF>& src,
                      uint32_t node, const int* ox, const bool* vx, const int* oy, const bool* vy, const int* oz,
                      const bool* vz, std::integer_sequence<int, qs...>) {
    (pull_dir<STORE, BUF, FAST, qs>(g, p, src, node, ox, vx, oy, vy, oz, vz), ...);
}

// Post-stream value of population q at node (x, y, z) by the streaming rule, for a functor that
// reads a neighbour. Its stale slot is read from the destination buffer, which is why such
// functors run in a launch before the one that rewrites those slots.
template <int STORE, bool BUF, int q>
LBM3_HD float pull_at(const Step3Params& p, const lbm::PopBuf<typename Store3<STORE>::T, BUF>& src, int x, int y, int z) {
    using T = typename Store3<STORE>::T;
    const int nxy = p.nx * p.ny;
    const uint32_t node = uint32_t(z * nxy + y * p.nx + x);
    int ox[3], oy[3], oz[3];
    bool vx[3], vy[3], vz[3];
    axis_pull(x, p.nx, p.px, 1, ox, vx);
    axis_pull(y, p.ny, p.py, p.nx, oy, vy);
    axis_pull(z, p.nz, p.pz, nxy, oz, vz);
    constexpr int a = CX[q] + 1, b = CY[q] + 1, c = CZ[q] + 1;
    const long long off = (long long)q * p.plane;
    const bool ok = vx[a] && vy[b] && vz[c] && !(q == 0 && p.stale_rest);
    T raw = src.ld(ok ? uint32_t(int(node) + ox[a] + oy[b] + oz[c]) : node, off);
    if (!ok) raw = static_cast<const T*>(p.dst)[off + node];
    return Store3<STORE>::dec(raw, q, p.rho0);
}

// The neighbour's unknown-direction values and the density and velocity used for f_eq_f.
template <int STORE, bool BUF, bool INLET, int q>
LBM3_HD void guo_neigh_dir(const Step3Params& p, const lbm::PopBuf<typename Store3<STORE>::T, BUF>& src, int xn, int y,
                           int z, float* gu, int& k, float& r, float& jx, float& jy, float& jz) {
    const float v = pull_at<STORE, BUF, q>(p, src, xn, y, z);
    if (INLET ? CX[q] > 0 : CX[q] < 0) gu[k++] = v;
    r += v;
    if (CX[q] > 0) jx += v;
    if (CX[q] < 0) jx -= v;
    if (CY[q] > 0) jy += v;
    if (CY[q] < 0) jy -= v;
    if (CZ[q] > 0) jz += v;
    if (CZ[q] < 0) jz -= v;
}
template <int STORE, bool BUF, bool INLET, int... qs>
LBM3_HD void guo_neigh_all(const Step3Params& p, const lbm::PopBuf<typename Store3<STORE>::T, BUF>& src, int xn, int y,
                           int z, float* gu, float& r, float& jx, float& jy, float& jz, std::integer_sequence<int, qs...>) {
    int k = 0;
    (guo_neigh_dir<STORE, BUF, INLET, qs>(p, src, xn, y, z, gu, k, r, jx, jy, jz), ...);
}
template <int STORE, bool BUF, bool INLET>
LBM3_HD void apply_guo(const Step3Params& p, const lbm::PopBuf<typename Store3<STORE>::T, BUF>& src, float* g, int x, int y,
                       int z) {
    float gu[9];
    float r = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
    guo_neigh_all<STORE, BUF, INLET>(p, src, INLET ? x + 1 : x - 1, y, z, gu, r, jx, jy, jz,
                                     std::make_integer_sequence<int, Q>{});
    if (p.guo_own) {
        r = jx = jy = jz = 0.0f;
        mom_all(g, r, jx, jy, jz, std::make_integer_sequence<int, Q>{});
    }
    GuoBc B;
    B.inlet = INLET;
    B.u_in = p.u_inlet;
    B.lin_ec = p.lin_ec;
    B.rf = r;
    B.uxf = jx / r;
    B.uyf = jy / r;
    B.uzf = jz / r;
    bc_guo3<INLET>(g, gu, B);
}

// ZG_OutflowBoundary<3>::apply (zeroGradientOutflow.cuh:8-59) on an x face: the populations
// entering the domain (c . n > 0, n the inward normal: c_x > 0 at x = 0, c_x < 0 at x = NX-1)
// take the interior neighbour's post-stream values; the others keep the node's own. The
// neighbour is pulled as a Guo node pulls it, with the same wavefront ordering.
template <int STORE, bool BUF, bool LEFT, int q>
LBM3_HD void zg_dir(const Step3Params& p, const lbm::PopBuf<typename Store3<STORE>::T, BUF>& src, float* g, int xn, int y,
                    int z) {
    if constexpr (LEFT ? CX[q] > 0 : CX[q] < 0) g[q] = pull_at<STORE, BUF, q>(p, src, xn, y, z);
}
template <int STORE, bool BUF, bool LEFT, int... qs>
LBM3_HD void zg_all(const Step3Params& p, const lbm::PopBuf<typename Store3<STORE>::T, BUF>& src, float* g, int xn, int y,
                    int z, std::integer_sequence<int, qs...>) {
    (zg_dir<STORE, BUF, LEFT, qs>(p, src, g, xn, y, z), ...);
}
template <int STORE, bool BUF, bool LEFT>
LBM3_HD void apply_zg(const Step3Params& p, const lbm::PopBuf<typename Store3<STORE>::T, BUF>& src, float* g, int x, int y,
                      int z) {
    zg_all<STORE, BUF, LEFT>(p, src, g, LEFT ? x + 1 : x - 1, y, z, std::make_integer_sequence<int, Q>{});
}

// A Guo node pulls its x +- 1 neighbour; on the ring of the plane (y or z at a non-periodic
// edge) that pull reads the neighbour's own stale slots in the destination buffer, which the
// neighbour's update rewrites. When both nodes are lanes of one wavefront that is safe: the
// wavefront issues its stores only after the loads they depend on, and both lanes store through
// the same instructions. The inlet (x = 0, neighbour x = 1) always shares a wavefront with its
// neighbour; the outlet (x = NX-1) does unless (NX-1) % 64 == 0, and then its ring nodes run in a
// launch of their own before the main launch.
LBM3_HD bool guo_ring(const Step3Params& p, int x, int y, int z) {
    if (!p.guo_ring_outlet || x != p.nx - 1) return false;
    return (!p.py && (y == 0 || y == p.ny - 1)) || (!p.pz && (z == 0 || z == p.nz - 1));
}

// One node of one time step: pull, boundary functor, macroscopics, Guo-forced BGK, MRT or CM,
// store. Requires 0 <= x < nx, 0 <= y < ny, 0 <= z < nz and nx*ny*nz < 2^31. GUO 1: the node may
// carry a flag that reads its x +- 1 neighbour (Guo inlet/outlet, ZG outflow on an x face); GUO 2:
// also a regularized flag (a variant of its own: its 27 equilibria and Pi spill registers).
template <int COLL, int STORE, bool BUF, int GUO = 0>
LBM3_HD void step3_node(const Step3Params& p, int x, int y, int z) {
    using St = Store3<STORE>;
    using T = typename St::T;
    const int nxy = p.nx * p.ny;
    const uint32_t node = uint32_t(z * nxy + y * p.nx + x);
    int ox[3], oy[3], oz[3];
    bool vx[3], vy[3], vz[3];
    axis_pull(x, p.nx, p.px, 1, ox, vx);
    axis_pull(y, p.ny, p.py, p.nx, oy, vy);
    axis_pull(z, p.nz, p.pz, nxy, oz, vz);
    const lbm::PopBuf<T, BUF> src(p.src), dst(p.dst);
    float g[Q];
    const bool lane_in = vx[0] && vx[2] && vy[0] && vy[2] && vz[0] && vz[2];
    if (wave_all(lane_in))
        pull_all<STORE, BUF, true>(g, p, src, node, ox, vx, oy, vy, oz, vz, std::make_integer_sequence<int, Q>{});
    else
        pull_all<STORE, BUF, false>(g, p, src, node, ox, vx, oy, vy, oz, vz, std::make_integer_sequence<int, Q>{});

    const uint8_t fl = p.flags ? p.flags[node] : uint8_t(LBM_FLUID);
    if (fl == LBM_BOUNCE_BACK) {
        const Box3 B{p.nx, p.ny, p.nz, p.px, p.py, p.pz};
        bc_bounce_back3(g, x, y, z, B);
    } else if (fl == LBM_CORNER_EDGE_BOUNCE_BACK) {
        const Box3 B{p.nx, p.ny, p.nz, p.px, p.py, p.pz};
        bc_corner_edge_bounce_back3(g, x, y, z, B);
    }
    if constexpr (GUO != 0) {
        if (fl == LBM_GUO_VELOCITY_INLET) apply_guo<STORE, BUF, true>(p, src, g, x, y, z);
        else if (fl == LBM_GUO_PRESSURE_OUTLET) apply_guo<STORE, BUF, false>(p, src, g, x, y, z);
        else if (fl == LBM_ZG_OUTFLOW) {
            if (x == 0) apply_zg<STORE, BUF, true>(p, src, g, x, y, z);
            else apply_zg<STORE, BUF, false>(p, src, g, x, y, z);
        } else if (GUO == 2 && (fl == LBM_REGULARIZED_BOUNCE_BACK || fl == LBM_REGULARIZED_INLET_LEFT)) {
            const Box3 B{p.nx, p.ny, p.nz, p.px, p.py, p.pz};
            const float ui = fl == LBM_REGULARIZED_INLET_LEFT ? p.u_inlet : 0.0f;
            bc_regularized3(g, x, y, z, B, ui, 0.0f, 0.0f, p.lin_ec);
        }
    }
    float fx = p.fx, fy = p.fy, fz = p.fz;
    if (p.force) {
        fx = p.force[node];
        fy = p.force[p.plane + node];
        fz = p.force[2 * p.plane + node];
    }
    const Macro3 m = macroscopics3(g, fx, fy, fz);
    if (p.out_rho) {
        p.out_rho[node] = m.rho;
        p.out_u[node] = m.ux;
        p.out_u[p.plane + node] = m.uy;
        p.out_u[2 * p.plane + node] = m.uz;
    }
    if (COLL == LBM_MRT) collide_mrt3(g, m, fx, fy, fz, p.mrt);
    else if (COLL == LBM_CM) collide_cm3(g, m, fx, fy, fz, p.omega);
    else collide_bgk3(g, m, fx, fy, fz, p.omega, p.lin_ec);
    if (p.nt) {
#pragma unroll
        for (int q = 0; q < Q; q++) dst.st_nt(node, (long long)q * p.plane, St::enc(g[q], q, p.rho0));
    } else {
#pragma unroll
        for (int q = 0; q < Q; q++) dst.st(node, (long long)q * p.plane, St::enc(g[q], q, p.rho0));
    }
}

}  // namespace lbm3
