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

#include <vector>

namespace lbm3 {

// Kernel arguments (passed by value).
struct Step3Params {
    const void* src;       // post-collision populations of step t-1, [27][plane]
    void* dst;             // receives step t; holds step t-2 (stale slots)
    const uint8_t* flags;  // per-node flag, or null (all FLUID); bit 7 (kPreLaunch) marks the
                           // nodes the pre-launch updates (see needs_pre_launch)
    const float* force;    // SoA [3][plane] force field, or null (uniform fx, fy, fz)
    int fbox[6];           // force is read inside [x0, x1) x [y0, y1) x [z0, z1) only
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
    int guo;               // flags that read a neighbour's populations present
    int corner_uz;         // LBM_QUIRK_CORNER3D_UZ
    int stale_rest;        // LBM_QUIRK_STALE_REST
    int nt;                // store with the non-temporal cache policy (buffers beyond the
                           // Infinity Cache)
    int region;            // 0 every node; 1 the nodes outside rbox; 2 the nodes of rbox only
                           // (the IBM split step)
    int rbox[6];           // [x0, x1) x [y0, y1) x [z0, z1)
    // The launch's node range: workgroups cover nbx x nby x nzl blocks from (ox, oy, oz), nodes
    // at x >= xe or y >= ye idle. Whole grid: 0, 0, 0, nx, ny, nz; region 2: the box.
    // Workgroup shape: 256 threads = wx wavefronts along x (64*wx consecutive nodes of a row)
    // times 4/wx rows; nbx, nby workgroups along x and y, nz along z.
    int wx;
    int nbx, nby;
    int ox, oy, oz, xe, ye, nzl;
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

// ZG_OutflowBoundary<3>::apply (zeroGradientOutflow.cuh:8-59): the first face in the order
// x = 0, x = NX-1, y = 0, y = NY-1, z = 0, z = NZ-1 gives the inward normal n (axis D, sign S);
// the populations entering the domain (c . n > 0) take the post-stream values of the node one
// step inward, the others keep the node's own; a node on no face is left as it is. One code
// path for every face: direction q is replaced when its component along the face's axis has
// the normal's sign (a per-direction conditional pull into a fixed element of g, so g stays in
// registers).
LBM3_HD int zg_face(const Step3Params& p, int x, int y, int z) {
    if (x == 0) return 0;
    if (x == p.nx - 1) return 1;
    if (y == 0) return 2;
    if (y == p.ny - 1) return 3;
    if (z == 0) return 4;
    if (z == p.nz - 1) return 5;
    return -1;
}
template <int STORE, bool BUF, int q>
LBM3_HD void zg_dir(const Step3Params& p, const lbm::PopBuf<typename Store3<STORE>::T, BUF>& src, float* g, int d, int s,
                    int xn, int yn, int zn) {
    const int c = d == 0 ? CX[q] : (d == 1 ? CY[q] : CZ[q]);
    if (c * s > 0) g[q] = pull_at<STORE, BUF, q>(p, src, xn, yn, zn);
}
template <int STORE, bool BUF, int... qs>
LBM3_HD void zg_all(const Step3Params& p, const lbm::PopBuf<typename Store3<STORE>::T, BUF>& src, float* g, int d, int s,
                    int xn, int yn, int zn, std::integer_sequence<int, qs...>) {
    (zg_dir<STORE, BUF, qs>(p, src, g, d, s, xn, yn, zn), ...);
}
template <int STORE, bool BUF>
LBM3_HD void apply_zg(const Step3Params& p, const lbm::PopBuf<typename Store3<STORE>::T, BUF>& src, float* g, int x, int y,
                      int z) {
    const int f = zg_face(p, x, y, z);
    if (f < 0) return;
    const int d = f >> 1, s = (f & 1) ? -1 : 1;
    zg_all<STORE, BUF>(p, src, g, d, s, x + (d == 0 ? s : 0), y + (d == 1 ? s : 0), z + (d == 2 ? s : 0),
                       std::make_integer_sequence<int, Q>{});
}

// Density and velocity sums of the post-stream populations of node (xn, yn, zn); uzq the z
// momentum read through the flat velocity table (ExtrapolatedCornerEdgeBoundary's quirk).
template <int STORE, bool BUF, int q>
LBM3_HD void mom_at_dir(const Step3Params& p, const lbm::PopBuf<typename Store3<STORE>::T, BUF>& src, int xn, int yn, int zn,
                        float& r, float& jx, float& jy, float& jz, float& jzq) {
    const float v = pull_at<STORE, BUF, q>(p, src, xn, yn, zn);
    r += v;
    if (CX[q] > 0) jx += v;
    if (CX[q] < 0) jx -= v;
    if (CY[q] > 0) jy += v;
    if (CY[q] < 0) jy -= v;
    if (CZ[q] > 0) jz += v;
    if (CZ[q] < 0) jz -= v;
    if constexpr (q + 6 < 3 * Q && flat_c(q + 6) != 0) jzq += float(flat_c(q + 6)) * v;
}
template <int STORE, bool BUF, int... qs>
LBM3_HD void mom_at(const Step3Params& p, const lbm::PopBuf<typename Store3<STORE>::T, BUF>& src, int xn, int yn, int zn,
                    float& r, float& jx, float& jy, float& jz, float& jzq, std::integer_sequence<int, qs...>) {
    (mom_at_dir<STORE, BUF, qs>(p, src, xn, yn, zn, r, jx, jy, jz, jzq), ...);
}

// RegularizedOutlet (x = NX-1): the x - 1 neighbour's density and velocity, then the functor.
template <int STORE, bool BUF>
LBM3_HD void apply_regularized_outlet(const Step3Params& p, const lbm::PopBuf<typename Store3<STORE>::T, BUF>& src,
                                      float* g, int x, int y, int z) {
    float r = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f, jzq = 0.0f;
    mom_at<STORE, BUF>(p, src, x - 1, y, z, r, jx, jy, jz, jzq, std::make_integer_sequence<int, Q>{});
    const float inv = 1.0f / r;
    bc_regularized_outlet3(g, r, jx * inv, jy * inv, jz * inv, p.lin_ec);
}

// The node one step inward on each axis where (x, y, z) is on a domain face (the reference's
// test is x == 0 / x == NX-1 whatever the periodicity).
LBM3_HD void inward_node(const Step3Params& p, int x, int y, int z, int& xi, int& yi, int& zi) {
    xi = x == 0 ? 1 : (x == p.nx - 1 ? p.nx - 2 : x);
    yi = y == 0 ? 1 : (y == p.ny - 1 ? p.ny - 2 : y);
    zi = z == 0 ? 1 : (z == p.nz - 1 ? p.nz - 2 : z);
}

// ExtrapolatedCornerEdgeBoundary: the inward node's density and velocity, then the functor.
template <int STORE, bool BUF>
LBM3_HD void apply_extrapolated(const Step3Params& p, const lbm::PopBuf<typename Store3<STORE>::T, BUF>& src, float* g,
                                int x, int y, int z) {
    int xi, yi, zi;
    inward_node(p, x, y, z, xi, yi, zi);
    float r = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f, jzq = 0.0f;
    mom_at<STORE, BUF>(p, src, xi, yi, zi, r, jx, jy, jz, jzq, std::make_integer_sequence<int, Q>{});
    const float inv = 1.0f / r;
    bc_extrapolated3(g, r, jx * inv, jy * inv, (p.corner_uz ? jzq : jz) * inv);
}

// A flag that reads a neighbour's post-stream populations pulls them by the streaming rule,
// which reads the neighbour's stale slots (sources outside a non-periodic domain) from the
// destination buffer; the neighbour's own update rewrites them. That is safe when both nodes are
// lanes of one wavefront (a row's 64-column group): the wavefront issues its stores after the
// loads they depend on. It is also safe when the neighbour has no stale slots (only the source
// buffer is read). Otherwise the node runs in a pre-launch before the main launch, which skips
// it (kPreLaunch in its flag). Neighbours read: Guo inlet x + 1, Guo and regularized outlets
// x - 1, ZG the node one step inward on its face; the corner-edge extrapolation's inward node
// has no stale slots (it is inside every non-periodic axis).
constexpr uint8_t kPreLaunch = 0x80;

LBM3_HD bool read_neighbour(const Step3Params& p, uint8_t fl, int x, int y, int z, int& xn, int& yn, int& zn) {
    xn = x;
    yn = y;
    zn = z;
    if (fl == LBM_GUO_VELOCITY_INLET) xn = x + 1;
    else if (fl == LBM_GUO_PRESSURE_OUTLET || fl == LBM_REGULARIZED_OUTLET) xn = x - 1;
    else if (fl == LBM_ZG_OUTFLOW) {
        const int f = zg_face(p, x, y, z);
        if (f < 0) return false;
        if (f < 2) xn = f == 0 ? 1 : p.nx - 2;
        else if (f < 4) yn = f == 2 ? 1 : p.ny - 2;
        else zn = f == 4 ? 1 : p.nz - 2;
    } else {
        return false;
    }
    return true;
}
LBM3_HD bool has_stale_slots(const Step3Params& p, int x, int y, int z) {
    return (!p.px && (x == 0 || x == p.nx - 1)) || (!p.py && (y == 0 || y == p.ny - 1)) ||
           (!p.pz && (z == 0 || z == p.nz - 1));
}
LBM3_HD bool needs_pre_launch(const Step3Params& p, uint8_t fl, int x, int y, int z) {
    int xn, yn, zn;
    if (!read_neighbour(p, fl, x, y, z, xn, yn, zn)) return false;
    if (!has_stale_slots(p, xn, yn, zn)) return false;
    return !(yn == y && zn == z && xn / 64 == x / 64);
}

// The pre-launch order (host). A node that reads a neighbour's stale slots must run before the
// neighbour's update: in a launch ahead of the main launch when the neighbour is in another
// wavefront (needs_pre_launch), and ahead of the neighbour's own pre-launch when the neighbour
// runs in one. levels[k] lists the nodes of level k + 1; the launches run from the last level to
// the first, then the main launch (level 0) skips every listed node. Returns null, or the reason
// no order exists (two nodes each reading the other's stale slots).
inline const char* pre_launch_levels(const Step3Params& p, const uint8_t* flags, std::vector<std::vector<int>>& levels) {
    levels.clear();
    const long long N = (long long)p.nx * p.ny * p.nz;
    std::vector<int> level(size_t(N), 0), reads(size_t(N), -1);
    for (long long n = 0; n < N; n++) {
        const int x = int(n % p.nx), y = int((n / p.nx) % p.ny), z = int(n / ((long long)p.nx * p.ny));
        int xn, yn, zn;
        const uint8_t fl = uint8_t(flags[n] & ~kPreLaunch);
        if (!read_neighbour(p, fl, x, y, z, xn, yn, zn) || !has_stale_slots(p, xn, yn, zn)) continue;
        reads[size_t(n)] = int((zn * (long long)p.ny + yn) * p.nx + xn);
        if (needs_pre_launch(p, fl, x, y, z)) level[size_t(n)] = 1;
    }
    // a reader of a pre-launched node runs one level earlier (chains are a few nodes long)
    for (int it = 0;; it++) {
        bool changed = false;
        for (long long n = 0; n < N; n++) {
            const int m = reads[size_t(n)];
            if (m < 0 || level[size_t(m)] == 0) continue;
            if (level[size_t(n)] <= level[size_t(m)]) {
                level[size_t(n)] = level[size_t(m)] + 1;
                changed = true;
            }
        }
        if (!changed) break;
        if (it > 16) return "boundary nodes read each other's stale slots in a cycle";
    }
    int top = 0;
    for (int v : level) top = v > top ? v : top;
    levels.assign(size_t(top), {});
    for (long long n = 0; n < N; n++)
        if (level[size_t(n)] > 0) levels[size_t(level[size_t(n)] - 1)].push_back(int(n));
    return nullptr;
}

// The post-stream populations of one node: pull, then the boundary functor of its flag.
// Requires 0 <= x < nx, 0 <= y < ny, 0 <= z < nz and nx*ny*nz < 2^31. GUO 1: the node may carry
// a flag that reads a neighbour (Guo inlet/outlet, ZG outflow); GUO 2: also a regularized or
// corner-edge extrapolation flag (a variant of its own: its 27 equilibria and Pi spill registers).
template <int STORE, bool BUF, int GUO>
LBM3_HD void post_stream3(const Step3Params& p, const lbm::PopBuf<typename Store3<STORE>::T, BUF>& src, int x, int y,
                          int z, uint32_t node, float* g) {
    const int nxy = p.nx * p.ny;
    int ox[3], oy[3], oz[3];
    bool vx[3], vy[3], vz[3];
    axis_pull(x, p.nx, p.px, 1, ox, vx);
    axis_pull(y, p.ny, p.py, p.nx, oy, vy);
    axis_pull(z, p.nz, p.pz, nxy, oz, vz);
    const bool lane_in = vx[0] && vx[2] && vy[0] && vy[2] && vz[0] && vz[2];
    if (wave_all(lane_in))
        pull_all<STORE, BUF, true>(g, p, src, node, ox, vx, oy, vy, oz, vz, std::make_integer_sequence<int, Q>{});
    else
        pull_all<STORE, BUF, false>(g, p, src, node, ox, vx, oy, vy, oz, vz, std::make_integer_sequence<int, Q>{});

    const uint8_t fl = p.flags ? uint8_t(p.flags[node] & ~kPreLaunch) : uint8_t(LBM_FLUID);
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
        else if (fl == LBM_ZG_OUTFLOW) apply_zg<STORE, BUF>(p, src, g, x, y, z);
        else if (GUO == 2 && (fl == LBM_REGULARIZED_BOUNCE_BACK || fl == LBM_REGULARIZED_INLET_LEFT)) {
            const Box3 B{p.nx, p.ny, p.nz, p.px, p.py, p.pz};
            const float ui = fl == LBM_REGULARIZED_INLET_LEFT ? p.u_inlet : 0.0f;
            bc_regularized3(g, x, y, z, B, ui, 0.0f, 0.0f, p.lin_ec);
        } else if (GUO == 2 && fl == LBM_REGULARIZED_OUTLET) {
            apply_regularized_outlet<STORE, BUF>(p, src, g, x, y, z);
        } else if (GUO == 2 && fl == LBM_EXTRAPOLATED_CORNER_EDGE) {
            apply_extrapolated<STORE, BUF>(p, src, g, x, y, z);
        }
    }
}

// The force on a node: the array inside the box fbox = [x0, x1) x [y0, y1) x [z0, z1) when one
// is set, the uniform force elsewhere.
LBM3_HD void node_force(const Step3Params& p, int x, int y, int z, uint32_t node, float& fx, float& fy, float& fz) {
    fx = p.fx;
    fy = p.fy;
    fz = p.fz;
    if (p.force && x >= p.fbox[0] && x < p.fbox[1] && y >= p.fbox[2] && y < p.fbox[3] && z >= p.fbox[4] &&
        z < p.fbox[5]) {
        fx = p.force[node];
        fy = p.force[p.plane + node];
        fz = p.force[2 * p.plane + node];
    }
}

// One node of one time step: post-stream populations, macroscopics, Guo-forced BGK, MRT or CM,
// store.
template <int COLL, int STORE, bool BUF, int GUO = 0>
LBM3_HD void step3_node(const Step3Params& p, int x, int y, int z) {
    using St = Store3<STORE>;
    using T = typename St::T;
    const uint32_t node = uint32_t(z * (p.nx * p.ny) + y * p.nx + x);
    const lbm::PopBuf<T, BUF> src(p.src), dst(p.dst);
    float g[Q];
    if constexpr (GUO >= 2) {
        post_stream3<STORE, BUF, GUO>(p, src, x, y, z, node, g);
    } else {
        // post_stream3 for GUO <= 1, written out: through the call the compiler kept g in
        // scratch (16 B plain, 112 B Guo variant) and the Guo kernel ran 1.7x slower; written
        // out, both keep g in registers. The regularized variant is the other way round (40 B
        // through the call, 160 B written out).
        const int nxy = p.nx * p.ny;
        int ox[3], oy[3], oz[3];
        bool vx[3], vy[3], vz[3];
        axis_pull(x, p.nx, p.px, 1, ox, vx);
        axis_pull(y, p.ny, p.py, p.nx, oy, vy);
        axis_pull(z, p.nz, p.pz, nxy, oz, vz);
        const bool lane_in = vx[0] && vx[2] && vy[0] && vy[2] && vz[0] && vz[2];
        if (wave_all(lane_in))
            pull_all<STORE, BUF, true>(g, p, src, node, ox, vx, oy, vy, oz, vz, std::make_integer_sequence<int, Q>{});
        else
            pull_all<STORE, BUF, false>(g, p, src, node, ox, vx, oy, vy, oz, vz, std::make_integer_sequence<int, Q>{});
        const uint8_t fl = p.flags ? uint8_t(p.flags[node] & ~kPreLaunch) : uint8_t(LBM_FLUID);
        if (fl == LBM_BOUNCE_BACK) {
            const Box3 B{p.nx, p.ny, p.nz, p.px, p.py, p.pz};
            bc_bounce_back3(g, x, y, z, B);
        } else if (fl == LBM_CORNER_EDGE_BOUNCE_BACK) {
            const Box3 B{p.nx, p.ny, p.nz, p.px, p.py, p.pz};
            bc_corner_edge_bounce_back3(g, x, y, z, B);
        }
        if constexpr (GUO == 1) {
            if (fl == LBM_GUO_VELOCITY_INLET) apply_guo<STORE, BUF, true>(p, src, g, x, y, z);
            else if (fl == LBM_GUO_PRESSURE_OUTLET) apply_guo<STORE, BUF, false>(p, src, g, x, y, z);
            else if (fl == LBM_ZG_OUTFLOW) apply_zg<STORE, BUF>(p, src, g, x, y, z);
        }
    }
    float fx, fy, fz;
    node_force(p, x, y, z, node, fx, fy, fz);
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
