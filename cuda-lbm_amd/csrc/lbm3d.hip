// lbm3d.hip — D3Q27 kernels and the C ABI of include/lbm3d_hip.h.
//
// Replaces LBM<3> (src/core/lbm.cuh:33-382) for BGK<3>, MRT<3> and CM<3>: its per-stage kernels
// and f_eq buffer become one fused pull-scheme kernel per step (lbm3d_step.hpp), launched as
// 256-thread workgroups holding whole x-rows up to 256 nodes. The step is HBM-bound: 27
// population reads and 27 writes per node update (216 B in FP32, 108 B in 16-bit storage)
// against ~330-370 VALU (BGK, MRT and CM without 27x27 matrices).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/lbm3d_hip.h"
#include "lbm3d_ibm.hpp"
#include "lbm_ibm_tables.hpp"

using namespace lbm3;

namespace {

thread_local std::string g_create_err3;

size_t elem_bytes(int storage) { return storage == LBM_F32 ? 4 : 2; }

// ---- kernels --------------------------------------------------------------------------

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2): physical
// workgroup b runs logical workgroup (b % 8) * (nb / 8) + b / 8, so consecutive logical
// workgroups (neighbouring row segments, which share cache lines at their ends) run on one
// XCD. The last nb % 8 workgroups keep their index. Eight wavefronts per SIMD (at most 64
// VGPRs, no scratch) measured 1-2% faster than the compiler's own 70-VGPR allocation.
template <int COLL, int STORE, bool BUF, int GUO>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void step3_kernel(Step3Params p) {
    const long long nb = (long long)p.nbx * p.nby * p.nzl;
    const long long b = blockIdx.x;
    const long long per = nb / 8;
    const long long L = b < per * 8 ? (b % 8) * per + b / 8 : b;
    int x, y, z;
    thread_node(p, L, int(threadIdx.x), x, y, z);
    x += p.ox;
    y += p.oy;
    z += p.oz;
    if (x >= p.xe || y >= p.ye) return;
    if (p.region == 1 && x >= p.rbox[0] && x < p.rbox[1] && y >= p.rbox[2] && y < p.rbox[3] && z >= p.rbox[4] &&
        z < p.rbox[5])
        return;
    if (GUO && p.flags && (p.flags[((long long)z * p.ny + y) * p.nx + x] & kPreLaunch)) return;
    step3_node<COLL, STORE, BUF, GUO>(p, x, y, z);
}

// The nodes marked kPreLaunch (needs_pre_launch: they read a neighbour's stale slots from
// another wavefront), listed in `nodes`. Runs before step3_kernel, which rewrites those slots
// and skips these nodes.
template <int COLL, int STORE, bool BUF, int GUO>
__global__ __launch_bounds__(256) void step3_pre_kernel(Step3Params p, const int* nodes, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int node = nodes[i];
    const int x = node % p.nx, y = (node / p.nx) % p.ny, z = node / (p.nx * p.ny);
    step3_node<COLL, STORE, BUF, GUO>(p, x, y, z);
}

// init_kernel<3> + init_node (init.cuh:10-43): both buffers = f_eq(rho, u), u AoS [N][3].
template <int STORE>
__global__ void init3_kernel(void* a, void* b, long long plane, long long N, float rho0, float lin_ec,
                             const float* rho, const float* u) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    using T = typename Store3<STORE>::T;
    float g[Q];
    for (int q = 0; q < Q; q++) g[q] = 0.0f;
    // f_eq is the BGK update at omega = 1 of any state with these moments and no force
    Macro3 m{rho[n], u[3 * n], u[3 * n + 1], u[3 * n + 2]};
    collide_bgk3(g, m, 0.0f, 0.0f, 0.0f, 1.0f, lin_ec);
    for (int q = 0; q < Q; q++) {
        const T v = Store3<STORE>::enc(g[q], q, rho0);
        static_cast<T*>(a)[q * plane + n] = v;
        static_cast<T*>(b)[q * plane + n] = v;
    }
}

// storage -> FP32 SoA [27][N]
template <int STORE>
__global__ void pops_to_f32_kernel(const void* s, long long plane, long long N, float rho0, float* out) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    using T = typename Store3<STORE>::T;
    for (int q = 0; q < Q; q++) out[q * N + n] = Store3<STORE>::dec(static_cast<const T*>(s)[q * plane + n], q, rho0);
}

// FP32 SoA [27][N] -> storage, both buffers
template <int STORE>
__global__ void pops_from_f32_kernel(void* a, void* b, long long plane, long long N, float rho0, const float* in) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    using T = typename Store3<STORE>::T;
    for (int q = 0; q < Q; q++) {
        const T v = Store3<STORE>::enc(in[q * N + n], q, rho0);
        static_cast<T*>(a)[q * plane + n] = v;
        static_cast<T*>(b)[q * plane + n] = v;
    }
}

// AoS [N][3] <-> SoA [3][plane]
__global__ void aos3_to_soa_kernel(const float* in, long long N, long long plane, float* out) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    for (int c = 0; c < 3; c++) out[c * plane + n] = in[3 * n + c];
}
__global__ void soa_to_aos3_kernel(const float* in, long long N, long long plane, float* out) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    for (int c = 0; c < 3; c++) out[3 * n + c] = in[c * plane + n];
}

// IBM pass kernels (lbm3d_ibm.hpp): one thread per band node or marker.
template <int STORE, bool BUF, int GUO>
__global__ __launch_bounds__(256) void ibm3_gather_kernel(Ibm3Params p) {
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b < p.n_band) ibm3_gather<STORE, BUF, GUO>(p, b);
}
// Four markers per workgroup, one per wavefront.
__global__ __launch_bounds__(256) void ibm3_interpolate_kernel(Ibm3Params p) {
    const int l = blockIdx.x * 4 + int(threadIdx.x >> 6), lane = int(threadIdx.x & 63);
    if (l >= p.n_markers) return;  // uniform per wavefront
    float s[4], t[4];
    ibm3_interp_partial(p, l, lane, 64, s);
    for (int off = 32; off > 0; off >>= 1) {
        for (int c = 0; c < 4; c++) t[c] = __shfl_xor(s[c], off, 64);
        ibm3_combine(s, t);
    }
    if (lane == 0) ibm3_interp_finish(p, l, s);
}
__global__ __launch_bounds__(256) void ibm3_spread_kernel(Ibm3Params p) {
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b < p.n_band) ibm3_spread(p, b);
}

// The step's force array outside the band: the external field (SoA) or the uniform force.
__global__ void force3_fill_kernel(float* out, const float* ext, long long N, long long plane, float fx, float fy, float fz) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const float f[3] = {fx, fy, fz};
    for (int c = 0; c < 3; c++) out[c * plane + n] = ext ? ext[c * plane + n] : f[c];
}

// Deterministic FP64 reduction: block b sums nodes b*256 + t + k*256*nb in a fixed order and
// a fixed LDS tree; the host adds the per-block partials in block order.
constexpr int kRedBlocks = 1024;
__global__ __launch_bounds__(256) void reduce3_kernel(int kind, const float* rho, const float* u, long long plane,
                                                      long long N, double* partials) {
    __shared__ double s[3][256];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (long long n = (long long)blockIdx.x * 256 + threadIdx.x; n < N; n += (long long)gridDim.x * 256) {
        const double r = rho[n];
        const double ux = u[n], uy = u[plane + n], uz = u[2 * plane + n];
        if (kind == LBM3D_REDUCE_MASS) {
            a0 += r;
        } else if (kind == LBM3D_REDUCE_MOMENTUM) {
            a0 += r * ux;
            a1 += r * uy;
            a2 += r * uz;
        } else {
            a0 += 0.5 * r * (ux * ux + uy * uy + uz * uz);
        }
    }
    s[0][threadIdx.x] = a0;
    s[1][threadIdx.x] = a1;
    s[2][threadIdx.x] = a2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w)
            for (int c = 0; c < 3; c++) s[c][threadIdx.x] += s[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; c++) partials[3 * blockIdx.x + c] = s[c][0];
}

int grid1(long long n) { return int((n + 255) / 256); }

}  // namespace

// ---- context ----------------------------------------------------------------------------

struct lbm3d_ctx {
    lbm3d_config cfg{};
    int device = 0;
    hipStream_t own_stream = nullptr, stream = nullptr;
    // IBM split steps: the step outside the band's box on aux_stream, beside the IBM pass and
    // the box's step on `stream` (lbm3d_ctx::one_step)
    hipStream_t aux_stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::vector<uint8_t> h_flags;      // host copy of the flags (empty: all FLUID)
    long long N = 0, plane = 0;
    void* buf[2] = {nullptr, nullptr};
    int cur = 0;                       // buffer holding the latest post-collision state
    uint8_t* d_flags = nullptr;        // null while every node is FLUID
    bool has_guo = false;              // flags that read a neighbour (or regularized flags) present
    bool has_reg = false;              // regularized or extrapolation flags (the GUO = 2 kernel variant)
    int* d_pre = nullptr;              // nodes of the pre-launches, in launch order (pre_launch_levels)
    int n_pre = 0;
    std::vector<int> pre_level_end;    // end of each pre-launch's nodes in d_pre
    float* d_force = nullptr;          // SoA [3][plane] or null
    struct Ibm3 {                      // immersed boundary (lbm3d_ibm_set_markers)
        bool active = false;
        bool force_valid = false;      // d_total holds the external force outside the band
        int n_markers = 0, n_band = 0, iters = 3;
        int box[6] = {0, 0, 0, 0, 0, 0};  // the band's bounding box [x0, x1) x [y0, y1) x [z0, z1)
        bool split_ok = false;         // the box may be stepped apart from the rest (ibm_split_ok)
        int *band_node = nullptr, *mk_ptr = nullptr, *mk_band = nullptr, *bn_ptr = nullptr, *bn_mk = nullptr;
        float *mk_w = nullptr, *bn_w = nullptr;
        float *brho = nullptr, *bu = nullptr, *bf = nullptr, *lag = nullptr, *lag_sum = nullptr;
        float* total = nullptr;        // the step's force array SoA [3][plane]
        std::vector<float> h_ds;
    } ibm;
    float* d_rho = nullptr;            // [plane]
    float* d_u = nullptr;              // SoA [3][plane]
    double* d_partials = nullptr;      // [kRedBlocks][3]
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    bool initialized = false, macros_valid = false;
    int64_t timestep = 0;
    std::string err;

    ~lbm3d_ctx() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (aux_stream) (void)hipStreamSynchronize(aux_stream);
        ibm_free();
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (aux_stream) (void)hipStreamDestroy(aux_stream);
        for (void* p : {buf[0], buf[1], (void*)d_flags, (void*)d_force, (void*)d_rho, (void*)d_u, (void*)d_partials, tmp,
                        (void*)d_pre})
            if (p) (void)hipFree(p);
        if (own_stream) (void)hipStreamDestroy(own_stream);
    }
    int fail(int code, const char* fmt, ...) {
        char b[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(b, sizeof b, fmt, ap);
        va_end(ap);
        err = b;
        return code;
    }
    int hip_fail(hipError_t e, const char* what) { return fail(LBM_E_HIP, "%s: %s", what, hipGetErrorString(e)); }
    int activate() {
        hipError_t e = hipSetDevice(device);
        return e == hipSuccess ? LBM_OK : hip_fail(e, "hipSetDevice");
    }
    int dev_alloc(void** p, size_t bytes, const char* what) {
        hipError_t e = hipMalloc(p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            *p = nullptr;
            return fail(LBM_E_NOMEM, "hipMalloc(%zu) for %s failed: %s", bytes, what, hipGetErrorString(e));
        }
        return LBM_OK;
    }
    int scratch(size_t bytes) {
        if (tmp_bytes >= bytes) return LBM_OK;
        if (tmp) {
            hipError_t e = hipStreamSynchronize(stream);
            if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
            (void)hipFree(tmp);
            tmp = nullptr;
            tmp_bytes = 0;
        }
        int rc = dev_alloc(&tmp, bytes, "staging");
        if (!rc) tmp_bytes = bytes;
        return rc;
    }
    // host (on_device = 0) or device source -> device `dst`, stream-ordered
    int upload(void* dst, const void* src, size_t bytes, int on_device) {
        hipError_t e = hipMemcpyAsync(dst, src, bytes, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync");
        if (!on_device) {
            e = hipStreamSynchronize(stream);
            if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
        }
        return LBM_OK;
    }
    int download(void* dst, const void* src, size_t bytes, int on_device) {
        hipError_t e = hipMemcpyAsync(dst, src, bytes, on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, stream);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync");
        if (!on_device) {
            e = hipStreamSynchronize(stream);
            if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
        }
        return LBM_OK;
    }
    float lin_ec() const { return (cfg.quirks & LBM_QUIRK_EQ3D_LINEAR) ? 1.0f : 3.0f; }
    void ibm_free() {
        for (void* q : {(void*)ibm.band_node, (void*)ibm.mk_ptr, (void*)ibm.mk_band, (void*)ibm.bn_ptr, (void*)ibm.bn_mk,
                        (void*)ibm.mk_w, (void*)ibm.bn_w, (void*)ibm.brho, (void*)ibm.bu, (void*)ibm.bf, (void*)ibm.lag,
                        (void*)ibm.lag_sum, (void*)ibm.total})
            if (q) (void)hipFree(q);
        ibm = Ibm3();
    }
    int ibm_pass(const Step3Params& sp, bool fits);
    void ibm_update_split();
    int one_step(bool write_macros);
};

// Whether an IBM step may step the band's box apart from the rest of the grid, concurrently
// with the IBM pass: the two launches then write disjoint nodes and read only the step's
// source, which holds when (1) no band node sits on a non-periodic face (whose gather would
// read stale slots the outside launch rewrites) and (2) the box grown by two nodes holds only
// FLUID and bounce-back nodes (no functor that reads a neighbour's populations, no
// pre-launched node, on either side of the box's faces).
void lbm3d_ctx::ibm_update_split() {
    ibm.split_ok = false;
    if (!ibm.active) return;
    const int n[3] = {cfg.nx, cfg.ny, cfg.nz};
    const int per[3] = {cfg.periodic_x, cfg.periodic_y, cfg.periodic_z};
    int lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        if (!per[a] && (ibm.box[2 * a] < 1 || ibm.box[2 * a + 1] > n[a] - 1)) return;
        lo[a] = std::max(0, ibm.box[2 * a] - 2);
        hi[a] = std::min(n[a], ibm.box[2 * a + 1] + 2);
    }
    if (!h_flags.empty())
        for (int z = lo[2]; z < hi[2]; z++)
            for (int y = lo[1]; y < hi[1]; y++)
                for (int x = lo[0]; x < hi[0]; x++) {
                    const uint8_t f = h_flags[(size_t(z) * size_t(cfg.ny) + size_t(y)) * size_t(cfg.nx) + size_t(x)];
                    if (f != LBM_FLUID && f != LBM_BOUNCE_BACK && f != LBM_CORNER_EDGE_BOUNCE_BACK) return;
                }
    // periodic axes: a box grown past an end wraps; check the wrapped slabs too
    for (int a = 0; a < 3; a++) {
        if (!per[a] || h_flags.empty()) continue;
        if (ibm.box[2 * a] - 2 >= 0 && ibm.box[2 * a + 1] + 2 <= n[a]) continue;
        return;  // not worth the special case: the split needs a box clear of the wrap
    }
    const char* v = std::getenv("LBM_IBM3_SPLIT");
    ibm.split_ok = !(v && v[0] == '0');
}

// The IBM pass of one step, ahead of the step launch on the same stream: it reads the step's
// source buffer and the stale slots of its destination before the step rewrites them.
int lbm3d_ctx::ibm_pass(const Step3Params& sp, bool fits) {
    if (!ibm.force_valid) {
        hipLaunchKernelGGL(force3_fill_kernel, dim3(grid1(N)), dim3(256), 0, stream, ibm.total, d_force, N, plane,
                           cfg.force[0], cfg.force[1], cfg.force[2]);
        ibm.force_valid = true;
    }
    Ibm3Params q{};
    q.sp = sp;
    q.n_markers = ibm.n_markers;
    q.n_band = ibm.n_band;
    q.clamp = (cfg.quirks & LBM_QUIRK_IBM_CLAMP) ? 1 : 0;
    q.band_node = ibm.band_node;
    q.mk_ptr = ibm.mk_ptr;
    q.mk_band = ibm.mk_band;
    q.mk_w = ibm.mk_w;
    q.bn_ptr = ibm.bn_ptr;
    q.bn_mk = ibm.bn_mk;
    q.bn_w = ibm.bn_w;
    q.ext = d_force;
    q.brho = ibm.brho;
    q.bu = ibm.bu;
    q.bf = ibm.bf;
    q.lag = ibm.lag;
    q.lag_sum = ibm.lag_sum;
    q.force_total = ibm.total;
    const dim3 gb(grid1(ibm.n_band)), gm(unsigned((ibm.n_markers + 3) / 4)), blk(256);
    const int G = has_reg ? 2 : (has_guo ? 1 : 0);
#define LBM3_GATHER(S, B)                                                                                    \
    if (G == 2) hipLaunchKernelGGL((ibm3_gather_kernel<S, B, 2>), gb, blk, 0, stream, q);                   \
    else if (G == 1) hipLaunchKernelGGL((ibm3_gather_kernel<S, B, 1>), gb, blk, 0, stream, q);              \
    else hipLaunchKernelGGL((ibm3_gather_kernel<S, B, 0>), gb, blk, 0, stream, q);
#define LBM3_GATHER_B(S)   \
    if (fits) {            \
        LBM3_GATHER(S, true)  \
    } else {               \
        LBM3_GATHER(S, false) \
    }
    switch (cfg.storage) {
        case LBM_F32: LBM3_GATHER_B(LBM_F32) break;
        case LBM_F16_SHIFTED: LBM3_GATHER_B(LBM_F16_SHIFTED) break;
        default: LBM3_GATHER_B(LBM_BF16_SHIFTED) break;
    }
#undef LBM3_GATHER_B
#undef LBM3_GATHER
    for (int it = 0; it < ibm.iters; it++) {
        q.first = it == 0;
        q.last = it == ibm.iters - 1;
        hipLaunchKernelGGL(ibm3_interpolate_kernel, gm, blk, 0, stream, q);
        hipLaunchKernelGGL(ibm3_spread_kernel, gb, blk, 0, stream, q);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? LBM_OK : hip_fail(e, "IBM launch");
}

int lbm3d_ctx::one_step(bool write_macros) {
    Step3Params p{};
    p.src = buf[cur];
    p.dst = buf[1 - cur];
    p.flags = d_flags;
    p.force = d_force;
    p.fbox[1] = cfg.nx;
    p.fbox[3] = cfg.ny;
    p.fbox[5] = cfg.nz;
    p.out_rho = write_macros ? d_rho : nullptr;
    p.out_u = write_macros ? d_u : nullptr;
    p.nx = cfg.nx;
    p.ny = cfg.ny;
    p.nz = cfg.nz;
    p.px = cfg.periodic_x != 0;
    p.py = cfg.periodic_y != 0;
    p.pz = cfg.periodic_z != 0;
    p.plane = plane;
    p.omega = cfg.omega;
    p.rho0 = cfg.rho0;
    p.fx = cfg.force[0];
    p.fy = cfg.force[1];
    p.fz = cfg.force[2];
    p.lin_ec = lin_ec();
    for (int k = 0; k < 5; k++) p.mrt.s[k] = cfg.S[4 + k];
    p.mrt.literal = (cfg.quirks & LBM_QUIRK_MRT3_LITERAL) ? 1 : 0;
    p.u_inlet = cfg.u_inlet;
    p.guo_own = (cfg.quirks & LBM_QUIRK_GUO3D_OWN_MOMENTS) ? 1 : 0;
    p.guo = has_guo ? 1 : 0;
    p.corner_uz = (cfg.quirks & LBM_QUIRK_CORNER3D_UZ) ? 1 : 0;
    p.stale_rest = (cfg.quirks & LBM_QUIRK_STALE_REST) ? 1 : 0;
    // Non-temporal stores keep the written populations from evicting the pulls' shared lines
    // once the two buffers no longer stay in the 256 MB Infinity Cache (+2% FP32, +4% FP16 at
    // 256^3); on resident grids they push every write to HBM (-20% FP32 at 64^3 and 96^3).
    // The switch sits at kNtBufferBytes, three quarters of the cache (the 2-D kernels' cache
    // threshold too), leaving room for the flags and the other data a step touches; grids
    // between it and the full cache (about 97^3-105^3 FP32, 122^3-133^3 FP16) were not timed
    // separately.
    constexpr double kNtBufferBytes = 192.0e6;
    p.nt = 2.0 * Q * double(plane) * double(elem_bytes(cfg.storage)) > kNtBufferBytes ? 1 : 0;
    p.wx = wave_cols(cfg.nx);
    p.nbx = (cfg.nx + 64 * p.wx - 1) / (64 * p.wx);
    p.nby = (cfg.ny + 4 / p.wx - 1) / (4 / p.wx);
    p.xe = cfg.nx;
    p.ye = cfg.ny;
    p.nzl = cfg.nz;
    const long long nblk = (long long)p.nbx * p.nby * cfg.nz;
    if (nblk > 0x7fffffffLL) return fail(LBM_E_INVALID, "grid needs too many workgroups");
    const bool fits = cfg.storage == LBM_F32 ? lbm::buf_fits<float>(Q, plane) : lbm::buf_fits<uint16_t>(Q, plane);
    // IBM split step: the nodes outside the band's box on aux_stream (they read neither the IBM
    // force nor anything the pass writes), beside the pass and then the box's nodes on `stream`
    const bool split = ibm.active && ibm.split_ok;
    hipStream_t outside_stream = stream;
    if (split) {
        if (!aux_stream) {
            hipError_t e = hipStreamCreateWithFlags(&aux_stream, hipStreamNonBlocking);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&ev_join, hipEventDisableTiming);
            if (e != hipSuccess) return hip_fail(e, "IBM stream set-up");
        }
        if (!ibm.force_valid) {
            hipLaunchKernelGGL(force3_fill_kernel, dim3(grid1(N)), dim3(256), 0, stream, ibm.total, d_force, N, plane,
                               cfg.force[0], cfg.force[1], cfg.force[2]);
            ibm.force_valid = true;
        }
        hipError_t e = hipEventRecord(ev_fork, stream);
        if (e != hipSuccess) return hip_fail(e, "IBM fork");
        // the pass is submitted first, so that its small launches start before the outside
        // step's workgroups fill the device
        int rc = ibm_pass(p, fits);
        if (rc) return rc;
        e = hipStreamWaitEvent(aux_stream, ev_fork, 0);
        if (e != hipSuccess) return hip_fail(e, "IBM fork");
        outside_stream = aux_stream;
    }
    if (ibm.active) {
        p.force = ibm.total;
        if (!d_force)  // the uniform force holds outside the band's box
            for (int k = 0; k < 6; k++) p.fbox[k] = ibm.box[k];
    }
    if (ibm.active && !split) {
        int rc = ibm_pass(p, fits);
        if (rc) return rc;
    }
    if (split) {
        p.region = 1;
        for (int k = 0; k < 6; k++) p.rbox[k] = ibm.box[k];
    }
    const dim3 grid{unsigned(nblk)}, block{256};
#define LBM3_PRE(C, S, B, G)                                                                          \
    for (size_t k = 0, beg = 0; k < pre_level_end.size(); beg = size_t(pre_level_end[k]), k++) {      \
        const int cnt = pre_level_end[k] - int(beg);                                                  \
        if (cnt > 0)                                                                                  \
            hipLaunchKernelGGL((step3_pre_kernel<C, S, B, G>), dim3(unsigned((cnt + 255) / 256)), block, 0, \
                               launch_stream, p, d_pre + beg, cnt);                                   \
    }
#define LBM3_LAUNCH_G(C, S, G)                                                                        \
    if (fits) {                                                                                       \
        LBM3_PRE(C, S, true, G)                                                                       \
        hipLaunchKernelGGL((step3_kernel<C, S, true, G>), grid_l, block, 0, launch_stream, p);                 \
    } else {                                                                                          \
        LBM3_PRE(C, S, false, G)                                                                      \
        hipLaunchKernelGGL((step3_kernel<C, S, false, G>), grid_l, block, 0, launch_stream, p);                \
    }
#define LBM3_LAUNCH(C, S)                                                                       \
    if (has_reg) {                                                                              \
        LBM3_LAUNCH_G(C, S, 2)                                                                  \
    } else if (has_guo) {                                                                       \
        LBM3_LAUNCH_G(C, S, 1)                                                                  \
    } else if (fits) {                                                                          \
        hipLaunchKernelGGL((step3_kernel<C, S, true, 0>), grid_l, block, 0, launch_stream, p);       \
    } else {                                                                                    \
        hipLaunchKernelGGL((step3_kernel<C, S, false, 0>), grid_l, block, 0, launch_stream, p);      \
    }
#define LBM3_LAUNCH_S(C)                                                                        \
    switch (cfg.storage) {                                                                      \
        case LBM_F32: LBM3_LAUNCH(C, LBM_F32) break;                                            \
        case LBM_F16_SHIFTED: LBM3_LAUNCH(C, LBM_F16_SHIFTED) break;                            \
        default: LBM3_LAUNCH(C, LBM_BF16_SHIFTED) break;                                        \
    }
    auto launch_all = [&](hipStream_t launch_stream, dim3 grid_l) {
        if (cfg.collision == LBM_MRT) {
            LBM3_LAUNCH_S(LBM_MRT)
        } else if (cfg.collision == LBM_CM) {
            LBM3_LAUNCH_S(LBM_CM)
        } else {
            LBM3_LAUNCH_S(LBM_BGK)
        }
    };
    launch_all(outside_stream, grid);
    if (split) {
        hipError_t e = hipEventRecord(ev_join, aux_stream);
        if (e != hipSuccess) return hip_fail(e, "IBM join");
        // the box's nodes (no pre-launched node is among them), after the pass: the same
        // workgroup shape over the box's rows
        const auto saved_pre = std::move(pre_level_end);
        pre_level_end.clear();
        p.region = 2;
        p.ox = ibm.box[0];
        p.oy = ibm.box[2];
        p.oz = ibm.box[4];
        p.xe = ibm.box[1];
        p.ye = ibm.box[3];
        p.nzl = ibm.box[5] - ibm.box[4];
        p.nbx = (ibm.box[1] - ibm.box[0] + 64 * p.wx - 1) / (64 * p.wx);
        p.nby = (ibm.box[3] - ibm.box[2] + 4 / p.wx - 1) / (4 / p.wx);
        launch_all(stream, dim3(unsigned((long long)p.nbx * p.nby * p.nzl)));
        pre_level_end = saved_pre;
        e = hipStreamWaitEvent(stream, ev_join, 0);
        if (e != hipSuccess) return hip_fail(e, "IBM join");
    }
#undef LBM3_LAUNCH_S
#undef LBM3_LAUNCH
#undef LBM3_LAUNCH_G
#undef LBM3_PRE
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "step3 launch");
    cur = 1 - cur;
    timestep++;
    if (write_macros) macros_valid = true;
    return LBM_OK;
}

// ---- C ABI --------------------------------------------------------------------------------

extern "C" {

int lbm3d_config_default(lbm3d_config* cfg, int32_t nx, int32_t ny, int32_t nz) {
    if (!cfg) return LBM_E_INVALID;
    std::memset(cfg, 0, sizeof *cfg);
    cfg->nx = nx;
    cfg->ny = ny;
    cfg->nz = nz;
    cfg->collision = LBM_BGK;
    cfg->omega = 1.0f;
    for (int i = 0; i < 27; i++) cfg->S[i] = 1.0f;
    cfg->rho0 = 1.0f;
    cfg->storage = LBM_F32;
    return LBM_OK;
}

int lbm3d_create(const lbm3d_config* cfg, lbm3d_ctx** out) {
    if (!cfg || !out) return LBM_E_INVALID;
    *out = nullptr;
    const lbm3d_config& k = *cfg;
    auto bad = [&](int code, const char* m) {
        g_create_err3 = m;
        return code;
    };
    if (k.nx < 3 || k.ny < 3 || k.nz < 3) return bad(LBM_E_INVALID, "grid must be at least 3x3x3");
    if ((long long)k.nx * k.ny * k.nz > (1LL << 31) - 64) return bad(LBM_E_INVALID, "grid exceeds 2^31 nodes");
    if (k.collision != LBM_BGK && k.collision != LBM_MRT && k.collision != LBM_CM) return bad(LBM_E_INVALID, "unknown collision operator");
    if (k.adapter == LBM_OPTIMAL_ADAPTER || k.adapter == LBM3D_EMPIRICAL_ADAPTER)
        return bad(LBM_E_UNSUPPORTED, "the reference's D3Q27 adapter operators (ACM, E-ACM) do not collide the node as shipped; CM<3> runs with NoAdapter");
    if (k.adapter != LBM_NO_ADAPTER) return bad(LBM_E_INVALID, "unknown adapter");
    if (k.collision != LBM_MRT && !(k.omega > 0.0f && k.omega < 2.0f)) return bad(LBM_E_INVALID, "BGK and CM need 0 < omega < 2 (tau > 0.5)");
    for (int i = 0; i < 27; i++)
        if (!std::isfinite(k.S[i])) return bad(LBM_E_INVALID, "relaxation rates must be finite");
    if (k.storage < LBM_F32 || k.storage > LBM_BF16_SHIFTED) return bad(LBM_E_INVALID, "unknown storage");
    if (!(k.rho0 > 0.0f)) return bad(LBM_E_INVALID, "rho0 must be positive");
    for (int c = 0; c < 3; c++)
        if (!std::isfinite(k.force[c])) return bad(LBM_E_INVALID, "force must be finite");
    if (k.quirks & ~(LBM_QUIRK_STALE_REST | LBM_QUIRK_EQ3D_LINEAR | LBM_QUIRK_MRT3_LITERAL | LBM_QUIRK_GUO3D_OWN_MOMENTS |
                     LBM_QUIRK_CORNER3D_UZ | LBM_QUIRK_IBM_CLAMP | LBM_QUIRK_IBM_NO_DS))
        return bad(LBM_E_UNSUPPORTED,
                   "only the STALE_REST, EQ3D_LINEAR, MRT3_LITERAL, GUO3D_OWN_MOMENTS, CORNER3D_UZ, IBM_CLAMP and IBM_NO_DS "
                   "quirks apply to D3Q27");
    if (!std::isfinite(k.u_inlet) || !(std::fabs(k.u_inlet) < 0.5f)) return bad(LBM_E_INVALID, "u_inlet must stay below 0.5");

    auto* c = new lbm3d_ctx();
    c->cfg = k;
    hipError_t e = hipGetDevice(&c->device);
    if (e != hipSuccess) {
        c->hip_fail(e, "hipGetDevice");
        g_create_err3 = c->err;
        delete c;
        return LBM_E_HIP;
    }
    c->N = (long long)k.nx * k.ny * k.nz;
    c->plane = (c->N + 63) / 64 * 64;
    e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        c->hip_fail(e, "hipStreamCreate");
        g_create_err3 = c->err;
        delete c;
        return LBM_E_HIP;
    }
    c->stream = c->own_stream;
    const size_t pb = size_t(c->plane) * Q * elem_bytes(k.storage);
    int rc = 0;
    rc = rc ? rc : c->dev_alloc(&c->buf[0], pb, "populations A");
    rc = rc ? rc : c->dev_alloc(&c->buf[1], pb, "populations B");
    rc = rc ? rc : c->dev_alloc(reinterpret_cast<void**>(&c->d_rho), size_t(c->plane) * 4, "rho");
    rc = rc ? rc : c->dev_alloc(reinterpret_cast<void**>(&c->d_u), size_t(c->plane) * 12, "u");
    rc = rc ? rc : c->dev_alloc(reinterpret_cast<void**>(&c->d_partials), kRedBlocks * 3 * sizeof(double), "partials");
    if (rc) {
        g_create_err3 = c->err;
        delete c;
        return rc;
    }
    *out = c;
    return LBM_OK;
}

int lbm3d_destroy(lbm3d_ctx* ctx) {
    delete ctx;
    return LBM_OK;
}

int lbm3d_set_stream(lbm3d_ctx* ctx, void* hip_stream) {
    if (!ctx) return LBM_E_INVALID;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "hipStreamSynchronize");
    ctx->stream = static_cast<hipStream_t>(hip_stream);
    return LBM_OK;
}

int lbm3d_set_flags(lbm3d_ctx* ctx, const uint8_t* flags, int on_device) {
    if (!ctx) return LBM_E_INVALID;
    int rc = ctx->activate();
    if (rc) return rc;
    if (!flags) {
        if (ctx->d_flags) {
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipFree(ctx->d_flags);
            ctx->d_flags = nullptr;
        }
        ctx->has_guo = false;
        ctx->has_reg = false;
        if (ctx->d_pre) {
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipFree(ctx->d_pre);
            ctx->d_pre = nullptr;
        }
        ctx->n_pre = 0;
        ctx->pre_level_end.clear();
        ctx->h_flags.clear();
        ctx->ibm_update_split();
        return LBM_OK;
    }
    std::vector<uint8_t> h(size_t(ctx->N));
    if (on_device) {
        hipError_t e = hipMemcpyAsync(h.data(), flags, h.size(), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "flags download");
    } else {
        std::memcpy(h.data(), flags, h.size());
    }
    bool any = false, guo = false, reg = false;
    const int nx = ctx->cfg.nx;
    for (size_t n = 0; n < h.size(); n++) {
        const int x = int(n % size_t(nx));
        if (h[n] == LBM_FLUID) continue;
        any = true;
        if (h[n] == LBM_BOUNCE_BACK || h[n] == LBM_CORNER_EDGE_BOUNCE_BACK) continue;
        if ((h[n] == LBM_GUO_VELOCITY_INLET && x == 0) ||
            ((h[n] == LBM_GUO_PRESSURE_OUTLET || h[n] == LBM_REGULARIZED_OUTLET) && x == nx - 1)) {
            if (ctx->cfg.periodic_x)
                return ctx->fail(LBM_E_INVALID, "Guo inlet/outlet and regularized outlet flags need a non-periodic x axis");
            guo = true;
            if (h[n] == LBM_REGULARIZED_OUTLET) reg = true;  // the GUO = 2 kernel variant
            continue;
        }
        if (h[n] == LBM_REGULARIZED_BOUNCE_BACK || h[n] == LBM_REGULARIZED_INLET_LEFT ||
            h[n] == LBM_EXTRAPOLATED_CORNER_EDGE) {
            guo = reg = true;  // the functors are compiled into the GUO = 2 kernel variant only
            continue;
        }
        if (h[n] == LBM_ZG_OUTFLOW) {
            guo = true;
            continue;
        }
        if (h[n] == LBM_GUO_VELOCITY_INLET || h[n] == LBM_GUO_PRESSURE_OUTLET || h[n] == LBM_REGULARIZED_OUTLET)
            return ctx->fail(LBM_E_UNSUPPORTED,
                             "flag %d at node %zu: Guo inlets are built at x = 0, Guo and regularized outlets at x = NX-1",
                             int(h[n]), n);
        return ctx->fail(LBM_E_UNSUPPORTED, "flag %d at node %zu: not built for D3Q27 (FLUID, BOUNCE_BACK, CORNER_EDGE_BOUNCE_BACK, REGULARIZED_BOUNCE_BACK, REGULARIZED_INLET_LEFT, REGULARIZED_OUTLET, EXTRAPOLATED_CORNER_EDGE, GUO_VELOCITY_INLET, GUO_PRESSURE_OUTLET, ZG_OUTFLOW are)", int(h[n]), n);
    }
    if (!any) return lbm3d_set_flags(ctx, nullptr, 0);
    // nodes reading a neighbour's stale slots that another wavefront (or an earlier launch)
    // rewrites run in pre-launches, in the order pre_launch_levels gives
    Step3Params q{};
    q.nx = ctx->cfg.nx;
    q.ny = ctx->cfg.ny;
    q.nz = ctx->cfg.nz;
    q.px = ctx->cfg.periodic_x != 0;
    q.py = ctx->cfg.periodic_y != 0;
    q.pz = ctx->cfg.periodic_z != 0;
    std::vector<std::vector<int>> levels;
    if (const char* why = pre_launch_levels(q, h.data(), levels)) return ctx->fail(LBM_E_UNSUPPORTED, "%s", why);
    std::vector<int> pre, level_end;
    for (size_t k = levels.size(); k-- > 0;) {  // the launch order: the last level first
        for (int n : levels[k]) {
            pre.push_back(n);
            h[size_t(n)] = uint8_t(h[size_t(n)] | kPreLaunch);
        }
        level_end.push_back(int(pre.size()));
    }
    ctx->has_guo = guo;
    ctx->has_reg = reg;
    if (ctx->d_pre) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(ctx->d_pre);
        ctx->d_pre = nullptr;
        ctx->n_pre = 0;
    }
    ctx->pre_level_end = level_end;
    if (!pre.empty()) {
        rc = ctx->dev_alloc(reinterpret_cast<void**>(&ctx->d_pre), pre.size() * sizeof(int), "pre-launch nodes");
        if (rc) return rc;
        rc = ctx->upload(ctx->d_pre, pre.data(), pre.size() * sizeof(int), 0);
        if (rc) return rc;
        ctx->n_pre = int(pre.size());
    }
    if (!ctx->d_flags) {
        rc = ctx->dev_alloc(reinterpret_cast<void**>(&ctx->d_flags), size_t(ctx->plane), "flags");
        if (rc) return rc;
    }
    rc = ctx->upload(ctx->d_flags, h.data(), h.size(), 0);
    if (rc) return rc;
    ctx->h_flags = std::move(h);  // with the pre-launch bits: such nodes never qualify for a split
    ctx->ibm_update_split();
    return LBM_OK;
}

int lbm3d_init_from_macros(lbm3d_ctx* ctx, const float* rho, const float* u, int on_device) {
    if (!ctx || !rho || !u) return LBM_E_INVALID;
    int rc = ctx->activate();
    if (rc) return rc;
    const float *dr = rho, *du = u;
    if (!on_device) {
        rc = ctx->scratch(size_t(ctx->N) * 16);
        if (rc) return rc;
        float* t = static_cast<float*>(ctx->tmp);
        rc = ctx->upload(t, rho, size_t(ctx->N) * 4, 0);
        rc = rc ? rc : ctx->upload(t + ctx->N, u, size_t(ctx->N) * 12, 0);
        if (rc) return rc;
        dr = t;
        du = t + ctx->N;
    }
    const dim3 g(grid1(ctx->N)), b(256);
    switch (ctx->cfg.storage) {
        case LBM_F32:
            hipLaunchKernelGGL(init3_kernel<LBM_F32>, g, b, 0, ctx->stream, ctx->buf[0], ctx->buf[1], ctx->plane, ctx->N, ctx->cfg.rho0, ctx->lin_ec(), dr, du);
            break;
        case LBM_F16_SHIFTED:
            hipLaunchKernelGGL(init3_kernel<LBM_F16_SHIFTED>, g, b, 0, ctx->stream, ctx->buf[0], ctx->buf[1], ctx->plane, ctx->N, ctx->cfg.rho0, ctx->lin_ec(), dr, du);
            break;
        default:
            hipLaunchKernelGGL(init3_kernel<LBM_BF16_SHIFTED>, g, b, 0, ctx->stream, ctx->buf[0], ctx->buf[1], ctx->plane, ctx->N, ctx->cfg.rho0, ctx->lin_ec(), dr, du);
            break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return ctx->hip_fail(e, "init3 launch");
    if (!on_device) {
        e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "hipStreamSynchronize");
    }
    ctx->cur = 0;
    ctx->timestep = 0;
    ctx->initialized = true;
    ctx->macros_valid = false;
    return LBM_OK;
}

int lbm3d_set_force_field(lbm3d_ctx* ctx, const float* force, int on_device) {
    if (!ctx) return LBM_E_INVALID;
    int rc = ctx->activate();
    if (rc) return rc;
    ctx->ibm.force_valid = false;
    if (!force) {
        if (ctx->d_force) {
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipFree(ctx->d_force);
            ctx->d_force = nullptr;
        }
        return LBM_OK;
    }
    if (!ctx->d_force) {
        rc = ctx->dev_alloc(reinterpret_cast<void**>(&ctx->d_force), size_t(ctx->plane) * 12, "force field");
        if (rc) return rc;
    }
    const float* in = force;
    if (!on_device) {
        rc = ctx->scratch(size_t(ctx->N) * 12);
        rc = rc ? rc : ctx->upload(ctx->tmp, force, size_t(ctx->N) * 12, 0);
        if (rc) return rc;
        in = static_cast<const float*>(ctx->tmp);
    }
    hipLaunchKernelGGL(aos3_to_soa_kernel, dim3(grid1(ctx->N)), dim3(256), 0, ctx->stream, in, ctx->N, ctx->plane, ctx->d_force);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return ctx->hip_fail(e, "force convert");
    if (!on_device) {
        e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "hipStreamSynchronize");
    }
    return LBM_OK;
}

int lbm3d_step(lbm3d_ctx* ctx, int32_t nsteps) {
    if (!ctx || nsteps < 0) return LBM_E_INVALID;
    if (!ctx->initialized) return ctx->fail(LBM_E_STATE, "lbm3d_step before lbm3d_init_from_macros / lbm3d_set_pops");
    int rc = ctx->activate();
    if (rc) return rc;
    for (int s = 0; s < nsteps; s++) {
        rc = ctx->one_step(s == nsteps - 1);
        if (rc) return rc;
    }
    return LBM_OK;
}

int lbm3d_sync(lbm3d_ctx* ctx) {
    if (!ctx) return LBM_E_INVALID;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? LBM_OK : ctx->hip_fail(e, "hipStreamSynchronize");
}

int lbm3d_get_timestep(const lbm3d_ctx* ctx, int64_t* t) {
    if (!ctx || !t) return LBM_E_INVALID;
    *t = ctx->timestep;
    return LBM_OK;
}

int lbm3d_get_macros(lbm3d_ctx* ctx, float* rho, float* u, int on_device) {
    if (!ctx) return LBM_E_INVALID;
    if (!ctx->macros_valid) return ctx->fail(LBM_E_STATE, "no macroscopic fields stored yet");
    int rc = ctx->activate();
    if (rc) return rc;
    if (rho) {
        rc = ctx->download(rho, ctx->d_rho, size_t(ctx->N) * 4, on_device);
        if (rc) return rc;
    }
    if (u) {
        float* o = u;
        if (!on_device) {
            rc = ctx->scratch(size_t(ctx->N) * 12);
            if (rc) return rc;
            o = static_cast<float*>(ctx->tmp);
        }
        hipLaunchKernelGGL(soa_to_aos3_kernel, dim3(grid1(ctx->N)), dim3(256), 0, ctx->stream, ctx->d_u, ctx->N, ctx->plane, o);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return ctx->hip_fail(e, "u convert");
        if (!on_device) return ctx->download(u, o, size_t(ctx->N) * 12, 0);
    }
    return LBM_OK;
}

int lbm3d_get_pops(lbm3d_ctx* ctx, float* f, int on_device) {
    if (!ctx || !f) return LBM_E_INVALID;
    if (!ctx->initialized) return ctx->fail(LBM_E_STATE, "no state yet");
    int rc = ctx->activate();
    if (rc) return rc;
    float* o = f;
    if (!on_device) {
        rc = ctx->scratch(size_t(ctx->N) * Q * 4);
        if (rc) return rc;
        o = static_cast<float*>(ctx->tmp);
    }
    const dim3 g(grid1(ctx->N)), b(256);
    const void* s = ctx->buf[ctx->cur];
    switch (ctx->cfg.storage) {
        case LBM_F32: hipLaunchKernelGGL(pops_to_f32_kernel<LBM_F32>, g, b, 0, ctx->stream, s, ctx->plane, ctx->N, ctx->cfg.rho0, o); break;
        case LBM_F16_SHIFTED: hipLaunchKernelGGL(pops_to_f32_kernel<LBM_F16_SHIFTED>, g, b, 0, ctx->stream, s, ctx->plane, ctx->N, ctx->cfg.rho0, o); break;
        default: hipLaunchKernelGGL(pops_to_f32_kernel<LBM_BF16_SHIFTED>, g, b, 0, ctx->stream, s, ctx->plane, ctx->N, ctx->cfg.rho0, o); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return ctx->hip_fail(e, "pops convert");
    if (!on_device) return ctx->download(f, o, size_t(ctx->N) * Q * 4, 0);
    return LBM_OK;
}

int lbm3d_set_pops(lbm3d_ctx* ctx, const float* f, int on_device) {
    if (!ctx || !f) return LBM_E_INVALID;
    int rc = ctx->activate();
    if (rc) return rc;
    const float* in = f;
    if (!on_device) {
        rc = ctx->scratch(size_t(ctx->N) * Q * 4);
        rc = rc ? rc : ctx->upload(ctx->tmp, f, size_t(ctx->N) * Q * 4, 0);
        if (rc) return rc;
        in = static_cast<const float*>(ctx->tmp);
    }
    const dim3 g(grid1(ctx->N)), b(256);
    switch (ctx->cfg.storage) {
        case LBM_F32: hipLaunchKernelGGL(pops_from_f32_kernel<LBM_F32>, g, b, 0, ctx->stream, ctx->buf[0], ctx->buf[1], ctx->plane, ctx->N, ctx->cfg.rho0, in); break;
        case LBM_F16_SHIFTED: hipLaunchKernelGGL(pops_from_f32_kernel<LBM_F16_SHIFTED>, g, b, 0, ctx->stream, ctx->buf[0], ctx->buf[1], ctx->plane, ctx->N, ctx->cfg.rho0, in); break;
        default: hipLaunchKernelGGL(pops_from_f32_kernel<LBM_BF16_SHIFTED>, g, b, 0, ctx->stream, ctx->buf[0], ctx->buf[1], ctx->plane, ctx->N, ctx->cfg.rho0, in); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return ctx->hip_fail(e, "pops convert");
    if (!on_device) {
        e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "hipStreamSynchronize");
    }
    ctx->cur = 0;
    ctx->initialized = true;
    ctx->macros_valid = false;
    return LBM_OK;
}

int lbm3d_reduce(lbm3d_ctx* ctx, int32_t kind, double* out) {
    if (!ctx || !out) return LBM_E_INVALID;
    if (kind < LBM3D_REDUCE_MASS || kind > LBM3D_REDUCE_KINETIC_ENERGY) return ctx->fail(LBM_E_INVALID, "unknown reduction %d", kind);
    if (!ctx->macros_valid) return ctx->fail(LBM_E_STATE, "no macroscopic fields stored yet");
    int rc = ctx->activate();
    if (rc) return rc;
    const long long need = (ctx->N + 255) / 256;
    const int nb = int(need < kRedBlocks ? need : kRedBlocks);
    hipLaunchKernelGGL(reduce3_kernel, dim3(nb), dim3(256), 0, ctx->stream, kind, ctx->d_rho, ctx->d_u, ctx->plane, ctx->N, ctx->d_partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return ctx->hip_fail(e, "reduce3 launch");
    std::vector<double> h(size_t(nb) * 3);
    rc = ctx->download(h.data(), ctx->d_partials, h.size() * sizeof(double), 0);
    if (rc) return rc;
    double s[3] = {0.0, 0.0, 0.0};
    for (int b = 0; b < nb; b++)
        for (int c = 0; c < 3; c++) s[c] += h[3 * size_t(b) + c];
    out[0] = s[0];
    if (kind == LBM3D_REDUCE_MOMENTUM) {
        out[1] = s[1];
        out[2] = s[2];
    }
    return LBM_OK;
}

int lbm3d_ibm_set_markers(lbm3d_ctx* ctx, const float* xyz, const float* ds, int32_t n, int32_t iters) {
    if (!ctx || n < 0 || (n > 0 && !xyz) || iters < 1) return LBM_E_INVALID;
    int rc = ctx->activate();
    if (rc) return rc;
    (void)hipStreamSynchronize(ctx->stream);
    ctx->ibm_free();
    if (n == 0) return LBM_OK;
    lbm::IbmTables3 T;
    const char* why = lbm::ibm_build_tables3(ctx->cfg.nx, ctx->cfg.ny, ctx->cfg.nz, (ctx->cfg.quirks & LBM_QUIRK_IBM_NO_DS) != 0,
                                             xyz, ds, n, T);
    if (why) return ctx->fail(LBM_E_INVALID, "%s", why);
    auto& I = ctx->ibm;
    const int nb = int(T.band_node.size());
    auto put = [&](auto** dst, const auto& v, const char* what) {
        using E = typename std::decay_t<decltype(v)>::value_type;
        int r = ctx->dev_alloc(reinterpret_cast<void**>(dst), std::max<size_t>(v.size(), 1) * sizeof(E), what);
        return r ? r : (v.empty() ? LBM_OK : ctx->upload(*dst, v.data(), v.size() * sizeof(E), 0));
    };
    rc = put(&I.band_node, T.band_node, "IBM band");
    rc = rc ? rc : put(&I.mk_ptr, T.mk_ptr, "IBM stencils");
    rc = rc ? rc : put(&I.mk_band, T.mk_band, "IBM stencils");
    rc = rc ? rc : put(&I.mk_w, T.mk_w, "IBM weights");
    rc = rc ? rc : put(&I.bn_ptr, T.bn_ptr, "IBM spread lists");
    rc = rc ? rc : put(&I.bn_mk, T.bn_mk, "IBM spread lists");
    rc = rc ? rc : put(&I.bn_w, T.bn_w, "IBM spread weights");
    rc = rc ? rc : ctx->dev_alloc(reinterpret_cast<void**>(&I.brho), size_t(nb) * 4, "IBM band rho");
    rc = rc ? rc : ctx->dev_alloc(reinterpret_cast<void**>(&I.bu), size_t(nb) * 12, "IBM band velocity");
    rc = rc ? rc : ctx->dev_alloc(reinterpret_cast<void**>(&I.bf), size_t(nb) * 12, "IBM band force");
    rc = rc ? rc : ctx->dev_alloc(reinterpret_cast<void**>(&I.lag), size_t(n) * 12, "IBM Lagrangian forces");
    rc = rc ? rc : ctx->dev_alloc(reinterpret_cast<void**>(&I.lag_sum), size_t(n) * 12, "IBM Lagrangian forces");
    rc = rc ? rc : ctx->dev_alloc(reinterpret_cast<void**>(&I.total), size_t(ctx->plane) * 12, "IBM force field");
    if (!rc) {
        hipError_t e = hipMemsetAsync(I.lag_sum, 0, size_t(n) * 12, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = ctx->hip_fail(e, "IBM reset");
    }
    if (rc) {
        ctx->ibm_free();
        return rc;
    }
    const int nx = ctx->cfg.nx, nxy = nx * ctx->cfg.ny;
    int box[6] = {nx, 0, ctx->cfg.ny, 0, ctx->cfg.nz, 0};
    for (int v : T.band_node) {
        const int z = v / nxy, y = (v - z * nxy) / nx, x = v - z * nxy - y * nx;
        box[0] = std::min(box[0], x);
        box[1] = std::max(box[1], x + 1);
        box[2] = std::min(box[2], y);
        box[3] = std::max(box[3], y + 1);
        box[4] = std::min(box[4], z);
        box[5] = std::max(box[5], z + 1);
    }
    std::copy(box, box + 6, I.box);
    I.n_markers = n;
    I.n_band = nb;
    I.iters = iters;
    I.h_ds.assign(size_t(n), 1.0f);
    if (ds) std::copy(ds, ds + n, I.h_ds.begin());
    I.force_valid = false;
    I.active = true;
    ctx->ibm_update_split();
    return LBM_OK;
}

int lbm3d_ibm_get_forces(lbm3d_ctx* ctx, float* lag_force, double* force_sum) {
    if (!ctx) return LBM_E_INVALID;
    if (!ctx->ibm.active) return ctx->fail(LBM_E_STATE, "no IBM markers set");
    int rc = ctx->activate();
    if (rc) return rc;
    const int n = ctx->ibm.n_markers;
    std::vector<float> h(3 * size_t(n));
    rc = ctx->download(h.data(), ctx->ibm.lag_sum, h.size() * sizeof(float), 0);
    if (rc) return rc;
    const bool no_ds = (ctx->cfg.quirks & LBM_QUIRK_IBM_NO_DS) != 0;
    double t[3] = {0.0, 0.0, 0.0};
    for (int l = 0; l < n; l++)
        for (int d = 0; d < 3; d++) {
            const float v = h[size_t(d) * n + l];
            if (lag_force) lag_force[3 * size_t(l) + d] = v;
            t[d] += double(v) * (no_ds ? 1.0 : double(ctx->ibm.h_ds[size_t(l)]));
        }
    if (force_sum)
        for (int d = 0; d < 3; d++) force_sum[d] = t[d];
    return LBM_OK;
}

const char* lbm3d_last_error(const lbm3d_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_err3.c_str(); }

}  // extern "C"
