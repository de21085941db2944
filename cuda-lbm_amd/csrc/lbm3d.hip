// lbm3d.hip — D3Q27 kernels and the C ABI of include/lbm3d_hip.h.
//
// Replaces LBM<3> (src/core/lbm.cuh:33-382) for BGK<3>, MRT<3> and CM<3>: its per-stage kernels
// and f_eq buffer become one fused pull-scheme kernel per step (lbm3d_step.hpp), launched as
// 256-thread workgroups holding whole x-rows up to 256 nodes. The step is HBM-bound: 27
// population reads and 27 writes per node update (216 B in FP32, 108 B in 16-bit storage)
// against ~330-370 VALU (BGK, MRT and CM without 27x27 matrices).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lbm3d_hip.h"
#include "lbm3d_step.hpp"

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
    const long long nb = (long long)p.nbx * p.nby * p.nz;
    const long long b = blockIdx.x;
    const long long per = nb / 8;
    const long long L = b < per * 8 ? (b % 8) * per + b / 8 : b;
    int x, y, z;
    thread_node(p, L, int(threadIdx.x), x, y, z);
    if (x >= p.nx || y >= p.ny) return;
    if (GUO && guo_ring(p, x, y, z)) return;
    step3_node<COLL, STORE, BUF, GUO>(p, x, y, z);
}

// The outlet's ring nodes (guo_ring) when they do not share a wavefront with their neighbour:
// thread i looks at node (side ? NX-1 : 0, y, z) of the 2 * NY * NZ plane nodes. Runs before
// step3_kernel, because these nodes read their x - 1 neighbours' stale slots, which that launch
// rewrites.
template <int COLL, int STORE, bool BUF, int GUO>
__global__ __launch_bounds__(256) void step3_ring_kernel(Step3Params p) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long per = (long long)p.ny * p.nz;
    if (i >= 2 * per) return;
    const int side = int(i / per);
    const long long r = i - side * per;
    const int x = side ? p.nx - 1 : 0, y = int(r % p.ny), z = int(r / p.ny);
    if (!guo_ring(p, x, y, z)) return;
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
    long long N = 0, plane = 0;
    void* buf[2] = {nullptr, nullptr};
    int cur = 0;                       // buffer holding the latest post-collision state
    uint8_t* d_flags = nullptr;        // null while every node is FLUID
    bool has_guo = false;              // Guo inlet/outlet, x-face ZG or regularized flags present
    bool has_reg = false;              // regularized flags present (the GUO = 2 kernel variant)
    float* d_force = nullptr;          // SoA [3][plane] or null
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
        for (void* p : {buf[0], buf[1], (void*)d_flags, (void*)d_force, (void*)d_rho, (void*)d_u, (void*)d_partials, tmp})
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
    int one_step(bool write_macros);
};

int lbm3d_ctx::one_step(bool write_macros) {
    Step3Params p{};
    p.src = buf[cur];
    p.dst = buf[1 - cur];
    p.flags = d_flags;
    p.force = d_force;
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
    p.guo_ring_outlet = (has_guo && (cfg.nx - 1) % 64 == 0) ? 1 : 0;
    p.stale_rest = (cfg.quirks & LBM_QUIRK_STALE_REST) ? 1 : 0;
    // Non-temporal stores keep the written populations from evicting the pulls' shared lines
    // once the two buffers exceed the 256 MB Infinity Cache (+2% FP32, +4% FP16 at 256^3);
    // on resident grids they push every write to HBM (-20% FP32 at 64^3 and 96^3).
    p.nt = 2.0 * Q * double(plane) * double(elem_bytes(cfg.storage)) > 192.0e6 ? 1 : 0;
    p.wx = wave_cols(cfg.nx);
    p.nbx = (cfg.nx + 64 * p.wx - 1) / (64 * p.wx);
    p.nby = (cfg.ny + 4 / p.wx - 1) / (4 / p.wx);
    const long long nblk = (long long)p.nbx * p.nby * cfg.nz;
    if (nblk > 0x7fffffffLL) return fail(LBM_E_INVALID, "grid needs too many workgroups");
    const bool fits = cfg.storage == LBM_F32 ? lbm::buf_fits<float>(Q, plane) : lbm::buf_fits<uint16_t>(Q, plane);
    const dim3 grid{unsigned(nblk)}, block{256};
    const long long nplane = 2LL * cfg.ny * cfg.nz;
    const dim3 pgrid{unsigned((nplane + 255) / 256)};
#define LBM3_LAUNCH_G(C, S, G)                                                                  \
    if (fits) {                                                                                 \
        if (p.guo_ring_outlet)                                                                  \
            hipLaunchKernelGGL((step3_ring_kernel<C, S, true, G>), pgrid, block, 0, stream, p); \
        hipLaunchKernelGGL((step3_kernel<C, S, true, G>), grid, block, 0, stream, p);           \
    } else {                                                                                    \
        if (p.guo_ring_outlet)                                                                  \
            hipLaunchKernelGGL((step3_ring_kernel<C, S, false, G>), pgrid, block, 0, stream, p); \
        hipLaunchKernelGGL((step3_kernel<C, S, false, G>), grid, block, 0, stream, p);          \
    }
#define LBM3_LAUNCH(C, S)                                                                       \
    if (has_reg) {                                                                              \
        LBM3_LAUNCH_G(C, S, 2)                                                                  \
    } else if (has_guo) {                                                                       \
        LBM3_LAUNCH_G(C, S, 1)                                                                  \
    } else if (fits) {                                                                          \
        hipLaunchKernelGGL((step3_kernel<C, S, true, 0>), grid, block, 0, stream, p);       \
    } else {                                                                                    \
        hipLaunchKernelGGL((step3_kernel<C, S, false, 0>), grid, block, 0, stream, p);      \
    }
#define LBM3_LAUNCH_S(C)                                                                        \
    switch (cfg.storage) {                                                                      \
        case LBM_F32: LBM3_LAUNCH(C, LBM_F32) break;                                            \
        case LBM_F16_SHIFTED: LBM3_LAUNCH(C, LBM_F16_SHIFTED) break;                            \
        default: LBM3_LAUNCH(C, LBM_BF16_SHIFTED) break;                                        \
    }
    if (cfg.collision == LBM_MRT) {
        LBM3_LAUNCH_S(LBM_MRT)
    } else if (cfg.collision == LBM_CM) {
        LBM3_LAUNCH_S(LBM_CM)
    } else {
        LBM3_LAUNCH_S(LBM_BGK)
    }
#undef LBM3_LAUNCH_S
#undef LBM3_LAUNCH
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
    if (k.quirks & ~(LBM_QUIRK_STALE_REST | LBM_QUIRK_EQ3D_LINEAR | LBM_QUIRK_MRT3_LITERAL | LBM_QUIRK_GUO3D_OWN_MOMENTS))
        return bad(LBM_E_UNSUPPORTED, "only the STALE_REST, EQ3D_LINEAR, MRT3_LITERAL and GUO3D_OWN_MOMENTS quirks apply to D3Q27");
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
        if ((h[n] == LBM_GUO_VELOCITY_INLET && x == 0) || (h[n] == LBM_GUO_PRESSURE_OUTLET && x == nx - 1)) {
            if (ctx->cfg.periodic_x) return ctx->fail(LBM_E_INVALID, "Guo inlet/outlet flags need a non-periodic x axis");
            guo = true;
            continue;
        }
        if (h[n] == LBM_REGULARIZED_BOUNCE_BACK || h[n] == LBM_REGULARIZED_INLET_LEFT) {
            guo = reg = true;  // the functor is compiled into the GUO = 2 kernel variant only
            continue;
        }
        if (h[n] == LBM_ZG_OUTFLOW && (x == 0 || x == nx - 1)) {
            if (ctx->cfg.periodic_x) return ctx->fail(LBM_E_INVALID, "ZG_OUTFLOW flags need a non-periodic x axis");
            guo = true;
            continue;
        }
        if (h[n] == LBM_GUO_VELOCITY_INLET || h[n] == LBM_GUO_PRESSURE_OUTLET)
            return ctx->fail(LBM_E_UNSUPPORTED, "flag %d at node %zu: Guo inlets are built at x = 0 and outlets at x = NX-1", int(h[n]), n);
        if (h[n] == LBM_ZG_OUTFLOW)
            return ctx->fail(LBM_E_UNSUPPORTED, "ZG_OUTFLOW at node %zu: built on the x = 0 and x = NX-1 faces only", n);
        return ctx->fail(LBM_E_UNSUPPORTED, "flag %d at node %zu: not built for D3Q27 (FLUID, BOUNCE_BACK, CORNER_EDGE_BOUNCE_BACK, REGULARIZED_BOUNCE_BACK, REGULARIZED_INLET_LEFT, GUO_VELOCITY_INLET, GUO_PRESSURE_OUTLET, ZG_OUTFLOW are)", int(h[n]), n);
    }
    if (!any) return lbm3d_set_flags(ctx, nullptr, 0);
    ctx->has_guo = guo;
    ctx->has_reg = reg;
    if (!ctx->d_flags) {
        rc = ctx->dev_alloc(reinterpret_cast<void**>(&ctx->d_flags), size_t(ctx->plane), "flags");
        if (rc) return rc;
    }
    return ctx->upload(ctx->d_flags, h.data(), h.size(), 0);
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

const char* lbm3d_last_error(const lbm3d_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_err3.c_str(); }

}  // extern "C"
