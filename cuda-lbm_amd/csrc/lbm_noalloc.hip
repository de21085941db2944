// lbm_noalloc.hip — lbm_reduce for callers that cannot allocate device memory.
//
// lbm_reduce stages its block partial sums (and a host reference field) in a buffer the
// context allocates on first use. lbm_reduce_noalloc runs the same kernels with the partials
// in `scratch`, a fixed file-scope device array: one per GPU, held until the process ends,
// shared by every call on that GPU (calls that overlap on one GPU race on it). The reference
// field of LBM_REDUCE_L2_VS_REF must already be device memory.
#include "lbm_ctx.hpp"

#include <mutex>

using namespace lbm;

namespace {

constexpr size_t kScratchDoubles = LBM_NOALLOC_SCRATCH_BYTES / sizeof(double);
static __device__ __align__(256) double scratch[kScratchDoubles];

// Device address of `scratch` per device, looked up once per device with that device
// current. The array's host-side name is only a handle for the runtime, never an address.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int scratch_address(lbm_ctx* ctx, double** out) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices)
        return ctx->fail(LBM_E_UNSUPPORTED, "device %d beyond the %d devices the scratch table covers", ctx->device,
                         kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[ctx->device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return ctx->hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[ctx->device] = p;
    }
    *out = static_cast<double*>(g_scratch_ptr[ctx->device]);
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_reduce_noalloc(lbm_ctx* ctx, int32_t kind, const float* ref_device, double* out, int out_on_device) {
    if (!ctx || !out) return LBM_E_INVALID;
    if (!ctx->macros_valid) return ctx->fail(LBM_E_STATE, "no macroscopic fields stored yet");
    int rc = ctx->activate();  // the scratch address below is the current device's
    if (rc) return rc;
    const hipMemcpyKind to_out = out_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    auto finish = [&](const void* src, size_t bytes, const char* what) -> int {
        hipError_t e = hipMemcpyAsync(out, src, bytes, to_out, ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, what);
        if (out_on_device) return LBM_OK;
        e = hipStreamSynchronize(ctx->stream);
        return e == hipSuccess ? LBM_OK : ctx->hip_fail(e, "hipStreamSynchronize");
    };
    if (kind == LBM_REDUCE_MOMENT_AVG) {
        if (!ctx->avg_valid) return ctx->fail(LBM_E_STATE, "no adapter averages computed yet");
        return finish(ctx->d_avg64, 3 * sizeof(double), "avg copy");
    }
    double* buf = nullptr;
    rc = scratch_address(ctx, &buf);
    if (rc) return rc;
    hipError_t e;
    if (kind == LBM_REDUCE_ROW_MEAN_UX) {
        if (size_t(ctx->cfg.ny) > kScratchDoubles)
            return ctx->fail(LBM_E_UNSUPPORTED, "ROW_MEAN_UX of %d rows exceeds the %zu-row scratch", ctx->cfg.ny,
                             kScratchDoubles);
        e = launch_row_mean_ux(ctx->d_ux, ctx->cfg.nx, ctx->cfg.ny, buf, ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "row mean launch");
        return finish(buf, size_t(ctx->cfg.ny) * sizeof(double), "row mean copy");
    }
    if (kind != LBM_REDUCE_MASS && kind != LBM_REDUCE_MOMENTUM && kind != LBM_REDUCE_L2_VS_REF)
        return ctx->fail(LBM_E_INVALID, "unknown reduction kind %d", int(kind));
    if (kind == LBM_REDUCE_L2_VS_REF && !ref_device)
        return ctx->fail(LBM_E_INVALID, "L2_VS_REF needs a reference field in device memory");
    const long long N = (long long)ctx->N;
    const long long np = reduce_partials_needed(N);
    if (size_t(np) * 3 + 4 > kScratchDoubles)
        return ctx->fail(LBM_E_UNSUPPORTED, "a grid of %lld nodes needs %lld partial sums; the scratch holds %zu", N,
                         np, (kScratchDoubles - 4) / 3);
    // every partial is written by the partials kernel and every output by the sum kernel:
    // nothing in the scratch has to start at zero
    double* part = buf;
    double* o3 = buf + 3 * np;
    e = launch_reduce(kind, ctx->d_rho, ctx->d_ux, ctx->d_uy, kind == LBM_REDUCE_L2_VS_REF ? ref_device : nullptr, N,
                      part, o3, ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "reduce launch");
    return finish(o3, (kind == LBM_REDUCE_MASS ? 1 : 2) * sizeof(double), "reduce copy");
}

}  // extern "C"
