// lbm_adjoint.hip — discrete adjoint of the fused step (no counterpart in the reference).
#include "lbm_ctx.hpp"

void lbm_ctx::adjoint_free() { adj = AdjointState(); }

extern "C" {

int lbm_adjoint_forward(lbm_ctx* ctx, int32_t, int32_t) {
    if (!ctx) return LBM_E_INVALID;
    return ctx->fail(LBM_E_UNSUPPORTED, "adjoint not built yet");
}

int lbm_adjoint_backward(lbm_ctx* ctx, const float*, float*, float*, int) {
    if (!ctx) return LBM_E_INVALID;
    return ctx->fail(LBM_E_UNSUPPORTED, "adjoint not built yet");
}

}  // extern "C"
