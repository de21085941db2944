// lbm_adjoint.hip — discrete adjoint of the fused step (no counterpart in the reference).
//
// Forward: the post-collision state f*_t is stored every K steps (checkpoints). Backward,
// segment by segment from the last: recompute the segment's K states from its checkpoint
// with the forward kernel, then run the adjoint steps t = end .. start+1:
//   lambda(x)[q] = mu_{t+1}(x + c_q)[q]            (reverse streaming: pull from downstream)
//   g(x)[q]      = f*_{t-1}(x - c_q)[q]            (forward gather, regathered)
//   mu_t(x)      = (dC/dg)^T lambda,  dL/dF(x) += (dC/dF)^T lambda
// seeded at t = T by the snapshot u_T = (sum g c + F/2)/rho. Fully periodic, FLUID-only
// domains with FP32 storage and BGK or MRT collision.
#include "lbm_adjoint.hpp"
#include "lbm_ctx.hpp"

#include <vector>

using namespace lbm;

namespace {

constexpr int kAdjBlock = 256;

struct AdjParams {
    int nx, ny;
    long long plane;
    const float* mu_next;   // mu_{t+1}, SoA planes
    float* mu_cur;          // mu_t
    const float* state;     // f*_{t-1}
    const float* force;     // AoS or null
    float fx, fy;
    float omega;
    float S[9];
    float* grad_force;      // AoS [N][2]
    const float* dLdu;      // AoS [N][2] (seed only)
    const float* rho0;      // initial density (u0 kernel)
    const float* u0x;
    const float* u0y;
    float* grad_u0;         // AoS
};

// Periodic neighbours of (x, y): columns x-1, x+1 and row offsets of rows y-1, y, y+1
// (32-bit element offsets; N < 2^31).
struct Nb {
    uint32_t xm, x, xp, rm, r0, rp;
};
__device__ __forceinline__ Nb neighbours(int x, int y, int nx, int ny) {
    Nb b;
    b.x = uint32_t(x);
    b.xm = uint32_t(x == 0 ? nx - 1 : x - 1);
    b.xp = uint32_t(x == nx - 1 ? 0 : x + 1);
    b.rm = uint32_t(y == 0 ? ny - 1 : y - 1) * uint32_t(nx);
    b.r0 = uint32_t(y) * uint32_t(nx);
    b.rp = uint32_t(y == ny - 1 ? 0 : y + 1) * uint32_t(nx);
    return b;
}
// Offset of the node at (x + s cx_q, y + s cy_q): s = -1 upstream (pull), +1 downstream.
template <int S>
__device__ __forceinline__ uint32_t nb_offset(const Nb& b, int q) {
    const int dx = S * CX[q], dy = S * CY[q];
    return (dy < 0 ? b.rm : (dy > 0 ? b.rp : b.r0)) + (dx < 0 ? b.xm : (dx > 0 ? b.xp : b.x));
}

__device__ __forceinline__ void gather_state(const AdjParams& p, const Nb& b, float* g) {
#pragma unroll
    for (int q = 0; q < Q; q++) g[q] = p.state[q * p.plane + nb_offset<-1>(b, q)];
}

// Uniform force, else the field (value selects: see node_force in lbm_step.hpp).
__device__ __forceinline__ void node_force(const AdjParams& p, uint32_t n, float& fx, float& fy) {
    const float* f = p.force;
    float vx = 0.0f, vy = 0.0f;
    if (f) {
        vx = f[2 * size_t(n)];
        vy = f[2 * size_t(n) + 1];
    }
    fx = f ? vx : p.fx;
    fy = f ? vy : p.fy;
}

// Workgroups cover rows in pieces of kAdjBlock nodes, ordered so that each of the 8 XCDs
// runs a contiguous range of rows (the y +- 1 neighbours then share its L2); the grid is
// padded to a multiple of 8 workgroups.
constexpr int kAdjXcds = 8;
__device__ __forceinline__ bool adj_node(const AdjParams& p, int& x, int& y, uint32_t& n) {
    const int bpr = (p.nx + kAdjBlock - 1) / kAdjBlock;
    const unsigned nb = gridDim.x;
    const unsigned lb = (blockIdx.x % kAdjXcds) * (nb / kAdjXcds) + blockIdx.x / kAdjXcds;
    y = int(lb / unsigned(bpr));
    x = int(lb - unsigned(y) * unsigned(bpr)) * kAdjBlock + int(threadIdx.x);
    if (y >= p.ny || x >= p.nx) return false;
    n = uint32_t(y) * uint32_t(p.nx) + uint32_t(x);
    return true;
}

__global__ __launch_bounds__(kAdjBlock) void adjoint_seed_kernel(AdjParams p) {
    int x, y;
    uint32_t n;
    if (!adj_node(p, x, y, n)) return;
    const Nb b = neighbours(x, y, p.nx, p.ny);
    float g[Q];
    gather_state(p, b, g);
    float fx, fy;
    node_force(p, n, fx, fy);
    AdjOut o = adjoint_snapshot(g, fx, fy, p.dLdu[2 * size_t(n)], p.dLdu[2 * size_t(n) + 1]);
#pragma unroll
    for (int q = 0; q < Q; q++) p.mu_cur[q * p.plane + n] = o.mu[q];
    p.grad_force[2 * size_t(n)] = o.gfx;
    p.grad_force[2 * size_t(n) + 1] = o.gfy;
}

template <int COLL>
__global__ __launch_bounds__(kAdjBlock) void adjoint_step_kernel(AdjParams p) {
    int x, y;
    uint32_t n;
    if (!adj_node(p, x, y, n)) return;
    const Nb b = neighbours(x, y, p.nx, p.ny);
    float lam[Q], g[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) lam[q] = p.mu_next[q * p.plane + nb_offset<1>(b, q)];
    gather_state(p, b, g);
    float fx, fy;
    node_force(p, n, fx, fy);
    AdjOut o = COLL == LBM_BGK ? adjoint_bgk(g, lam, fx, fy, p.omega) : adjoint_mrt(g, lam, fx, fy, p.S);
#pragma unroll
    for (int q = 0; q < Q; q++) p.mu_cur[q * p.plane + n] = o.mu[q];
    p.grad_force[2 * size_t(n)] += o.gfx;
    p.grad_force[2 * size_t(n) + 1] += o.gfy;
}

__global__ __launch_bounds__(kAdjBlock) void adjoint_u0_kernel(AdjParams p) {
    int x, y;
    uint32_t n;
    if (!adj_node(p, x, y, n)) return;
    const Nb b = neighbours(x, y, p.nx, p.ny);
    float lam[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) lam[q] = p.mu_next[q * p.plane + nb_offset<1>(b, q)];
    float gx, gy;
    adjoint_init_u(lam, p.rho0[n], p.u0x[n], p.u0y[n], gx, gy);
    p.grad_u0[2 * size_t(n)] = gx;
    p.grad_u0[2 * size_t(n) + 1] = gy;
}

dim3 node_grid(int nx, int ny) {
    const unsigned long long nb = (unsigned long long)((nx + kAdjBlock - 1) / kAdjBlock) * unsigned(ny);
    return dim3(unsigned((nb + kAdjXcds - 1) / kAdjXcds * kAdjXcds));
}

}  // namespace

void lbm_ctx::adjoint_free() {
    for (float* c : adj.checkpoints)
        if (c) (void)hipFree(c);
    void* ptrs[] = {adj.d_seg, adj.d_mu[0], adj.d_mu[1], adj.d_grad_force, adj.d_u0, adj.d_rho_init};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    adj = AdjointState();
}

extern "C" {

int lbm_adjoint_forward(lbm_ctx* ctx, int32_t nsteps, int32_t every) {
    if (!ctx || nsteps < 1 || every < 1) return LBM_E_INVALID;
    if (!ctx->initialized) return ctx->fail(LBM_E_STATE, "adjoint forward before initialisation");
    const lbm_config& c = ctx->cfg;
    if (!(c.collision == LBM_BGK || c.collision == LBM_MRT) || c.storage != LBM_F32 || !c.periodic_x ||
        !c.periodic_y || ctx->d_flags || ctx->ibm.active || (c.quirks & (LBM_QUIRK_STALE_REST | LBM_QUIRK_MRT_FORCE_ROWS)))
        return ctx->fail(LBM_E_UNSUPPORTED,
                         "the adjoint covers BGK/MRT, FP32 storage, fully periodic FLUID-only domains without IBM or quirks");
    int rc = ctx->activate();
    if (rc) return rc;
    AdjointState& A = ctx->adj;
    const size_t pb = size_t(ctx->plane) * 9 * sizeof(float);
    const int ncp = (nsteps + every - 1) / every;
    A.ready = false;
    if (A.checkpoints.size() != size_t(ncp)) {  // same count: the buffers are reused as they are
        (void)hipStreamSynchronize(ctx->stream);
        ctx->adjoint_free();
        A.checkpoints.assign(size_t(ncp), nullptr);
        for (int i = 0; i < ncp; i++) {
            rc = ctx->dev_alloc(reinterpret_cast<void**>(&A.checkpoints[size_t(i)]), pb, "adjoint checkpoint");
            if (rc) {
                ctx->adjoint_free();
                return rc;
            }
        }
    }
    const bool from_init = ctx->timestep == 0 && ctx->init_from_macros && ctx->macros_valid;
    if (from_init) {
        if (!A.d_u0) {
            rc = ctx->dev_alloc(reinterpret_cast<void**>(&A.d_rho_init), ctx->N * 4, "adjoint rho0");
            rc = rc ? rc : ctx->dev_alloc(reinterpret_cast<void**>(&A.d_u0), ctx->N * 8, "adjoint u0");
            if (rc) {
                ctx->adjoint_free();
                return rc;
            }
        }
        (void)hipMemcpyAsync(A.d_rho_init, ctx->d_rho, ctx->N * 4, hipMemcpyDeviceToDevice, ctx->stream);
        (void)hipMemcpyAsync(A.d_u0, ctx->d_ux, ctx->N * 4, hipMemcpyDeviceToDevice, ctx->stream);
        (void)hipMemcpyAsync(A.d_u0 + ctx->N, ctx->d_uy, ctx->N * 4, hipMemcpyDeviceToDevice, ctx->stream);
    }
    const bool pairs = ctx->two_step_ok();
    for (int t = 0; t < nsteps;) {
        if (t % every == 0) {
            hipError_t e = hipMemcpyAsync(A.checkpoints[size_t(t / every)], ctx->buf[ctx->cur], pb, hipMemcpyDeviceToDevice,
                                          ctx->stream);
            if (e != hipSuccess) return ctx->hip_fail(e, "checkpoint copy");
        }
        // a pair of steps per launch unless a checkpoint falls between them
        if (pairs && t + 2 <= nsteps && (t + 1) % every != 0) {
            rc = ctx->two_step(t + 2 == nsteps);
            t += 2;
        } else {
            rc = ctx->one_step(t == nsteps - 1);
            t += 1;
        }
        if (rc) return rc;
    }
    A.nsteps = nsteps;
    A.every = every;
    A.has_u0 = from_init;
    A.ready = true;
    return LBM_OK;
}

int lbm_adjoint_backward(lbm_ctx* ctx, const float* dL_du, float* dL_dforce, float* dL_du0, int on_device) {
    if (!ctx || !dL_du) return LBM_E_INVALID;
    AdjointState& A = ctx->adj;
    if (!A.ready) return ctx->fail(LBM_E_STATE, "lbm_adjoint_backward before lbm_adjoint_forward");
    if (dL_du0 && !A.has_u0) return ctx->fail(LBM_E_STATE, "dL/du0 needs a forward run starting from lbm_init_from_macros");
    int rc = ctx->activate();
    if (rc) return rc;
    const size_t N = ctx->N;
    const size_t pb = size_t(ctx->plane) * 9 * sizeof(float);
    const int K = A.every, T = A.nsteps;
    if (A.d_seg && A.seg_states != K) {  // checkpoint spacing changed since the last backward
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(A.d_seg);
        A.d_seg = nullptr;
    }
    if (!A.d_seg) {
        rc = ctx->dev_alloc(reinterpret_cast<void**>(&A.d_seg), pb * size_t(K), "adjoint segment states");
        if (rc) return rc;
        A.seg_states = K;
    }
    if (!A.d_mu[0]) {
        rc = ctx->dev_alloc(reinterpret_cast<void**>(&A.d_mu[0]), pb, "adjoint mu");
        rc = rc ? rc : ctx->dev_alloc(reinterpret_cast<void**>(&A.d_mu[1]), pb, "adjoint mu");
        rc = rc ? rc : ctx->dev_alloc(reinterpret_cast<void**>(&A.d_grad_force), N * 8, "adjoint dL/dF");
        if (rc) return rc;
    }
    const float* dldu = dL_du;
    float* gu0 = nullptr;
    if (!on_device) {
        rc = ctx->scratch(N * 16);
        if (rc) return rc;
        rc = ctx->upload(ctx->tmp, dL_du, N * 8, 0);
        if (rc) return rc;
        dldu = static_cast<const float*>(ctx->tmp);
        gu0 = static_cast<float*>(ctx->tmp) + 2 * N;
    } else {
        gu0 = dL_du0;
    }
    AdjParams p{};
    p.nx = ctx->cfg.nx;
    p.ny = ctx->cfg.ny;
    p.plane = ctx->plane;
    p.force = ctx->d_force_field;
    p.fx = ctx->cfg.force[0];
    p.fy = ctx->cfg.force[1];
    p.omega = ctx->cfg.omega;
    for (int k = 0; k < 9; k++) p.S[k] = ctx->cfg.S[k];
    p.grad_force = A.d_grad_force;
    p.dLdu = dldu;
    const dim3 nb = node_grid(ctx->cfg.nx, ctx->cfg.ny);
    int cur = 0;  // mu[cur] receives mu_t
    const int nseg = (T + K - 1) / K;
    float* seg = A.d_seg;
    const size_t pf = size_t(ctx->plane) * 9;
    for (int s = nseg - 1; s >= 0; s--) {
        const int t0 = s * K;                        // checkpoint state f*_{t0}
        const int len = (T - t0) < K ? (T - t0) : K; // adjoint steps t0+1 .. t0+len
        hipError_t e = hipMemcpyAsync(seg, A.checkpoints[size_t(s)], pb, hipMemcpyDeviceToDevice, ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "segment restore");
        // recompute f*_{t0+1} .. f*_{t0+len-1}: pairs of steps per launch, the launch also
        // storing its intermediate state; an odd remainder as one step
        const bool pairs = ctx->cfg.steps_per_launch != 1;
        for (int i = 1; i < len;) {
            StepParams sp = ctx->step_params(ctx->cur, false);
            sp.src = seg + size_t(i - 1) * pf;
            if (pairs && i + 1 < len) {
                sp.mid = seg + size_t(i) * pf;
                sp.dst = seg + size_t(i + 1) * pf;
                e = launch_step2(ctx->cfg.collision, LBM_F32, sp, step2_policy(ctx->plane, LBM_F32, ctx->cfg.collision), ctx->stream);
                i += 2;
            } else {
                sp.dst = seg + size_t(i) * pf;
                e = launch_step(ctx->cfg.collision, LBM_F32, 0, sp, step_policy(ctx->plane, LBM_F32), nullptr, ctx->stream);
                i += 1;
            }
            if (e != hipSuccess) return ctx->hip_fail(e, "segment recompute");
        }
        for (int t = t0 + len; t >= t0 + 1; t--) {
            p.state = seg + size_t(t - 1 - t0) * pf;
            p.mu_cur = A.d_mu[cur];
            p.mu_next = A.d_mu[1 - cur];
            if (t == T) hipLaunchKernelGGL(adjoint_seed_kernel, nb, dim3(kAdjBlock), 0, ctx->stream, p);
            else if (ctx->cfg.collision == LBM_BGK)
                hipLaunchKernelGGL(adjoint_step_kernel<LBM_BGK>, nb, dim3(kAdjBlock), 0, ctx->stream, p);
            else
                hipLaunchKernelGGL(adjoint_step_kernel<LBM_MRT>, nb, dim3(kAdjBlock), 0, ctx->stream, p);
            e = hipGetLastError();
            if (e != hipSuccess) return ctx->hip_fail(e, "adjoint step launch");
            cur = 1 - cur;
        }
    }
    // mu_1 is in mu[1 - cur]
    if (dL_du0) {
        p.mu_next = A.d_mu[1 - cur];
        p.rho0 = A.d_rho_init;
        p.u0x = A.d_u0;
        p.u0y = A.d_u0 + N;
        p.grad_u0 = gu0;
        hipLaunchKernelGGL(adjoint_u0_kernel, nb, dim3(kAdjBlock), 0, ctx->stream, p);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return ctx->hip_fail(e, "adjoint u0 launch");
    }
    hipError_t e;
    if (on_device) {
        if (dL_dforce) {
            e = hipMemcpyAsync(dL_dforce, A.d_grad_force, N * 8, hipMemcpyDeviceToDevice, ctx->stream);
            if (e != hipSuccess) return ctx->hip_fail(e, "gradient copy");
        }
        return LBM_OK;
    }
    if (dL_dforce) {
        e = hipMemcpyAsync(dL_dforce, A.d_grad_force, N * 8, hipMemcpyDeviceToHost, ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "gradient copy");
    }
    if (dL_du0) {
        e = hipMemcpyAsync(dL_du0, gu0, N * 8, hipMemcpyDeviceToHost, ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "gradient copy");
    }
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "hipStreamSynchronize");
    return LBM_OK;
}

}  // extern "C"
