// lbm_capi.hip — the extern "C" boundary of liblbm_hip.so (declared in include/lbm_hip.h).
#include "lbm_ctx.hpp"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace lbm;

namespace {

constexpr uint8_t kSupported[] = {LBM_FLUID,
                                  LBM_BOUNCE_BACK,
                                  LBM_ZOU_HE_TOP,
                                  LBM_ZOU_HE_LEFT,
                                  LBM_ZG_OUTFLOW,
                                  LBM_REGULARIZED_INLET_TOP,
                                  LBM_REGULARIZED_BOUNCE_BACK,
                                  LBM_REGULARIZED_BOUNCE_BACK_CORNER,
                                  LBM_MOVING_WALL_BOUNCE_BACK,
                                  LBM_CYLINDER,
                                  LBM_PRESSURE_OUTLET};

bool flag_supported(uint8_t f) {
    for (uint8_t s : kSupported)
        if (s == f) return true;
    return false;
}

size_t storage_bytes(int storage) { return storage == LBM_F32 ? 4 : 2; }

thread_local std::string g_create_err;

}  // namespace

int lbm_ctx::fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    err = buf;
    return code;
}

int lbm_ctx::hip_fail(hipError_t e, const char* what) {
    return fail(LBM_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

int lbm_ctx::activate() {
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    return LBM_OK;
}

int lbm_ctx::dev_alloc(void** p, size_t bytes, const char* what) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(LBM_E_NOMEM, "hipMalloc(%zu bytes) for %s failed: %s", bytes, what, hipGetErrorString(e));
    }
    return LBM_OK;
}

// Copy `bytes` from a host or device pointer to device memory `dst` on the stream.
int lbm_ctx::upload(void* dst, const void* src, size_t bytes, int on_device) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync (upload)");
    if (!on_device) {
        e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    }
    return LBM_OK;
}

int lbm_ctx::scratch(size_t bytes) {
    if (bytes <= tmp_bytes) return LBM_OK;
    if (tmp) {
        (void)hipStreamSynchronize(stream);
        (void)hipFree(tmp);
        tmp = nullptr;
        tmp_bytes = 0;
    }
    int rc = dev_alloc(&tmp, bytes, "staging buffer");
    if (rc) return rc;
    tmp_bytes = bytes;
    return LBM_OK;
}

int lbm_ctx::ensure_partials(long long n) {
    if (n <= partials_n) return LBM_OK;
    if (partials) {
        (void)hipStreamSynchronize(stream);
        (void)hipFree(partials);
        partials = nullptr;
    }
    int rc = dev_alloc(reinterpret_cast<void**>(&partials), size_t(n) * 3 * sizeof(double), "reduction partials");
    if (rc) return rc;
    partials_n = n;
    return LBM_OK;
}

StepParams lbm_ctx::step_params(int src_idx, bool write_macros) const {
    StepParams p{};
    p.nx = cfg.nx;
    p.ny = cfg.ny;
    p.px = cfg.periodic_x;
    p.py = cfg.periodic_y;
    p.plane = plane;
    p.src = buf[src_idx];
    p.dst = buf[1 - src_idx];
    p.flags = d_flags;
    p.row_class = d_row_class;
    p.omega = cfg.omega;
    for (int k = 0; k < 9; k++) p.S[k] = cfg.S[k];
    p.rho0 = cfg.rho0;
    p.fx = cfg.force[0];
    p.fy = cfg.force[1];
    p.force_field = d_force_field;
    p.ibm_force = ibm.active ? ibm.d_force_total : nullptr;
    p.ibm_x0 = ibm.x0;
    p.ibm_x1 = ibm.x1;
    p.ibm_y0 = ibm.y0;
    p.ibm_y1 = ibm.y1;
    p.forced = (cfg.force[0] != 0.0f || cfg.force[1] != 0.0f || d_force_field || ibm.active) ? 1 : 0;
    p.wall_ux = cfg.wall_u[0];
    p.wall_uy = cfg.wall_u[1];
    p.u_lid = cfg.u_lid;
    p.quirks = cfg.quirks;
    p.avg = (cfg.collision == LBM_CM && cfg.adapter == LBM_OPTIMAL_ADAPTER) ? d_avg : nullptr;
    p.out_rho = write_macros ? d_rho : nullptr;
    p.out_ux = write_macros ? d_ux : nullptr;
    p.out_uy = write_macros ? d_uy : nullptr;
    return p;  // ntx and nt are set by launch_step from the launch policy
}

int lbm_ctx::adapter_pass(int src_idx) {
    StepParams p = step_params(src_idx, false);
    LaunchPolicy lp = step_policy(plane, cfg.storage);
    long long nb = step_grid_blocks(p, lp);
    int rc = ensure_partials(nb);
    if (rc) return rc;
    hipError_t e = launch_step(cfg.collision, cfg.storage, 1, p, lp, partials, stream);
    if (e != hipSuccess) return hip_fail(e, "moments pass launch");
    e = launch_finalize_avg(partials, nb, (long long)N, d_avg, d_avg64, stream);
    if (e != hipSuccess) return hip_fail(e, "finalize launch");
    avg_valid = true;
    return LBM_OK;
}

// CM<OptimalAdapter> with same-step averages (the reference's order: the averages of a
// step's post-stream state feed that step's collision) in one launch per step: the
// look-ahead strip kernel collides step t+1 with averages computed by the launch before and
// sums the moments of step t+2 on the way. Needs the strip machinery's conditions: local
// boundary functors, no stale-rest quirk, no IBM body.
bool lbm_ctx::ahead_ok() const {
    const char* v = std::getenv("LBM_ADAPTER_LOOKAHEAD");
    if (v && v[0] == '0') return false;
    return cfg.collision == LBM_CM && cfg.adapter == LBM_OPTIMAL_ADAPTER && cfg.adapter_mode == LBM_ADAPTER_SAME_STEP &&
           flags_local && !ibm.active && !(cfg.quirks & LBM_QUIRK_STALE_REST) && !edge_in_store;
}

int lbm_ctx::ahead_step(bool write_macros) {
    hipError_t e;
    int rc;
    if (!ahead_valid) {  // this step's averages from a moments pass over its post-stream state
        StepParams p = step_params(cur, false);
        LaunchPolicy lp = step_policy(plane, cfg.storage);
        long long nb = step_grid_blocks(p, lp);
        rc = ensure_partials(nb);
        if (rc) return rc;
        e = launch_step(cfg.collision, cfg.storage, 1, p, lp, partials, stream);
        if (e != hipSuccess) return hip_fail(e, "moments pass launch");
        e = launch_finalize_avg(partials, nb, (long long)N, d_ahead, d_ahead64, stream);
        if (e != hipSuccess) return hip_fail(e, "finalize launch");
    }
    if (n_edge > 0 && !ahead_edges) {  // state t-1 at the edge nodes: buf[1-cur] -> store
        e = launch_edge_copy(cfg.storage, buf[1 - cur], d_edge[edge_cur], plane, cfg.nx, cfg.ny, cfg.periodic_x,
                             cfg.periodic_y, 1, stream);
        if (e != hipSuccess) return hip_fail(e, "edge save launch");
    }
    StepParams p = step_params(cur, write_macros);
    p.mid = p.dst;  // phase A's own nodes are the launch's output
    p.avg = d_ahead;
    p.edge_prev = d_edge[edge_cur];
    p.edge_next = d_edge[1 - edge_cur];
    p.n_edge = n_edge;
    const Step2Policy lp = ahead_policy(plane, cfg.storage);
    const long long np = ahead_partials_needed(p, lp);
    rc = ensure_partials(np);
    if (rc) return rc;
    e = launch_step_ahead(cfg.collision, cfg.storage, p, lp, partials, stream);
    if (e != hipSuccess) return hip_fail(e, "look-ahead step launch");
    e = launch_finalize_ahead(partials, np, (long long)N, d_ahead, d_ahead64, d_avg, d_avg64, stream);
    if (e != hipSuccess) return hip_fail(e, "finalize launch");
    if (n_edge > 0) edge_cur = 1 - edge_cur;
    ahead_edges = n_edge > 0;
    ahead_valid = true;
    avg_valid = true;
    cur = 1 - cur;
    timestep++;
    if (write_macros) macros_valid = true;
    return LBM_OK;
}

// After a two-step launch, put the edge nodes' previous state back into buf[1-cur], where
// one-step launches (and the IBM gather) read unstreamed slots.
int lbm_ctx::restore_edges() {
    if (!edge_in_store) return LBM_OK;
    hipError_t e = launch_edge_copy(cfg.storage, buf[1 - cur], d_edge[edge_cur], plane, cfg.nx, cfg.ny, cfg.periodic_x,
                                    cfg.periodic_y, 0, stream);
    if (e != hipSuccess) return hip_fail(e, "edge restore launch");
    edge_in_store = false;
    return LBM_OK;
}

bool lbm_ctx::two_step_ok() const {
    const bool adapt = cfg.collision == LBM_CM && cfg.adapter == LBM_OPTIMAL_ADAPTER;
    return cfg.steps_per_launch != 1 && !adapt && !ibm.active && !(cfg.quirks & LBM_QUIRK_STALE_REST) && flags_local;
}

// Two steps from buf[cur] into buf[1-cur] in one launch (the same arithmetic as two
// one-step launches).
int lbm_ctx::two_step(bool write_macros) {
    hipError_t e;
    ahead_valid = ahead_edges = false;
    if (n_edge > 0 && !edge_in_store) {  // state t-1 at the edge nodes: buf[1-cur] -> store
        e = launch_edge_copy(cfg.storage, buf[1 - cur], d_edge[edge_cur], plane, cfg.nx, cfg.ny, cfg.periodic_x,
                             cfg.periodic_y, 1, stream);
        if (e != hipSuccess) return hip_fail(e, "edge save launch");
    }
    StepParams p = step_params(cur, write_macros);
    p.edge_prev = d_edge[edge_cur];
    p.edge_next = d_edge[1 - edge_cur];
    p.n_edge = n_edge;
    e = launch_step2(cfg.collision, cfg.storage, p, step2_policy(plane, cfg.storage, cfg.collision), stream);
    if (e != hipSuccess) return hip_fail(e, "two-step launch");
    if (n_edge > 0) {
        edge_cur = 1 - edge_cur;
        edge_in_store = true;
    }
    cur = 1 - cur;
    timestep += 2;
    if (write_macros) macros_valid = true;
    return LBM_OK;
}

// One step of an IBM run whose box avoids non-periodic edges, in two concurrent parts: the
// step outside the IBM box on aux_stream (it reads neither the IBM force nor anything the
// IBM pass writes) and, on `stream`, the IBM pass followed by the box's step. Both parts
// read buf[cur] and write disjoint nodes of buf[1-cur]; `stream` waits for aux_stream
// before anything later runs.
int lbm_ctx::split_ibm_step(bool write_macros) {
    hipError_t e;
    if (!aux_stream) {
        e = hipStreamCreateWithFlags(&aux_stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev_join, hipEventDisableTiming);
        if (e != hipSuccess) return hip_fail(e, "IBM stream set-up");
    }
    StepParams p = step_params(cur, write_macros);
    LaunchPolicy lp = step_policy(plane, cfg.storage);
    // The IBM pass (one workgroup) is submitted before the outside step, so that it has its
    // compute unit before the step's workgroups fill the device.
    e = hipEventRecord(ev_fork, stream);
    if (e != hipSuccess) return hip_fail(e, "IBM fork");
    int rc = ibm_pass(cur);
    if (rc) return rc;
    e = hipStreamWaitEvent(aux_stream, ev_fork, 0);
    if (e != hipSuccess) return hip_fail(e, "IBM fork");
    p.region = 1;
    e = launch_step(cfg.collision, cfg.storage, 0, p, lp, partials, aux_stream);
    if (e != hipSuccess) return hip_fail(e, "step launch (outside the IBM box)");
    e = hipEventRecord(ev_join, aux_stream);
    if (e != hipSuccess) return hip_fail(e, "IBM join");
    p.region = 2;
    e = launch_step(cfg.collision, cfg.storage, 0, p, lp, partials, stream);
    if (e != hipSuccess) {
        (void)hipStreamWaitEvent(stream, ev_join, 0);
        return hip_fail(e, "step launch (IBM box)");
    }
    e = hipStreamWaitEvent(stream, ev_join, 0);
    if (e != hipSuccess) return hip_fail(e, "IBM join");
    cur = 1 - cur;
    timestep++;
    if (write_macros) macros_valid = true;
    return LBM_OK;
}

void lbm_ctx::ibm_graph_free() {
    if (ibm_graph) (void)hipGraphExecDestroy(ibm_graph);
    ibm_graph = nullptr;
    ibm_graph_cur = -1;
    ibm_graph_stream = nullptr;
}

// Graph replay applies once a direct split step has run (the box's external force written,
// the internal stream created) and on a stream that can be captured (not the null stream).
// LBM_HIP_GRAPHS=0 in the environment turns it off.
bool lbm_ctx::ibm_graph_ok() const {
    const char* v = std::getenv("LBM_HIP_GRAPHS");
    const bool enabled = !(v && v[0] == '0');
    const bool adapt = cfg.collision == LBM_CM && cfg.adapter == LBM_OPTIMAL_ADAPTER;
    return enabled && ibm.active && ibm.split_ok && !adapt && ibm.box_valid && aux_stream && stream &&
           !edge_in_store && !ibm_fused_ok();
}

// Two split IBM steps through the graph, capturing it first when there is none for this
// buffer parity and stream. Leaves cur as it was (two steps) and advances timestep by 2.
int lbm_ctx::ibm_graph_pair() {
    if (!ibm_graph || ibm_graph_cur != cur || ibm_graph_stream != stream) {
        ibm_graph_free();
        const int c0 = cur;
        const int64_t t0 = timestep;
        hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed);
        if (e != hipSuccess) return hip_fail(e, "IBM step capture");
        int rc = split_ibm_step(false);
        if (!rc) rc = split_ibm_step(false);
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(stream, &g);
        cur = c0;  // the capture recorded the two steps without running them
        timestep = t0;
        if (rc || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            (void)hipGetLastError();
            return rc ? rc : hip_fail(e, "IBM step capture");
        }
        e = hipGraphInstantiate(&ibm_graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            ibm_graph = nullptr;
            return hip_fail(e, "IBM graph instantiate");
        }
        ibm_graph_cur = c0;
        ibm_graph_stream = stream;
    }
    hipError_t e = hipGraphLaunch(ibm_graph, stream);
    if (e != hipSuccess) return hip_fail(e, "IBM graph launch");
    timestep += 2;
    return LBM_OK;
}

// One step from buf[cur] into buf[1-cur].
int lbm_ctx::one_step(bool write_macros) {
    int rc = restore_edges();
    if (rc) return rc;
    if (ahead_ok()) return ahead_step(write_macros);
    ahead_valid = ahead_edges = false;
    bool adapt = cfg.collision == LBM_CM && cfg.adapter == LBM_OPTIMAL_ADAPTER;
    if (ibm.active && ibm.split_ok && !adapt) return ibm_fused_ok() ? fused_ibm_step(write_macros) : split_ibm_step(write_macros);
    if (ibm.active) {
        rc = ibm_pass(cur);
        if (rc) return rc;
    }
    bool lagged = adapt && cfg.adapter_mode == LBM_ADAPTER_LAGGED;
    if (adapt && (!lagged || !avg_valid)) {
        rc = adapter_pass(cur);
        if (rc) return rc;
    }
    StepParams p = step_params(cur, write_macros);
    LaunchPolicy lp = step_policy(plane, cfg.storage);
    int mode = lagged ? 2 : 0;
    long long nb = step_grid_blocks(p, lp);
    if (mode == 2) {
        rc = ensure_partials(nb);
        if (rc) return rc;
    }
    hipError_t e = launch_step(cfg.collision, cfg.storage, mode, p, lp, partials, stream);
    if (e != hipSuccess) return hip_fail(e, "step launch");
    if (mode == 2) {
        e = launch_finalize_avg(partials, nb, (long long)N, d_avg, d_avg64, stream);
        if (e != hipSuccess) return hip_fail(e, "finalize launch");
    }
    cur = 1 - cur;
    timestep++;
    if (write_macros) macros_valid = true;
    return LBM_OK;
}

lbm_ctx::~lbm_ctx() {
    (void)hipSetDevice(device);
    if (own_stream) (void)hipStreamSynchronize(own_stream);
    if (stream != own_stream) (void)hipStreamSynchronize(stream);
    if (aux_stream) (void)hipStreamSynchronize(aux_stream);
    void* ptrs[] = {buf[0],  buf[1], d_flags, d_row_class, d_force_field, d_rho,     d_ux,      d_uy,
                    d_avg,   d_avg64, d_ahead, d_ahead64, partials,     tmp,       d_edge[0], d_edge[1]};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    ibm_graph_free();
    ibm_free();
    adjoint_free();
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (aux_stream) (void)hipStreamDestroy(aux_stream);
    if (own_stream) (void)hipStreamDestroy(own_stream);
}

// =====================================================================================
extern "C" {

const char* lbm_version(void) { return "lbm-mi355x 0.1 (gfx950)"; }

// With a NULL context: the reason the calling thread's last lbm_create failed.
const char* lbm_last_error(const lbm_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }

int lbm_config_default(lbm_config* cfg, int32_t nx, int32_t ny) {
    if (!cfg) return LBM_E_INVALID;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->nx = nx;
    cfg->ny = ny;
    cfg->collision = LBM_BGK;
    cfg->adapter = LBM_NO_ADAPTER;
    cfg->adapter_mode = LBM_ADAPTER_SAME_STEP;
    float tau = 3.0f * (1.0f / 6.0f) + 0.5f;  // viscosity_to_tau(1/6), scenario.cuh:36-38
    cfg->omega = 1.0f / tau;
    const float S[9] = {0.0f, cfg->omega, cfg->omega, 0.0f, cfg->omega, 0.0f, cfg->omega, cfg->omega, cfg->omega};
    for (int k = 0; k < 9; k++) cfg->S[k] = S[k];
    cfg->rho0 = 1.0f;
    cfg->storage = LBM_F32;
    cfg->u_lid = 0.1f;  // ScenarioTrait::u_max (scenario.cuh:59)
    return LBM_OK;
}

int lbm_create(const lbm_config* cfg, lbm_ctx** out) {
    if (!cfg || !out) return LBM_E_INVALID;
    *out = nullptr;
    auto* c = new lbm_ctx();
    c->cfg = *cfg;
    hipError_t e = hipGetDevice(&c->device);
    if (e != hipSuccess) {
        c->hip_fail(e, "hipGetDevice");
        g_create_err = c->err;
        delete c;
        return LBM_E_HIP;
    }
    const lbm_config& k = *cfg;
    auto bad = [&](const char* m) {
        delete c;
        g_create_err = m;
        return LBM_E_INVALID;
    };
    if (k.nx < 3 || k.ny < 3) return bad("grid must be at least 3x3");
    if ((long long)k.nx * k.ny > (1LL << 31) - 1) return bad("grid exceeds 2^31 nodes");
    if (k.collision < LBM_BGK || k.collision > LBM_CM) return bad("unknown collision operator");
    if (k.adapter != LBM_NO_ADAPTER && !(k.adapter == LBM_OPTIMAL_ADAPTER && k.collision == LBM_CM))
        return bad("adapters apply to CM only");
    if (k.adapter_mode != LBM_ADAPTER_SAME_STEP && k.adapter_mode != LBM_ADAPTER_LAGGED) return bad("bad adapter_mode");
    if (k.storage < LBM_F32 || k.storage > LBM_BF16_SHIFTED) return bad("unknown storage");
    // init.cuh:62-64: viscosity > 0 and tau > 0.5, i.e. 0 < omega < 2; u_max < 0.5.
    if (k.collision == LBM_BGK && !(k.omega > 0.0f && k.omega < 2.0f)) return bad("BGK needs 0 < omega < 2 (tau > 0.5)");
    for (int i = 0; i < 9; i++)
        if (!std::isfinite(k.S[i])) return bad("relaxation rates must be finite");
    if (!(std::fabs(k.u_lid) < 0.5f) || !(std::fabs(k.wall_u[0]) < 0.5f) || !(std::fabs(k.wall_u[1]) < 0.5f))
        return bad("boundary velocities must stay below 0.5 (init.cuh:64)");
    if (!(k.rho0 > 0.0f)) return bad("rho0 must be positive");
    if (!std::isfinite(k.force[0]) || !std::isfinite(k.force[1])) return bad("force must be finite");

    c->N = size_t(k.nx) * size_t(k.ny);
    c->plane = (long long)((c->N + 63) / 64 * 64);
    e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        c->hip_fail(e, "hipStreamCreate");
        g_create_err = c->err;
        delete c;
        return LBM_E_HIP;
    }
    c->stream = c->own_stream;
    size_t pb = size_t(c->plane) * 9 * storage_bytes(k.storage);
    int rc = 0;
    rc = rc ? rc : c->dev_alloc(&c->buf[0], pb, "populations A");
    rc = rc ? rc : c->dev_alloc(&c->buf[1], pb, "populations B");
    rc = rc ? rc : c->dev_alloc(reinterpret_cast<void**>(&c->d_rho), c->N * 4, "rho");
    rc = rc ? rc : c->dev_alloc(reinterpret_cast<void**>(&c->d_ux), c->N * 4, "ux");
    rc = rc ? rc : c->dev_alloc(reinterpret_cast<void**>(&c->d_uy), c->N * 4, "uy");
    rc = rc ? rc : c->dev_alloc(reinterpret_cast<void**>(&c->d_avg), 4 * sizeof(float), "adapter averages");
    rc = rc ? rc : c->dev_alloc(reinterpret_cast<void**>(&c->d_avg64), 4 * sizeof(double), "adapter averages");
    rc = rc ? rc : c->dev_alloc(reinterpret_cast<void**>(&c->d_ahead), 4 * sizeof(float), "adapter averages");
    rc = rc ? rc : c->dev_alloc(reinterpret_cast<void**>(&c->d_ahead64), 4 * sizeof(double), "adapter averages");
    c->n_edge = edge_count(k.nx, k.ny, k.periodic_x, k.periodic_y);
    if (c->n_edge > 0) {
        const size_t eb = size_t(c->n_edge) * 9 * storage_bytes(k.storage);
        rc = rc ? rc : c->dev_alloc(&c->d_edge[0], eb, "edge store");
        rc = rc ? rc : c->dev_alloc(&c->d_edge[1], eb, "edge store");
    }
    if (rc) {
        g_create_err = c->err;
        delete c;
        return rc;
    }
    *out = c;
    return LBM_OK;
}

int lbm_destroy(lbm_ctx* ctx) {
    delete ctx;
    return LBM_OK;
}

int lbm_set_stream(lbm_ctx* ctx, void* s) {
    if (!ctx) return LBM_E_INVALID;
    if (static_cast<hipStream_t>(s) != ctx->stream) ctx->ibm_graph_free();
    // NULL is the HIP null stream (PyTorch's default stream reports handle 0), not a request
    // for the context's own stream: mapping it to the own non-blocking stream would let the
    // library race with the caller's work.
    ctx->stream = static_cast<hipStream_t>(s);
    return LBM_OK;
}

int lbm_set_flags(lbm_ctx* ctx, const uint8_t* flags, int on_device) {
    if (!ctx || !flags) return LBM_E_INVALID;
    ctx->ibm_graph_free();  // the captured IBM steps read the flags and the force field
    ctx->adj.ready = false;  // the adjoint's recorded forward no longer matches the state
    ctx->ahead_valid = ctx->ahead_edges = false;
    int rc = ctx->activate();
    if (rc) return rc;
    const int nx = ctx->cfg.nx, ny = ctx->cfg.ny;
    std::vector<uint8_t> h(ctx->N);
    if (on_device) {
        hipError_t e = hipMemcpy(h.data(), flags, ctx->N, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return ctx->hip_fail(e, "hipMemcpy (flags)");
    } else {
        std::memcpy(h.data(), flags, ctx->N);
    }
    std::vector<uint8_t> rowc(size_t(ny), 0);
    bool any = false, local = true;
    for (int y = 0; y < ny; y++) {
        for (int x = 0; x < nx; x++) {
            uint8_t f = h[size_t(y) * nx + x];
            if (!flag_supported(f))
                return ctx->fail(LBM_E_UNSUPPORTED, "BC flag %d at node (%d,%d) is not supported", int(f), x, y);
            if (f == LBM_FLUID) continue;
            any = true;
            if (flag_reads_neighbour(f)) local = false;
            if (f == LBM_PRESSURE_OUTLET) {
                // reads the post-stream populations of (x - 1, y); as for ZG_OUTFLOW, none of
                // them may be an unstreamed slot (another thread writes that slot)
                if (x == 0 && !ctx->cfg.periodic_x)
                    return ctx->fail(LBM_E_UNSUPPORTED, "PRESSURE_OUTLET at (%d,%d) has no node to its left", x, y);
                const int ix = x == 0 ? nx - 1 : x - 1;
                for (int i = 0; i < 9; i++) {
                    int xs = ix - CX[i], ys = y - CY[i];
                    bool okx = (xs >= 0 && xs < nx) || ctx->cfg.periodic_x;
                    bool oky = (ys >= 0 && ys < ny) || ctx->cfg.periodic_y;
                    if (!okx || !oky)
                        return ctx->fail(LBM_E_UNSUPPORTED,
                                         "PRESSURE_OUTLET at (%d,%d): the node to its left has an unstreamed population", x, y);
                }
            }
            if (f == LBM_ZG_OUTFLOW) {
                // The interior neighbour's unknown populations must all stream from inside
                // the domain: a stale slot of another node would be raced by its writer.
                int n0 = 0, n1 = 0, ix = x, iy = y;
                if (x == 0) { n0 = 1; ix = 1; }
                else if (x == nx - 1) { n0 = -1; ix = nx - 2; }
                else if (y == 0) { n1 = 1; iy = 1; }
                else if (y == ny - 1) { n1 = -1; iy = ny - 2; }
                else continue;
                for (int i = 0; i < 9; i++) {
                    if (CX[i] * n0 + CY[i] * n1 <= 0) continue;
                    int xs = ix - CX[i], ys = iy - CY[i];
                    bool okx = (xs >= 0 && xs < nx) || ctx->cfg.periodic_x;
                    bool oky = (ys >= 0 && ys < ny) || ctx->cfg.periodic_y;
                    if (!okx || !oky)
                        return ctx->fail(LBM_E_UNSUPPORTED,
                                         "ZG_OUTFLOW at (%d,%d): its interior neighbour has an unstreamed population", x, y);
                }
            }
            if (f == LBM_REGULARIZED_BOUNCE_BACK_CORNER && (ctx->cfg.quirks & LBM_QUIRK_STALE_REST))
                return ctx->fail(LBM_E_UNSUPPORTED, "REGULARIZED_BOUNCE_BACK_CORNER with LBM_QUIRK_STALE_REST");
        }
    }
    if (ctx->d_flags) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(ctx->d_flags);
        (void)hipFree(ctx->d_row_class);
        ctx->d_flags = nullptr;
        ctx->d_row_class = nullptr;
    }
    ctx->flags_local = local;
    if (!any) return LBM_OK;
    row_classes(h.data(), nx, ny, rowc.data());
    rc = ctx->dev_alloc(reinterpret_cast<void**>(&ctx->d_flags), ctx->N, "flags");
    if (rc) return rc;
    rc = ctx->dev_alloc(reinterpret_cast<void**>(&ctx->d_row_class), size_t(ny), "row classes");
    if (rc) return rc;
    rc = ctx->upload(ctx->d_flags, h.data(), ctx->N, 0);
    if (rc) return rc;
    return ctx->upload(ctx->d_row_class, rowc.data(), size_t(ny), 0);
}

int lbm_init_from_macros(lbm_ctx* ctx, const float* rho, const float* u, int on_device) {
    if (!ctx || !rho || !u) return LBM_E_INVALID;
    ctx->adj.ready = false;  // the adjoint's recorded forward no longer matches the state
    ctx->ahead_valid = ctx->ahead_edges = false;
    int rc = ctx->activate();
    if (rc) return rc;
    const float *drho = rho, *du = u;
    if (!on_device) {
        rc = ctx->scratch(ctx->N * 12);
        if (rc) return rc;
        float* t = static_cast<float*>(ctx->tmp);
        rc = ctx->upload(t, rho, ctx->N * 4, 0);
        if (rc) return rc;
        rc = ctx->upload(t + ctx->N, u, ctx->N * 8, 0);
        if (rc) return rc;
        drho = t;
        du = t + ctx->N;
    }
    InitParams ip{drho, du, ctx->buf[0], ctx->buf[1], ctx->plane, (long long)ctx->N, ctx->cfg.rho0};
    hipError_t e = launch_init(ctx->cfg.storage, ip, ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "init launch");
    e = hipMemcpyAsync(ctx->d_rho, drho, ctx->N * 4, hipMemcpyDeviceToDevice, ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "rho copy");
    e = launch_aos_to_soa(du, (long long)ctx->N, ctx->d_ux, ctx->d_uy, ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "u split");
    if (!on_device) {
        e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "hipStreamSynchronize");
    }
    ctx->cur = 0;
    ctx->edge_in_store = false;
    ctx->timestep = 0;
    ctx->initialized = true;
    ctx->macros_valid = true;
    ctx->avg_valid = false;
    ctx->init_from_macros = true;
    return LBM_OK;
}

int lbm_set_force_field(lbm_ctx* ctx, const float* force, int on_device) {
    if (!ctx) return LBM_E_INVALID;
    ctx->ibm_graph_free();  // the captured IBM steps read the flags and the force field
    ctx->adj.ready = false;  // the adjoint's recorded forward no longer matches the state
    ctx->ahead_valid = ctx->ahead_edges = false;
    int rc = ctx->activate();
    if (rc) return rc;
    ctx->ibm.box_valid = false;  // the IBM box's external force changes with the field
    if (!force) {
        if (ctx->d_force_field) {
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipFree(ctx->d_force_field);
            ctx->d_force_field = nullptr;
        }
        return LBM_OK;
    }
    if (!ctx->d_force_field) {
        rc = ctx->dev_alloc(reinterpret_cast<void**>(&ctx->d_force_field), ctx->N * 8, "force field");
        if (rc) return rc;
    }
    return ctx->upload(ctx->d_force_field, force, ctx->N * 8, on_device);
}

int lbm_step(lbm_ctx* ctx, int32_t nsteps) {
    if (!ctx || nsteps < 0) return LBM_E_INVALID;
    ctx->adj.ready = false;  // the adjoint's recorded forward no longer matches the state
    if (!ctx->initialized) return ctx->fail(LBM_E_STATE, "lbm_step before lbm_init_from_macros / lbm_set_pops");
    int rc = ctx->activate();
    if (rc) return rc;
    int i = 0;
    if (ctx->two_step_ok())
        for (; i + 2 <= nsteps; i += 2) {
            rc = ctx->two_step(i + 2 == nsteps);
            if (rc) return rc;
        }
    for (; i < nsteps; i++) {
        // IBM split steps: pairs by graph replay, the last step direct (it stores the fields)
        if (i + 3 <= nsteps && ctx->ibm_graph_ok()) {
            rc = ctx->ibm_graph_pair();
            if (rc) return rc;
            i++;
            continue;
        }
        rc = ctx->one_step(i == nsteps - 1);
        if (rc) return rc;
    }
    return LBM_OK;
}

int lbm_sync(lbm_ctx* ctx) {
    if (!ctx) return LBM_E_INVALID;
    int rc = ctx->activate();
    if (rc) return rc;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "hipStreamSynchronize");
    return LBM_OK;
}

int lbm_get_timestep(const lbm_ctx* ctx, int64_t* t) {
    if (!ctx || !t) return LBM_E_INVALID;
    *t = ctx->timestep;
    return LBM_OK;
}

int lbm_get_steps_per_launch(const lbm_ctx* ctx, int32_t* steps) {
    if (!ctx || !steps) return LBM_E_INVALID;
    *steps = ctx->two_step_ok() ? 2 : 1;
    return LBM_OK;
}

int lbm_get_macros(lbm_ctx* ctx, float* rho, float* u, int on_device) {
    if (!ctx) return LBM_E_INVALID;
    if (!ctx->macros_valid) return ctx->fail(LBM_E_STATE, "no macroscopic fields stored yet");
    int rc = ctx->activate();
    if (rc) return rc;
    hipError_t e;
    if (on_device) {
        if (rho) {
            e = hipMemcpyAsync(rho, ctx->d_rho, ctx->N * 4, hipMemcpyDeviceToDevice, ctx->stream);
            if (e != hipSuccess) return ctx->hip_fail(e, "rho copy");
        }
        if (u) {
            e = launch_macros_to_aos(ctx->d_ux, ctx->d_uy, (long long)ctx->N, u, ctx->stream);
            if (e != hipSuccess) return ctx->hip_fail(e, "u interleave");
        }
        return LBM_OK;
    }
    if (rho) {
        e = hipMemcpyAsync(rho, ctx->d_rho, ctx->N * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "rho copy");
    }
    if (u) {
        rc = ctx->scratch(ctx->N * 8);
        if (rc) return rc;
        e = launch_macros_to_aos(ctx->d_ux, ctx->d_uy, (long long)ctx->N, static_cast<float*>(ctx->tmp), ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "u interleave");
        e = hipMemcpyAsync(u, ctx->tmp, ctx->N * 8, hipMemcpyDeviceToHost, ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "u copy");
    }
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "hipStreamSynchronize");
    return LBM_OK;
}

int lbm_get_pops(lbm_ctx* ctx, float* f, int on_device) {
    if (!ctx || !f) return LBM_E_INVALID;
    if (!ctx->initialized) return ctx->fail(LBM_E_STATE, "no state yet");
    int rc = ctx->activate();
    if (rc) return rc;
    float* out = f;
    if (!on_device) {
        rc = ctx->scratch(ctx->N * 36);
        if (rc) return rc;
        out = static_cast<float*>(ctx->tmp);
    }
    hipError_t e = launch_pops_to_f32(ctx->cfg.storage, ctx->buf[ctx->cur], ctx->plane, (long long)ctx->N, ctx->cfg.rho0, out, ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "pops convert");
    if (!on_device) {
        e = hipMemcpyAsync(f, out, ctx->N * 36, hipMemcpyDeviceToHost, ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "pops copy");
        e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "hipStreamSynchronize");
    }
    return LBM_OK;
}

int lbm_set_pops(lbm_ctx* ctx, const float* f, int on_device) {
    if (!ctx || !f) return LBM_E_INVALID;
    ctx->adj.ready = false;  // the adjoint's recorded forward no longer matches the state
    ctx->ahead_valid = ctx->ahead_edges = false;
    int rc = ctx->activate();
    if (rc) return rc;
    const float* in = f;
    if (!on_device) {
        rc = ctx->scratch(ctx->N * 36);
        if (rc) return rc;
        rc = ctx->upload(ctx->tmp, f, ctx->N * 36, 0);
        if (rc) return rc;
        in = static_cast<const float*>(ctx->tmp);
    }
    // Both buffers get the state (the stale slots of the first step read it, as after init).
    hipError_t e = launch_pops_from_f32(ctx->cfg.storage, ctx->buf[0], ctx->buf[1], ctx->plane, (long long)ctx->N, ctx->cfg.rho0, in, ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "pops convert");
    if (!on_device) {
        e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "hipStreamSynchronize");
    }
    ctx->cur = 0;
    ctx->edge_in_store = false;
    ctx->initialized = true;
    ctx->macros_valid = false;
    ctx->avg_valid = false;
    ctx->init_from_macros = false;
    return LBM_OK;
}

int lbm_reduce(lbm_ctx* ctx, int32_t kind, const float* ref, double* out) {
    if (!ctx || !out) return LBM_E_INVALID;
    if (!ctx->macros_valid) return ctx->fail(LBM_E_STATE, "no macroscopic fields stored yet");
    int rc = ctx->activate();
    if (rc) return rc;
    hipError_t e;
    const long long N = (long long)ctx->N;
    if (kind == LBM_REDUCE_MOMENT_AVG) {
        if (!ctx->avg_valid) return ctx->fail(LBM_E_STATE, "no adapter averages computed yet");
        e = hipMemcpyAsync(out, ctx->d_avg64, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "avg copy");
        e = hipStreamSynchronize(ctx->stream);
        return e == hipSuccess ? LBM_OK : ctx->hip_fail(e, "hipStreamSynchronize");
    }
    if (kind == LBM_REDUCE_ROW_MEAN_UX) {
        rc = ctx->scratch(size_t(ctx->cfg.ny) * sizeof(double));
        if (rc) return rc;
        e = launch_row_mean_ux(ctx->d_ux, ctx->cfg.nx, ctx->cfg.ny, static_cast<double*>(ctx->tmp), ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "row mean launch");
        e = hipMemcpyAsync(out, ctx->tmp, size_t(ctx->cfg.ny) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "row mean copy");
        e = hipStreamSynchronize(ctx->stream);
        return e == hipSuccess ? LBM_OK : ctx->hip_fail(e, "hipStreamSynchronize");
    }
    if (kind != LBM_REDUCE_MASS && kind != LBM_REDUCE_MOMENTUM && kind != LBM_REDUCE_L2_VS_REF)
        return ctx->fail(LBM_E_INVALID, "unknown reduction kind %d", int(kind));
    if (kind == LBM_REDUCE_L2_VS_REF && !ref) return ctx->fail(LBM_E_INVALID, "L2_VS_REF needs a reference field");
    long long np = reduce_partials_needed(N);
    size_t need = size_t(np) * 3 * sizeof(double) + 4 * sizeof(double) + (ref ? size_t(N) * 8 : 0);
    rc = ctx->scratch(need);
    if (rc) return rc;
    double* part = static_cast<double*>(ctx->tmp);
    double* o3 = part + 3 * np;
    float* dref = nullptr;
    if (kind == LBM_REDUCE_L2_VS_REF) {
        dref = reinterpret_cast<float*>(o3 + 4);
        rc = ctx->upload(dref, ref, size_t(N) * 8, 0);
        if (rc) return rc;
    }
    e = launch_reduce(kind, ctx->d_rho, ctx->d_ux, ctx->d_uy, dref, N, part, o3, ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "reduce launch");
    double h[3];
    e = hipMemcpyAsync(h, o3, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "reduce copy");
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "hipStreamSynchronize");
    out[0] = h[0];
    if (kind != LBM_REDUCE_MASS) out[1] = h[1];
    return LBM_OK;
}

}  // extern "C"
