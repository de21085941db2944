// lbm_ctx.hpp — the context behind the opaque `lbm_ctx*` of the C ABI.
//
// One context = one solver instance: the two population buffers (pull scheme, A/B), the
// boundary flags, the macroscopic outputs of the last step and the optional IBM and
// adjoint state, all device memory of the GPU that was current at lbm_create, and one
// stream. Mirrors what `LBM<dim>` owns in the reference (lbm.cuh:36-43, 85-90) without
// process-global symbols.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/lbm_hip.h"
#include "lbm_kernels.hpp"

// Immersed-boundary state (IBMManager<2>, IBMManager.cuh:31-51), rebuilt for a band of
// Eulerian nodes: the union of the markers' stencils.
struct IbmState {
    bool active = false;
    int n_markers = 0, n_band = 0, iters = 3;
    int x0 = 0, x1 = 0, y0 = 0, y1 = 0;     // box holding the band (exclusive ends)
    float* d_force_total = nullptr;        // AoS [N][2]: total node force inside the box
    // device tables
    float* d_xy = nullptr;                 // [n][2]
    float* d_ds = nullptr;                 // [n]
    int* d_band_node = nullptr;            // [n_band] global node index
    int* d_mk_ptr = nullptr;               // [n+1] marker -> stencil entries
    int* d_mk_band = nullptr;              // band index per entry
    float* d_mk_w = nullptr;               // kernel weight per entry
    int* d_bn_ptr = nullptr;               // [n_band+1] band node -> (marker, weight) entries
    int* d_bn_mk = nullptr;
    float* d_bn_w = nullptr;
    int* d_box_map = nullptr;              // band index or -1 per cell of the box
    float* d_lag_force = nullptr;          // [n][2] summed over the iterations of the last step
    double* d_force_sum = nullptr;         // [2]
    bool box_valid = false;                // box cells outside the band hold the external force
    bool split_ok = false;                 // the box avoids non-periodic domain edges
    int n_ent = 0;                         // stencil (= spread) table entries
    size_t shmem = 0;                      // LDS of a pass reading its tables from memory
    size_t shmem_tables = 0;               // LDS of a pass with the tables staged, or 0 if too large
    int* d_mkf_band = nullptr;             // fixed-width lists, slot-major (lbm_ibm_tables.hpp)
    float* d_mkf_w = nullptr;
    int* d_bnf_mk = nullptr;
    float* d_bnf_w = nullptr;
    int kspread = 0;                       // spread slots per band entry (8 or 16)
    bool fast = false;                     // the pass keeps its lists in registers
    size_t shmem_fast = 0;                 // its LDS
    bool fused = false;                    // one launch per step can run the pass (lbm_ibm.hip)
    size_t shmem_fused = 0;                // the LDS of that launch
    int fused_wgs = 0;                     // its IBM workgroups (each with whole box rows)
    void* d_fused_args = nullptr;          // its constant IBM arguments
    std::vector<float> h_ds;               // arc lengths (host copy, for the force sum)
};

// Discrete-adjoint state: checkpoints of the forward populations.
struct AdjointState {
    int nsteps = 0, every = 0;
    std::vector<float*> checkpoints;       // FP32 SoA [9][plane] each
    float* d_seg = nullptr;                // states of one segment, [seg_states][9][plane]
    int seg_states = 0;
    float* d_mu[2] = {nullptr, nullptr};   // adjoint of the pre-collision state
    float* d_grad_force = nullptr;         // AoS [N][2]
    float* d_u0 = nullptr;                 // initial velocity SoA [2][N] (for dL/du0)
    float* d_rho_init = nullptr;
    bool has_u0 = false;                   // forward started from lbm_init_from_macros
    bool ready = false;
};

struct lbm_ctx {
    lbm_config cfg{};
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    // IBM steps run the step outside the IBM box on aux_stream, concurrently with the IBM
    // pass and the box's step on `stream` (fork and join through the two events)
    hipStream_t aux_stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // Two split IBM steps (buf[ibm_graph_cur] -> other -> back) captured once as a HIP graph
    // and replayed: the seven runtime calls of a split step cost more host time than the GPU
    // spends on the step. Dropped whenever anything the steps read changes.
    hipGraphExec_t ibm_graph = nullptr;
    int ibm_graph_cur = -1;
    hipStream_t ibm_graph_stream = nullptr;
    size_t N = 0;
    long long plane = 0;
    void* buf[2] = {nullptr, nullptr};
    int cur = 0;                       // buffer holding the latest post-collision state
    // Two-step launches: edge stores of the non-periodic edge nodes ([9][n_edge] each).
    // edge_in_store: the state one step before buf[cur] at those nodes lives in
    // d_edge[edge_cur] rather than in buf[1-cur] (true after a two-step launch).
    void* d_edge[2] = {nullptr, nullptr};
    int n_edge = 0;
    int edge_cur = 0;
    bool edge_in_store = false;
    bool flags_local = true;           // no flag needs a neighbour's populations
    uint8_t* d_flags = nullptr;
    uint8_t* d_row_class = nullptr;
    float* d_force_field = nullptr;
    float *d_rho = nullptr, *d_ux = nullptr, *d_uy = nullptr;
    float* d_avg = nullptr;
    double* d_avg64 = nullptr;
    // CM<OptimalAdapter>, same-step averages: the averages of the next step, computed by the
    // previous look-ahead launch (valid while nothing else changed the state)
    float* d_ahead = nullptr;
    double* d_ahead64 = nullptr;
    bool ahead_valid = false;
    bool ahead_edges = false;  // d_edge[edge_cur] holds the edge nodes' state one step back
    double* partials = nullptr;
    long long partials_n = 0;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    bool initialized = false, macros_valid = false, avg_valid = false, init_from_macros = false;
    int64_t timestep = 0;
    IbmState ibm;
    AdjointState adj;
    std::string err;

    ~lbm_ctx();
    int fail(int code, const char* fmt, ...);
    int hip_fail(hipError_t e, const char* what);
    int activate();
    int dev_alloc(void** p, size_t bytes, const char* what);
    int upload(void* dst, const void* src, size_t bytes, int on_device);
    int scratch(size_t bytes);
    int ensure_partials(long long n);
    lbm::StepParams step_params(int src_idx, bool write_macros) const;
    int adapter_pass(int src_idx);
    bool ahead_ok() const;
    int ahead_step(bool write_macros);
    int one_step(bool write_macros);
    int split_ibm_step(bool write_macros);
    bool ibm_fused_ok() const;
    int fused_ibm_step(bool write_macros);
    bool ibm_graph_ok() const;
    int ibm_graph_pair();
    void ibm_graph_free();
    bool two_step_ok() const;
    int two_step(bool write_macros);
    int restore_edges();
    // lbm_ibm.hip
    int ibm_pass(int src_idx);
    void ibm_free();
    // lbm_adjoint.hip
    void adjoint_free();
};
