/*
 * lbm_hip.h — C ABI of liblbm_hip.so, the MI355X (gfx950) D2Q9 lattice-Boltzmann library.
 *
 * The reference (Carabalone/cuda-lbm) has no C ABI: its boundary is the compile-time
 * C++ template API of `LBM<dim>` (src/core/lbm.cuh:33-382) driven by `main()`
 * (src/main.cu:91-165) and configured by a `ScenarioTrait` (src/scenarios/scenario.cuh:22-78).
 * Each entry point below names the reference member or stage it replaces.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative LBM_E* code on failure; the
 *     message is available from lbm_last_error(ctx). Nothing calls exit() (the reference's
 *     checkCudaErrors does: src/util/utility.cu:4-12).
 *   - A context owns all of its device memory and one HIP stream. Contexts share no global
 *     state (the reference keeps its constants in process-wide __constant__ symbols,
 *     src/core/lbm.cu:8-19), so several contexts may live in one process.
 *   - Grids are NX x NY, node = y*NX + x (src/core/lbm_constants.cuh:343-350).
 *   - Host-facing vector fields (u, force) are AoS [node][2]; populations exchanged through
 *     lbm_get_pops/lbm_set_pops are SoA [q][node] FP32 (src/core/lbm_constants.cuh:320-323).
 *   - `on_device` = 1 means the pointer is device memory visible to the context's GPU;
 *     the call is then stream-ordered on the context's stream and does not synchronise.
 */
#ifndef LBM_HIP_H
#define LBM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------- */
#define LBM_OK              0
#define LBM_E_INVALID     (-1)   /* bad argument or configuration            */
#define LBM_E_UNSUPPORTED (-2)   /* valid reference feature not built here   */
#define LBM_E_HIP         (-3)   /* a HIP runtime call failed                */
#define LBM_E_NOMEM       (-4)   /* device allocation failed                 */
#define LBM_E_STATE       (-5)   /* call out of order (e.g. step before init) */

/* ---- collision operators (src/core/collision/{BGK,MRT,CM}) ----------------------- */
#define LBM_BGK 0   /* BGK<2>  src/core/collision/BGK/BGK.cuh:13-51            */
#define LBM_MRT 1   /* MRT<2>  src/core/collision/MRT/MRT.cu:4-76              */
#define LBM_CM  2   /* CM<2,A> src/core/collision/CM/CM.cuh:27-250             */

/* ---- higher-order relaxation adapters for CM (src/core/collision/adapters.cuh) --- */
#define LBM_NO_ADAPTER      0   /* adapters.cuh:16-22  */
#define LBM_OPTIMAL_ADAPTER 1   /* adapters.cuh:48-111 */

/* OptimalAdapter domain averages: same-step (exact reference order, one extra read pass)
 * or lagged by one step (no extra pass). */
#define LBM_ADAPTER_SAME_STEP 0
#define LBM_ADAPTER_LAGGED    1

/* ---- population storage in HBM ------------------------------------------------- */
#define LBM_F32          0   /* FP32 populations, 72 B per node update             */
#define LBM_F16_SHIFTED  1   /* FP16 of f_i - w_i*rho0, FP32 collision, 36 B        */
#define LBM_BF16_SHIFTED 2   /* BF16 of f_i - w_i*rho0, FP32 collision, 36 B        */

/* ---- boundary flags: reference names and order, src/core/lbm_constants.cuh:377-397 --- */
enum lbm_bc_flag {
    LBM_FLUID = 0,
    LBM_BOUNCE_BACK,                   /* bbDomainBoundary.cuh:15-51                   */
    LBM_ZOU_HE_TOP,                    /* zouHeInflow.cuh:35-68                        */
    LBM_ZOU_HE_LEFT,                   /* zouHeInflow.cuh:9-32                         */
    LBM_ZOU_HE_TOP_LEFT_TOP_INFLOW,    /* not implemented in the reference either      */
    LBM_ZOU_HE_TOP_RIGHT_TOP_INFLOW,   /* not implemented in the reference either      */
    LBM_CYLINDER,                      /* cylinderBoundary.cuh:22-34 (full-way BB)     */
    LBM_ZG_OUTFLOW,                    /* zeroGradientOutflow.cuh:9-59                 */
    LBM_PRESSURE_OUTLET,               /* pressureOutlet.cuh:7-42                      */
    LBM_REGULARIZED_INLET_TOP,         /* regularizedInlet.cuh:15-69                   */
    LBM_REGULARIZED_INLET_LEFT,
    LBM_REGULARIZED_BOUNCE_BACK,       /* regularizedBounceBack.cuh:13-100             */
    LBM_REGULARIZED_BOUNCE_BACK_CORNER,/* regularizedBounceBack.cuh:107-222            */
    LBM_EXTRAPOLATED_CORNER_EDGE,
    LBM_CORNER_EDGE_BOUNCE_BACK,
    LBM_PRESSURE_INLET_LEFT,
    LBM_GUO_VELOCITY_INLET,
    LBM_GUO_PRESSURE_OUTLET,
    LBM_REGULARIZED_OUTLET,
    LBM_MOVING_WALL_BOUNCE_BACK        /* appended: Ladd moving-wall bounce-back; the
                                          reference's lid is REGULARIZED_INLET_TOP    */
};

/* ---- reference-literal behaviours (bitmask in lbm_config.quirks) -------------------
 * Default (0) is the corrected behaviour. Setting a bit reproduces the reference's code
 * as written, so each quirk's effect can be measured. */
#define LBM_QUIRK_STALE_REST     (1u << 0) /* rest population not streamed: streaming.cu:9   */
#define LBM_QUIRK_ZOUHE_RHO      (1u << 1) /* Zou-He left density uses f2 for f3: zouHeInflow.cuh:15-16 */
#define LBM_QUIRK_MRT_FORCE_ROWS (1u << 2) /* MRT forcing rows 4/5 swapped: MRT.cu:17-20      */
#define LBM_QUIRK_IBM_2X2        (1u << 3) /* 2x2 stencil of the 4-point kernel: IBM_impl.cu:19-20 */
#define LBM_QUIRK_IBM_CLAMP      (1u << 4) /* Lagrangian force / corrected u clamped > 1e-8: IBM_impl.cuh:22-24,42-44 */
#define LBM_QUIRK_IBM_NO_DS      (1u << 5) /* spreading without marker arc length: IBM_impl.cu:143-144 */

typedef struct lbm_config {
    int32_t  nx, ny;
    int32_t  collision;      /* LBM_BGK | LBM_MRT | LBM_CM                                  */
    int32_t  adapter;        /* LBM_NO_ADAPTER | LBM_OPTIMAL_ADAPTER (CM only)              */
    int32_t  adapter_mode;   /* LBM_ADAPTER_SAME_STEP | LBM_ADAPTER_LAGGED                  */
    float    omega;          /* BGK rate 1/tau (ScenarioTrait::omega, scenario.cuh:36-38)    */
    float    S[9];           /* MRT/CM rates in the operator's own moment order (scenario.cuh:47-57) */
    float    rho0;           /* reference density of the shifted storage and moving walls   */
    int32_t  storage;        /* LBM_F32 | LBM_F16_SHIFTED | LBM_BF16_SHIFTED                */
    int32_t  periodic_x;     /* streaming wraps in x (the reference's PERIODIC_X)            */
    int32_t  periodic_y;
    float    force[2];       /* uniform Guo body force, used where no force field is set    */
    float    wall_u[2];      /* velocity of LBM_MOVING_WALL_BOUNCE_BACK nodes               */
    float    u_lid;          /* Scenario::u_max as used by REGULARIZED_INLET_TOP / ZOU_HE_*  */
    uint32_t quirks;         /* LBM_QUIRK_* bitmask                                         */
    int32_t  steps_per_launch; /* 0: two time steps per launch wherever supported (the same
                                  algorithm; agrees with 1 to FP32 rounding), 1: one step per
                                  launch                                                       */
    int32_t  reserved[7];
} lbm_config;

typedef struct lbm_ctx lbm_ctx;

/* Fill `cfg` with defaults: BGK, tau = 0.5 + 3/6 (nu = 1/6, scenario.cuh:36-38),
 * S = DEFAULT_MRT_S_MATRIX(omega) (scenario.cuh:10-20), rho0 = 1, FP32, not periodic. */
int lbm_config_default(lbm_config* cfg, int32_t nx, int32_t ny);

/* Replaces LBM<2>::allocate (lbm.cuh:92-125) + send_consts (lbm.cuh:63-82).
 * Allocates two population buffers, flags and macroscopic outputs on the current device. */
int lbm_create(const lbm_config* cfg, lbm_ctx** out);
/* Replaces LBM<2>::free / ~LBM (lbm.cuh:127-140, 379-381). NULL is accepted. */
int lbm_destroy(lbm_ctx* ctx);

/* Run the context's work on `hip_stream` (a hipStream_t) instead of the non-blocking stream
 * the context creates for itself, e.g. PyTorch's current stream. NULL selects the HIP null
 * stream (PyTorch's default stream reports handle 0). */
int lbm_set_stream(lbm_ctx* ctx, void* hip_stream);

/* Replaces setup_boundary_flags (boundaries.cuh:170-197): one BC flag (enum lbm_bc_flag)
 * per node. Unsupported flags are rejected with LBM_E_UNSUPPORTED. */
int lbm_set_flags(lbm_ctx* ctx, const uint8_t* flags, int on_device);

/* Replaces init_kernel + init_node (init.cuh:10-43): f = f_back = f_eq(rho, u).
 * rho: [N]; u: AoS [N][2]. Resets the step counter. */
int lbm_init_from_macros(lbm_ctx* ctx, const float* rho, const float* u, int on_device);

/* Replaces reset_forces + Init::apply_forces (macroscopics.cuh:13-48) for a
 * non-uniform force: AoS [N][2]. NULL removes the field (the uniform cfg.force applies). */
int lbm_set_force_field(lbm_ctx* ctx, const float* force, int on_device);

/* One call = `nsteps` iterations of the reference's loop body (main.cu:96-114):
 * stream, boundaries, macroscopics, forces, IBM, Guo correction, equilibrium, collision,
 * fused into one pull-scheme kernel per step (plus the IBM / adapter passes when active),
 * or into one kernel per pair of steps where steps_per_launch allows it (an odd remainder
 * runs as a single step). Asynchronous on the context's stream. The last step also stores
 * rho and u. */
int lbm_step(lbm_ctx* ctx, int32_t nsteps);
int lbm_sync(lbm_ctx* ctx);
int lbm_get_timestep(const lbm_ctx* ctx, int64_t* t);
/* Time steps per kernel launch that lbm_step uses for the current configuration, flags and
 * bodies: 2 (two-step launches) or 1. No counterpart in the reference, which launches eight
 * kernels per step; for profiling and bandwidth accounting. */
int lbm_get_steps_per_launch(const lbm_ctx* ctx, int32_t* steps);

/* Replaces update_macroscopics (lbm.cuh:148-154): density [N] and Guo-corrected
 * velocity AoS [N][2] of the last step's post-stream state. Either pointer may be NULL. */
int lbm_get_macros(lbm_ctx* ctx, float* rho, float* u, int on_device);

/* Checkpoint / resume of the raw state: FP32 SoA [9][N] (the reduced storages convert). */
int lbm_get_pops(lbm_ctx* ctx, float* f, int on_device);
int lbm_set_pops(lbm_ctx* ctx, const float* f, int on_device);

/* ---- validation / conservation reductions (FP64 accumulation, deterministic) ------ */
#define LBM_REDUCE_MASS        0 /* out[0] = sum rho                                      */
#define LBM_REDUCE_MOMENTUM    1 /* out[0..1] = sum rho*u                                 */
#define LBM_REDUCE_L2_VS_REF   2 /* out[0] = sum |u-ref|^2, out[1] = sum |ref|^2; ref AoS [N][2] */
#define LBM_REDUCE_ROW_MEAN_UX 3 /* out[y] = mean_x u_x(x,y), NY values (Poiseuille profile) */
#define LBM_REDUCE_MOMENT_AVG  4 /* out[0..2] = <rho>, <rho|u|>, <|Pi|> (adapter inputs)   */
/* Runs on the last stored macroscopic fields; `ref` is host memory. Synchronises. */
int lbm_reduce(lbm_ctx* ctx, int32_t kind, const float* ref, double* out);
/* lbm_reduce without any allocation: no counterpart in the reference. For callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc). Same kernels and results as lbm_reduce; the partial sums live in a fixed
 * per-GPU array of LBM_NOALLOC_SCRATCH_BYTES, held until the process ends and shared by every
 * call on that GPU: calls that overlap on one GPU race. `ref_device` (L2_VS_REF) is device
 * memory. Limits: N <= 715,784,192 nodes (43,688 partial sums), NY <= 131,072 rows
 * (ROW_MEAN_UX); larger grids return LBM_E_UNSUPPORTED. out_on_device = 1: `out` is device
 * memory, written asynchronously on the context's stream (no synchronisation); 0: host
 * memory, synchronises. */
#define LBM_NOALLOC_SCRATCH_BYTES (1u << 20)
int lbm_reduce_noalloc(lbm_ctx* ctx, int32_t kind, const float* ref_device, double* out, int out_on_device);

/* ---- immersed boundary (src/IBM, multi-direct forcing IBMManager.cuh:222-252) ------
 * Static markers in lattice units, AoS [n][2]; ds: per-marker arc length (NULL = 1).
 * iters = ITER_MAX (IBMManager.cuh:8, 3 in the reference). n = 0 removes the body.
 * Each later lbm_step runs the IBM pass per step on the context's stream: on grids that stay
 * in the Infinity Cache as one launch that also steps the grid; otherwise the pass and the
 * markers' box are stepped there while the rest of the grid is stepped on an internal stream
 * that the context's stream joins before lbm_step returns. */
int lbm_ibm_set_markers(lbm_ctx* ctx, const float* xy, const float* ds, int32_t n, int32_t iters);
/* Lagrangian forces of the last step, AoS [n][2] (may be NULL), and the total force the
 * markers apply to the fluid, summed in FP64 (force_sum[2], may be NULL). Synchronises. */
int lbm_ibm_get_forces(lbm_ctx* ctx, float* lag_force, double* force_sum);

/* ---- discrete adjoint (no counterpart in the reference) -----------------------------
 * Forward: from the current state run `nsteps` steps, keeping an FP32 checkpoint of the
 * populations every `checkpoint_every` steps. Only BGK or MRT with FP32 storage, fully
 * periodic, FLUID-only flags. The forward's final macroscopic fields are stored as by
 * lbm_step. */
int lbm_adjoint_forward(lbm_ctx* ctx, int32_t nsteps, int32_t checkpoint_every);
/* Backward: given dL/du of the final velocity snapshot (AoS [N][2]), accumulate
 * dL/dforce (AoS [N][2]) and, when the initial state came from lbm_init_from_macros,
 * dL/du0 of the initial velocity (AoS [N][2]). Output pointers may be NULL. Recomputes
 * each segment from its checkpoint. */
int lbm_adjoint_backward(lbm_ctx* ctx, const float* dL_du, float* dL_dforce, float* dL_du0,
                         int on_device);

/* ---- output files (host arrays; no device work, no context) --------------------------
 * Replace LBM::save_vtk (lbm.cuh:269-343): <dir>/vtk/sim_data_%06d.vti, VTK XML ImageData
 * version 1.0 with UInt64 headers, appended raw Float32 Density [N] and Velocity [N][3]
 * (dim = 2: u AoS [N][2], written with u_z = 0 and nz = 1; dim = 3: u AoS [N][3]). The
 * Velocity offset is the one the appended layout implies (8 + 4N); flags =
 * LBM_VTI_REFERENCE_OFFSET writes the reference's 4N, which omits the density block's
 * length word. */
#define LBM_VTI_REFERENCE_OFFSET (1u << 0)
int lbm_write_vti(const char* dir, int64_t timestep, int32_t nx, int32_t ny, int32_t nz, int32_t dim, const float* rho,
                  const float* u, uint32_t flags);
/* Replaces LBM::save_macroscopics (lbm.cuh:173-204): <dir>/density/density_<t>.bin (N floats)
 * and <dir>/velocity/velocity_<t>.bin (N x dim floats, AoS). */
int lbm_write_macros(const char* dir, int64_t timestep, int64_t n, int32_t dim, const float* rho, const float* u);
/* Replaces LBM<3>::save_midplane_slice (lbm.cuh:206-267): the z = nz/2 plane of rho [nz][ny][nx]
 * and u AoS [nz][ny][nx][3] into <dir>/midplane_density/density_<t>.bin and
 * <dir>/midplane_velocity/velocity_<t>.bin. */
int lbm_write_midplane(const char* dir, int64_t timestep, int32_t nx, int32_t ny, int32_t nz, const float* rho,
                       const float* u);

const char* lbm_last_error(const lbm_ctx* ctx);
const char* lbm_version(void);

#ifdef __cplusplus
}
#endif

#endif /* LBM_HIP_H */
