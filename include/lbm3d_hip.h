/*
 * lbm3d_hip.h — C ABI of the D3Q27 path of liblbm_hip.so (MI355X, gfx950).
 *
 * The reference's default build is D3Q27 (src/defines.hpp:4): `LBM<3>` (src/core/lbm.cuh:33-382)
 * with 27 SoA population planes, node = z*NX*NY + y*NX + x (src/core/lbm_constants.cuh:351-358),
 * stepped by main() (src/main.cu:96-114). This header is the 3-D counterpart of lbm_hip.h and
 * keeps its conventions: every function returns 0 or a negative LBM_E* code (lbm_hip.h) with
 * the message in lbm3d_last_error(ctx); a context owns its device memory and one stream;
 * host-facing vector fields are AoS [node][3]; populations crossing the boundary are FP32 SoA
 * [27][node]; `on_device` = 1 marks a device pointer (the call is then stream-ordered and does
 * not synchronise).
 *
 * Built: BGK<3> with Guo forcing (src/core/collision/BGK/BGK.cuh:55-96), MRT<3> in the De Rosis
 * raw-moment basis (src/core/collision/MRT/MRT.cu:79-456, basis src/core/lbm_constants.cuh:218-246),
 * CM<3, NoAdapter> (src/core/collision/CM/CM_D3Q27.cuh:1-105), the FLUID, BOUNCE_BACK and
 * CORNER_EDGE_BOUNCE_BACK flags (src/core/boundaries/boundaries.cuh:89-138, BounceBack<3>
 * src/functors/boundaryConditions/bbDomainBoundary.cuh:52-90, EdgeCornerBounceBack
 * src/functors/boundaryConditions/edgeCornerBounceBack.cuh:10-49), ZG_OUTFLOW on every domain
 * face (ZG_OutflowBoundary<3>, zeroGradientOutflow.cuh:8-59), REGULARIZED_BOUNCE_BACK,
 * REGULARIZED_INLET_LEFT (regularizedBoundary3D.cuh:10-137), REGULARIZED_OUTLET (RegularizedOutlet,
 * regularizedOutlet.cuh), EXTRAPOLATED_CORNER_EDGE (ExtrapolatedCornerEdgeBoundary<3>,
 * cornerBoundaries.cuh:6-36), GUO_VELOCITY_INLET at
 * x = 0 and GUO_PRESSURE_OUTLET at x = NX-1 (guoInlet3D.cuh, guoOutlet3D.cuh), runtime periodic
 * axes, FP32 or 16-bit shifted population storage. The CM<3> adapters (ACM_D3Q27.cuh,
 * E_ACM_D3Q27.cuh) are rejected with LBM_E_UNSUPPORTED.
 */
#ifndef LBM3D_HIP_H
#define LBM3D_HIP_H

#include <stdint.h>

#include "lbm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Reference-literal 3-D equilibrium (bit in lbm3d_config.quirks). The reference's De Rosis
 * form (src/core/equilibrium/equilibrium.cu:70-96) gives the 12 edge and 8 corner directions
 * a linear coefficient 1 where the second-order equilibrium has 3 (1/c_s^2), so its first
 * moment is (17/27) rho u instead of rho u. Default (0): the second-order equilibrium. */
#define LBM_QUIRK_EQ3D_LINEAR (1u << 6)

/* Reference-literal MRT<3> post-collision moments (bit in lbm3d_config.quirks), as
 * compute_m_post_component writes them (src/core/collision/MRT/MRT.cu:155-201): the momentum
 * moments are rho*u with no forcing term (:168-171), so a step adds F/2 of momentum instead of F,
 * and the equilibrium of moment 19 (cx^2 cy^2 - cx^2 cz^2) carries (2 ux^2 + 1) where the moment
 * of the equilibrium has (3 ux^2 + 1) (:188). Default (0): rho*u + F/2 (the (1 - S/2) F rule
 * with S = 1 applied to every other moment) and (3 ux^2 + 1). The reference's active path,
 * compute_m_post_combined (MRT.cu:203-395), leaves moments 8-26 unset and is not restated. */
#define LBM_QUIRK_MRT3_LITERAL (1u << 7)

/* Reference-literal Guo boundary moments (bit in lbm3d_config.quirks). GuoVelocityInlet3D and
 * GuoPressureOutlet3D (src/functors/boundaryConditions/guoInlet3D.cuh:8-60, guoOutlet3D.cuh:6-66)
 * name the interior neighbour fluid_node but compute its rho_f, u_f from the boundary node's
 * own post-stream populations (unknown ones included); only f_neq = f(neighbour) - f_eq(rho_f,
 * u_f) reads the neighbour. Default (0): rho_f, u_f of the neighbour. */
#define LBM_QUIRK_GUO3D_OWN_MOMENTS (1u << 8)

/* Reference-literal velocity of ExtrapolatedCornerEdgeBoundary<3> (bit in lbm3d_config.quirks):
 * src/functors/boundaryConditions/cornerBoundaries.cuh:23 sums u_z with C[3*2+i], the flat
 * velocity table read at 6 + i, instead of C[3*i+2]. Default (0): the interior node's u_z. */
#define LBM_QUIRK_CORNER3D_UZ (1u << 9)

/* CM<3> adapters (src/core/collision/adapters.cuh): LBM_NO_ADAPTER is built. LBM_OPTIMAL_ADAPTER
 * (CM<3, OptimalAdapter>, src/core/collision/CM/ACM_D3Q27.cuh) and LBM3D_EMPIRICAL_ADAPTER
 * (E_ACM_D3Q27.cuh) are rejected with LBM_E_UNSUPPORTED: as shipped, both assign instead of
 * accumulate their central moments over the directions (ACM_D3Q27.cuh:30-52, E_ACM_D3Q27.cuh:40-65)
 * and write their 27 results to f[0..26] of the whole population array instead of the node's
 * own populations (ACM_D3Q27.cuh:100-126, E_ACM_D3Q27.cuh:103-129), so there is no operator to
 * restate. */
#define LBM3D_EMPIRICAL_ADAPTER 2

typedef struct lbm3d_config {
    int32_t  nx, ny, nz;
    int32_t  collision;      /* LBM_BGK | LBM_MRT | LBM_CM                                    */
    int32_t  adapter;        /* LBM_NO_ADAPTER (CM<3> adapters are not built)                  */
    float    omega;          /* BGK and CM rate 1/tau (ScenarioTrait::omega, src/scenarios/scenario.cuh:36-38) */
    float    S[27];          /* MRT<3> rates in the basis order (ScenarioTrait::S); S[4..8], the
                                deviatoric second moments, are the ones MRT<3> reads            */
    float    rho0;           /* reference density of the shifted storage                      */
    int32_t  storage;        /* LBM_F32 | LBM_F16_SHIFTED | LBM_BF16_SHIFTED                  */
    int32_t  periodic_x;     /* streaming wraps on the axis (PERIODIC_X/Y/Z of               */
    int32_t  periodic_y;     /*   src/core/streaming/streaming.cu:45-53, never defined there) */
    int32_t  periodic_z;
    float    force[3];       /* uniform Guo body force, used where no force field is set      */
    float    u_inlet;        /* Scenario::u_max of GUO_VELOCITY_INLET and REGULARIZED_INLET_LEFT
                                (boundaries.cuh:110-125)                                        */
    uint32_t quirks;         /* LBM_QUIRK_STALE_REST | LBM_QUIRK_EQ3D_LINEAR | LBM_QUIRK_MRT3_LITERAL
                                | LBM_QUIRK_GUO3D_OWN_MOMENTS | LBM_QUIRK_CORNER3D_UZ            */
    int32_t  reserved[8];
} lbm3d_config;

typedef struct lbm3d_ctx lbm3d_ctx;

/* Defaults: BGK, no adapter, tau = 1 (omega = 1), S = 1 everywhere, rho0 = 1, FP32, not periodic,
 * no force. */
int lbm3d_config_default(lbm3d_config* cfg, int32_t nx, int32_t ny, int32_t nz);

/* Replaces LBM<3>::allocate (lbm.cuh:92-125) + send_consts (lbm.cuh:63-82). */
int lbm3d_create(const lbm3d_config* cfg, lbm3d_ctx** out);
/* Replaces LBM<3>::free (lbm.cuh:127-140). NULL is accepted. */
int lbm3d_destroy(lbm3d_ctx* ctx);
/* Run on `hip_stream` (a hipStream_t; NULL = the HIP null stream). */
int lbm3d_set_stream(lbm3d_ctx* ctx, void* hip_stream);

/* Replaces setup_boundary_flags (boundaries.cuh:170-197) for 3-D: one flag per node, LBM_FLUID,
 * LBM_BOUNCE_BACK, LBM_CORNER_EDGE_BOUNCE_BACK (acts on nodes with two or more domain faces, like
 * FLUID elsewhere), LBM_REGULARIZED_BOUNCE_BACK (u = 0) and LBM_REGULARIZED_INLET_LEFT
 * (u = (u_inlet, 0, 0)) on any face node (like FLUID inside), LBM_EXTRAPOLATED_CORNER_EDGE (nodes
 * with two or more domain faces), LBM_ZG_OUTFLOW (any face node; the first face of x = 0,
 * x = NX-1, y = 0, y = NY-1, z = 0, z = NZ-1 that holds the node names its inward neighbour),
 * LBM_GUO_VELOCITY_INLET (x = 0 only), LBM_GUO_PRESSURE_OUTLET or LBM_REGULARIZED_OUTLET (x = NX-1
 * only); Guo and regularized-outlet flags need a non-periodic x axis; other flags give
 * LBM_E_UNSUPPORTED. NULL = all FLUID. ZG, outlet and extrapolated nodes read a neighbour's
 * post-stream populations, including slots the neighbour's own update overwrites: a neighbour
 * in the same wavefront is pulled before any store; other such nodes run in launches of their
 * own first, ordered so that each reads its neighbour before that neighbour's launch. Flags that
 * read each other in a cycle across wavefronts give LBM_E_UNSUPPORTED. */
int lbm3d_set_flags(lbm3d_ctx* ctx, const uint8_t* flags, int on_device);

/* Replaces init_kernel<3> + init_node (src/core/init/init.cuh:10-43): f = f_back = f_eq.
 * rho [N]; u AoS [N][3]. Resets the step counter. */
int lbm3d_init_from_macros(lbm3d_ctx* ctx, const float* rho, const float* u, int on_device);

/* Replaces reset_forces + Init::apply_forces (macroscopics.cuh:13-48) for a non-uniform
 * force, AoS [N][3]. NULL removes the field (the uniform cfg.force applies). */
int lbm3d_set_force_field(lbm3d_ctx* ctx, const float* force, int on_device);

/* `nsteps` iterations of main.cu:96-114 (stream, boundaries, macroscopics, forces, Guo
 * correction, equilibrium, BGK, MRT or CM collision), one fused pull-scheme kernel per step.
 * Asynchronous; the last step also stores rho and the corrected u. */
int lbm3d_step(lbm3d_ctx* ctx, int32_t nsteps);
int lbm3d_sync(lbm3d_ctx* ctx);
int lbm3d_get_timestep(const lbm3d_ctx* ctx, int64_t* t);

/* Replaces update_macroscopics (lbm.cuh:148-154): rho [N], corrected u AoS [N][3] of the
 * last step's post-stream state. Either pointer may be NULL. */
int lbm3d_get_macros(lbm3d_ctx* ctx, float* rho, float* u, int on_device);

/* Raw state, FP32 SoA [27][N]. lbm3d_set_pops writes both buffers (as init does). */
int lbm3d_get_pops(lbm3d_ctx* ctx, float* f, int on_device);
int lbm3d_set_pops(lbm3d_ctx* ctx, const float* f, int on_device);

/* Reductions over the last stored macroscopic fields, FP64, deterministic. Synchronises. */
#define LBM3D_REDUCE_MASS           0 /* out[0] = sum rho                                      */
#define LBM3D_REDUCE_MOMENTUM       1 /* out[0..2] = sum rho*u                                 */
#define LBM3D_REDUCE_KINETIC_ENERGY 2 /* out[0] = sum 0.5*rho*|u|^2; compute_kinetic_energy_kernel
                                         (src/scenarios/taylorGreen3D/postprocessing/
                                         kineticEnergy.cuh:7-42) before its /(N*0.05^2)  */
int lbm3d_reduce(lbm3d_ctx* ctx, int32_t kind, double* out);

/* ---- immersed boundary, 3-D (multi-direct forcing IBMManager<3>::multi_direct,
 * IBMManager.cuh:222-252; interpolation and spreading IBM_impl.cu:53-118, 156-196) ----------
 * Replaces IBMManager<3>::init_and_dispatch + send_to_gpu (IBMManager.cuh:54-110) for static
 * markers in lattice units, AoS [n][3], in the caller's order (the reference Morton-sorts a
 * body's markers first, IBMManager.cuh:57-66: pass them sorted to reproduce its order); ds:
 * per-marker area (NULL = 1; ignored under LBM_QUIRK_IBM_NO_DS). iters = ITER_MAX
 * (IBMManager.cuh:8, 3 in the reference); LBM_QUIRK_IBM_CLAMP applies as in 2-D. Each step
 * interpolates the post-stream rho and u over the 4x4x4 Peskin stencils, iterates the
 * forcing on the band of stencil nodes and adds the spread force to the external force there.
 * n = 0 removes the body. When the band's box is clear of non-periodic faces and of every flag
 * but bounce-back, the nodes outside the box are stepped on an internal stream beside the IBM
 * pass, joined to the context's stream before lbm3d_step returns. */
int lbm3d_ibm_set_markers(lbm3d_ctx* ctx, const float* xyz, const float* ds, int32_t n, int32_t iters);
/* Lagrangian forces summed over the iterations of the last step, AoS [n][3] (may be NULL), and
 * the total force the markers apply to the fluid, sum_l F_l ds_l in FP64 (force_sum[3], may be
 * NULL). Synchronises. */
int lbm3d_ibm_get_forces(lbm3d_ctx* ctx, float* lag_force, double* force_sum);

const char* lbm3d_last_error(const lbm3d_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* LBM3D_HIP_H */
