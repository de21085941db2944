// lbm_kernels.hpp — launchers of the D2Q9 kernels (internal to liblbm_hip.so).
#pragma once

#include <hip/hip_runtime.h>

#include "lbm_step.hpp"

namespace lbm {

constexpr int kXcds = 8;  // MI355X: 8 XCDs, workgroups dealt round-robin

struct InitParams {
    const float* rho;  // [N]
    const float* u;    // AoS [N][2]
    void* buf0;
    void* buf1;
    long long plane, N;
    float rho0;
};

// Launch shape of the fused step: workgroup size, XCD-contiguous workgroup order,
// non-temporal stores and nodes per thread, chosen by whether the two population buffers
// fit in the Infinity Cache.
struct LaunchPolicy {
    int block = 256;
    int remap = 0;
    int nt = 0;
    int seg = kSegWide;
};
LaunchPolicy step_policy(long long plane, int storage);
long long step_grid_blocks(const StepParams& p, const LaunchPolicy& lp);
// mode: 0 step, 1 moments pass, 2 step + moments (CM only). Sets p.ntx and p.nt from lp.
hipError_t launch_step(int coll, int storage, int mode, const StepParams& p, const LaunchPolicy& lp, double* partials,
                       hipStream_t s);
// Two steps per launch (temporal blocking), in one of two forms: tiles (ty = 6 or 14 rows of
// kT2W nodes per workgroup) or marching strips (strip = 1: w = 4 or 8 wavefronts per
// workgroup, h rows per strip). XCD remap of the workgroup order. The caller sets
// p.edge_prev / edge_next / n_edge when a non-periodic axis exists.
struct Step2Policy {
    int ty = 6;
    int remap = 1;
    int strip = 0;
    int w = 8;
    int h = 64;
};
Step2Policy step2_policy(long long plane, int storage, int coll);
long long step2_grid_blocks(const StepParams& p, const Step2Policy& lp);
hipError_t launch_step2(int coll, int storage, const StepParams& p, const Step2Policy& lp, hipStream_t s);
// to_store = 1: edge nodes of `buf` -> store; 0: store -> edge nodes of `buf`.
hipError_t launch_edge_copy(int storage, void* buf, void* store, long long plane, int nx, int ny, int px, int py,
                            int to_store, hipStream_t s);
hipError_t launch_finalize_avg(const double* partials, long long n, long long N, float* avg, double* avg64, hipStream_t s);
// One CM<OptimalAdapter> step per launch (marching strips of 4 wavefronts) that also sums the
// next step's adapter moments into partials (one triple per strip): p.avg holds the averages
// of this step, p.mid == p.dst, edge stores as for launch_step2 (the next store receives the
// source state of the edge nodes). Then launch_finalize_ahead: avg <- ahead, ahead <- means.
Step2Policy ahead_policy(long long plane, int storage);
long long ahead_partials_needed(const StepParams& p, const Step2Policy& lp);
hipError_t launch_step_ahead(int coll, int storage, const StepParams& p, const Step2Policy& lp, double* partials,
                             hipStream_t s);
hipError_t launch_finalize_ahead(const double* partials, long long n, long long N, float* ahead, double* ahead64, float* avg,
                                 double* avg64, hipStream_t s);
hipError_t launch_init(int storage, const InitParams& p, hipStream_t s);
hipError_t launch_pops_to_f32(int storage, const void* buf, long long plane, long long N, float rho0, float* out, hipStream_t s);
hipError_t launch_pops_from_f32(int storage, void* b0, void* b1, long long plane, long long N, float rho0, const float* in, hipStream_t s);
hipError_t launch_macros_to_aos(const float* ux, const float* uy, long long N, float* u, hipStream_t s);
hipError_t launch_aos_to_soa(const float* u, long long N, float* ux, float* uy, hipStream_t s);
long long reduce_partials_needed(long long N);
hipError_t launch_reduce(int kind, const float* rho, const float* ux, const float* uy, const float* ref, long long N,
                         double* partials, double* out3, hipStream_t s);
hipError_t launch_row_mean_ux(const float* ux, int nx, int ny, double* out, hipStream_t s);

}  // namespace lbm
