// lbm.cuh — host driver `LBM<2>` for reference-style scenarios, over the C ABI of
// liblbm_hip.so.
//
// The reference's LBM<dim> (src/core/lbm.cuh:33-382) owns device buffers and runs eight
// stage kernels per step. Here the stages are one fused kernel inside the library, so the
// per-step surface is step<Scenario>(n); the rest of the public surface scenario code and
// validators touch is kept: allocate/init, h_rho/h_u (AoS u), timestep/update_ts,
// update_macroscopics(), compute_error<Scenario>(), save_vtk(t).
//
// init<Scenario>() evaluates the scenario's Init functor in a small kernel compiled in the
// including translation unit (so __device__ functors work), its Boundary functor on the
// host (as setup_boundary_flags does, boundaries.cuh:170-197), and hands the result to the
// library. Forces from Init::apply_forces are evaluated once and held static.
#ifndef LBM_H
#define LBM_H

#include <hip/hip_runtime.h>

#include <cstdio>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../../include/lbm_hip.h"
#include "../scenarios/scenario.cuh"
#include "lbm_constants.cuh"

namespace lbm_compat {

template <class S, class = void>
struct periodic_x : std::false_type {};
template <class S>
struct periodic_x<S, std::void_t<decltype(S::periodic_x)>> : std::integral_constant<bool, S::periodic_x> {};
template <class S, class = void>
struct periodic_y : std::false_type {};
template <class S>
struct periodic_y<S, std::void_t<decltype(S::periodic_y)>> : std::integral_constant<bool, S::periodic_y> {};
template <class S, class = void>
struct storage_of : std::integral_constant<int, LBM_F32> {};
template <class S>
struct storage_of<S, std::void_t<decltype(S::storage)>> : std::integral_constant<int, S::storage> {};

template <class InitT>
__global__ void init_fields_kernel(float* rho, float* u, float* force, InitT init, int n) {
    int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= n) return;
    InitT f = init;
    f(rho, u, force, node);
}

inline void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
inline void check_lbm(lbm_ctx* c, int rc, const char* what) {
    if (rc != LBM_OK) throw std::runtime_error(std::string(what) + ": " + lbm_last_error(c));
}

}  // namespace lbm_compat

template <int dim>
class LBM;

template <>
class LBM<2> {
public:
    std::vector<float> h_rho;
    std::vector<float> h_u;
    int timestep = 0, update_ts = 0;
    lbm_ctx* ctx = nullptr;

    LBM() = default;
    LBM(const LBM&) = delete;
    LBM& operator=(const LBM&) = delete;
    ~LBM() { free(); }

    void free() {
        if (ctx) lbm_destroy(ctx);
        ctx = nullptr;
    }

    // LBM<dim>::allocate (lbm.cuh:92-125): host mirrors and the scenario's bodies.
    template <typename Scenario>
    void allocate() {
        h_rho.assign(size_t(NX) * NY, 0.0f);
        h_u.assign(size_t(NX) * NY * 2, 0.0f);
        Scenario::add_bodies();
    }

    template <typename Scenario>
    void init() {
        using namespace lbm_compat;
        static_assert(Scenario::CollisionOp::adapter >= 0, "EmpiricalAdapter has no D2Q9 form");
        const int n = NX * NY;
        // constants the functors may read (send_consts, lbm.cuh:63-82)
        float v = Scenario::viscosity, t = Scenario::tau, o = Scenario::omega;
        check(hipMemcpyToSymbol(HIP_SYMBOL(vis), &v, sizeof(float)), "vis");
        check(hipMemcpyToSymbol(HIP_SYMBOL(tau), &t, sizeof(float)), "tau");
        check(hipMemcpyToSymbol(HIP_SYMBOL(omega), &o, sizeof(float)), "omega");
        check(hipMemcpyToSymbol(HIP_SYMBOL(WEIGHTS), h_weights, sizeof(h_weights)), "weights");
        check(hipMemcpyToSymbol(HIP_SYMBOL(C), h_C, sizeof(h_C)), "C");
        check(hipMemcpyToSymbol(HIP_SYMBOL(OPP), h_OPP, sizeof(h_OPP)), "OPP");
        // Init functor per node (init_kernel, init.cuh:29-43)
        float *d_rho = nullptr, *d_u = nullptr, *d_f = nullptr;
        check(hipMalloc(&d_rho, size_t(n) * 4), "hipMalloc");
        check(hipMalloc(&d_u, size_t(n) * 8), "hipMalloc");
        check(hipMalloc(&d_f, size_t(n) * 8), "hipMalloc");
        check(hipMemset(d_u, 0, size_t(n) * 8), "hipMemset");
        check(hipMemset(d_f, 0, size_t(n) * 8), "hipMemset");
        auto init = Scenario::init();
        hipLaunchKernelGGL(init_fields_kernel<decltype(init)>, dim3((n + 255) / 256), dim3(256), 0, 0, d_rho, d_u, d_f,
                           init, n);
        check(hipGetLastError(), "init kernel");
        check(hipDeviceSynchronize(), "init kernel");
        std::vector<float> hf(size_t(n) * 2);
        check(hipMemcpy(hf.data(), d_f, hf.size() * 4, hipMemcpyDeviceToHost), "force copy");
        bool uniform = true;
        for (size_t i = 2; i < hf.size() && uniform; i += 2) uniform = hf[i] == hf[0] && hf[i + 1] == hf[1];
        // configuration from the scenario's static members
        lbm_config c;
        lbm_config_default(&c, NX, NY);
        c.collision = Scenario::CollisionOp::kind;
        c.adapter = Scenario::CollisionOp::adapter;
        c.omega = Scenario::omega;
        for (int k = 0; k < 9; k++) c.S[k] = Scenario::S[k];
        c.storage = storage_of<Scenario>::value;
        c.periodic_x = periodic_x<Scenario>::value;
        c.periodic_y = periodic_y<Scenario>::value;
        c.u_lid = Scenario::u_max;
        c.wall_u[0] = Scenario::u_max;
        c.wall_u[1] = 0.0f;
        if (uniform) {
            c.force[0] = hf[0];
            c.force[1] = hf[1];
        }
        free();
        int rc = lbm_create(&c, &ctx);
        if (rc != LBM_OK) throw std::runtime_error(std::string("lbm_create: ") + lbm_last_error(nullptr));
        // Boundary functor on the host (setup_boundary_flags, boundaries.cuh:170-197)
        std::vector<uint8_t> flags(static_cast<size_t>(n));
        auto boundary = Scenario::boundary();
        for (int y = 0; y < NY; y++)
            for (int x = 0; x < NX; x++) flags[size_t(y) * NX + x] = uint8_t(boundary(x, y));
        check_lbm(ctx, lbm_set_flags(ctx, flags.data(), 0), "lbm_set_flags");
        if (!uniform) check_lbm(ctx, lbm_set_force_field(ctx, d_f, 1), "lbm_set_force_field");
        check_lbm(ctx, lbm_init_from_macros(ctx, d_rho, d_u, 1), "lbm_init_from_macros");
        check_lbm(ctx, lbm_sync(ctx), "lbm_sync");
        (void)hipFree(d_rho);
        (void)hipFree(d_u);
        (void)hipFree(d_f);
        // immersed bodies (IBMManager::init_and_dispatch, IBMManager.cuh:54-110)
        std::vector<float> xy;
        for (const IBMBody& b : Scenario::IBM_bodies)
            for (int i = 0; i < 2 * b.num_points; i++) xy.push_back(b.points[i]);
        if (!xy.empty())
            check_lbm(ctx, lbm_ibm_set_markers(ctx, xy.data(), nullptr, int(xy.size() / 2), 3), "lbm_ibm_set_markers");
        timestep = update_ts = 0;
    }

    // increase_ts (lbm.cuh:142-146)
    template <typename Scenario>
    void increase_ts() {
        timestep++;
        Scenario::update_ts(float(timestep));
    }

    // n iterations of the time loop (main.cu:96-114), fused.
    template <typename Scenario>
    void step(int n = 1) {
        lbm_compat::check_lbm(ctx, lbm_step(ctx, n), "lbm_step");
        timestep += n;
        Scenario::update_ts(float(timestep));
    }

    // update_macroscopics (lbm.cuh:148-154): host copies of rho and AoS u.
    void update_macroscopics() {
        lbm_compat::check_lbm(ctx, lbm_get_macros(ctx, h_rho.data(), h_u.data(), 0), "lbm_get_macros");
        update_ts = timestep;
    }

    template <typename Scenario>
    float compute_error() {
        if constexpr (Scenario::has_analytical_solution) {
            return Scenario::compute_error(*this);
        } else {
            return 0.0f;
        }
    }

    // save_macroscopics (lbm.cuh:173-204): raw float32 dumps
    // <dir>/density/density_<t>.bin (N values) and <dir>/velocity/velocity_<t>.bin (AoS u,
    // 2N values: the "aos" layout of the reference's graphics/visualize.py reader).
    void save_macroscopics(int t, const std::string& dir = "output") {
        update_macroscopics();
        if (lbm_write_macros(dir.c_str(), t, (int64_t)NX * NY, 2, h_rho.data(), h_u.data()) != LBM_OK)
            throw std::runtime_error("save_macroscopics: cannot write the dumps under " + dir);
    }

    // save_vtk (lbm.cuh:269-343): <dir>/vtk/sim_data_%06d.vti, VTK ImageData (UInt64
    // headers) with appended raw Float32 "Density" and 3-component "Velocity". vtk_flags =
    // LBM_VTI_REFERENCE_OFFSET writes the reference's Velocity offset (without the density
    // block's length word).
    void save_vtk(int t, const std::string& dir = "output", uint32_t vtk_flags = 0) {
        update_macroscopics();
        if (lbm_write_vti(dir.c_str(), t, NX, NY, 1, 2, h_rho.data(), h_u.data(), vtk_flags) != LBM_OK)
            throw std::runtime_error("save_vtk: cannot write " + dir + "/vtk");
    }
};

#endif  // LBM_H
