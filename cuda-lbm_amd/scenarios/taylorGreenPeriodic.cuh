// Taylor-Green vortex on a fully periodic grid, D2Q9 BGK (the benchmark's plain case:
// 128^2, nu = 0.01, u0 = 0.04), in the ScenarioTrait API of scenarios/scenario.cuh.
#ifndef TAYLOR_GREEN_PERIODIC_CUH
#define TAYLOR_GREEN_PERIODIC_CUH

#include <cmath>

#include "scenarios/scenario.cuh"

struct TGPeriodicInit {
    float nu, u0;
    __host__ __device__ TGPeriodicInit(float nu_, float u0_) : nu(nu_), u0(u0_) {}
    __host__ __device__ void apply_forces(float*, float*, float* force, int node) {
        force[get_vec_index(node, 0)] = 0.0f;
        force[get_vec_index(node, 1)] = 0.0f;
    }
    // decaying vortex at t = 0 on cell centres, p = cs^2 (rho - 1)
    __device__ void operator()(float* rho, float* u, float* force, int node) {
        const float x = float(node % NX) + 0.5f, y = float(node / NX) + 0.5f;
        const float kx = 2.0f * float(M_PI) / NX, ky = 2.0f * float(M_PI) / NY;
        u[get_vec_index(node, 0)] = -u0 * sqrtf(ky / kx) * cosf(kx * x) * sinf(ky * y);
        u[get_vec_index(node, 1)] = u0 * sqrtf(kx / ky) * sinf(kx * x) * cosf(ky * y);
        const float p = -0.25f * u0 * u0 * ((ky / kx) * cosf(2.0f * kx * x) + (kx / ky) * cosf(2.0f * ky * y));
        rho[node] = 1.0f + 3.0f * p;
        apply_forces(rho, u, force, node);
    }
};

struct AllFluid {
    int operator()(int, int) const { return BC_flag::FLUID; }
};

struct TGPeriodicValidation {
    float u0, nu, t;
    void operator()(int xn, int yn, float& ux, float& uy) const {
        const double x = xn + 0.5, y = yn + 0.5, kx = 2 * M_PI / NX, ky = 2 * M_PI / NY;
        const double decay = std::exp(-t * nu * (kx * kx + ky * ky));
        ux = float(-u0 * std::sqrt(ky / kx) * std::cos(kx * x) * std::sin(ky * y) * decay);
        uy = float(u0 * std::sqrt(kx / ky) * std::sin(kx * x) * std::cos(ky * y) * decay);
    }
};

struct TaylorGreenPeriodic : ScenarioTrait<TGPeriodicInit, AllFluid, TGPeriodicValidation, BGK<2>> {
    static constexpr float u_max = 0.04f;
    static constexpr float viscosity = 0.01f;
    static constexpr float tau = viscosity_to_tau(viscosity);
    static constexpr float omega = 1.0f / tau;
    static constexpr float S[quadratures] = DEFAULT_MRT_S_MATRIX(omega);
    static constexpr bool periodic_x = true;
    static constexpr bool periodic_y = true;
    static const char* name() { return "TaylorGreenPeriodic"; }
    static InitType init() { return InitType(viscosity, u_max); }
    static BoundaryType boundary() { return BoundaryType(); }
    static ValidationType validation() { return ValidationType{u_max, viscosity, t}; }
    // relative L2 error of u in percent (taylorGreenScenario.cuh:59-88)
    template <typename LBMSolver>
    static float compute_error(LBMSolver& solver) {
        if (solver.update_ts < solver.timestep) solver.update_macroscopics();
        const auto v = validation();
        double e = 0.0, r = 0.0;
        for (int y = 0; y < NY; y++)
            for (int x = 0; x < NX; x++) {
                float ax, ay;
                v(x, y, ax, ay);
                const int i = (y * NX + x) * dimensions;
                const double dx = solver.h_u[i] - ax, dy = solver.h_u[i + 1] - ay;
                e += dx * dx + dy * dy;
                r += double(ax) * ax + double(ay) * ay;
            }
        return float(std::sqrt(e / r) * 100.0);
    }
};

#endif
