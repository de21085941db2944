// lbm_run — C++ driver for a compile-time scenario (the role of src/main.cu): allocate,
// init, run the fused steps through liblbm_hip.so and print the scenario's validation.
// Build with -DRUN_TAYLOR_GREEN or -DRUN_LID_DRIVEN, or -DSCENARIO_HEADER='"path"'
// -DSCENARIO_TYPE=Name, and -DNX=.. -DNY=.. .
// Usage: lbm_run [steps [output_dir]]; with an output directory the final fields are also
// written as .bin dumps and a .vti file (save_macroscopics / save_vtk).
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "core/lbm.cuh"

#if defined(RUN_TAYLOR_GREEN)
#include "taylorGreenPeriodic.cuh"
using Scenario = TaylorGreenPeriodic;
#elif defined(RUN_LID_DRIVEN)
#include "lidDrivenMovingWall.cuh"
using Scenario = LidDrivenMovingWall;
#elif defined(SCENARIO_HEADER)
// any scenario header, included from where it lies (e.g. one of the reference's own)
#include SCENARIO_HEADER
using Scenario = SCENARIO_TYPE;
#else
#error "choose a scenario: -DRUN_TAYLOR_GREEN or -DRUN_LID_DRIVEN"
#endif

int main(int argc, char** argv) {
    const int steps = argc > 1 ? std::atoi(argv[1]) : 1000;
    const char* out_dir = argc > 2 ? argv[2] : nullptr;
    try {
        LBM<dimensions> lbm;
        lbm.allocate<Scenario>();
        lbm.init<Scenario>();
        auto t0 = std::chrono::steady_clock::now();
        lbm.step<Scenario>(steps);
        lbm.update_macroscopics();
        double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("scenario %s grid %dx%d steps %d wall %.3f s (%.1f MLUPS incl. launch)\n", Scenario::name(), NX, NY,
                    steps, s, double(NX) * NY * steps / s / 1e6);
        if constexpr (Scenario::has_analytical_solution) {
            std::printf("validation error %.6f %%\n", lbm.compute_error<Scenario>());
        } else {
            std::printf("centre u_x(y):");
            for (int y = 0; y < NY; y += NY / 8) std::printf(" %.5f", lbm.h_u[2 * (y * NX + NX / 2)]);
            std::printf("\n");
        }
        if (out_dir) {
            lbm.save_macroscopics(lbm.timestep, out_dir);
            lbm.save_vtk(lbm.timestep, out_dir);
            std::printf("wrote %s/density/density_%d.bin, velocity_%d.bin, vtk/sim_data_%06d.vti\n", out_dir,
                        lbm.timestep, lbm.timestep, lbm.timestep);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "lbm_run: %s\n", e.what());
        return 1;
    }
    return 0;
}
