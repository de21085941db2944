// lbm3d_emu.cpp — host emulation of the D3Q27 step of liblbm_hip.so (test infrastructure).
//
// Runs the product's per-thread body (cuda-lbm_amd/csrc/lbm3d_step.hpp, __host__ __device__)
// over the same (x, y, z) grid the kernel covers, with population planes packed back to back
// (plane = N, no padding) so that any out-of-range index lands outside an exact-size
// allocation. Built two ways:
//   liblbm3d_emu.so — driven from the CPU tests (parity of the kernel logic vs the oracle);
//   lbm3d_emu_asan  — AddressSanitizer executable running a matrix of small cases.
#include "../../cuda-lbm_amd/csrc/lbm3d_step.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace lbm3;

namespace {

template <int COLL, int STORE>
int run3(const lbm3d_config& c, const float* f_in, const uint8_t* flags, const float* force_aos, int nsteps,
         float* f_out, float* rho_out, float* u_aos_out) {
    using St = Store3<STORE>;
    using T = typename St::T;
    const long long N = (long long)c.nx * c.ny * c.nz;
    std::vector<T> a(size_t(Q * N)), b(size_t(Q * N));
    for (int q = 0; q < Q; q++)
        for (long long n = 0; n < N; n++) a[size_t(q * N + n)] = b[size_t(q * N + n)] = St::enc(f_in[q * N + n], q, c.rho0);
    std::vector<float> force;
    if (force_aos) {
        force.resize(size_t(3 * N));
        for (long long n = 0; n < N; n++)
            for (int k = 0; k < 3; k++) force[size_t(k * N + n)] = force_aos[3 * n + k];
    }
    std::vector<float> rho(static_cast<size_t>(N)), u(static_cast<size_t>(3 * N));
    Step3Params p{};
    p.flags = flags;
    p.force = force_aos ? force.data() : nullptr;
    p.nx = c.nx;
    p.ny = c.ny;
    p.nz = c.nz;
    p.px = c.periodic_x != 0;
    p.py = c.periodic_y != 0;
    p.pz = c.periodic_z != 0;
    p.plane = N;
    p.omega = c.omega;
    p.rho0 = c.rho0;
    p.fx = c.force[0];
    p.fy = c.force[1];
    p.fz = c.force[2];
    p.lin_ec = (c.quirks & LBM_QUIRK_EQ3D_LINEAR) ? 1.0f : 3.0f;
    for (int k = 0; k < 5; k++) p.mrt.s[k] = c.S[4 + k];
    p.mrt.literal = (c.quirks & LBM_QUIRK_MRT3_LITERAL) ? 1 : 0;
    p.u_inlet = c.u_inlet;
    p.guo_own = (c.quirks & LBM_QUIRK_GUO3D_OWN_MOMENTS) ? 1 : 0;
    bool guo = false;
    if (flags)
        for (long long n = 0; n < N; n++) guo = guo || flags[n] == LBM_GUO_VELOCITY_INLET || flags[n] == LBM_GUO_PRESSURE_OUTLET || flags[n] == LBM_ZG_OUTFLOW ||
                           flags[n] == LBM_REGULARIZED_BOUNCE_BACK || flags[n] == LBM_REGULARIZED_INLET_LEFT;
    p.guo = guo ? 1 : 0;
    p.guo_ring_outlet = (guo && (c.nx - 1) % 64 == 0) ? 1 : 0;
    p.stale_rest = (c.quirks & LBM_QUIRK_STALE_REST) ? 1 : 0;
    p.wx = wave_cols(c.nx);
    p.nbx = (c.nx + 64 * p.wx - 1) / (64 * p.wx);
    p.nby = (c.ny + 4 / p.wx - 1) / (4 / p.wx);
    T* cur = a.data();
    T* nxt = b.data();
    for (int s = 0; s < nsteps; s++) {
        p.src = cur;
        p.dst = nxt;
        const bool last = s == nsteps - 1;
        p.out_rho = last ? rho.data() : nullptr;
        p.out_u = last ? u.data() : nullptr;
        // the kernel's grid: every workgroup (in the kernel's XCD order) of 256 threads, idle
        // lanes return; each node must be visited exactly once
        const long long nb = (long long)p.nbx * p.nby * p.nz, per = nb / 8;
        std::vector<uint8_t> seen(size_t(N), 0);
        if (guo) {  // the ring launch first, as on the GPU
            const long long pn = (long long)p.ny * p.nz;
            for (long long i = 0; i < 2 * pn; i++) {
                const int side = int(i / pn);
                const long long r = i - side * pn;
                const int x = side ? p.nx - 1 : 0, y = int(r % p.ny), z = int(r / p.ny);
                if (!guo_ring(p, x, y, z)) continue;
                if (seen[size_t((z * p.ny + y) * p.nx + x)]++) std::abort();
                step3_node<COLL, STORE, false, 2>(p, x, y, z);
            }
        }
        for (long long b = 0; b < nb; b++) {
            const long long L = b < per * 8 ? (b % 8) * per + b / 8 : b;
            // a wavefront's loads precede its stores: within each wavefront the Guo and ZG lanes (which
            // read a neighbouring lane's stale slots) run before the other lanes
            for (int wv = 0; wv < 4; wv++)
                for (int pass = 0; pass < 2; pass++)
                    for (int t = 64 * wv; t < 64 * wv + 64; t++) {
                        int x, y, z;
                        thread_node(p, L, t, x, y, z);
                        if (x >= p.nx || y >= p.ny) continue;
                        if (guo && guo_ring(p, x, y, z)) continue;
                        if (z < 0 || z >= p.nz || x < 0 || y < 0) std::abort();
                        const uint8_t fl = flags ? flags[(z * p.ny + y) * p.nx + x] : uint8_t(LBM_FLUID);
                        const bool gl = fl == LBM_GUO_VELOCITY_INLET || fl == LBM_GUO_PRESSURE_OUTLET || fl == LBM_ZG_OUTFLOW;
                        if ((pass == 0) != gl) continue;
                        if (seen[size_t((z * p.ny + y) * p.nx + x)]++) std::abort();
                        if (guo) step3_node<COLL, STORE, false, 2>(p, x, y, z);
                        else step3_node<COLL, STORE, false>(p, x, y, z);
                    }
        }
        for (long long n = 0; n < N; n++)
            if (seen[size_t(n)] != 1) std::abort();
        T* t = cur;
        cur = nxt;
        nxt = t;
    }
    for (int q = 0; q < Q; q++)
        for (long long n = 0; n < N; n++) f_out[q * N + n] = St::dec(cur[size_t(q * N + n)], q, c.rho0);
    if (nsteps > 0) {
        if (rho_out) std::memcpy(rho_out, rho.data(), size_t(N) * 4);
        if (u_aos_out)
            for (long long n = 0; n < N; n++)
                for (int k = 0; k < 3; k++) u_aos_out[3 * n + k] = u[size_t(k * N + n)];
    }
    return 0;
}

}  // namespace

extern "C" {

// f_in / f_out: FP32 SoA [27][N]; both buffers start from f_in (as lbm3d_set_pops does).
int emu3d_run(const lbm3d_config* c, const float* f_in, const uint8_t* flags, const float* force_aos, int nsteps,
              float* f_out, float* rho_out, float* u_aos_out) {
    if (c->collision == LBM_CM) {
        switch (c->storage) {
            case LBM_F32: return run3<LBM_CM, LBM_F32>(*c, f_in, flags, force_aos, nsteps, f_out, rho_out, u_aos_out);
            case LBM_F16_SHIFTED: return run3<LBM_CM, LBM_F16_SHIFTED>(*c, f_in, flags, force_aos, nsteps, f_out, rho_out, u_aos_out);
            default: return run3<LBM_CM, LBM_BF16_SHIFTED>(*c, f_in, flags, force_aos, nsteps, f_out, rho_out, u_aos_out);
        }
    }
    if (c->collision == LBM_MRT) {
        switch (c->storage) {
            case LBM_F32: return run3<LBM_MRT, LBM_F32>(*c, f_in, flags, force_aos, nsteps, f_out, rho_out, u_aos_out);
            case LBM_F16_SHIFTED: return run3<LBM_MRT, LBM_F16_SHIFTED>(*c, f_in, flags, force_aos, nsteps, f_out, rho_out, u_aos_out);
            default: return run3<LBM_MRT, LBM_BF16_SHIFTED>(*c, f_in, flags, force_aos, nsteps, f_out, rho_out, u_aos_out);
        }
    }
    switch (c->storage) {
        case LBM_F32: return run3<LBM_BGK, LBM_F32>(*c, f_in, flags, force_aos, nsteps, f_out, rho_out, u_aos_out);
        case LBM_F16_SHIFTED: return run3<LBM_BGK, LBM_F16_SHIFTED>(*c, f_in, flags, force_aos, nsteps, f_out, rho_out, u_aos_out);
        default: return run3<LBM_BGK, LBM_BF16_SHIFTED>(*c, f_in, flags, force_aos, nsteps, f_out, rho_out, u_aos_out);
    }
}

}  // extern "C"

#ifdef EMU_MAIN
// Bounds matrix: every access of the step body on exact-size buffers under ASan, over grid
// shapes below, at and above one wavefront row, every periodicity pattern, bounce-back on
// all faces, a force field, every storage and both quirks.
int main() {
    const int shapes[][3] = {{3, 3, 3}, {5, 4, 3}, {64, 3, 4}, {65, 5, 3}, {130, 6, 4}, {7, 9, 11}, {193, 3, 5}, {260, 5, 3},
                             {129, 4, 3}};
    int cases = 0;
    for (const auto& s : shapes)
        for (int per = 0; per < 8; per++)
            for (int st = 0; st <= 2; st++)
                for (int var = 0; var < 6; var++) {
                    if (var == 5 && (per & 1)) continue;  // Guo flags need a non-periodic x axis
                    lbm3d_config c;
                    std::memset(&c, 0, sizeof c);
                    c.nx = s[0];
                    c.ny = s[1];
                    c.nz = s[2];
                    c.collision = var == 3 ? LBM_MRT : (var == 4 || var == 5 ? LBM_CM : LBM_BGK);
                    c.u_inlet = 0.03f;
                    c.omega = 1.3f;
                    for (int k = 0; k < 27; k++) c.S[k] = k >= 4 && k <= 8 ? 1.3f : 1.0f;
                    c.rho0 = 1.0f;
                    c.storage = st;
                    c.periodic_x = per & 1;
                    c.periodic_y = (per >> 1) & 1;
                    c.periodic_z = (per >> 2) & 1;
                    c.force[0] = 1e-5f;
                    c.quirks = var == 2 ? (LBM_QUIRK_STALE_REST | LBM_QUIRK_EQ3D_LINEAR)
                                        : (var == 3 ? LBM_QUIRK_MRT3_LITERAL : (var == 5 && st == 1 ? LBM_QUIRK_GUO3D_OWN_MOMENTS : 0));
                    const long long N = (long long)c.nx * c.ny * c.nz;
                    std::vector<float> f(static_cast<size_t>(Q * N)), out(static_cast<size_t>(Q * N));
                    std::vector<float> rho(static_cast<size_t>(N)), u(static_cast<size_t>(3 * N));
                    for (int q = 0; q < Q; q++)
                        for (long long n = 0; n < N; n++) f[size_t(q * N + n)] = w3(q) * (1.0f + 0.01f * float((n * 7 + q) % 13));
                    std::vector<uint8_t> flags(size_t(N), 0);
                    for (int z = 0; z < c.nz; z++)
                        for (int y = 0; y < c.ny; y++)
                            for (int x = 0; x < c.nx; x++)
                                if (x == 0 || y == 0 || z == 0 || x == c.nx - 1 || y == c.ny - 1 || z == c.nz - 1)
                                    flags[size_t((z * c.ny + y) * c.nx + x)] =
                                        var == 5 && x == 0 ? LBM_GUO_VELOCITY_INLET
                                                           : (var == 5 && x == c.nx - 1 ? LBM_GUO_PRESSURE_OUTLET : LBM_BOUNCE_BACK);
                    std::vector<float> force(size_t(3 * N), 2e-6f);
                    emu3d_run(&c, f.data(), var >= 1 ? flags.data() : nullptr, var == 1 || var >= 3 ? force.data() : nullptr, 3,
                              out.data(), rho.data(), u.data());
                    cases++;
                }
    std::printf("emu3d bounds matrix: %d cases ok\n", cases);
    return 0;
}
#endif
