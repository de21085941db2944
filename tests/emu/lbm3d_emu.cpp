// lbm3d_emu.cpp — host emulation of the D3Q27 step of liblbm_hip.so (test infrastructure).
//
// Runs the product's per-thread body (cuda-lbm_amd/csrc/lbm3d_step.hpp, __host__ __device__)
// over the same (x, y, z) grid the kernel covers, with population planes packed back to back
// (plane = N, no padding) so that any out-of-range index lands outside an exact-size
// allocation. Built two ways:
//   liblbm3d_emu.so — driven from the CPU tests (parity of the kernel logic vs the oracle),
//                     with the IBM pass (lbm3d_ibm.hpp) over the same band tables;
//   lbm3d_emu_asan  — AddressSanitizer executable running a matrix of small cases.
#include "../../cuda-lbm_amd/csrc/lbm3d_ibm.hpp"
#include "../../cuda-lbm_amd/csrc/lbm_ibm_tables.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace lbm3;

namespace {

struct Markers {
    const float* xyz = nullptr;
    const float* ds = nullptr;
    int n = 0, iters = 3;
    float* lag_out = nullptr;  // AoS [n][3], the last step's summed Lagrangian forces
};

template <int COLL, int STORE>
int run3(const lbm3d_config& c, const float* f_in, const uint8_t* flags, const float* force_aos, int nsteps,
         float* f_out, float* rho_out, float* u_aos_out, const Markers& mk = Markers()) {
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
    std::vector<uint8_t> fl;  // the device copy's flags, with the pre-launch marks
    if (flags) fl.assign(flags, flags + N);
    p.flags = flags ? fl.data() : nullptr;
    p.force = force_aos ? force.data() : nullptr;
    p.fbox[1] = c.nx;
    p.fbox[3] = c.ny;
    p.fbox[5] = c.nz;
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
        for (long long n = 0; n < N; n++)
            guo = guo || flags[n] == LBM_GUO_VELOCITY_INLET || flags[n] == LBM_GUO_PRESSURE_OUTLET || flags[n] == LBM_ZG_OUTFLOW ||
                  flags[n] == LBM_REGULARIZED_BOUNCE_BACK || flags[n] == LBM_REGULARIZED_INLET_LEFT ||
                  flags[n] == LBM_REGULARIZED_OUTLET || flags[n] == LBM_EXTRAPOLATED_CORNER_EDGE;
    p.guo = guo ? 1 : 0;
    p.corner_uz = (c.quirks & LBM_QUIRK_CORNER3D_UZ) ? 1 : 0;
    p.stale_rest = (c.quirks & LBM_QUIRK_STALE_REST) ? 1 : 0;
    p.wx = wave_cols(c.nx);
    p.nbx = (c.nx + 64 * p.wx - 1) / (64 * p.wx);
    p.nby = (c.ny + 4 / p.wx - 1) / (4 / p.wx);
    // the pre-launch nodes in launch order, marked as lbm3d_set_flags marks them
    std::vector<long long> pre;
    if (flags) {
        std::vector<std::vector<int>> levels;
        if (pre_launch_levels(p, flags, levels)) return -2;
        for (size_t k = levels.size(); k-- > 0;)
            for (int n : levels[k]) {
                pre.push_back(n);
                fl[size_t(n)] = uint8_t(fl[size_t(n)] | kPreLaunch);
            }
    }
    // IBM: the library's tables, band arrays of exact size, the step's force array (plane = N)
    lbm::IbmTables3 tb;
    std::vector<float> brho, bu, bf, lag, lag_sum, total;
    int box[6] = {c.nx, 0, c.ny, 0, c.nz, 0};
    if (mk.n > 0) {
        if (lbm::ibm_build_tables3(c.nx, c.ny, c.nz, (c.quirks & LBM_QUIRK_IBM_NO_DS) != 0, mk.xyz, mk.ds, mk.n, tb)) return -3;
        const size_t nb = tb.band_node.size();
        brho.resize(nb);
        bu.resize(3 * nb);
        bf.resize(3 * nb);
        lag.resize(3 * size_t(mk.n));
        lag_sum.resize(3 * size_t(mk.n));
        total.resize(size_t(3 * N));
        for (long long n = 0; n < N; n++)
            for (int k = 0; k < 3; k++) total[size_t(k * N + n)] = force_aos ? force[size_t(k * N + n)] : c.force[k];
        for (int v : tb.band_node) {
            const int z = v / (c.nx * c.ny), y = (v / c.nx) % c.ny, x = v % c.nx;
            box[0] = std::min(box[0], x);
            box[1] = std::max(box[1], x + 1);
            box[2] = std::min(box[2], y);
            box[3] = std::max(box[3], y + 1);
            box[4] = std::min(box[4], z);
            box[5] = std::max(box[5], z + 1);
        }
    }
    T* cur = a.data();
    T* nxt = b.data();
    for (int s = 0; s < nsteps; s++) {
        p.src = cur;
        p.dst = nxt;
        if (mk.n > 0) {  // the IBM pass, as lbm3d_ctx::ibm_pass launches it
            Ibm3Params q{};
            q.sp = p;
            q.sp.force = force_aos ? force.data() : nullptr;
            q.n_markers = mk.n;
            q.n_band = int(tb.band_node.size());
            q.clamp = (c.quirks & LBM_QUIRK_IBM_CLAMP) ? 1 : 0;
            q.band_node = tb.band_node.data();
            q.mk_ptr = tb.mk_ptr.data();
            q.mk_band = tb.mk_band.data();
            q.mk_w = tb.mk_w.data();
            q.bn_ptr = tb.bn_ptr.data();
            q.bn_mk = tb.bn_mk.data();
            q.bn_w = tb.bn_w.data();
            q.ext = force_aos ? force.data() : nullptr;
            q.brho = brho.data();
            q.bu = bu.data();
            q.bf = bf.data();
            q.lag = lag.data();
            q.lag_sum = lag_sum.data();
            q.force_total = total.data();
            for (int b = 0; b < q.n_band; b++) {
                if (guo) ibm3_gather<STORE, false, 2>(q, b);
                else ibm3_gather<STORE, false, 0>(q, b);
            }
            for (int it = 0; it < mk.iters; it++) {
                q.first = it == 0;
                q.last = it == mk.iters - 1;
                for (int l = 0; l < mk.n; l++) {  // the wavefront's lanes and butterfly, lane by lane
                    float sums[64][4], next[64][4];
                    for (int lane = 0; lane < 64; lane++) ibm3_interp_partial(q, l, lane, 64, sums[lane]);
                    for (int off = 32; off > 0; off >>= 1) {
                        for (int lane = 0; lane < 64; lane++) {
                            for (int c = 0; c < 4; c++) next[lane][c] = sums[lane][c];
                            ibm3_combine(next[lane], sums[lane ^ off]);
                        }
                        std::memcpy(sums, next, sizeof sums);
                    }
                    ibm3_interp_finish(q, l, sums[0]);
                }
                for (int b = 0; b < q.n_band; b++) ibm3_spread(q, b);
            }
            p.force = total.data();
            if (!force_aos) std::copy(box, box + 6, p.fbox);
        }
        const bool last = s == nsteps - 1;
        p.out_rho = last ? rho.data() : nullptr;
        p.out_u = last ? u.data() : nullptr;
        // the kernel's grid: every workgroup (in the kernel's XCD order) of 256 threads, idle
        // lanes return; each node must be visited exactly once
        const long long nb = (long long)p.nbx * p.nby * p.nz, per = nb / 8;
        std::vector<uint8_t> seen(size_t(N), 0);
        for (long long n : pre) {  // the pre-launch first, as on the GPU
            const int x = int(n % p.nx), y = int((n / p.nx) % p.ny), z = int(n / ((long long)p.nx * p.ny));
            if (seen[size_t(n)]++) std::abort();
            step3_node<COLL, STORE, false, 2>(p, x, y, z);
        }
        // A wavefront issues its stores after all the loads they depend on (SIMD: every lane's
        // pulls, the neighbour pulls of divergent branches included, precede the common stores).
        // Emulated per wavefront: each lane runs on a shadow of the destination whose wavefront
        // nodes hold their values from before the wavefront, and its results are then copied to
        // the destination. (Reads of other wavefronts' stale slots are excluded by the
        // pre-launch rule.)
        std::vector<T> shadow(nxt, nxt + size_t(Q) * N);
        std::vector<long long> wnodes;
        std::vector<T> before;
        for (long long b = 0; b < nb; b++) {
            const long long L = b < per * 8 ? (b % 8) * per + b / 8 : b;
            for (int wv = 0; wv < 4; wv++) {
                wnodes.clear();
                for (int t = 64 * wv; t < 64 * wv + 64; t++) {
                    int x, y, z;
                    thread_node(p, L, t, x, y, z);
                    if (x >= p.nx || y >= p.ny) continue;
                    if (z < 0 || z >= p.nz || x < 0 || y < 0) std::abort();
                    const long long n = (z * (long long)p.ny + y) * p.nx + x;
                    const uint8_t f = flags ? fl[size_t(n)] : uint8_t(LBM_FLUID);
                    if (guo && (f & kPreLaunch)) continue;
                    wnodes.push_back(n);
                }
                before.resize(wnodes.size() * Q);
                for (size_t k = 0; k < wnodes.size(); k++)
                    for (int q = 0; q < Q; q++) before[k * Q + q] = shadow[size_t(q * N + wnodes[k])] = nxt[q * N + wnodes[k]];
                Step3Params ps = p;
                ps.dst = shadow.data();
                for (size_t k = 0; k < wnodes.size(); k++) {
                    const long long n = wnodes[k];
                    const int x = int(n % p.nx), y = int((n / p.nx) % p.ny), z = int(n / ((long long)p.nx * p.ny));
                    if (seen[size_t(n)]++) std::abort();
                    if (guo) step3_node<COLL, STORE, false, 2>(ps, x, y, z);
                    else step3_node<COLL, STORE, false>(ps, x, y, z);
                    for (int q = 0; q < Q; q++) {  // the lane's stores, then the shadow back to "before"
                        nxt[q * N + n] = shadow[size_t(q * N + n)];
                        shadow[size_t(q * N + n)] = before[k * Q + q];
                    }
                }
                for (size_t k = 0; k < wnodes.size(); k++)
                    for (int q = 0; q < Q; q++) shadow[size_t(q * N + wnodes[k])] = nxt[q * N + wnodes[k]];
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
    if (mk.lag_out)
        for (int l = 0; l < mk.n; l++)
            for (int d = 0; d < 3; d++) mk.lag_out[3 * l + d] = lag_sum[size_t(d) * mk.n + l];
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

// emu3d_run with static IBM markers AoS [n][3] (ds may be null); lag_out AoS [n][3].
int emu3d_run_ibm(const lbm3d_config* c, const float* f_in, const uint8_t* flags, const float* force_aos,
                  const float* xyz, const float* ds, int n, int iters, int nsteps, float* f_out, float* rho_out,
                  float* u_aos_out, float* lag_out) {
    Markers mk;
    mk.xyz = xyz;
    mk.ds = ds;
    mk.n = n;
    mk.iters = iters;
    mk.lag_out = lag_out;
    if (c->collision == LBM_CM) {
        switch (c->storage) {
            case LBM_F32: return run3<LBM_CM, LBM_F32>(*c, f_in, flags, force_aos, nsteps, f_out, rho_out, u_aos_out, mk);
            case LBM_F16_SHIFTED: return run3<LBM_CM, LBM_F16_SHIFTED>(*c, f_in, flags, force_aos, nsteps, f_out, rho_out, u_aos_out, mk);
            default: return run3<LBM_CM, LBM_BF16_SHIFTED>(*c, f_in, flags, force_aos, nsteps, f_out, rho_out, u_aos_out, mk);
        }
    }
    if (c->collision == LBM_MRT) {
        switch (c->storage) {
            case LBM_F32: return run3<LBM_MRT, LBM_F32>(*c, f_in, flags, force_aos, nsteps, f_out, rho_out, u_aos_out, mk);
            case LBM_F16_SHIFTED: return run3<LBM_MRT, LBM_F16_SHIFTED>(*c, f_in, flags, force_aos, nsteps, f_out, rho_out, u_aos_out, mk);
            default: return run3<LBM_MRT, LBM_BF16_SHIFTED>(*c, f_in, flags, force_aos, nsteps, f_out, rho_out, u_aos_out, mk);
        }
    }
    switch (c->storage) {
        case LBM_F32: return run3<LBM_BGK, LBM_F32>(*c, f_in, flags, force_aos, nsteps, f_out, rho_out, u_aos_out, mk);
        case LBM_F16_SHIFTED: return run3<LBM_BGK, LBM_F16_SHIFTED>(*c, f_in, flags, force_aos, nsteps, f_out, rho_out, u_aos_out, mk);
        default: return run3<LBM_BGK, LBM_BF16_SHIFTED>(*c, f_in, flags, force_aos, nsteps, f_out, rho_out, u_aos_out, mk);
    }
}

}  // extern "C"

#ifdef EMU_MAIN
// Bounds matrix: every access of the step body and the IBM pass on exact-size buffers under
// ASan, over grid shapes below, at and above one wavefront row, every periodicity pattern,
// bounce-back on all faces, a force field, every storage and the quirks.
int main() {
    const int shapes[][3] = {{3, 3, 3}, {5, 4, 3}, {64, 3, 4}, {65, 5, 3}, {130, 6, 4}, {7, 9, 11}, {193, 3, 5}, {260, 5, 3},
                             {129, 4, 3}};
    int cases = 0;
    for (const auto& s : shapes)
        for (int per = 0; per < 8; per++)
            for (int st = 0; st <= 2; st++)
                for (int var = 0; var < 10; var++) {
                    // Guo and regularized-outlet flags need a non-periodic x axis
                    if ((var == 5 || var == 6) && (per & 1)) continue;
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
                    // 6: regularized inlet and outlet; 7: ZG on every face node; 8: corner-edge
                    // extrapolation on edges and corners (its velocity quirk with FP16 storage)
                    for (int z = 0; z < c.nz; z++)
                        for (int y = 0; y < c.ny; y++)
                            for (int x = 0; x < c.nx; x++) {
                                const int faces = (x == 0) + (y == 0) + (z == 0) + (x == c.nx - 1) + (y == c.ny - 1) + (z == c.nz - 1);
                                uint8_t& fv = flags[size_t((z * c.ny + y) * c.nx + x)];
                                if (var == 6 && x == 0) fv = LBM_REGULARIZED_INLET_LEFT;
                                if (var == 6 && x == c.nx - 1) fv = LBM_REGULARIZED_OUTLET;
                                if (var == 7 && faces > 0) fv = LBM_ZG_OUTFLOW;
                                if (var == 8 && faces >= 2) fv = LBM_EXTRAPOLATED_CORNER_EDGE;
                            }
                    if (var == 8 && st == 1) c.quirks |= LBM_QUIRK_CORNER3D_UZ;
                    std::vector<float> force(size_t(3 * N), 2e-6f);
                    // 9: IBM markers at the grid's corners, centre and beyond its faces, with walls
                    // (bounce-back on every face) and a force field
                    const float mk[] = {0.2f, 0.3f, 0.1f, float(c.nx) - 0.6f, float(c.ny) - 0.4f, float(c.nz) - 0.7f,
                                        0.5f * float(c.nx), 0.5f * float(c.ny), 0.5f * float(c.nz), -1.5f, 1.0f, 1.0f,
                                        float(c.nx) + 1.2f, float(c.ny) + 0.5f, 0.0f};
                    std::vector<float> lag(15);
                    const int rc = var == 9 ? emu3d_run_ibm(&c, f.data(), flags.data(), force.data(), mk, nullptr, 5, 2, 3,
                                                            out.data(), rho.data(), u.data(), lag.data())
                                            : emu3d_run(&c, f.data(), var >= 1 ? flags.data() : nullptr,
                                                        var == 1 || var >= 3 ? force.data() : nullptr, 3, out.data(), rho.data(),
                                                        u.data());
                    if (rc != 0) {
                        std::printf("emu3d_run rejected shape %dx%dx%d per %d var %d\n", c.nx, c.ny, c.nz, per, var);
                        return 1;
                    }
                    cases++;
                }
    std::printf("emu3d bounds matrix: %d cases ok\n", cases);
    return 0;
}
#endif
