// lbm_emu.cpp — host emulation of liblbm_hip.so's step sequence (test infrastructure).
//
// Runs the product's own per-thread body (cuda-lbm_amd/csrc/lbm_step.hpp, a __host__
// __device__ function) over the same segment grid on the CPU, with population planes
// packed back to back (plane = N, no padding) so that any out-of-range index the kernel
// would form lands outside an exact-size allocation. Built two ways:
//   liblbm_emu.so  — driven from the CPU tests (parity of the kernel logic vs the oracle);
//   lbm_emu_asan   — AddressSanitizer executable running a matrix of small cases (bounds
//                    check of every access the kernel body makes, before any GPU run).
#include "../../cuda-lbm_amd/csrc/lbm_step.hpp"
#include "../../cuda-lbm_amd/csrc/lbm_adjoint.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

using namespace lbm;

namespace {

// A value no kernel stores: NaN in the storage format (poisons unwritten tile entries).
template <class T>
T poison() {
    if constexpr (sizeof(T) == 4) {
        float f;
        uint32_t b = 0x7fc00000u;
        std::memcpy(&f, &b, 4);
        return T(f);
    } else {
        return T(0x7e00);  // FP16 NaN; a BF16 NaN as well (exponent all ones, mantissa != 0)
    }
}

// One two-step launch: every tile's phase A into a fresh poisoned tile, then its phase B.
// reverse = 1 runs the tiles in reverse order (results must not depend on tile order).
template <int COLL, int STORE, int TY>
void launch2_emu(const StepParams& p, int reverse) {
    using T = typename Elem<STORE>::T;
    const int tiles_x = (p.nx + kT2W - 1) / kT2W, tiles_y = (p.ny + TY - 1) / TY;
    const long long nt = (long long)tiles_x * tiles_y;
    std::vector<T> tile(size_t(Q) * Tile2<TY>::NH);
    for (long long k = 0; k < nt; k++) {
        const long long lb = reverse ? nt - 1 - k : k;
        const int by = int(lb / tiles_x), bx = int(lb - (long long)by * tiles_x);
        std::fill(tile.begin(), tile.end(), poison<T>());
        const int x0 = bx * kT2W;
        const bool interior_x = interior_cols(p, x0);
        for (int t = 0; t < Tile2<TY>::THREADS; t++)
            step2_phase_a<COLL, STORE, TY, true>(p, tile.data(), t, x0, by * TY, interior_x);
        for (int t = 0; t < Tile2<TY>::THREADS; t++) step2_phase_b<COLL, STORE, TY, true>(p, tile.data(), t, x0, by * TY);
    }
}

// One marching-strip two-step launch (W wavefronts, H rows per strip), chunk by chunk: all
// phase-A rows of a chunk, then its phase-B rows, on a poisoned ring.
template <int COLL, int STORE, int W>
void launch2s_emu(const StepParams& p, int H, int reverse) {
    using T = typename Elem<STORE>::T;
    constexpr int R = W + 2;
    const int strips_x = (p.nx + kT2W - 1) / kT2W, strips_y = (p.ny + H - 1) / H;
    const long long ns = (long long)strips_x * strips_y;
    std::vector<T> ring(size_t(Q) * R * kT2L);
    for (long long k0 = 0; k0 < ns; k0++) {
        const long long lb = reverse ? ns - 1 - k0 : k0;
        const int sy = int(lb / strips_x), sx = int(lb - (long long)sy * strips_x);
        const int x0 = sx * kT2W, y0 = sy * H;
        const int hs = p.ny - y0 < H ? p.ny - y0 : H;
        const int K = (hs + 1) / W + 1;
        std::fill(ring.begin(), ring.end(), poison<T>());
        const bool interior_x = interior_cols(p, x0);
        for (int k = 0; k < K; k++) {
            for (int w = 0; w < W; w++)
                for (int l = 0; l < kT2L; l++) {
                    const int i = k * W + w;
                    int gx, gy;
                    if (i > hs + 1) continue;
                    if (!strip_node(p, x0, y0, l, i, gx, gy)) {  // outside a non-periodic domain
                        step2_fill_halo<STORE, true>(p, ring.data() + (i % R) * kT2L + l, R * kT2L, x0 - 1 + l, y0 - 1 + i);
                        continue;
                    }
                    float g[Q];
                    step2_gather<STORE, true>(p, gx, gy, g, interior_x && interior_row(p, y0 - 1 + i));
                    step2s_phase_a<COLL, STORE, true>(p, ring.data(), R, i % R, l, i, x0, y0, hs, gx, gy, g);
                }
            for (int w = 0; w < W; w++)
                for (int l = 0; l < kT2L; l++) {
                    const int j = k * W - 2 + w;
                    if (j >= 0 && j < hs) step2s_phase_b<COLL, STORE, true>(p, ring.data(), R, l, j, x0, y0);
                }
        }
    }
}

// One look-ahead launch (CM<OptimalAdapter>, same-step averages): marching strips of W
// wavefronts and H rows whose phase A is the step (own nodes into p.mid = p.dst) and whose
// phase B sums the next step's moments, one triple per strip into partials (logical order).
template <int COLL, int STORE, int W>
void launch_ahead_emu(const StepParams& p, int H, int reverse, std::vector<double>& partials) {
    using T = typename Elem<STORE>::T;
    constexpr int R = W + 2;
    const int strips_x = (p.nx + kT2W - 1) / kT2W, strips_y = (p.ny + H - 1) / H;
    const long long ns = (long long)strips_x * strips_y;
    partials.assign(size_t(3 * ns), 0.0);
    std::vector<T> ring(size_t(Q) * R * kT2L);
    for (long long k0 = 0; k0 < ns; k0++) {
        const long long lb = reverse ? ns - 1 - k0 : k0;
        const int sy = int(lb / strips_x), sx = int(lb - (long long)sy * strips_x);
        const int x0 = sx * kT2W, y0 = sy * H;
        const int hs = p.ny - y0 < H ? p.ny - y0 : H;
        const int K = (hs + 1) / W + 1;
        std::fill(ring.begin(), ring.end(), poison<T>());
        const bool interior_x = interior_cols(p, x0);
        MomAcc acc{0, 0, 0};
        for (int k = 0; k < K; k++) {
            for (int w = 0; w < W; w++)
                for (int l = 0; l < kT2L; l++) {
                    const int i = k * W + w;
                    int gx, gy;
                    if (i > hs + 1) continue;
                    if (!strip_node(p, x0, y0, l, i, gx, gy)) {
                        step2_fill_halo<STORE, true>(p, ring.data() + (i % R) * kT2L + l, R * kT2L, x0 - 1 + l, y0 - 1 + i);
                        continue;
                    }
                    float g[Q];
                    step2_gather<STORE, true>(p, gx, gy, g, interior_x && interior_row(p, y0 - 1 + i));
                    step2s_phase_a<COLL, STORE, true, true>(p, ring.data(), R, i % R, l, i, x0, y0, hs, gx, gy, g);
                }
            for (int w = 0; w < W; w++)
                for (int l = 0; l < kT2L; l++) {
                    const int j = k * W - 2 + w;
                    if (j >= 0 && j < hs) step2s_phase_b_moments<STORE, true>(p, ring.data(), R, l, j, x0, y0, acc);
                }
        }
        partials[size_t(3 * lb)] = acc.rho;
        partials[size_t(3 * lb + 1)] = acc.jmag;
        partials[size_t(3 * lb + 2)] = acc.pimag;
    }
}

template <int COLL, int STORE, int SEG>
int run_t(const lbm_config& c, const uint8_t* flags, const float* rho, const float* u, const float* force,
          int nsteps, float* f_out, float* rho_out, float* u_out, double* avg_out, int two_ty, int reverse) {
    using St = Elem<STORE>;
    using T = typename St::T;
    const int nx = c.nx, ny = c.ny;
    const long long N = (long long)nx * ny;
    T* buf[2];
    buf[0] = new T[9 * N];
    buf[1] = new T[9 * N];
    for (long long n = 0; n < N; n++) {
        float r = rho[n], ux = u[2 * n], uy = u[2 * n + 1];
        float usq15 = 1.5f * (ux * ux + uy * uy);
        for (int q = 0; q < Q; q++) {
            float fe = feq(q, r, ux, uy, usq15);
            St::store1(buf[0] + q * N + n, q, c.rho0, fe);
            St::store1(buf[1] + q * N + n, q, c.rho0, fe);
        }
    }
    uint8_t* fl = nullptr;
    uint8_t* rc = nullptr;
    bool any = false;
    if (flags) {
        for (long long n = 0; n < N; n++) any = any || flags[n] != LBM_FLUID;
    }
    if (any) {
        fl = new uint8_t[N];
        rc = new uint8_t[ny];
        std::memcpy(fl, flags, N);
        row_classes(fl, nx, ny, rc);
    }
    float* ff = nullptr;
    if (force) {
        ff = new float[2 * N];
        std::memcpy(ff, force, 2 * N * sizeof(float));
    }
    float* out_r = new float[N];
    float* out_x = new float[N];
    float* out_y = new float[N];
    float avg[3] = {0.f, 0.f, 0.f};
    double avg64[3] = {0, 0, 0};
    const bool adapt = c.collision == LBM_CM && c.adapter == LBM_OPTIMAL_ADAPTER;
    bool local = true;
    if (fl)
        for (long long n = 0; n < N; n++)
            local = local && !flag_reads_neighbour(fl[n]);
    const bool two = two_ty != 0 && !adapt && !(c.quirks & LBM_QUIRK_STALE_REST) && local;
    // CM<OptimalAdapter> with same-step averages: look-ahead launches where the library uses them
    const bool ahead = two_ty < 0 && adapt && c.adapter_mode == LBM_ADAPTER_SAME_STEP &&
                       !(c.quirks & LBM_QUIRK_STALE_REST) && local;
    bool ahead_valid = false, ahead_edges = false;
    float ahead_avg[3] = {0.f, 0.f, 0.f};
    double ahead64[3] = {0, 0, 0};
    std::vector<double> partials;
    const int n_edge = edge_count(nx, ny, c.periodic_x, c.periodic_y);
    std::vector<T> edge[2] = {std::vector<T>(size_t(9) * n_edge), std::vector<T>(size_t(9) * n_edge)};
    int edge_cur = 0;
    bool in_store = false;
    int cur = 0;
    for (int s = 0; s < nsteps;) {
        const bool pair = two && s + 2 <= nsteps;
        if (pair && n_edge > 0 && !in_store) {  // state t-1 at the edge nodes -> store
            for (int e = 0; e < n_edge; e++) {
                int x, y;
                edge_node(e, nx, ny, c.periodic_y, x, y);
                for (int q = 0; q < Q; q++) edge[edge_cur][size_t(q) * n_edge + e] = buf[1 - cur][q * N + (long long)y * nx + x];
            }
        }
        if (!pair && in_store) {  // back into the buffer one-step launches read
            for (int e = 0; e < n_edge; e++) {
                int x, y;
                edge_node(e, nx, ny, c.periodic_y, x, y);
                if (edge_slot(x, y, nx, ny, c.periodic_x, c.periodic_y) != e) continue;
                for (int q = 0; q < Q; q++) buf[1 - cur][q * N + (long long)y * nx + x] = edge[edge_cur][size_t(q) * n_edge + e];
            }
            in_store = false;
        }
        StepParams p{};
        p.nx = nx;
        p.ny = ny;
        p.px = c.periodic_x;
        p.py = c.periodic_y;
        p.plane = N;
        p.src = buf[cur];
        p.dst = buf[1 - cur];
        p.flags = fl;
        p.row_class = rc;
        p.omega = c.omega;
        for (int k = 0; k < 9; k++) p.S[k] = c.S[k];
        p.rho0 = c.rho0;
        p.fx = c.force[0];
        p.fy = c.force[1];
        p.force_field = ff;
        p.forced = (c.force[0] != 0.f || c.force[1] != 0.f || ff) ? 1 : 0;
        p.wall_ux = c.wall_u[0];
        p.wall_uy = c.wall_u[1];
        p.u_lid = c.u_lid;
        p.quirks = c.quirks;
        p.avg = adapt ? avg : nullptr;
        bool last = s + (pair ? 2 : 1) == nsteps;
        p.out_rho = last ? out_r : nullptr;
        p.out_ux = last ? out_x : nullptr;
        p.out_uy = last ? out_y : nullptr;
        if (pair) {
            p.edge_prev = n_edge ? edge[edge_cur].data() : nullptr;
            p.edge_next = n_edge ? edge[1 - edge_cur].data() : nullptr;
            p.n_edge = n_edge;
            if (two_ty < 0) {  // marching strips: -(W * 1000 + H)
                const int W = (-two_ty) / 1000, H = (-two_ty) % 1000;
                if (W == 8) launch2s_emu<COLL, STORE, 8>(p, H, reverse);
                else launch2s_emu<COLL, STORE, 4>(p, H, reverse);
            } else if (two_ty == 14) {
                launch2_emu<COLL, STORE, 14>(p, reverse);
            } else {
                launch2_emu<COLL, STORE, 6>(p, reverse);
            }
            if (n_edge > 0) {
                edge_cur = 1 - edge_cur;
                in_store = true;
            }
            cur = 1 - cur;
            s += 2;
            continue;
        }
        p.ntx = (nx + SEG - 1) / SEG;
        const long long nseg = (long long)p.ntx * ny;
        if (ahead) {
            if (!ahead_valid) {  // this step's averages from a moments pass
                MomAcc tot{0, 0, 0};
                for (long long sg = 0; sg < nseg; sg++) step_segment<LBM_BGK, STORE, 1, SEG>(p, sg, tot);
                ahead64[0] = tot.rho / double(N);
                ahead64[1] = tot.jmag / double(N);
                ahead64[2] = tot.pimag / double(N);
                for (int k = 0; k < 3; k++) ahead_avg[k] = float(ahead64[k]);
            }
            if (n_edge > 0 && !ahead_edges) {  // state t-1 at the edge nodes -> store
                for (int e = 0; e < n_edge; e++) {
                    int x, y;
                    edge_node(e, nx, ny, c.periodic_y, x, y);
                    for (int q = 0; q < Q; q++)
                        edge[edge_cur][size_t(q) * n_edge + e] = buf[1 - cur][q * N + (long long)y * nx + x];
                }
            }
            p.mid = p.dst;
            p.avg = ahead_avg;
            p.edge_prev = n_edge ? edge[edge_cur].data() : nullptr;
            p.edge_next = n_edge ? edge[1 - edge_cur].data() : nullptr;
            p.n_edge = n_edge;
            const int H = (-two_ty) % 1000;
            launch_ahead_emu<COLL, STORE, 4>(p, H, reverse, partials);
            // finalize: the averages used become the reported ones, the sums the next ones
            for (int k = 0; k < 3; k++) {
                avg[k] = ahead_avg[k];
                avg64[k] = ahead64[k];
            }
            double a = 0, b = 0, cc = 0;
            for (size_t i = 0; i < partials.size(); i += 3) {
                a += partials[i];
                b += partials[i + 1];
                cc += partials[i + 2];
            }
            ahead64[0] = a / double(N);
            ahead64[1] = b / double(N);
            ahead64[2] = cc / double(N);
            for (int k = 0; k < 3; k++) ahead_avg[k] = float(ahead64[k]);
            if (n_edge > 0) edge_cur = 1 - edge_cur;
            ahead_edges = n_edge > 0;
            ahead_valid = true;
            cur = 1 - cur;
            s += 1;
            continue;
        }
        if (adapt) {
            MomAcc tot{0, 0, 0};
            for (long long sg = 0; sg < nseg; sg++) step_segment<LBM_BGK, STORE, 1, SEG>(p, sg, tot);
            avg64[0] = tot.rho / double(N);
            avg64[1] = tot.jmag / double(N);
            avg64[2] = tot.pimag / double(N);
            for (int k = 0; k < 3; k++) avg[k] = float(avg64[k]);
        }
        MomAcc dummy{0, 0, 0};
        for (long long sg = 0; sg < nseg; sg++) step_segment<COLL, STORE, 0, SEG>(p, sg, dummy);
        cur = 1 - cur;
        s += 1;
    }
    if (f_out)
        for (int q = 0; q < Q; q++)
            for (long long n = 0; n < N; n++) f_out[q * N + n] = St::load1(buf[cur] + q * N + n, q, c.rho0);
    if (nsteps > 0) {
        for (long long n = 0; n < N; n++) {
            if (rho_out) rho_out[n] = out_r[n];
            if (u_out) {
                u_out[2 * n] = out_x[n];
                u_out[2 * n + 1] = out_y[n];
            }
        }
    }
    if (avg_out)
        for (int k = 0; k < 3; k++) avg_out[k] = avg64[k];
    delete[] buf[0];
    delete[] buf[1];
    delete[] fl;
    delete[] rc;
    delete[] ff;
    delete[] out_r;
    delete[] out_x;
    delete[] out_y;
    return 0;
}

template <int STORE, int SEG>
int run_w(const lbm_config& c, const uint8_t* flags, const float* rho, const float* u, const float* force, int n,
          float* f, float* r, float* uu, double* a, int two_ty, int rev) {
    if (c.collision == LBM_BGK) return run_t<LBM_BGK, STORE, SEG>(c, flags, rho, u, force, n, f, r, uu, a, two_ty, rev);
    if (c.collision == LBM_MRT) return run_t<LBM_MRT, STORE, SEG>(c, flags, rho, u, force, n, f, r, uu, a, two_ty, rev);
    return run_t<LBM_CM, STORE, SEG>(c, flags, rho, u, force, n, f, r, uu, a, two_ty, rev);
}

template <int STORE>
int run_s(const lbm_config& c, int seg, const uint8_t* flags, const float* rho, const float* u, const float* force,
          int n, float* f, float* r, float* uu, double* a, int two_ty, int rev) {
    if (seg == kSegWide) return run_w<STORE, kSegWide>(c, flags, rho, u, force, n, f, r, uu, a, two_ty, rev);
    if (seg == kSegNarrow<STORE>)
        return run_w<STORE, kSegNarrow<STORE>>(c, flags, rho, u, force, n, f, r, uu, a, two_ty, rev);
    return -1;
}

}  // namespace

// seg: nodes per thread of one-step launches, the wide (4) or the storage's narrow width
// (FP32 1, 16-bit 2); 0 selects the wide one. two_ty: 0 one step per launch; 6 or 14: two-step
// tile launches with that tile height, -(W * 1000 + H): two-step marching strips of W
// wavefronts and H rows; both where the configuration allows them (as the library does).
// reverse: run two-step tiles or strips in reverse order.
extern "C" int emu_run(const lbm_config* c, int seg, int two_ty, int reverse, const uint8_t* flags, const float* rho,
                       const float* u, const float* force, int nsteps, float* f_out, float* rho_out, float* u_out,
                       double* avg_out) {
    if (seg == 0) seg = kSegWide;
    if (two_ty < 0) {
        const int W = (-two_ty) / 1000, H = (-two_ty) % 1000;
        if ((W != 4 && W != 8) || H < 1) return -1;
    } else if (two_ty != 0 && two_ty != 6 && two_ty != 14) {
        return -1;
    }
    if (c->storage == LBM_F16_SHIFTED)
        return run_s<LBM_F16_SHIFTED>(*c, seg, flags, rho, u, force, nsteps, f_out, rho_out, u_out, avg_out, two_ty,
                                      reverse);
    if (c->storage == LBM_BF16_SHIFTED)
        return run_s<LBM_BF16_SHIFTED>(*c, seg, flags, rho, u, force, nsteps, f_out, rho_out, u_out, avg_out, two_ty,
                                       reverse);
    return run_s<LBM_F32>(*c, seg, flags, rho, u, force, nsteps, f_out, rho_out, u_out, avg_out, two_ty, reverse);
}

// One one-step launch over every node (f_full) against the same step split at an IBM box
// (x range a multiple of 4 nodes, or ending at nx): every node outside it, then the box
// alone (f_split, on a poisoned destination). FP32 storage, periodic, all FLUID, with a
// force field inside the box standing in for the IBM force. seg: nodes per thread (1 or 4).
template <int COLL, int SEG>
void region_run(StepParams p, float* dst_full, float* dst_split, long long N, int bx0, int bx1, int by0, int by1) {
    MomAcc dummy{0.0, 0.0, 0.0};
    p.ntx = (p.nx + SEG - 1) / SEG;
    const long long nseg = (long long)p.ntx * p.ny;
    p.dst = dst_full;
    p.region = 0;
    for (long long sg = 0; sg < nseg; sg++) step_segment<COLL, LBM_F32, 0, SEG>(p, sg, dummy);
    p.dst = dst_split;
    p.region = 1;
    for (long long sg = 0; sg < nseg; sg++) step_segment<COLL, LBM_F32, 0, SEG>(p, sg, dummy);
    p.region = 2;
    const long long nbox = (long long)((bx1 - bx0 + SEG - 1) / SEG) * (by1 - by0);
    for (long long sg = 0; sg < nbox; sg++) step_segment<COLL, LBM_F32, 0, SEG>(p, sg, dummy);
    (void)N;
}

extern "C" int emu_region_split(const lbm_config* c, int seg, int bx0, int bx1, int by0, int by1, const float* f_in,
                                const float* box_force, float* f_full, float* f_split) {
    if (c->storage != LBM_F32 || !c->periodic_x || !c->periodic_y) return -1;
    if (bx0 % 4 != 0 || (bx1 % 4 != 0 && bx1 != c->nx) || bx0 >= bx1 || by0 >= by1 || bx1 > c->nx || by1 > c->ny)
        return -2;
    const long long N = (long long)c->nx * c->ny;
    std::vector<float> a(f_in, f_in + 9 * N), full(size_t(9) * N, poison<float>()), split(size_t(9) * N, poison<float>());
    std::vector<float> ibm(box_force, box_force + 2 * N);
    StepParams p{};
    p.nx = c->nx;
    p.ny = c->ny;
    p.px = 1;
    p.py = 1;
    p.plane = N;
    p.src = a.data();
    p.omega = c->omega;
    for (int k = 0; k < 9; k++) p.S[k] = c->S[k];
    p.rho0 = c->rho0;
    p.fx = c->force[0];
    p.fy = c->force[1];
    p.forced = 1;
    p.quirks = c->quirks;
    p.ibm_force = ibm.data();
    p.ibm_x0 = bx0;
    p.ibm_x1 = bx1;
    p.ibm_y0 = by0;
    p.ibm_y1 = by1;
    auto go = [&](auto coll_tag) {
        constexpr int COLL = decltype(coll_tag)::value;
        if (seg == 4) region_run<COLL, 4>(p, full.data(), split.data(), N, bx0, bx1, by0, by1);
        else region_run<COLL, 1>(p, full.data(), split.data(), N, bx0, bx1, by0, by1);
    };
    if (c->collision == LBM_BGK) go(std::integral_constant<int, LBM_BGK>{});
    else if (c->collision == LBM_MRT) go(std::integral_constant<int, LBM_MRT>{});
    else go(std::integral_constant<int, LBM_CM>{});
    std::memcpy(f_full, full.data(), 9 * N * sizeof(float));
    std::memcpy(f_split, split.data(), 9 * N * sizeof(float));
    return 0;
}

// One two-step launch from FP32 populations f_in ([9][N], fully periodic, all FLUID, FP32
// storage) with the intermediate-state output on: f_mid receives state t+1, f_out state t+2.
// form 0: tiles of 6 rows; 1: strips of 4 wavefronts x 8 rows.
extern "C" int emu_two_step_mid(const lbm_config* c, int form, const float* f_in, float* f_mid, float* f_out) {
    if (c->storage != LBM_F32 || !c->periodic_x || !c->periodic_y) return -1;
    const long long N = (long long)c->nx * c->ny;
    std::vector<float> a(f_in, f_in + 9 * N), b(size_t(9) * N, poison<float>()), m(size_t(9) * N, poison<float>());
    StepParams p{};
    p.nx = c->nx;
    p.ny = c->ny;
    p.px = 1;
    p.py = 1;
    p.plane = N;
    p.src = a.data();
    p.dst = b.data();
    p.omega = c->omega;
    for (int k = 0; k < 9; k++) p.S[k] = c->S[k];
    p.rho0 = c->rho0;
    p.fx = c->force[0];
    p.fy = c->force[1];
    p.forced = (c->force[0] != 0.f || c->force[1] != 0.f) ? 1 : 0;
    p.quirks = c->quirks;
    p.mid = m.data();
    if (form == 0) {
        if (c->collision == LBM_BGK) launch2_emu<LBM_BGK, LBM_F32, 6>(p, 0);
        else if (c->collision == LBM_MRT) launch2_emu<LBM_MRT, LBM_F32, 6>(p, 0);
        else launch2_emu<LBM_CM, LBM_F32, 6>(p, 0);
    } else {
        if (c->collision == LBM_BGK) launch2s_emu<LBM_BGK, LBM_F32, 4>(p, 8, 0);
        else if (c->collision == LBM_MRT) launch2s_emu<LBM_MRT, LBM_F32, 4>(p, 8, 0);
        else launch2s_emu<LBM_CM, LBM_F32, 4>(p, 8, 0);
    }
    std::memcpy(f_mid, m.data(), 9 * N * sizeof(float));
    std::memcpy(f_out, b.data(), 9 * N * sizeof(float));
    return 0;
}

// The product's node adjoints (lbm_adjoint.hpp) on the host: mode 0 BGK, 1 MRT, 2 snapshot
// seed (lam = dL/du in lam[0..1]), 3 initial-state feq (g[0] = rho0, F = u0).
extern "C" int emu_adjoint_node(int mode, const lbm_config* c, const float* g, const float* lam, const float* F,
                                float* mu, float* gF) {
    AdjOut o{};
    if (mode == 0) o = adjoint_bgk(g, lam, F[0], F[1], c->omega);
    else if (mode == 1) o = adjoint_mrt(g, lam, F[0], F[1], c->S);
    else if (mode == 2) o = adjoint_snapshot(g, F[0], F[1], lam[0], lam[1]);
    else {
        adjoint_init_u(lam, g[0], F[0], F[1], o.gfx, o.gfy);
    }
    for (int q = 0; q < Q; q++) mu[q] = o.mu[q];
    gF[0] = o.gfx;
    gF[1] = o.gfy;
    return 0;
}

#ifdef EMU_MAIN
// Bounds matrix: every collision x storage x boundary layout on small and ragged grids.
int main() {
    const int sizes[][2] = {{3, 3}, {5, 4}, {8, 8}, {13, 7}, {16, 9}, {33, 17}, {64, 5}, {70, 11}, {130, 9}};
    int cases = 0;
    for (auto& sz : sizes) {
        for (int layout = 0; layout < 6; layout++) {
            for (int coll = 0; coll < 3; coll++) {
                for (int storage = 0; storage < 3; storage++) {
                    for (int adapt = 0; adapt < (coll == LBM_CM ? 2 : 1); adapt++) {
                        lbm_config c;
                        std::memset(&c, 0, sizeof(c));
                        c.nx = sz[0];
                        c.ny = sz[1];
                        c.collision = coll;
                        c.adapter = adapt;
                        c.omega = 1.2f;
                        for (int k = 0; k < 9; k++) c.S[k] = 1.1f;
                        c.rho0 = 1.0f;
                        c.storage = storage;
                        c.u_lid = 0.05f;
                        c.wall_u[0] = 0.04f;
                        c.force[0] = 1e-5f;
                        const long long N = (long long)c.nx * c.ny;
                        std::vector<uint8_t> fl(N, LBM_FLUID);
                        for (int y = 0; y < c.ny; y++)
                            for (int x = 0; x < c.nx; x++) {
                                uint8_t& f = fl[(long long)y * c.nx + x];
                                bool xe = x == 0 || x == c.nx - 1, ye = y == 0 || y == c.ny - 1;
                                switch (layout) {
                                    case 0: c.periodic_x = c.periodic_y = 1; break;  // all fluid, periodic
                                    case 1:  // channel: periodic x, BB walls
                                        c.periodic_x = 1;
                                        if (ye) f = LBM_BOUNCE_BACK;
                                        break;
                                    case 2:  // cavity: moving lid + BB
                                        if (y == c.ny - 1 && !xe) f = LBM_MOVING_WALL_BOUNCE_BACK;
                                        else if (xe || ye) f = LBM_BOUNCE_BACK;
                                        break;
                                    case 3:  // regularized cavity
                                        if (xe && ye) f = LBM_REGULARIZED_BOUNCE_BACK_CORNER;
                                        else if (y == c.ny - 1) f = LBM_REGULARIZED_INLET_TOP;
                                        else if (xe || ye) f = LBM_REGULARIZED_BOUNCE_BACK;
                                        break;
                                    case 4:  // inflow / outflow channel
                                        if (ye) f = LBM_BOUNCE_BACK;
                                        else if (x == 0) f = LBM_ZOU_HE_LEFT;
                                        else if (x == c.nx - 1) f = LBM_ZG_OUTFLOW;
                                        else if (x == c.nx / 2 && y == c.ny / 2) f = LBM_BOUNCE_BACK;
                                        break;
                                    case 5:  // pressure outlet, full-way bounce-back block
                                        if (ye) f = LBM_BOUNCE_BACK;
                                        else if (x == 0) f = LBM_ZOU_HE_LEFT;
                                        else if (x == c.nx - 1) f = LBM_PRESSURE_OUTLET;
                                        else if (x == c.nx / 3 && y == c.ny / 2) f = LBM_CYLINDER;
                                        break;
                                }
                            }
                        std::vector<float> rho(N), u(2 * N), F(2 * N);
                        for (long long n = 0; n < N; n++) {
                            rho[n] = 1.0f + 0.001f * float(n % 7);
                            u[2 * n] = 0.01f * float(n % 5) - 0.02f;
                            u[2 * n + 1] = 0.01f * float(n % 3) - 0.01f;
                            F[2 * n] = 1e-6f * float(n % 4);
                            F[2 * n + 1] = -1e-6f * float(n % 3);
                        }
                        std::vector<float> f(9 * N), r(N), uu(2 * N);
                        std::vector<double> a(3);
                        const int segs[2] = {kSegWide, storage == LBM_F32 ? 1 : 2};
                        for (int seg : segs) {
                            if (emu_run(&c, seg, 0, 0, fl.data(), rho.data(), u.data(), (cases & 1) ? F.data() : nullptr,
                                        3, f.data(), r.data(), uu.data(), a.data()) != 0) {
                                std::printf("emu_run rejected seg %d\n", seg);
                                return 1;
                            }
                            cases++;
                        }
                        for (int ty : {6, 14, -4003, -8016}) {  // two-step tiles and strips (5 steps)
                            if (emu_run(&c, 0, ty, cases & 1, fl.data(), rho.data(), u.data(),
                                        (cases & 2) ? F.data() : nullptr, 5, f.data(), r.data(), uu.data(), a.data()) != 0) {
                                std::printf("emu_run rejected two-step ty %d\n", ty);
                                return 1;
                            }
                            cases++;
                        }
                    }
                }
            }
        }
    }
    // split IBM steps: every node outside the box, then the box, on exact-size buffers
    for (int coll = 0; coll < 3; coll++)
        for (int sg : {1, 4}) {
            const int boxes[][6] = {{45, 19, 8, 24, 5, 14}, {45, 19, 0, 8, 0, 3}, {45, 19, 36, 45, 10, 19},
                                    {9, 4, 4, 9, 0, 4}, {130, 9, 64, 128, 2, 7}};
            for (auto& b : boxes) {
                lbm_config c;
                std::memset(&c, 0, sizeof(c));
                c.nx = b[0];
                c.ny = b[1];
                c.collision = coll;
                c.omega = 1.2f;
                for (int k = 0; k < 9; k++) c.S[k] = 1.1f;
                c.rho0 = 1.0f;
                c.periodic_x = c.periodic_y = 1;
                c.force[0] = 1e-5f;
                const long long N = (long long)c.nx * c.ny;
                std::vector<float> f(size_t(9) * N), F(size_t(2) * N, 1e-5f), o1(size_t(9) * N), o2(size_t(9) * N);
                for (long long n = 0; n < 9 * N; n++) f[size_t(n)] = wgt(int(n / N)) * (1.0f + 1e-3f * float(n % 7));
                if (emu_region_split(&c, sg, b[2], b[3], b[4], b[5], f.data(), F.data(), o1.data(), o2.data()) != 0 ||
                    std::memcmp(o1.data(), o2.data(), o1.size() * sizeof(float)) != 0) {
                    printf("region split mismatch: %dx%d box %d..%d x %d..%d seg %d coll %d\n", b[0], b[1], b[2], b[3],
                           b[4], b[5], sg, coll);
                    return 1;
                }
                cases++;
            }
        }
    std::printf("emu bounds matrix: %d cases ok\n", cases);
    return 0;
}
#endif
