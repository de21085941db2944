// lbm_ibm.hip — immersed-boundary multi-direct forcing (IBMManager<2>::multi_direct,
// IBMManager.cuh:222-252) for static markers.
//
// The reference runs five whole-domain kernels per iteration (memset, interpolate,
// Lagrangian force, atomic spread, correct, accumulate) plus a D2D copy of u: ~180 B per
// node per step for a force that is non-zero on a few thousand nodes. Here the Eulerian
// side is restricted to the *band* — the union of the markers' kernel stencils — whose
// post-stream macroscopics, corrected velocities and forces live in LDS of a single
// workgroup for all iterations. Spreading is a gather by band node over a precomputed
// (marker, weight) list in marker order, so the sum is deterministic and equals the
// reference's scatter order. The result is the total node force on the band's bounding
// box, which the fused step reads there. Box cells outside the band hold the external force,
// which changes only with the force field: they are written on the first pass after the
// markers or the field change, and later passes write the band alone.
#include "lbm_ctx.hpp"
#include "lbm_ibm_tables.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <vector>

using namespace lbm;

namespace {

constexpr int kIbmBlock = 1024;

struct IbmParams {
    NodeArgs na;
    int storage;
    int n_markers, n_band, iters, clamp;
    int n_ent;                    // stencil entries (= spread entries)
    int x0, y0, bw, bh;
    int write_box;                // 1: every box cell, 0: the band's cells only
    float fx, fy;                 // uniform force
    const float* force_field;     // AoS or null
    const int* band_node;
    const int* mk_ptr;
    const int* mk_band;
    const float* mk_w;
    const int* bn_ptr;
    const int* bn_mk;
    const float* bn_w;
    const int* box_map;           // band index or -1 per box cell
    float* force_total;           // AoS [N][2]
    float* lag_force;             // AoS [n][2], summed over iterations
    const int* mkf_band;          // fixed-width lists, slot-major (lbm_ibm_tables.hpp)
    const float* mkf_w;
    const int* bnf_mk;
    const float* bnf_w;
};

// TABLES_LDS: the interpolation and spreading tables are staged in LDS once per pass, so
// the iterations run from LDS only (used whenever they fit beside the band data).
template <int STORE, bool TABLES_LDS>
__global__ __launch_bounds__(kIbmBlock) void ibm_kernel(IbmParams p) {
    extern __shared__ float lds[];
    const int nb = p.n_band, nm = p.n_markers;
    float* s_rho = lds;
    float* s_ux = s_rho + nb;
    float* s_uy = s_ux + nb;
    float* s_fx = s_uy + nb;   // accumulated force: F_ext + sum of iterations
    float* s_fy = s_fx + nb;
    float* s_lx = s_fy + nb;   // Lagrangian force of the current iteration
    float* s_ly = s_lx + nm;
    float* s_sx = s_ly + nm;   // summed over iterations
    float* s_sy = s_sx + nm;
    int* s_node = reinterpret_cast<int*>(s_sy + nm);  // global node of each band entry
    const int nx = p.na.nx;
    const int* mk_ptr = p.mk_ptr;
    const int* mk_band = p.mk_band;
    const float* mk_w = p.mk_w;
    const int* bn_ptr = p.bn_ptr;
    const int* bn_mk = p.bn_mk;
    const float* bn_w = p.bn_w;
    if constexpr (TABLES_LDS) {
        int* t_mk_ptr = s_node + nb;
        int* t_bn_ptr = t_mk_ptr + nm + 1;
        int* t_mk_band = t_bn_ptr + nb + 1;
        int* t_bn_mk = t_mk_band + p.n_ent;
        float* t_mk_w = reinterpret_cast<float*>(t_bn_mk + p.n_ent);
        float* t_bn_w = t_mk_w + p.n_ent;
        for (int i = threadIdx.x; i <= nm; i += kIbmBlock) t_mk_ptr[i] = p.mk_ptr[i];
        for (int i = threadIdx.x; i <= nb; i += kIbmBlock) t_bn_ptr[i] = p.bn_ptr[i];
        for (int i = threadIdx.x; i < p.n_ent; i += kIbmBlock) {
            t_mk_band[i] = p.mk_band[i];
            t_bn_mk[i] = p.bn_mk[i];
            t_mk_w[i] = p.mk_w[i];
            t_bn_w[i] = p.bn_w[i];
        }
        mk_ptr = t_mk_ptr;
        mk_band = t_mk_band;
        mk_w = t_mk_w;
        bn_ptr = t_bn_ptr;
        bn_mk = t_bn_mk;
        bn_w = t_bn_w;
    }

    // post-stream macroscopics of the band (uncorrected_macroscopics_kernel<2>,
    // macroscopics.cu:5-38) and the external force (reset_forces, macroscopics.cuh:13-48)
    for (int b = threadIdx.x; b < nb; b += kIbmBlock) {
        int node = p.band_node[b];
        s_node[b] = node;
        int y = node / nx, x = node - y * nx;
        Pop9 g = gather_node_inl<STORE>(p.na, x, y);
        float r = 0.f, jx = 0.f, jy = 0.f;
#pragma unroll
        for (int i = 0; i < Q; i++) {
            r += g.v[i];
            jx += g.v[i] * float(CX[i]);
            jy += g.v[i] * float(CY[i]);
        }
        float inv = 1.0f / r;
        s_rho[b] = r;
        s_ux[b] = jx * inv;
        s_uy[b] = jy * inv;
        float fx = p.fx, fy = p.fy;
        if (p.force_field) {
            fx = p.force_field[2LL * node];
            fy = p.force_field[2LL * node + 1];
        }
        s_fx[b] = fx;
        s_fy[b] = fy;
    }
    for (int l = threadIdx.x; l < nm; l += kIbmBlock) s_sx[l] = s_sy[l] = 0.f;
    __syncthreads();

    for (int it = 0; it < p.iters; it++) {
        // interpolate_velocities_kernel<2> + compute_lagrangian_kernel (IBM_impl.cu:5-51,
        // IBM_impl.cuh:7-26): u_L, rho_L over the stencil, F_L = 2 rho_L (0 - u_L)
        for (int l = threadIdx.x; l < nm; l += kIbmBlock) {
            float ux = 0.f, uy = 0.f, r = 0.f;
            for (int e = mk_ptr[l]; e < mk_ptr[l + 1]; e++) {
                float k = mk_w[e];
                int b = mk_band[e];
                r += k * s_rho[b];
                ux += k * s_ux[b];
                uy += k * s_uy[b];
            }
            float lx = 2.0f * r * (0.0f - ux), ly = 2.0f * r * (0.0f - uy);
            if (p.clamp) {
                lx = lx > 1e-8f ? lx : 0.0f;
                ly = ly > 1e-8f ? ly : 0.0f;
            }
            s_lx[l] = lx;
            s_ly[l] = ly;
            s_sx[l] += lx;
            s_sy[l] += ly;
        }
        __syncthreads();
        // spread_forces_kernel<2> (IBM_impl.cu:120-154) as a gather, then
        // correct_velocities_kernel and accumulate_forces_kernel (IBM_impl.cuh:28-68)
        for (int b = threadIdx.x; b < nb; b += kIbmBlock) {
            float fx = 0.f, fy = 0.f;
            for (int e = bn_ptr[b]; e < bn_ptr[b + 1]; e++) {
                float w = bn_w[e];
                int l = bn_mk[e];
                fx += w * s_lx[l];
                fy += w * s_ly[l];
            }
            float cx = s_ux[b] + fx / (2.0f * s_rho[b]);
            float cy = s_uy[b] + fy / (2.0f * s_rho[b]);
            if (p.clamp) {
                cx = cx > 1e-8f ? cx : 0.0f;
                cy = cy > 1e-8f ? cy : 0.0f;
            }
            s_ux[b] = cx;
            s_uy[b] = cy;
            s_fx[b] += fx;
            s_fy[b] += fy;
        }
        __syncthreads();
    }

    // total force on the band's box: band nodes from LDS, the rest the external force
    if (!p.write_box) {
        for (int b = threadIdx.x; b < nb; b += kIbmBlock) {
            const long long node = s_node[b];
            p.force_total[2 * node] = s_fx[b];
            p.force_total[2 * node + 1] = s_fy[b];
        }
    }
    const int ncell = p.write_box ? p.bw * p.bh : 0;
    for (int c = threadIdx.x; c < ncell; c += kIbmBlock) {
        int yy = p.y0 + c / p.bw, xx = p.x0 + c % p.bw;
        long long node = (long long)yy * nx + xx;
        int b = p.box_map[c];
        float fx, fy;
        if (b >= 0) {
            fx = s_fx[b];
            fy = s_fy[b];
        } else {
            fx = p.fx;
            fy = p.fy;
            if (p.force_field) {
                fx = p.force_field[2 * node];
                fy = p.force_field[2 * node + 1];
            }
        }
        p.force_total[2 * node] = fx;
        p.force_total[2 * node + 1] = fy;
    }
    for (int l = threadIdx.x; l < nm; l += kIbmBlock) {
        p.lag_force[2 * l] = s_sx[l];
        p.lag_force[2 * l + 1] = s_sy[l];
    }
}

// The same pass with every list in registers: thread t owns marker t (its kIbmSlots stencil
// slots) and band entries t + r * kIbmBlock, r < NBR (their K spread slots, velocity, density
// and accumulated force). The lists are read once per pass, the iterations touch LDS only for
// what other threads read (band velocities and densities, Lagrangian forces), and every slot
// loop is unrolled, so its LDS reads are issued together rather than one dependent pair per
// entry. The sums run over the same entries in the same order as ibm_kernel's (padding slots
// add +0), so both kernels give the same forces. Needs n_markers <= kIbmBlock and
// n_band <= NBR * kIbmBlock.
template <int STORE, int K, int NBR>
__global__ __launch_bounds__(kIbmBlock) void ibm_fast_kernel(IbmParams p) {
    // band entry b: (rho, ux, uy, -) at s_mac[b]; marker l: its force of the iteration at s_lag[l]
    extern __shared__ float4 lds4[];
    float4* s_mac = lds4;
    float2* s_lag = reinterpret_cast<float2*>(s_mac + p.n_band);
    const int nb = p.n_band, nm = p.n_markers, t = threadIdx.x;
    const int nx = p.na.nx;

    // the lists, as LDS element pointers (one register per slot, nothing left to derive from
    // them inside the iterations), loaded first: their loads overlap the band gather
    const bool has_mk = t < nm;
    const float4* mb[kIbmSlots];
    float mw[kIbmSlots];
#pragma unroll
    for (int k = 0; k < kIbmSlots; k++) {
        mb[k] = s_mac + (has_mk ? p.mkf_band[k * nm + t] : 0);
        mw[k] = has_mk ? p.mkf_w[k * nm + t] : 0.0f;
    }
    const float2* bm[NBR][K];
    float bw[NBR][K];
#pragma unroll
    for (int r = 0; r < NBR; r++) {
        const int b = t + r * kIbmBlock;
        const bool has = b < nb;
#pragma unroll
        for (int k = 0; k < K; k++) {
            bm[r][k] = s_lag + (has ? p.bnf_mk[k * nb + b] : 0);
            bw[r][k] = has ? p.bnf_w[k * nb + b] : 0.0f;
        }
    }
    // post-stream macroscopics of the band (uncorrected_macroscopics_kernel<2>,
    // macroscopics.cu:5-38) and the external force (reset_forces, macroscopics.cuh:13-48)
    int node[NBR];
    float rho[NBR], ux[NBR], uy[NBR], fxa[NBR], fya[NBR];
#pragma unroll
    for (int r = 0; r < NBR; r++) {
        const int b = t + r * kIbmBlock;
        rho[r] = 1.0f;
        ux[r] = uy[r] = fxa[r] = fya[r] = 0.0f;
        node[r] = 0;
        if (b < nb) {
            node[r] = p.band_node[b];
            const int y = node[r] / nx, x = node[r] - y * nx;
            Pop9 g = gather_node_inl<STORE>(p.na, x, y);
            float rr = 0.f, jx = 0.f, jy = 0.f;
#pragma unroll
            for (int i = 0; i < Q; i++) {
                rr += g.v[i];
                jx += g.v[i] * float(CX[i]);
                jy += g.v[i] * float(CY[i]);
            }
            const float inv = 1.0f / rr;
            rho[r] = rr;
            ux[r] = jx * inv;
            uy[r] = jy * inv;
            s_mac[b] = make_float4(rr, ux[r], uy[r], 0.0f);
            fxa[r] = p.fx;
            fya[r] = p.fy;
            if (p.force_field) {
                fxa[r] = p.force_field[2LL * node[r]];
                fya[r] = p.force_field[2LL * node[r] + 1];
            }
        }
    }
    __syncthreads();

    float sx = 0.0f, sy = 0.0f;
    for (int it = 0; it < p.iters; it++) {
        // interpolate_velocities_kernel<2> + compute_lagrangian_kernel (IBM_impl.cu:5-51,
        // IBM_impl.cuh:7-26)
        if (has_mk) {
            float r = 0.f, vxs = 0.f, vys = 0.f;
#pragma unroll
            for (int k = 0; k < kIbmSlots; k++) {
                const float4 m = *mb[k];
                r += mw[k] * m.x;
                vxs += mw[k] * m.y;
                vys += mw[k] * m.z;
            }
            float lx = 2.0f * r * (0.0f - vxs), ly = 2.0f * r * (0.0f - vys);
            if (p.clamp) {
                lx = lx > 1e-8f ? lx : 0.0f;
                ly = ly > 1e-8f ? ly : 0.0f;
            }
            s_lag[t] = make_float2(lx, ly);
            sx += lx;
            sy += ly;
        }
        __syncthreads();
        // spread_forces_kernel<2> (IBM_impl.cu:120-154) as a gather, then
        // correct_velocities_kernel and accumulate_forces_kernel (IBM_impl.cuh:28-68)
#pragma unroll
        for (int r = 0; r < NBR; r++) {
            const int b = t + r * kIbmBlock;
            if (b < nb) {
                float fx = 0.f, fy = 0.f;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const float2 l = *bm[r][k];
                    fx += bw[r][k] * l.x;
                    fy += bw[r][k] * l.y;
                }
                float cx = ux[r] + fx / (2.0f * rho[r]);
                float cy = uy[r] + fy / (2.0f * rho[r]);
                if (p.clamp) {
                    cx = cx > 1e-8f ? cx : 0.0f;
                    cy = cy > 1e-8f ? cy : 0.0f;
                }
                ux[r] = cx;
                uy[r] = cy;
                s_mac[b].y = cx;
                s_mac[b].z = cy;
                fxa[r] += fx;
                fya[r] += fy;
            }
        }
        __syncthreads();
    }

    if (has_mk) {
        p.lag_force[2 * t] = sx;
        p.lag_force[2 * t + 1] = sy;
    }
    if (!p.write_box) {
#pragma unroll
        for (int r = 0; r < NBR; r++) {
            if (t + r * kIbmBlock < nb) {
                p.force_total[2LL * node[r]] = fxa[r];
                p.force_total[2LL * node[r] + 1] = fya[r];
            }
        }
        return;
    }
    // every box cell: band entries from LDS (the iterations are over), the rest the external force
#pragma unroll
    for (int r = 0; r < NBR; r++) {
        const int b = t + r * kIbmBlock;
        if (b < nb) s_mac[b] = make_float4(fxa[r], fya[r], 0.0f, 0.0f);
    }
    __syncthreads();
    const int ncell = p.bw * p.bh;
    for (int c = t; c < ncell; c += kIbmBlock) {
        const int yy = p.y0 + c / p.bw, xx = p.x0 + c % p.bw;
        const long long nd = (long long)yy * nx + xx;
        const int b = p.box_map[c];
        float fx, fy;
        if (b >= 0) {
            fx = s_mac[b].x;
            fy = s_mac[b].y;
        } else {
            fx = p.fx;
            fy = p.fy;
            if (p.force_field) {
                fx = p.force_field[2 * nd];
                fy = p.force_field[2 * nd + 1];
            }
        }
        p.force_total[2 * nd] = fx;
        p.force_total[2 * nd + 1] = fy;
    }
}

// ---- one launch per IBM step ---------------------------------------------------------------
//
// The first n_ibm workgroups of the launch each run the whole IBM pass (the sums of
// ibm_fast_kernel and ibm_kernel: same entries, same order) and then step a share of the IBM
// box, whole rows of it with at most one segment per thread, reading the forces they wrote for
// those rows; the other workgroups step the nodes outside the box. The box's step needs
// nothing from other workgroups and the outside step reads no IBM force, so no workgroup waits
// for another: the pass runs beside the outside step, on one stream, with no cross-queue
// dependency per step. Workgroups are 512 threads (two band entries and one marker per
// thread). The launch's dynamic LDS is allocated to every workgroup, and above ~24 KB it
// costs the outside step occupancy, so the pass keeps only what threads share there: band
// state (12 B per entry), Lagrangian forces (8 B per marker) and the spread lists' 16-bit
// marker indices (CSR); the marker lists sit in registers and the spread weights are read from
// memory (cached) in each iteration. The pass's constant arguments sit in device memory,
// leaving the launch's scalar registers to the step.
constexpr int kIbmStepBlock = 512;
constexpr int kIbmStepRounds = 2;  // band entries per thread
constexpr size_t kIbmStepLdsMax = 24 * 1024;

// Waves per SIMD requested for the launch: the outside step's register budget (80 VGPRs for
// FP32 populations; the 16-bit step needs about 108).
template <int STORE>
constexpr int kIbmStepWaves = STORE == LBM_F32 ? 6 : 4;

struct IbmStepArgs {
    const int* band_node;
    const int* mkf_band;
    const float* mkf_w;
    const int* bn_ptr;
    const int* bn_mk;
    const float* bn_w;
    float* force_total;
    float* lag_force;
    int n_band, n_markers, n_ent, iters;
    int clamp, y0, bh, rows_per;
};

// LDS of the pass, in bytes: band state (rho, ux, uy), Lagrangian forces (8-byte aligned),
// spread marker indices.
inline size_t ibm_step_lds(int nb, int nm, int ne) {
    return (size_t(nb) * 12 + 7) / 8 * 8 + size_t(nm) * 8 + size_t(ne) * 2;
}

template <int STORE>
__device__ __forceinline__ void ibm_step_pass(const StepParams& sp, const IbmStepArgs* __restrict__ ia, int row0,
                                              int row1) {
    constexpr int K = 8;  // longest spread list
    extern __shared__ float2 lds2[];
    const int nb = ia->n_band, nm = ia->n_markers, ne = ia->n_ent, t = threadIdx.x;
    float* s_mac = reinterpret_cast<float*>(lds2);
    float2* s_lag = lds2 + (size_t(nb) * 3 + 1) / 2;
    unsigned short* s_bmk = reinterpret_cast<unsigned short*>(s_lag + nm);
    const int nx = sp.nx;
    const NodeArgs na = node_args(sp);
    // band nodes, their spread entries [e0, e1) and weights
    int node[kIbmStepRounds], e0[kIbmStepRounds], e1[kIbmStepRounds];
#pragma unroll
    for (int r = 0; r < kIbmStepRounds; r++) {
        const int b = t + r * kIbmStepBlock;
        node[r] = b < nb ? ia->band_node[b] : 0;
        e0[r] = b < nb ? ia->bn_ptr[b] : 0;
        e1[r] = b < nb ? ia->bn_ptr[b + 1] : 0;
    }
    // the spread marker indices into LDS, eight per thread in flight
    constexpr int SU = 8;
    for (int base = 0; base < ne; base += SU * kIbmStepBlock) {
        int m[SU];
#pragma unroll
        for (int u = 0; u < SU; u++) {
            const int e = base + u * kIbmStepBlock + t;
            m[u] = e < ne ? ia->bn_mk[e] : 0;
        }
#pragma unroll
        for (int u = 0; u < SU; u++) {
            const int e = base + u * kIbmStepBlock + t;
            if (e < ne) s_bmk[e] = static_cast<unsigned short>(m[u]);
        }
    }
    // post-stream macroscopics of the band and the external force (as in ibm_kernel): the
    // pulls of both entries issued together, the boundary functor only where a band node is
    // not plain fluid
    Pop9 g[kIbmStepRounds];
    uint8_t fl[kIbmStepRounds];
#pragma unroll
    for (int r = 0; r < kIbmStepRounds; r++) {
        const int y = node[r] / nx, x = node[r] - y * nx;
#pragma unroll
        for (int q = 0; q < Q; q++) g[r].v[q] = gather<STORE>(na, q, x, y);
        fl[r] = na.flags ? na.flags[node[r]] : uint8_t(LBM_FLUID);
    }
    float fxa[kIbmStepRounds], fya[kIbmStepRounds];
#pragma unroll
    for (int r = 0; r < kIbmStepRounds; r++) {
        const int b = t + r * kIbmStepBlock;
        fxa[r] = fya[r] = 0.0f;
        if (b < nb) {
            if (fl[r] != LBM_FLUID) {  // rare: the out-of-line general path
                const int y = node[r] / nx, x = node[r] - y * nx;
                g[r] = gather_node<STORE>(na, x, y);
            }
            float rr = 0.f, jx = 0.f, jy = 0.f;
#pragma unroll
            for (int i = 0; i < Q; i++) {
                rr += g[r].v[i];
                jx += g[r].v[i] * float(CX[i]);
                jy += g[r].v[i] * float(CY[i]);
            }
            const float inv = 1.0f / rr;
            s_mac[3 * b] = rr;
            s_mac[3 * b + 1] = jx * inv;
            s_mac[3 * b + 2] = jy * inv;
            fxa[r] = sp.fx;
            fya[r] = sp.fy;
            if (sp.force_field) {
                fxa[r] = sp.force_field[2LL * node[r]];
                fya[r] = sp.force_field[2LL * node[r] + 1];
            }
        }
    }
    // the marker lists, as LDS element pointers
    const bool has_mk = t < nm;
    const float* mb[kIbmSlots];
    float mw[kIbmSlots];
#pragma unroll
    for (int k = 0; k < kIbmSlots; k++) {
        mb[k] = s_mac + 3 * (has_mk ? ia->mkf_band[k * nm + t] : 0);
        mw[k] = has_mk ? ia->mkf_w[k * nm + t] : 0.0f;
    }
    __syncthreads();
    const int clamp = ia->clamp;
    float sx = 0.0f, sy = 0.0f;
    for (int it = 0; it < ia->iters; it++) {
        // the spread weights, read from memory in every iteration (registers and LDS hold the
        // rest); issued first, so that they arrive during the interpolation
        float wsp[kIbmStepRounds][K];
#pragma unroll
        for (int r = 0; r < kIbmStepRounds; r++)
#pragma unroll
            for (int k = 0; k < K; k++) wsp[r][k] = e0[r] + k < e1[r] ? ia->bn_w[e0[r] + k] : 0.0f;
        if (has_mk) {
            float r = 0.f, vxs = 0.f, vys = 0.f;
#pragma unroll
            for (int k = 0; k < kIbmSlots; k++) {
                r += mw[k] * mb[k][0];
                vxs += mw[k] * mb[k][1];
                vys += mw[k] * mb[k][2];
            }
            float lx = 2.0f * r * (0.0f - vxs), ly = 2.0f * r * (0.0f - vys);
            if (clamp) {
                lx = lx > 1e-8f ? lx : 0.0f;
                ly = ly > 1e-8f ? ly : 0.0f;
            }
            s_lag[t] = make_float2(lx, ly);
            sx += lx;
            sy += ly;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kIbmStepRounds; r++) {
            const int b = t + r * kIbmStepBlock;
            if (b < nb) {
                float fx = 0.f, fy = 0.f;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    if (e0[r] + k < e1[r]) {
                        const float2 lg = s_lag[s_bmk[e0[r] + k]];
                        fx += wsp[r][k] * lg.x;
                        fy += wsp[r][k] * lg.y;
                    }
                }
                const float rho = s_mac[3 * b];
                float cx = s_mac[3 * b + 1] + fx / (2.0f * rho);
                float cy = s_mac[3 * b + 2] + fy / (2.0f * rho);
                if (clamp) {
                    cx = cx > 1e-8f ? cx : 0.0f;
                    cy = cy > 1e-8f ? cy : 0.0f;
                }
                s_mac[3 * b + 1] = cx;
                s_mac[3 * b + 2] = cy;
                fxa[r] += fx;
                fya[r] += fy;
            }
        }
        __syncthreads();
    }
    // the forces of this workgroup's box rows; the Lagrangian sums from workgroup 0
    const int y0 = ia->y0;
    float* ft = ia->force_total;
#pragma unroll
    for (int r = 0; r < kIbmStepRounds; r++) {
        const int b = t + r * kIbmStepBlock;
        if (b < nb) {
            const int row = node[r] / nx - y0;
            if (row >= row0 && row < row1) {
                ft[2LL * node[r]] = fxa[r];
                ft[2LL * node[r] + 1] = fya[r];
            }
        }
    }
    if (blockIdx.x == 0 && has_mk) {
        ia->lag_force[2 * t] = sx;
        ia->lag_force[2 * t + 1] = sy;
    }
}

// sp.region is 1 (the step outside the box); the IBM workgroups step their box rows as
// region 2, one segment per thread.
template <int COLL, int STORE>
__global__ __launch_bounds__(kIbmStepBlock) __attribute__((amdgpu_waves_per_eu(kIbmStepWaves<STORE>))) void ibm_step_kernel(
    StepParams sp, const IbmStepArgs* __restrict__ ia, int n_ibm, long long nb_padded, int remap) {
    constexpr int SEG = kSegNarrow<STORE>;
    MomAcc acc{0.0, 0.0, 0.0};
    if (int(blockIdx.x) < n_ibm) {
        const int rows_per = ia->rows_per, bh = ia->bh;
        const int row0 = int(blockIdx.x) * rows_per;
        const int row1 = row0 + rows_per < bh ? row0 + rows_per : bh;
        ibm_step_pass<STORE>(sp, ia, row0, row1);
        __syncthreads();  // this workgroup's box forces are written
        const int bs = (sp.ibm_x1 - sp.ibm_x0 + SEG - 1) / SEG;
        const int seg = row0 * bs + int(threadIdx.x);
        if (seg < row1 * bs) {
            sp.region = 2;
            step_segment<COLL, STORE, 0, SEG>(sp, seg, acc);
        }
        return;
    }
    const long long i = (long long)blockIdx.x - n_ibm;
    const long long lb = remap ? (i % kXcds) * (nb_padded / kXcds) + i / kXcds : i;
    step_segment<COLL, STORE, 0, SEG>(sp, lb * kIbmStepBlock + threadIdx.x, acc);
}

template <class T>
int put(lbm_ctx* c, T** dst, const std::vector<T>& v, const char* what) {
    size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(T);
    int rc = c->dev_alloc(reinterpret_cast<void**>(dst), bytes, what);
    if (rc) return rc;
    if (v.empty()) return LBM_OK;
    return c->upload(*dst, v.data(), v.size() * sizeof(T), 0);
}

// Raise the dynamic-LDS limit of the six IBM kernel instances to the device's per-workgroup
// maximum, once per process (a process-wide attribute of each kernel: set it once, with one
// value, rather than per marker set). Returns whether the limit is in place.
bool allow_dynamic_lds(int max_bytes) {
    static std::once_flag once;
    static bool ok = false;
    std::call_once(once, [max_bytes] {
        const void* k[] = {reinterpret_cast<const void*>(&ibm_kernel<LBM_F32, true>),
                           reinterpret_cast<const void*>(&ibm_kernel<LBM_F16_SHIFTED, true>),
                           reinterpret_cast<const void*>(&ibm_kernel<LBM_BF16_SHIFTED, true>),
                           reinterpret_cast<const void*>(&ibm_kernel<LBM_F32, false>),
                           reinterpret_cast<const void*>(&ibm_kernel<LBM_F16_SHIFTED, false>),
                           reinterpret_cast<const void*>(&ibm_kernel<LBM_BF16_SHIFTED, false>)};
        bool all = true;
        for (const void* f : k)
            if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, max_bytes) != hipSuccess) all = false;
        (void)hipGetLastError();
        ok = all;
    });
    return ok;
}

}  // namespace

// The one-launch IBM step applies to cache-resident grids (narrow segments) once the box's
// external force is in place; LBM_IBM_FUSED=0 in the environment selects the two-stream
// split step instead.
bool lbm_ctx::ibm_fused_ok() const {
    const char* v = std::getenv("LBM_IBM_FUSED");
    if (v && v[0] == '0') return false;
    const bool adapt = cfg.collision == LBM_CM && cfg.adapter == LBM_OPTIMAL_ADAPTER;
    const int narrow = cfg.storage == LBM_F32 ? kSegNarrow<LBM_F32> : kSegNarrow<LBM_F16_SHIFTED>;
    return ibm.active && ibm.fused && ibm.split_ok && ibm.box_valid && !adapt &&
           step_policy(plane, cfg.storage).seg == narrow;
}

int lbm_ctx::fused_ibm_step(bool write_macros) {
    StepParams sp = step_params(cur, write_macros);
    sp.region = 1;
    const int seg = cfg.storage == LBM_F32 ? kSegNarrow<LBM_F32> : kSegNarrow<LBM_F16_SHIFTED>;
    sp.ntx = (sp.nx + seg - 1) / seg;
    sp.nt = 0;
    const long long nseg = (long long)sp.ntx * sp.ny;
    long long nb = (nseg + kIbmStepBlock - 1) / kIbmStepBlock;
    nb = (nb + kXcds - 1) / kXcds * kXcds;
    const long long grid = nb + ibm.fused_wgs;
    if (grid > 0x7fffffffLL) return fail(LBM_E_UNSUPPORTED, "grid too large for one IBM step launch");
    const int st = cfg.storage == LBM_F16_SHIFTED ? 1 : (cfg.storage == LBM_BF16_SHIFTED ? 2 : 0);
    const int co = cfg.collision == LBM_BGK ? 0 : (cfg.collision == LBM_MRT ? 1 : 2);
    static const void* const kern[3][3] = {
        {reinterpret_cast<const void*>(&ibm_step_kernel<LBM_BGK, LBM_F32>),
         reinterpret_cast<const void*>(&ibm_step_kernel<LBM_MRT, LBM_F32>),
         reinterpret_cast<const void*>(&ibm_step_kernel<LBM_CM, LBM_F32>)},
        {reinterpret_cast<const void*>(&ibm_step_kernel<LBM_BGK, LBM_F16_SHIFTED>),
         reinterpret_cast<const void*>(&ibm_step_kernel<LBM_MRT, LBM_F16_SHIFTED>),
         reinterpret_cast<const void*>(&ibm_step_kernel<LBM_CM, LBM_F16_SHIFTED>)},
        {reinterpret_cast<const void*>(&ibm_step_kernel<LBM_BGK, LBM_BF16_SHIFTED>),
         reinterpret_cast<const void*>(&ibm_step_kernel<LBM_MRT, LBM_BF16_SHIFTED>),
         reinterpret_cast<const void*>(&ibm_step_kernel<LBM_CM, LBM_BF16_SHIFTED>)}};
    const IbmStepArgs* ia = static_cast<const IbmStepArgs*>(ibm.d_fused_args);
    int n_ibm = ibm.fused_wgs, remap = 1;
    long long nb_arg = nb;
    void* args[] = {&sp, &ia, &n_ibm, &nb_arg, &remap};
    hipError_t e = hipLaunchKernel(kern[st][co], dim3(unsigned(grid)), dim3(kIbmStepBlock), args, ibm.shmem_fused, stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "IBM step launch");
    cur = 1 - cur;
    timestep++;
    if (write_macros) macros_valid = true;
    return LBM_OK;
}

void lbm_ctx::ibm_free() {
    ibm_graph_free();
    void* ptrs[] = {ibm.d_force_total, ibm.d_xy, ibm.d_ds, ibm.d_band_node, ibm.d_mk_ptr, ibm.d_mk_band, ibm.d_mk_w,
                    ibm.d_bn_ptr, ibm.d_bn_mk, ibm.d_bn_w, ibm.d_lag_force, ibm.d_force_sum, ibm.d_box_map,
                    ibm.d_mkf_band, ibm.d_mkf_w, ibm.d_bnf_mk, ibm.d_bnf_w, ibm.d_fused_args};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    ibm = IbmState();
}

int lbm_ctx::ibm_pass(int src_idx) {
    StepParams sp = step_params(src_idx, false);
    IbmParams p{};
    p.na = node_args(sp);
    p.storage = cfg.storage;
    p.n_markers = ibm.n_markers;
    p.n_band = ibm.n_band;
    p.iters = ibm.iters;
    p.clamp = (cfg.quirks & LBM_QUIRK_IBM_CLAMP) ? 1 : 0;
    p.x0 = ibm.x0;
    p.y0 = ibm.y0;
    p.bw = ibm.x1 - ibm.x0;
    p.bh = ibm.y1 - ibm.y0;
    p.fx = cfg.force[0];
    p.fy = cfg.force[1];
    p.force_field = d_force_field;
    p.band_node = ibm.d_band_node;
    p.mk_ptr = ibm.d_mk_ptr;
    p.mk_band = ibm.d_mk_band;
    p.mk_w = ibm.d_mk_w;
    p.bn_ptr = ibm.d_bn_ptr;
    p.bn_mk = ibm.d_bn_mk;
    p.bn_w = ibm.d_bn_w;
    p.box_map = ibm.d_box_map;
    p.force_total = ibm.d_force_total;
    p.lag_force = ibm.d_lag_force;
    p.write_box = ibm.box_valid ? 0 : 1;
    p.n_ent = ibm.n_ent;
    p.mkf_band = ibm.d_mkf_band;
    p.mkf_w = ibm.d_mkf_w;
    p.bnf_mk = ibm.d_bnf_mk;
    p.bnf_w = ibm.d_bnf_w;
    if (ibm.fast) {
        if (ibm.shmem_fast != size_t(ibm.n_band) * sizeof(float4) + size_t(ibm.n_markers) * sizeof(float2))
            return fail(LBM_E_STATE, "IBM pass LDS size inconsistent with its band");
        const void* k = nullptr;
        const int st = cfg.storage == LBM_F16_SHIFTED ? 1 : (cfg.storage == LBM_BF16_SHIFTED ? 2 : 0);
        const int v = ibm.kspread == 16 ? 2 : (ibm.n_band > kIbmBlock ? 1 : 0);
        static const void* const fast[3][3] = {
            {reinterpret_cast<const void*>(&ibm_fast_kernel<LBM_F32, 8, 1>),
             reinterpret_cast<const void*>(&ibm_fast_kernel<LBM_F32, 8, 2>),
             reinterpret_cast<const void*>(&ibm_fast_kernel<LBM_F32, 16, 1>)},
            {reinterpret_cast<const void*>(&ibm_fast_kernel<LBM_F16_SHIFTED, 8, 1>),
             reinterpret_cast<const void*>(&ibm_fast_kernel<LBM_F16_SHIFTED, 8, 2>),
             reinterpret_cast<const void*>(&ibm_fast_kernel<LBM_F16_SHIFTED, 16, 1>)},
            {reinterpret_cast<const void*>(&ibm_fast_kernel<LBM_BF16_SHIFTED, 8, 1>),
             reinterpret_cast<const void*>(&ibm_fast_kernel<LBM_BF16_SHIFTED, 8, 2>),
             reinterpret_cast<const void*>(&ibm_fast_kernel<LBM_BF16_SHIFTED, 16, 1>)}};
        k = fast[st][v];
        void* args[] = {&p};
        hipError_t e = hipLaunchKernel(k, dim3(1), dim3(kIbmBlock), args, ibm.shmem_fast, stream);
        if (e == hipSuccess) e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "IBM launch");
        ibm.box_valid = true;
        return LBM_OK;
    }
    const bool tl = ibm.shmem_tables > 0;
    size_t sh = tl ? ibm.shmem_tables : ibm.shmem;
    if (cfg.storage == LBM_F16_SHIFTED) {
        if (tl) hipLaunchKernelGGL((ibm_kernel<LBM_F16_SHIFTED, true>), dim3(1), dim3(kIbmBlock), sh, stream, p);
        else hipLaunchKernelGGL((ibm_kernel<LBM_F16_SHIFTED, false>), dim3(1), dim3(kIbmBlock), sh, stream, p);
    } else if (cfg.storage == LBM_BF16_SHIFTED) {
        if (tl) hipLaunchKernelGGL((ibm_kernel<LBM_BF16_SHIFTED, true>), dim3(1), dim3(kIbmBlock), sh, stream, p);
        else hipLaunchKernelGGL((ibm_kernel<LBM_BF16_SHIFTED, false>), dim3(1), dim3(kIbmBlock), sh, stream, p);
    } else {
        if (tl) hipLaunchKernelGGL((ibm_kernel<LBM_F32, true>), dim3(1), dim3(kIbmBlock), sh, stream, p);
        else hipLaunchKernelGGL((ibm_kernel<LBM_F32, false>), dim3(1), dim3(kIbmBlock), sh, stream, p);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "IBM launch");
    ibm.box_valid = true;
    return LBM_OK;
}

extern "C" {

int lbm_ibm_set_markers(lbm_ctx* ctx, const float* xy, const float* ds, int32_t n, int32_t iters) {
    if (!ctx || n < 0 || (n > 0 && !xy) || iters < 1) return LBM_E_INVALID;
    int rc = ctx->activate();
    if (rc) return rc;
    (void)hipStreamSynchronize(ctx->stream);
    ctx->ibm_free();
    if (n == 0) return LBM_OK;
    IbmTables T;
    const char* why = ibm_build_tables(ctx->cfg.nx, ctx->cfg.ny, (ctx->cfg.quirks & LBM_QUIRK_IBM_2X2) != 0,
                                       (ctx->cfg.quirks & LBM_QUIRK_IBM_NO_DS) != 0, xy, ds, n, T);
    if (why) return ctx->fail(LBM_E_INVALID, "%s", why);
    const int nb = int(T.band_node.size());
    const int nx = ctx->cfg.nx, ny = ctx->cfg.ny;
    size_t shmem = (size_t(5) * nb + size_t(4) * n) * sizeof(float) + size_t(nb) * sizeof(int);
    int maxsh = 0;
    (void)hipDeviceGetAttribute(&maxsh, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device);
    if (shmem > size_t(maxsh))
        return ctx->fail(LBM_E_UNSUPPORTED, "IBM band of %d nodes and %d markers needs %zu B of LDS (max %d)", nb, n,
                         shmem, maxsh);
    IbmState& I = ctx->ibm;
    std::vector<float> xyv(xy, xy + 2 * size_t(n)), dsv(size_t(n), 1.0f);
    if (ds)
        for (int l = 0; l < n; l++) dsv[size_t(l)] = ds[l];
    rc = put(ctx, &I.d_xy, xyv, "IBM markers");
    rc = rc ? rc : put(ctx, &I.d_ds, dsv, "IBM ds");
    rc = rc ? rc : put(ctx, &I.d_band_node, T.band_node, "IBM band");
    rc = rc ? rc : put(ctx, &I.d_mk_ptr, T.mk_ptr, "IBM stencils");
    rc = rc ? rc : put(ctx, &I.d_mk_band, T.mk_band, "IBM stencils");
    rc = rc ? rc : put(ctx, &I.d_mk_w, T.mk_w, "IBM weights");
    rc = rc ? rc : put(ctx, &I.d_bn_ptr, T.bn_ptr, "IBM spread lists");
    rc = rc ? rc : put(ctx, &I.d_bn_mk, T.bn_mk, "IBM spread lists");
    rc = rc ? rc : put(ctx, &I.d_bn_w, T.bn_w, "IBM spread weights");
    rc = rc ? rc : put(ctx, &I.d_box_map, T.box_map, "IBM box map");
    // the register-resident pass: one marker per thread, at most two band entries per thread
    // (eight spread slots each) or one (sixteen)
    const bool fast = T.kspread > 0 && n <= kIbmBlock && nb <= (T.kspread == 8 ? 2 : 1) * kIbmBlock;
    if (fast) {
        rc = rc ? rc : put(ctx, &I.d_mkf_band, T.mkf_band, "IBM stencils");
        rc = rc ? rc : put(ctx, &I.d_mkf_w, T.mkf_w, "IBM weights");
        rc = rc ? rc : put(ctx, &I.d_bnf_mk, T.bnf_mk, "IBM spread lists");
        rc = rc ? rc : put(ctx, &I.d_bnf_w, T.bnf_w, "IBM spread weights");
    }
    rc = rc ? rc : ctx->dev_alloc(reinterpret_cast<void**>(&I.d_force_total), ctx->N * 8, "IBM force field");
    rc = rc ? rc : ctx->dev_alloc(reinterpret_cast<void**>(&I.d_lag_force), size_t(n) * 8, "IBM Lagrangian forces");
    if (rc) {
        ctx->ibm_free();
        return rc;
    }
    (void)hipMemsetAsync(I.d_lag_force, 0, size_t(n) * 8, ctx->stream);
    (void)hipStreamSynchronize(ctx->stream);
    I.n_markers = n;
    I.n_band = nb;
    I.iters = iters;
    I.x0 = T.x0;
    I.x1 = T.x1;
    I.y0 = T.y0;
    I.y1 = T.y1;
    I.shmem = shmem;
    I.split_ok = (ctx->cfg.periodic_x || (I.x0 > 0 && I.x1 < nx)) && (ctx->cfg.periodic_y || (I.y0 > 0 && I.y1 < ny));
    I.n_ent = int(T.mk_band.size());
    // the tables beside the band data: mk_ptr, bn_ptr, mk_band, bn_mk (ints), mk_w, bn_w
    const size_t tables = (size_t(n) + 1 + size_t(nb) + 1 + 4 * T.mk_band.size()) * 4;
    I.shmem_tables = shmem + tables <= size_t(maxsh) && allow_dynamic_lds(maxsh) ? shmem + tables : 0;
    I.fast = fast && std::getenv("LBM_IBM_CSR") == nullptr;
    I.kspread = T.kspread;
    I.shmem_fast = size_t(nb) * sizeof(float4) + size_t(n) * sizeof(float2);
    // the one-launch step: at most eight spread entries per band node, two band entries and one
    // marker per thread, LDS within the budget that keeps the outside step's occupancy, box
    // rows that fit one workgroup's segments
    const int seg = ctx->cfg.storage == LBM_F32 ? kSegNarrow<LBM_F32> : kSegNarrow<LBM_F16_SHIFTED>;
    const int bs = (T.x1 - T.x0 + seg - 1) / seg, bh = T.y1 - T.y0;
    I.shmem_fused = ibm_step_lds(nb, n, int(T.bn_mk.size()));
    I.fused = fast && T.kspread == 8 && nb <= kIbmStepRounds * kIbmStepBlock && n <= kIbmStepBlock &&
              T.bn_mk.size() < 65536 && I.shmem_fused <= kIbmStepLdsMax && bs <= kIbmStepBlock;
    if (I.fused) {
        const int rows_per = kIbmStepBlock / bs;
        I.fused_wgs = (bh + rows_per - 1) / rows_per;
        IbmStepArgs a{};
        a.band_node = I.d_band_node;
        a.mkf_band = I.d_mkf_band;
        a.mkf_w = I.d_mkf_w;
        a.bn_ptr = I.d_bn_ptr;
        a.bn_mk = I.d_bn_mk;
        a.bn_w = I.d_bn_w;
        a.force_total = I.d_force_total;
        a.lag_force = I.d_lag_force;
        a.n_band = nb;
        a.n_markers = n;
        a.n_ent = int(T.bn_mk.size());
        a.iters = iters;
        a.clamp = (ctx->cfg.quirks & LBM_QUIRK_IBM_CLAMP) ? 1 : 0;
        a.y0 = T.y0;
        a.bh = bh;
        a.rows_per = rows_per;
        rc = ctx->dev_alloc(&I.d_fused_args, sizeof(a), "IBM step arguments");
        rc = rc ? rc : ctx->upload(I.d_fused_args, &a, sizeof(a), 0);
        if (rc) {
            ctx->ibm_free();
            return rc;
        }
    }
    I.h_ds = dsv;
    I.active = true;
    return LBM_OK;
}

int lbm_ibm_get_forces(lbm_ctx* ctx, float* lag_force, double* force_sum) {
    if (!ctx) return LBM_E_INVALID;
    if (!ctx->ibm.active) return ctx->fail(LBM_E_STATE, "no IBM markers set");
    int rc = ctx->activate();
    if (rc) return rc;
    const int n = ctx->ibm.n_markers;
    std::vector<float> h(2 * size_t(n));
    hipError_t e = hipMemcpyAsync(h.data(), ctx->ibm.d_lag_force, 2 * size_t(n) * sizeof(float), hipMemcpyDeviceToHost,
                                  ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "IBM force copy");
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "hipStreamSynchronize");
    if (lag_force) std::copy(h.begin(), h.end(), lag_force);
    if (force_sum) {
        const bool no_ds = (ctx->cfg.quirks & LBM_QUIRK_IBM_NO_DS) != 0;
        double tx = 0.0, ty = 0.0;
        for (int l = 0; l < n; l++) {
            double d = no_ds ? 1.0 : double(ctx->ibm.h_ds[size_t(l)]);
            tx += double(h[2 * size_t(l)]) * d;
            ty += double(h[2 * size_t(l) + 1]) * d;
        }
        force_sum[0] = tx;
        force_sum[1] = ty;
    }
    return LBM_OK;
}

}  // extern "C"
