// lbm_kernels.hip — gfx950 kernels of the D2Q9 step and their launchers.
#include "lbm_kernels.hpp"

#include <hip/hip_runtime.h>

namespace lbm {

constexpr int kBlock = 256;      // workgroup of the auxiliary kernels (4 wave64)
constexpr int kMaxStepBlock = 512;

// Workgroup i runs on XCD i % 8. With remap, each XCD gets a contiguous run of logical
// workgroups, so the cache lines that neighbouring row segments share stay in one XCD's L2
// (measured: 25% faster on an 8192^2 grid).
__device__ __forceinline__ long long logical_block(long long nb_padded, int remap) {
    long long i = blockIdx.x;
    if (!remap) return i;
    return (i % kXcds) * (nb_padded / kXcds) + i / kXcds;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Deterministic block sum of three doubles; result valid in thread 0.
__device__ __forceinline__ void block_sum3(double& a, double& b, double& c) {
    __shared__ double part[3][kMaxStepBlock / 64];
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        part[0][wv] = a;
        part[1][wv] = b;
        part[2][wv] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = b = c = 0.0;
        for (int w = 0; w < int(blockDim.x) / 64; w++) {
            a += part[0][w];
            b += part[1][w];
            c += part[2][w];
        }
    }
}

template <int COLL, int STORE, int MODE, int SEG>
__global__ __launch_bounds__(kMaxStepBlock) void fused_step_kernel(StepParams p, double* partials, long long nb_padded,
                                                                   int remap) {
    long long lb = logical_block(nb_padded, remap);
    MomAcc acc{0.0, 0.0, 0.0};
    step_segment<COLL, STORE, MODE, SEG>(p, lb * blockDim.x + threadIdx.x, acc);
    if constexpr (MODE != 0) {
        block_sum3(acc.rho, acc.jmag, acc.pimag);
        if (threadIdx.x == 0) {
            partials[3 * lb + 0] = acc.rho;
            partials[3 * lb + 1] = acc.jmag;
            partials[3 * lb + 2] = acc.pimag;
        }
    }
}

// Two time steps per launch; see step2_phase_a / step2_phase_b. Tiles serve cache-resident
// grids only, so their arrays always fit one buffer resource.
template <int COLL, int STORE, int TY>
__global__ __launch_bounds__(Tile2<TY>::THREADS) void fused_step2_kernel(StepParams p, long long nb_padded, int remap,
                                                                         int tiles_x, long long n_tiles) {
    using T = typename Elem<STORE>::T;
    __shared__ T tile[Q * Tile2<TY>::NH];
    const long long lb = logical_block(nb_padded, remap);
    if (lb >= n_tiles) return;  // whole workgroup: remap padding
    const int by = int(lb / tiles_x), bx = int(lb - (long long)by * tiles_x);
    const int x0 = bx * kT2W, y0 = by * TY;
    step2_phase_a<COLL, STORE, TY, true>(p, tile, threadIdx.x, x0, y0, interior_cols(p, x0));
    __syncthreads();
    step2_phase_b<COLL, STORE, TY, true>(p, tile, threadIdx.x, x0, y0);
}

// Two time steps per launch, marching strips; see step2s_phase_a / step2s_phase_b. BUF:
// population arrays addressed through buffer resources (arrays below 4 GiB).
// Minimum wavefronts per SIMD requested for the strip kernels (a register cap: 80 VGPRs at
// 6, 72 at 7). 8192^2, us per step at the best strip height, against the unconstrained
// kernel: FP32 BGK 508 vs 516, MRT 509 vs 510, CM 522 vs 546 (at 6); FP16 BGK 317 vs 335,
// MRT 343 vs 365, CM 392 vs 410 (at 7, 16-bit MRT/CM spilling 8-10 VGPRs to scratch).
template <int STORE>
constexpr int kStripWavesPerEu = STORE == LBM_F32 ? 6 : 7;

template <int COLL, int STORE, int W, bool BUF>
__global__ __launch_bounds__(kT2L * W) __attribute__((amdgpu_waves_per_eu(kStripWavesPerEu<STORE>)))
void fused_step2s_kernel(StepParams p, long long nb_padded, int remap, int strips_x, long long n_strips, int H) {
    using T = typename Elem<STORE>::T;
    constexpr int R = W + 2;  // ring rows
    __shared__ T ring[Q * R * kT2L];
    const long long lb = logical_block(nb_padded, remap);
    if (lb >= n_strips) return;  // whole workgroup: remap padding
    const int sy = int(lb / strips_x), sx = int(lb - (long long)sy * strips_x);
    const int x0 = sx * kT2W, y0 = sy * H;
    const int hs = p.ny - y0 < H ? p.ny - y0 : H;  // output rows of this strip
    // the wavefront's row: uniform, so the ring slots and row tests below are scalar work
    const int l = threadIdx.x & (kT2L - 1), w = __builtin_amdgcn_readfirstlane(int(threadIdx.x / kT2L));
    const int K = (hs + 1) / W + 1;  // chunks: phase-A rows 0 .. hs + 1, phase-B rows 0 .. hs - 1
    const bool interior_x = interior_cols(p, x0);
    int i = w, gx = 0, gy = 0;
    bool va = i <= hs + 1 && strip_node(p, x0, y0, l, i, gx, gy);
    StepParams pl = p;
    for (int k = 0; k < K; k++) {
        // Opaque per iteration: values derived from these (plane offsets, pull shifts,
        // buffer resources, uniform flags) are then rebuilt inside the march with scalar
        // ALU instead of being held across it, which exceeds the SGPR file and spills to
        // VGPR lanes.
        asm volatile("" : "+s"(pl.plane), "+s"(pl.nx), "+s"(pl.n_edge), "+s"(pl.src), "+s"(pl.dst), "+s"(pl.mid),
                     "+s"(pl.edge_prev), "+s"(pl.edge_next));
        asm volatile("" : "+s"(pl.px), "+s"(pl.py), "+s"(pl.ny), "+s"(pl.flags), "+s"(pl.forced), "+s"(pl.force_field),
                     "+s"(pl.ibm_force), "+s"(pl.out_rho), "+s"(pl.quirks));
        if (va) {
            float g[Q];
            step2_gather<STORE, BUF>(pl, gx, gy, g, interior_x && interior_row(pl, y0 - 1 + i));
            step2s_phase_a<COLL, STORE, BUF>(pl, ring, R, i % R, l, i, x0, y0, hs, gx, gy, g);
        } else if (i <= hs + 1) {  // outside a non-periodic domain
            step2_fill_halo<STORE, BUF>(pl, ring + (i % R) * kT2L + l, R * kT2L, x0 - 1 + l, y0 - 1 + i);
        }
        __syncthreads();
        const int in = (k + 1) * W + w;
        int nxg = 0, nyg = 0;
        const bool vn = k + 1 < K && in <= hs + 1 && strip_node(p, x0, y0, l, in, nxg, nyg);
        const int j = k * W - 2 + w;
        if (j >= 0 && j < hs) step2s_phase_b<COLL, STORE, BUF>(pl, ring, R, l, j, x0, y0);
        __syncthreads();
        i = in;
        va = vn;
        gx = nxg;
        gy = nyg;
    }
}

// One step per launch with the OptimalAdapter's averages of the next step computed on the
// way (look-ahead): the marching-strip machinery of fused_step2s_kernel with phase A as the
// step itself (own nodes stored to p.mid = the destination) and phase B reduced to the
// moment sums of the following step's post-stream state. Each workgroup writes its sums to
// partials[logical block]; finalize_ahead_kernel turns them into the next step's averages.
template <int COLL, int STORE, int W, bool BUF>
__global__ __launch_bounds__(kT2L * W) __attribute__((amdgpu_waves_per_eu(kStripWavesPerEu<STORE>)))
void fused_step_ahead_kernel(StepParams p, long long nb_padded, int remap, int strips_x, long long n_strips, int H,
                             double* partials) {
    using T = typename Elem<STORE>::T;
    constexpr int R = W + 2;  // ring rows
    __shared__ T ring[Q * R * kT2L];
    const long long lb = logical_block(nb_padded, remap);
    if (lb >= n_strips) return;  // whole workgroup: remap padding
    const int sy = int(lb / strips_x), sx = int(lb - (long long)sy * strips_x);
    const int x0 = sx * kT2W, y0 = sy * H;
    const int hs = p.ny - y0 < H ? p.ny - y0 : H;
    const int l = threadIdx.x & (kT2L - 1), w = __builtin_amdgcn_readfirstlane(int(threadIdx.x / kT2L));
    const int K = (hs + 1) / W + 1;
    const bool interior_x = interior_cols(p, x0);
    int i = w, gx = 0, gy = 0;
    bool va = i <= hs + 1 && strip_node(p, x0, y0, l, i, gx, gy);
    MomAcc acc{0.0, 0.0, 0.0};
    StepParams pl = p;
    for (int k = 0; k < K; k++) {
        asm volatile("" : "+s"(pl.plane), "+s"(pl.nx), "+s"(pl.n_edge), "+s"(pl.src), "+s"(pl.dst), "+s"(pl.mid),
                     "+s"(pl.edge_prev), "+s"(pl.edge_next));
        asm volatile("" : "+s"(pl.px), "+s"(pl.py), "+s"(pl.ny), "+s"(pl.flags), "+s"(pl.forced), "+s"(pl.force_field),
                     "+s"(pl.ibm_force), "+s"(pl.out_rho), "+s"(pl.quirks));
        if (va) {
            float g[Q];
            step2_gather<STORE, BUF>(pl, gx, gy, g, interior_x && interior_row(pl, y0 - 1 + i));
            step2s_phase_a<COLL, STORE, BUF, true>(pl, ring, R, i % R, l, i, x0, y0, hs, gx, gy, g);
        } else if (i <= hs + 1) {
            step2_fill_halo<STORE, BUF>(pl, ring + (i % R) * kT2L + l, R * kT2L, x0 - 1 + l, y0 - 1 + i);
        }
        __syncthreads();
        const int in = (k + 1) * W + w;
        int nxg = 0, nyg = 0;
        const bool vn = k + 1 < K && in <= hs + 1 && strip_node(p, x0, y0, l, in, nxg, nyg);
        const int j = k * W - 2 + w;
        if (j >= 0 && j < hs) step2s_phase_b_moments<STORE, BUF>(pl, ring, R, l, j, x0, y0, acc);
        __syncthreads();
        i = in;
        va = vn;
        gx = nxg;
        gy = nyg;
    }
    block_sum3(acc.rho, acc.jmag, acc.pimag);
    if (threadIdx.x == 0) {
        partials[3 * lb + 0] = acc.rho;
        partials[3 * lb + 1] = acc.jmag;
        partials[3 * lb + 2] = acc.pimag;
    }
}

// Copy the populations of the non-periodic edge nodes between a population buffer and an
// edge store ([9][n_edge], edge_slot order). to_store = 1: buffer -> store (every slot);
// 0: store -> buffer (each node from its first slot only).
template <int STORE>
__global__ __launch_bounds__(kBlock) void edge_copy_kernel(void* buf, void* store, long long plane, int nx, int ny, int px,
                                                         int py, int n_edge, int to_store) {
    using T = typename Elem<STORE>::T;
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= n_edge) return;
    int x, y;
    edge_node(e, nx, ny, py, x, y);
    const long long node = (long long)y * nx + x;
    T* b = static_cast<T*>(buf);
    T* st = static_cast<T*>(store);
    if (to_store) {
#pragma unroll
        for (int q = 0; q < Q; q++) st[(long long)q * n_edge + e] = b[q * plane + node];
    } else if (edge_slot(x, y, nx, ny, px, py) == e) {
#pragma unroll
        for (int q = 0; q < Q; q++) b[q * plane + node] = st[(long long)q * n_edge + e];
    }
}

// Sum `n` partial triples in a fixed order and store the means (sum / N) as floats, the
// MomentInfo the adapter reads (lbm.cuh:25-31, macroscopics.cuh:116-118).
__global__ __launch_bounds__(kBlock) void finalize_avg_kernel(const double* partials, long long n, double inv_nodes,
                                                            float* avg, double* avg64) {
    double a = 0.0, b = 0.0, c = 0.0;
    for (long long i = threadIdx.x; i < n; i += kBlock) {
        a += partials[3 * i];
        b += partials[3 * i + 1];
        c += partials[3 * i + 2];
    }
    block_sum3(a, b, c);
    if (threadIdx.x == 0) {
        avg[0] = float(a * inv_nodes);
        avg[1] = float(b * inv_nodes);
        avg[2] = float(c * inv_nodes);
        if (avg64) {
            avg64[0] = a * inv_nodes;
            avg64[1] = b * inv_nodes;
            avg64[2] = c * inv_nodes;
        }
    }
}

// After a look-ahead launch: the averages that launch collided with (`ahead`) become the
// reported ones (`avg`), and the launch's partial sums become the next step's `ahead`.
__global__ __launch_bounds__(kBlock) void finalize_ahead_kernel(const double* partials, long long n, double inv_nodes,
                                                              float* ahead, double* ahead64, float* avg, double* avg64) {
    if (threadIdx.x < 3) {
        avg[threadIdx.x] = ahead[threadIdx.x];
        avg64[threadIdx.x] = ahead64[threadIdx.x];
    }
    double a = 0.0, b = 0.0, c = 0.0;
    for (long long i = threadIdx.x; i < n; i += kBlock) {
        a += partials[3 * i];
        b += partials[3 * i + 1];
        c += partials[3 * i + 2];
    }
    block_sum3(a, b, c);  // its barrier orders the copies above before the stores below
    if (threadIdx.x == 0) {
        ahead[0] = float(a * inv_nodes);
        ahead[1] = float(b * inv_nodes);
        ahead[2] = float(c * inv_nodes);
        ahead64[0] = a * inv_nodes;
        ahead64[1] = b * inv_nodes;
        ahead64[2] = c * inv_nodes;
    }
}

// init_kernel + init_node (init.cuh:10-43): f = f_back = f_eq(rho, u).
template <int STORE>
__global__ __launch_bounds__(kBlock) void init_kernel(InitParams p) {
    using St = Elem<STORE>;
    using T = typename St::T;
    long long n = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (n >= p.N) return;
    float r = p.rho[n], ux = p.u[2 * n], uy = p.u[2 * n + 1];
    float usq15 = 1.5f * (ux * ux + uy * uy);
    T* a = static_cast<T*>(p.buf0);
    T* b = static_cast<T*>(p.buf1);
#pragma unroll
    for (int q = 0; q < Q; q++) {
        float fe = feq(q, r, ux, uy, usq15);
        St::store1(a + q * p.plane + n, q, p.rho0, fe);
        St::store1(b + q * p.plane + n, q, p.rho0, fe);
    }
}

// FP32 SoA <-> storage.
template <int STORE>
__global__ __launch_bounds__(kBlock) void pops_to_f32_kernel(const void* buf, long long plane, long long N, float rho0, float* out) {
    using St = Elem<STORE>;
    long long n = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (n >= N) return;
    const typename St::T* b = static_cast<const typename St::T*>(buf);
#pragma unroll
    for (int q = 0; q < Q; q++) out[q * N + n] = St::load1(b + q * plane + n, q, rho0);
}
template <int STORE>
__global__ __launch_bounds__(kBlock) void pops_from_f32_kernel(void* buf0, void* buf1, long long plane, long long N, float rho0, const float* in) {
    using St = Elem<STORE>;
    long long n = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (n >= N) return;
    typename St::T* a = static_cast<typename St::T*>(buf0);
    typename St::T* b = static_cast<typename St::T*>(buf1);
#pragma unroll
    for (int q = 0; q < Q; q++) {
        float v = in[q * N + n];
        St::store1(a + q * plane + n, q, rho0, v);
        if (b) St::store1(b + q * plane + n, q, rho0, v);
    }
}

__global__ __launch_bounds__(kBlock) void macros_to_aos_kernel(const float* ux, const float* uy, long long N, float* u) {
    long long n = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (n >= N) return;
    u[2 * n] = ux[n];
    u[2 * n + 1] = uy[n];
}

__global__ __launch_bounds__(kBlock) void aos_to_soa_kernel(const float* u, long long N, float* ux, float* uy) {
    long long n = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (n >= N) return;
    ux[n] = u[2 * n];
    uy[n] = u[2 * n + 1];
}

// Validation reductions over the stored macroscopic fields, FP64, fixed order:
// each workgroup sums a contiguous range, then one workgroup sums the partials.
__global__ __launch_bounds__(kBlock) void reduce_fields_kernel(int kind, const float* rho, const float* ux, const float* uy,
                                                             const float* ref, long long N, long long chunk,
                                                             double* partials) {
    long long beg = (long long)blockIdx.x * chunk, end = beg + chunk < N ? beg + chunk : N;
    double a = 0.0, b = 0.0, c = 0.0;
    for (long long n = beg + threadIdx.x; n < end; n += kBlock) {
        double r = rho[n], vx = ux[n], vy = uy[n];
        if (kind == LBM_REDUCE_MASS) {
            a += r;
        } else if (kind == LBM_REDUCE_MOMENTUM) {
            a += r * vx;
            b += r * vy;
        } else {  // L2 vs reference
            double ex = vx - double(ref[2 * n]), ey = vy - double(ref[2 * n + 1]);
            a += ex * ex + ey * ey;
            b += double(ref[2 * n]) * double(ref[2 * n]) + double(ref[2 * n + 1]) * double(ref[2 * n + 1]);
        }
    }
    block_sum3(a, b, c);
    if (threadIdx.x == 0) {
        partials[3 * blockIdx.x] = a;
        partials[3 * blockIdx.x + 1] = b;
        partials[3 * blockIdx.x + 2] = c;
    }
}
__global__ __launch_bounds__(kBlock) void sum_partials_kernel(const double* partials, long long n, double* out) {
    double a = 0.0, b = 0.0, c = 0.0;
    for (long long i = threadIdx.x; i < n; i += kBlock) {
        a += partials[3 * i];
        b += partials[3 * i + 1];
        c += partials[3 * i + 2];
    }
    block_sum3(a, b, c);
    if (threadIdx.x == 0) {
        out[0] = a;
        out[1] = b;
        out[2] = c;
    }
}
// Row means of u_x (the Poiseuille validator's x-average, poiseuilleScenario.cuh:62-74):
// one workgroup per row, FP64.
__global__ __launch_bounds__(kBlock) void row_mean_ux_kernel(const float* ux, int nx, double* out) {
    int y = blockIdx.x;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int x = threadIdx.x; x < nx; x += kBlock) a += double(ux[(long long)y * nx + x]);
    block_sum3(a, b, c);
    if (threadIdx.x == 0) out[y] = a / double(nx);
}

// ---- launchers ----------------------------------------------------------------------------

// Measured on MI355X (BGK, periodic, back-to-back launches; us per step):
//   FP32 1024^2: 1 node/thread, 64-thread groups, remap 18.1 (4 nodes, 256, identity 26.8)
//   FP32 2048^2: 1 node, 128, remap, nt 68.9 (4 nodes, 256, remap, nt 79.4)
//   FP32 8192^2: 4 nodes, 128, remap, nt 944 (1 node: 998)
//   FP16 1024^2: 2 nodes, 256, remap 16.9 (4 nodes: 21.4)
//   FP16 2048^2: 2 nodes, 64, remap 44.2;  FP16 8192^2: 2 nodes, 128, remap, nt 611 (4 nodes: 675)
// Small grids are latency-bound and want many wavefronts (narrow segments); only the FP32
// grids far beyond the Infinity Cache prefer 16-byte vectors per thread.
LaunchPolicy step_policy(long long plane, int storage) {
    const bool f32 = storage == LBM_F32;
    const double bytes = 2.0 * 9.0 * double(plane) * (f32 ? 4.0 : 2.0);
    const bool resident = bytes <= 192.0e6;  // fits the 256 MB Infinity Cache with room to spare
    LaunchPolicy lp;
    lp.remap = 1;
    if (resident) {
        lp.block = f32 ? 64 : 256;
        lp.nt = 0;
        lp.seg = f32 ? kSegNarrow<LBM_F32> : kSegNarrow<LBM_F16_SHIFTED>;
    } else if (f32 && plane > (1LL << 23)) {  // beyond 8.4M nodes (1.2 GB of FP32 populations)
        lp.block = 128;
        lp.nt = 1;
        lp.seg = kSegWide;
    } else {
        lp.block = 128;
        lp.nt = 1;
        lp.seg = f32 ? kSegNarrow<LBM_F32> : kSegNarrow<LBM_F16_SHIFTED>;
    }
    return lp;
}

long long step_grid_blocks(const StepParams& p, const LaunchPolicy& lp) {
    long long ntx = (p.nx + lp.seg - 1) / lp.seg;
    long long nseg = ntx * p.ny;
    long long nb = (nseg + lp.block - 1) / lp.block;
    return lp.remap ? (nb + kXcds - 1) / kXcds * kXcds : nb;
}

template <int COLL, int STORE, int MODE, int SEG>
static hipError_t launch_step_t(StepParams p, const LaunchPolicy& lp, double* partials, hipStream_t s) {
    long long nb = step_grid_blocks(p, lp);
    if (p.region == 2) {  // the IBM box only
        const long long nseg = (long long)((p.ibm_x1 - p.ibm_x0 + SEG - 1) / SEG) * (p.ibm_y1 - p.ibm_y0);
        nb = (nseg + lp.block - 1) / lp.block;
        if (lp.remap) nb = (nb + kXcds - 1) / kXcds * kXcds;
    }
    p.ntx = (p.nx + SEG - 1) / SEG;
    p.nt = lp.nt;
    hipLaunchKernelGGL((fused_step_kernel<COLL, STORE, MODE, SEG>), dim3(unsigned(nb)), dim3(lp.block), 0, s, p, partials,
                       nb, lp.remap);
    return hipGetLastError();
}

template <int STORE, int SEG>
static hipError_t launch_step_w(int coll, int mode, const StepParams& p, const LaunchPolicy& lp, double* partials,
                                hipStream_t s) {
    if (mode == 1) return launch_step_t<LBM_BGK, STORE, 1, SEG>(p, lp, partials, s);
    if (mode == 2) return launch_step_t<LBM_CM, STORE, 2, SEG>(p, lp, partials, s);
    if (coll == LBM_BGK) return launch_step_t<LBM_BGK, STORE, 0, SEG>(p, lp, partials, s);
    if (coll == LBM_MRT) return launch_step_t<LBM_MRT, STORE, 0, SEG>(p, lp, partials, s);
    return launch_step_t<LBM_CM, STORE, 0, SEG>(p, lp, partials, s);
}

template <int STORE>
static hipError_t launch_step_s(int coll, int mode, const StepParams& p, const LaunchPolicy& lp, double* partials,
                                hipStream_t s) {
    if (lp.seg == kSegWide) return launch_step_w<STORE, kSegWide>(coll, mode, p, lp, partials, s);
    if (lp.seg == kSegNarrow<STORE>) return launch_step_w<STORE, kSegNarrow<STORE>>(coll, mode, p, lp, partials, s);
    return hipErrorInvalidValue;
}

hipError_t launch_step(int coll, int storage, int mode, const StepParams& p, const LaunchPolicy& lp, double* partials,
                       hipStream_t s) {
    if (lp.block < 64 || lp.block > kMaxStepBlock || lp.block % 64 != 0) return hipErrorInvalidValue;
    // a split step needs an IBM box whose x range is whole segments, and no partial sums
    if (p.region != 0 && (mode != 0 || !p.ibm_force || p.ibm_x0 % 4 != 0 || (p.ibm_x1 % 4 != 0 && p.ibm_x1 != p.nx) ||
                          p.ibm_x0 >= p.ibm_x1 || p.ibm_y0 >= p.ibm_y1))
        return hipErrorInvalidValue;
    if (storage == LBM_F16_SHIFTED) return launch_step_s<LBM_F16_SHIFTED>(coll, mode, p, lp, partials, s);
    if (storage == LBM_BF16_SHIFTED) return launch_step_s<LBM_BF16_SHIFTED>(coll, mode, p, lp, partials, s);
    return launch_step_s<LBM_F32>(coll, mode, p, lp, partials, s);
}

long long step2_grid_blocks(const StepParams& p, const Step2Policy& lp) {
    const int rows = lp.strip ? lp.h : lp.ty;
    const long long tiles = (long long)((p.nx + kT2W - 1) / kT2W) * ((p.ny + rows - 1) / rows);
    return lp.remap ? (tiles + kXcds - 1) / kXcds * kXcds : tiles;
}

template <int COLL, int STORE, int W, bool BUF>
static hipError_t launch_step2s_t(const StepParams& p, const Step2Policy& lp, hipStream_t s) {
    const int strips_x = (p.nx + kT2W - 1) / kT2W;
    const long long n_strips = (long long)strips_x * ((p.ny + lp.h - 1) / lp.h);
    const long long nb = step2_grid_blocks(p, lp);
    hipLaunchKernelGGL((fused_step2s_kernel<COLL, STORE, W, BUF>), dim3(unsigned(nb)), dim3(kT2L * W), 0, s, p, nb,
                       lp.remap, strips_x, n_strips, lp.h);
    return hipGetLastError();
}

template <int STORE, int W, bool BUF>
static hipError_t launch_step2s_c(int coll, const StepParams& p, const Step2Policy& lp, hipStream_t s) {
    if (coll == LBM_BGK) return launch_step2s_t<LBM_BGK, STORE, W, BUF>(p, lp, s);
    if (coll == LBM_MRT) return launch_step2s_t<LBM_MRT, STORE, W, BUF>(p, lp, s);
    return launch_step2s_t<LBM_CM, STORE, W, BUF>(p, lp, s);
}

template <int COLL, int STORE, int TY>
static hipError_t launch_step2_t(const StepParams& p, const Step2Policy& lp, hipStream_t s) {
    const int tiles_x = (p.nx + kT2W - 1) / kT2W;
    const long long n_tiles = (long long)tiles_x * ((p.ny + TY - 1) / TY);
    const long long nb = step2_grid_blocks(p, lp);
    hipLaunchKernelGGL((fused_step2_kernel<COLL, STORE, TY>), dim3(unsigned(nb)), dim3(Tile2<TY>::THREADS), 0, s, p, nb,
                       lp.remap, tiles_x, n_tiles);
    return hipGetLastError();
}

template <int STORE, int TY>
static hipError_t launch_step2_c(int coll, const StepParams& p, const Step2Policy& lp, hipStream_t s) {
    if (coll == LBM_BGK) return launch_step2_t<LBM_BGK, STORE, TY>(p, lp, s);
    if (coll == LBM_MRT) return launch_step2_t<LBM_MRT, STORE, TY>(p, lp, s);
    return launch_step2_t<LBM_CM, STORE, TY>(p, lp, s);
}

template <int STORE>
static hipError_t launch_step2_s(int coll, const StepParams& p, const Step2Policy& lp, hipStream_t s) {
    using T = typename Elem<STORE>::T;
    // population buffers (and the intermediate-state buffer) span 9 planes; edge stores are
    // far smaller
    const bool fits = buf_fits<T>(Q, p.plane);
    if (lp.strip) {
        if (lp.h < 1) return hipErrorInvalidValue;
        if (lp.w == 4) return fits ? launch_step2s_c<STORE, 4, true>(coll, p, lp, s) : launch_step2s_c<STORE, 4, false>(coll, p, lp, s);
        if (lp.w == 8 && fits) return launch_step2s_c<STORE, 8, true>(coll, p, lp, s);
        return hipErrorInvalidValue;
    }
    if (!fits) return hipErrorInvalidValue;  // tiles serve cache-resident grids only
    if (lp.ty == 6) return launch_step2_c<STORE, 6>(coll, p, lp, s);
    if (lp.ty == 14) return launch_step2_c<STORE, 14>(coll, p, lp, s);
    return hipErrorInvalidValue;
}

hipError_t launch_step2(int coll, int storage, const StepParams& p, const Step2Policy& lp, hipStream_t s) {
    if (p.n_edge != edge_count(p.nx, p.ny, p.px, p.py)) return hipErrorInvalidValue;
    if (p.n_edge > 0 && (!p.edge_prev || !p.edge_next)) return hipErrorInvalidValue;
    if (storage == LBM_F16_SHIFTED) return launch_step2_s<LBM_F16_SHIFTED>(coll, p, lp, s);
    if (storage == LBM_BF16_SHIFTED) return launch_step2_s<LBM_BF16_SHIFTED>(coll, p, lp, s);
    return launch_step2_s<LBM_F32>(coll, p, lp, s);
}

// Measured on MI355X (us per step, two-step launches, periodic grids; population bytes =
// both buffers; strips of 4 wavefronts x H rows, 6 (FP32) / 7 (16-bit) wavefronts per SIMD):
//   FP32 BGK        tiles   H=12   H=16   H=24   H=32      CM: tiles  H=12   H=16
//   1448^2  151 MB   16.4   18.0   17.0   16.1   16.6          19.1   17.5   17.4
//   2048^2  302 MB   39.3   35.3   35.0   37.3   36.5          44.0   37.3   36.5
//   2896^2  604 MB   77.1   69.8   72.3   78.9   80.4          87.8   71.6   72.4
//   4096^2  1.2 GB  151.5  134.6  139.3  149.5  153.8         170.1  137.2  138.4
//   5792^2  2.4 GB  297.8  258.5  271.3  302.8  308.7         337.5  270.4  270.4
//   8192^2  4.8 GB  580.6  508.0  522.5  591.3  615.7         659.3  531.8  521.6
//   (MRT follows BGK within 2 %.)
//   16-bit BGK      tiles   H=16   H=24   H=32   H=48   H=64   (MRT, CM best H in brackets)
//   1024^2   38 MB    9.0   10.3     -      -      -      -     (tiles, tiles)
//   1280^2   59 MB   12.7   12.2     -      -      -      -     (16, 16)
//   1448^2   75 MB   16.4   16.6   16.4   18.1   21.5   27.3   (16, 16)
//   2048^2  151 MB   30.8   26.1   25.9   26.3   27.6   31.5   (16, 16)
//   2896^2  302 MB   65.6   47.8   47.7   48.4   50.7   56.9   (32, 32)
//   4096^2  604 MB  132.8   94.1   90.2   88.3   92.8   92.9   (32, 32)
//   5792^2  1.2 GB  258.4  177.2  168.3  164.6  167.5  171.1   (48, 48)
//   8192^2  2.4 GB  508.7  354.1  328.3  324.2  320.8  321.9   (64, 64)
// Tiles win only for grids well inside the 256 MB Infinity Cache (1024^2: 9.5 us against
// 10.6 for strips), where a strip has too few rows for the march to pay. FP32 strips are
// bound by load latency (short strips: more workgroups in flight), 16-bit strips by VALU
// (longer strips: fewer recomputed halo rows).
Step2Policy step2_policy(long long plane, int storage, int coll) {
    const double bytes = 2.0 * 9.0 * double(plane) * (storage == LBM_F32 ? 4.0 : 2.0);
    Step2Policy lp;
    lp.remap = 1;
    lp.strip = 1;
    lp.w = 4;
    if (storage == LBM_F32) {
        if (bytes <= 100.0e6 || (bytes <= 256.0e6 && coll != LBM_CM)) {
            lp.strip = 0;
            lp.ty = 6;
        } else {
            lp.h = coll == LBM_CM ? 16 : 12;
        }
    } else {
        if (bytes <= 48.0e6) {
            lp.strip = 0;
            lp.ty = 6;
        } else {
            lp.h = bytes <= 200.0e6 ? 16 : (bytes <= 700.0e6 ? 32 : (bytes <= 1.6e9 ? 48 : 64));
        }
    }
    return lp;
}

hipError_t launch_edge_copy(int storage, void* buf, void* store, long long plane, int nx, int ny, int px, int py,
                            int to_store, hipStream_t s) {
    const int n_edge = edge_count(nx, ny, px, py);
    if (n_edge == 0) return hipSuccess;
    const unsigned nb = unsigned((n_edge + kBlock - 1) / kBlock);
    if (storage == LBM_F16_SHIFTED)
        hipLaunchKernelGGL(edge_copy_kernel<LBM_F16_SHIFTED>, dim3(nb), dim3(kBlock), 0, s, buf, store, plane, nx, ny, px, py, n_edge, to_store);
    else if (storage == LBM_BF16_SHIFTED)
        hipLaunchKernelGGL(edge_copy_kernel<LBM_BF16_SHIFTED>, dim3(nb), dim3(kBlock), 0, s, buf, store, plane, nx, ny, px, py, n_edge, to_store);
    else
        hipLaunchKernelGGL(edge_copy_kernel<LBM_F32>, dim3(nb), dim3(kBlock), 0, s, buf, store, plane, nx, ny, px, py, n_edge, to_store);
    return hipGetLastError();
}

template <int COLL, int STORE, bool BUF>
static hipError_t launch_ahead_t(const StepParams& p, const Step2Policy& lp, double* partials, hipStream_t s) {
    const int strips_x = (p.nx + kT2W - 1) / kT2W;
    const long long n_strips = (long long)strips_x * ((p.ny + lp.h - 1) / lp.h);
    const long long nb = step2_grid_blocks(p, lp);
    hipLaunchKernelGGL((fused_step_ahead_kernel<COLL, STORE, 4, BUF>), dim3(unsigned(nb)), dim3(kT2L * 4), 0, s, p, nb,
                       lp.remap, strips_x, n_strips, lp.h, partials);
    return hipGetLastError();
}

template <int STORE>
static hipError_t launch_ahead_s(int coll, const StepParams& p, const Step2Policy& lp, double* partials, hipStream_t s) {
    using T = typename Elem<STORE>::T;
    const bool fits = buf_fits<T>(Q, p.plane);
    (void)coll;  // the OptimalAdapter applies to CM only
    return fits ? launch_ahead_t<LBM_CM, STORE, true>(p, lp, partials, s) : launch_ahead_t<LBM_CM, STORE, false>(p, lp, partials, s);
}

long long ahead_partials_needed(const StepParams& p, const Step2Policy& lp) {
    return (long long)((p.nx + kT2W - 1) / kT2W) * ((p.ny + lp.h - 1) / lp.h);
}

Step2Policy ahead_policy(long long plane, int storage) {
    Step2Policy lp = step2_policy(plane, storage, LBM_CM);
    lp.strip = 1;
    lp.w = 4;
    if (lp.h < 1 || lp.h > 64) lp.h = 16;
    if (storage == LBM_F32 && lp.h > 16) lp.h = 16;
    return lp;
}

hipError_t launch_step_ahead(int coll, int storage, const StepParams& p, const Step2Policy& lp, double* partials,
                             hipStream_t s) {
    if (coll != LBM_CM || !lp.strip || lp.w != 4 || lp.h < 1 || !partials || !p.avg || p.mid != p.dst)
        return hipErrorInvalidValue;
    if (p.n_edge != edge_count(p.nx, p.ny, p.px, p.py)) return hipErrorInvalidValue;
    if (p.n_edge > 0 && (!p.edge_prev || !p.edge_next)) return hipErrorInvalidValue;
    if (storage == LBM_F16_SHIFTED) return launch_ahead_s<LBM_F16_SHIFTED>(coll, p, lp, partials, s);
    if (storage == LBM_BF16_SHIFTED) return launch_ahead_s<LBM_BF16_SHIFTED>(coll, p, lp, partials, s);
    return launch_ahead_s<LBM_F32>(coll, p, lp, partials, s);
}

hipError_t launch_finalize_ahead(const double* partials, long long n, long long N, float* ahead, double* ahead64, float* avg,
                                 double* avg64, hipStream_t s) {
    hipLaunchKernelGGL(finalize_ahead_kernel, dim3(1), dim3(kBlock), 0, s, partials, n, 1.0 / double(N), ahead, ahead64, avg,
                       avg64);
    return hipGetLastError();
}

hipError_t launch_finalize_avg(const double* partials, long long n, long long N, float* avg, double* avg64, hipStream_t s) {
    hipLaunchKernelGGL(finalize_avg_kernel, dim3(1), dim3(kBlock), 0, s, partials, n, 1.0 / double(N), avg, avg64);
    return hipGetLastError();
}

static unsigned blocks_for(long long n) { return unsigned((n + kBlock - 1) / kBlock); }

hipError_t launch_init(int storage, const InitParams& p, hipStream_t s) {
    unsigned nb = blocks_for(p.N);
    if (storage == LBM_F16_SHIFTED) hipLaunchKernelGGL(init_kernel<LBM_F16_SHIFTED>, dim3(nb), dim3(kBlock), 0, s, p);
    else if (storage == LBM_BF16_SHIFTED) hipLaunchKernelGGL(init_kernel<LBM_BF16_SHIFTED>, dim3(nb), dim3(kBlock), 0, s, p);
    else hipLaunchKernelGGL(init_kernel<LBM_F32>, dim3(nb), dim3(kBlock), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_pops_to_f32(int storage, const void* buf, long long plane, long long N, float rho0, float* out, hipStream_t s) {
    unsigned nb = blocks_for(N);
    if (storage == LBM_F16_SHIFTED) hipLaunchKernelGGL(pops_to_f32_kernel<LBM_F16_SHIFTED>, dim3(nb), dim3(kBlock), 0, s, buf, plane, N, rho0, out);
    else if (storage == LBM_BF16_SHIFTED) hipLaunchKernelGGL(pops_to_f32_kernel<LBM_BF16_SHIFTED>, dim3(nb), dim3(kBlock), 0, s, buf, plane, N, rho0, out);
    else hipLaunchKernelGGL(pops_to_f32_kernel<LBM_F32>, dim3(nb), dim3(kBlock), 0, s, buf, plane, N, rho0, out);
    return hipGetLastError();
}

hipError_t launch_pops_from_f32(int storage, void* b0, void* b1, long long plane, long long N, float rho0, const float* in, hipStream_t s) {
    unsigned nb = blocks_for(N);
    if (storage == LBM_F16_SHIFTED) hipLaunchKernelGGL(pops_from_f32_kernel<LBM_F16_SHIFTED>, dim3(nb), dim3(kBlock), 0, s, b0, b1, plane, N, rho0, in);
    else if (storage == LBM_BF16_SHIFTED) hipLaunchKernelGGL(pops_from_f32_kernel<LBM_BF16_SHIFTED>, dim3(nb), dim3(kBlock), 0, s, b0, b1, plane, N, rho0, in);
    else hipLaunchKernelGGL(pops_from_f32_kernel<LBM_F32>, dim3(nb), dim3(kBlock), 0, s, b0, b1, plane, N, rho0, in);
    return hipGetLastError();
}

hipError_t launch_macros_to_aos(const float* ux, const float* uy, long long N, float* u, hipStream_t s) {
    hipLaunchKernelGGL(macros_to_aos_kernel, dim3(blocks_for(N)), dim3(kBlock), 0, s, ux, uy, N, u);
    return hipGetLastError();
}

hipError_t launch_aos_to_soa(const float* u, long long N, float* ux, float* uy, hipStream_t s) {
    hipLaunchKernelGGL(aos_to_soa_kernel, dim3(blocks_for(N)), dim3(kBlock), 0, s, u, N, ux, uy);
    return hipGetLastError();
}

long long reduce_partials_needed(long long N) {
    long long chunk = 64LL * kBlock;
    return (N + chunk - 1) / chunk;
}

hipError_t launch_reduce(int kind, const float* rho, const float* ux, const float* uy, const float* ref, long long N,
                         double* partials, double* out3, hipStream_t s) {
    long long chunk = 64LL * kBlock;
    long long nb = reduce_partials_needed(N);
    hipLaunchKernelGGL(reduce_fields_kernel, dim3(unsigned(nb)), dim3(kBlock), 0, s, kind, rho, ux, uy, ref, N, chunk, partials);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(kBlock), 0, s, partials, nb, out3);
    return hipGetLastError();
}

hipError_t launch_row_mean_ux(const float* ux, int nx, int ny, double* out, hipStream_t s) {
    hipLaunchKernelGGL(row_mean_ux_kernel, dim3(unsigned(ny)), dim3(kBlock), 0, s, ux, nx, out);
    return hipGetLastError();
}

}  // namespace lbm
