// mfma_mrt.cpp — the MFMA experiment for the MRT moment transforms (D2Q9), measured against
// the VALU form. Not part of liblbm_hip.so: the product's MRT keeps the VALU form because this
// experiment shows no gain (numbers and the rocprofv3 summary: profiles/mfma_mrt_experiment.txt).
//
// Both kernels run the same MRT collision on SoA FP32 populations f[9][N] (no streaming, so
// the two differ only in how the transforms are done): m = M f, m* = m - S (m - m_eq(m)),
// f* = M^-1 m*, with the Lallemand-Luo basis order of the reference (lbm_constants.cuh:33-55;
// moments rho, e, eps, jx, qx, jy, qy, pxx, pxy).
//   mrt_valu: one node per lane, both 9 x 9 products as dense FMAs in registers.
//   mrt_mfma: 16 nodes per MFMA block, D = M F with v_mfma_f32_16x16x4f32 (M padded to 16 x 16,
//             F = 9 population rows of 16 nodes, K in steps of 4), the B operand loaded straight
//             from the population planes in the MFMA lane layout (lane l: row 4k + l/16, node
//             l % 16); the equilibrium needs rho, jx, jy of each node, fetched by cross-lane
//             shuffles; the moments are moved into the B layout of the inverse product by
//             shuffles; f* = M^-1 m* by MFMA; 4 stores per lane.
// usage: mfma_mrt [n_side] [reps]   (N = n_side^2 nodes; default 1024, 20)
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#define CHECK(x)                                                                              \
    do {                                                                                      \
        hipError_t e_ = (x);                                                                  \
        if (e_ != hipSuccess) {                                                               \
            std::printf("%s:%d %s\n", __FILE__, __LINE__, hipGetErrorString(e_));             \
            std::exit(1);                                                                     \
        }                                                                                     \
    } while (0)

namespace {

constexpr int Q = 9;
// velocity set (rest, E, N, W, S, NE, NW, SW, SE) and the moment basis rows
constexpr int CX[Q] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int CY[Q] = {0, 0, 1, 0, -1, 1, 1, -1, -1};

__constant__ float cM[Q * Q];
__constant__ float cMinv[Q * Q];
__constant__ float cS[Q];

void basis(float* M) {
    for (int q = 0; q < Q; q++) {
        const float cx = float(CX[q]), cy = float(CY[q]), c2 = cx * cx + cy * cy;
        const float row[Q] = {1.0f,
                              -4.0f + 3.0f * c2,
                              4.0f - 10.5f * c2 + 4.5f * c2 * c2,
                              cx,
                              (-5.0f + 3.0f * c2) * cx,
                              cy,
                              (-5.0f + 3.0f * c2) * cy,
                              cx * cx - cy * cy,
                              cx * cy};
        for (int i = 0; i < Q; i++) M[i * Q + q] = row[i];
    }
}
void invert(const float* M, float* R) {  // Gauss-Jordan in double
    double a[Q][2 * Q];
    for (int i = 0; i < Q; i++)
        for (int j = 0; j < 2 * Q; j++) a[i][j] = j < Q ? M[i * Q + j] : (j - Q == i ? 1.0 : 0.0);
    for (int c = 0; c < Q; c++) {
        int p = c;
        for (int r = c + 1; r < Q; r++)
            if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        for (int j = 0; j < 2 * Q; j++) std::swap(a[c][j], a[p][j]);
        const double d = a[c][c];
        for (int j = 0; j < 2 * Q; j++) a[c][j] /= d;
        for (int r = 0; r < Q; r++)
            if (r != c) {
                const double f = a[r][c];
                for (int j = 0; j < 2 * Q; j++) a[r][j] -= f * a[c][j];
            }
    }
    for (int i = 0; i < Q; i++)
        for (int j = 0; j < Q; j++) R[i * Q + j] = float(a[i][Q + j]);
}

// equilibrium moment of row i from rho, jx, jy
__device__ __forceinline__ float meq(int i, float rho, float jx, float jy) {
    const float inv = 1.0f / rho, j2 = (jx * jx + jy * jy) * inv;
    switch (i) {
        case 0: return rho;
        case 1: return -2.0f * rho + 3.0f * j2;
        case 2: return rho - 3.0f * j2;
        case 3: return jx;
        case 4: return -jx;
        case 5: return jy;
        case 6: return -jy;
        case 7: return (jx * jx - jy * jy) * inv;
        case 8: return jx * jy * inv;
        default: return 0.0f;
    }
}

__global__ __launch_bounds__(256) void mrt_valu(const float* __restrict__ f, float* __restrict__ out, long long N) {
    const long long n = (long long)blockIdx.x * 256 + threadIdx.x;
    if (n >= N) return;
    float g[Q], m[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) g[q] = f[q * N + n];
#pragma unroll
    for (int i = 0; i < Q; i++) {
        float s = 0.0f;
#pragma unroll
        for (int q = 0; q < Q; q++) s += cM[i * Q + q] * g[q];
        m[i] = s;
    }
    const float rho = m[0], jx = m[3], jy = m[5];
#pragma unroll
    for (int i = 0; i < Q; i++) m[i] -= cS[i] * (m[i] - meq(i, rho, jx, jy));
#pragma unroll
    for (int q = 0; q < Q; q++) {
        float s = 0.0f;
#pragma unroll
        for (int i = 0; i < Q; i++) s += cMinv[q * Q + i] * m[i];
        out[q * N + n] = s;
    }
}

typedef float f4 __attribute__((ext_vector_type(4)));

// 64 nodes per wavefront (4 blocks of 16), 4 wavefronts per workgroup. N % 64 == 0.
__global__ __launch_bounds__(256) void mrt_mfma(const float* __restrict__ f, float* __restrict__ out, long long N) {
    const int lane = int(threadIdx.x & 63), grp = lane >> 4, col = lane & 15;
    const long long wave = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const long long base = wave * 64;
    if (base >= N) return;
    // A operands: lane holds A[i = col][k = 4 kb + grp]
    float aM[3], aMi[3];
#pragma unroll
    for (int kb = 0; kb < 3; kb++) {
        const int k = 4 * kb + grp;
        aM[kb] = (col < Q && k < Q) ? cM[col * Q + k] : 0.0f;
        aMi[kb] = (col < Q && k < Q) ? cMinv[col * Q + k] : 0.0f;
    }
#pragma unroll
    for (int blk = 0; blk < 4; blk++) {
        const long long node = base + 16 * blk + col;
        // forward: D = M F; B[k][n] = f[4 kb + grp][node]
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < 3; kb++) {
            const int q = 4 * kb + grp;
            const float b = q < Q ? f[q * N + node] : 0.0f;
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(aM[kb], b, acc, 0, 0, 0);
        }
        // lane holds moments rows 4 grp + j of node col; rho, jx (group 0 rows 0, 3), jy (group 1 row 1)
        const float rho = __shfl(acc[0], col, 64), jx = __shfl(acc[3], col, 64), jy = __shfl(acc[1], 16 + col, 64);
        f4 ms;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int i = 4 * grp + j;
            ms[j] = i < Q ? acc[j] - cS[i] * (acc[j] - meq(i, rho, jx, jy)) : 0.0f;
        }
        // inverse: B[k][n] = m*[4 kb + grp][col], held by lane 16 kb + col in its register grp
        f4 res = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < 3; kb++) {
            const int src = 16 * kb + col;
            const float v0 = __shfl(ms[0], src, 64), v1 = __shfl(ms[1], src, 64), v2 = __shfl(ms[2], src, 64),
                        v3 = __shfl(ms[3], src, 64);
            const float b = grp == 0 ? v0 : (grp == 1 ? v1 : (grp == 2 ? v2 : v3));
            res = __builtin_amdgcn_mfma_f32_16x16x4f32(aMi[kb], b, res, 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int q = 4 * grp + j;
            if (q < Q) out[q * N + node] = res[j];
        }
    }
}

}  // namespace

int main(int argc, char** argv) {
    const int side = argc > 1 ? std::atoi(argv[1]) : 1024;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 20;
    const long long N = (long long)side * side;
    if (side <= 0 || N % 64 != 0 || reps <= 0) {
        std::printf("n_side^2 must be a positive multiple of 64\n");
        return 2;
    }
    float M[Q * Q], Mi[Q * Q], S[Q] = {0.0f, 1.1f, 1.2f, 0.0f, 1.4f, 0.0f, 1.4f, 1.7f, 1.7f};
    basis(M);
    invert(M, Mi);
    CHECK(hipMemcpyToSymbol(HIP_SYMBOL(cM), M, sizeof M));
    CHECK(hipMemcpyToSymbol(HIP_SYMBOL(cMinv), Mi, sizeof Mi));
    CHECK(hipMemcpyToSymbol(HIP_SYMBOL(cS), S, sizeof S));
    const float w[Q] = {4.f / 9, 1.f / 9, 1.f / 9, 1.f / 9, 1.f / 9, 1.f / 36, 1.f / 36, 1.f / 36, 1.f / 36};
    std::vector<float> h(size_t(Q) * N);
    unsigned r = 12345u;
    for (int q = 0; q < Q; q++)
        for (long long n = 0; n < N; n++) {
            r = r * 1664525u + 1013904223u;
            h[size_t(q * N + n)] = w[q] * (1.0f + 0.05f * (float(r >> 8) / 16777216.0f - 0.5f));
        }
    float *df, *da, *db;
    CHECK(hipMalloc(&df, h.size() * 4));
    CHECK(hipMalloc(&da, h.size() * 4));
    CHECK(hipMalloc(&db, h.size() * 4));
    CHECK(hipMemcpy(df, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    const dim3 gv(unsigned((N + 255) / 256)), gm(unsigned((N / 64 + 3) / 4)), blk(256);
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    float tv = 0.0f, tm = 0.0f;
    for (int pass = 0; pass < 2; pass++) {  // pass 0 warms up
        CHECK(hipEventRecord(e0));
        for (int i = 0; i < reps; i++) hipLaunchKernelGGL(mrt_valu, gv, blk, 0, 0, df, da, N);
        CHECK(hipEventRecord(e1));
        CHECK(hipEventSynchronize(e1));
        CHECK(hipEventElapsedTime(&tv, e0, e1));
        CHECK(hipEventRecord(e0));
        for (int i = 0; i < reps; i++) hipLaunchKernelGGL(mrt_mfma, gm, blk, 0, 0, df, db, N);
        CHECK(hipEventRecord(e1));
        CHECK(hipEventSynchronize(e1));
        CHECK(hipEventElapsedTime(&tm, e0, e1));
    }
    CHECK(hipGetLastError());
    std::vector<float> a(h.size()), b(h.size());
    CHECK(hipMemcpy(a.data(), da, a.size() * 4, hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(b.data(), db, b.size() * 4, hipMemcpyDeviceToHost));
    double md = 0.0;
    for (size_t i = 0; i < a.size(); i++) md = std::fmax(md, std::fabs(double(a[i]) - b[i]) / std::fabs(double(a[i]) + 1e-30));
    const double bytes = 2.0 * Q * 4.0 * double(N);
    std::printf("nodes %lld reps %d\n", N, reps);
    std::printf("mrt_valu %.2f us/launch  %.1f GB/s\n", 1000.0 * tv / reps, bytes / (1e-3 * tv / reps) / 1e9);
    std::printf("mrt_mfma %.2f us/launch  %.1f GB/s\n", 1000.0 * tm / reps, bytes / (1e-3 * tm / reps) / 1e9);
    std::printf("max relative difference %.3e\n", md);
    CHECK(hipFree(df));
    CHECK(hipFree(da));
    CHECK(hipFree(db));
    return 0;
}
