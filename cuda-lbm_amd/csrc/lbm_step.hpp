// lbm_step.hpp — per-thread body of the fused D2Q9 pull-scheme step.
//
// One thread owns a segment of V consecutive nodes of one row (V = 4 for FP32 storage,
// 8 for FP16/BF16, i.e. 16 bytes per population plane). For each of the nine planes it
// pulls the segment's post-stream values from the upstream row/column of the source
// buffer — one possibly misaligned 16-byte load, which gfx950 serves as a single
// global_load_dwordx4 — collides in FP32 registers and writes 16 aligned bytes per plane
// to the destination buffer. That replaces the reference's stream, boundaries,
// macroscopics, correction, equilibrium and collide kernels (main.cu:98-114) and their
// f_eq buffer with one read and one write of each population.
//
// Segments touching a non-periodic edge, a boundary flag or a ragged row end take the
// general path: a per-node gather that applies the reference's streaming rule exactly —
// a population whose upstream node lies outside a non-periodic domain is not streamed and
// keeps the value the destination slot already holds (the push kernel of streaming.cu:20-25
// leaves that slot untouched) — followed by the boundary functor of the node's flag.
//
// The body is __host__ __device__: lbm_kernels.hip launches it on the GPU, and the CPU
// tests run the same function over the same grid in a host emulation.
#pragma once

#include "lbm_node.hpp"

#include <string.h>

namespace lbm {

// Kernel arguments (passed by value).
struct StepParams {
    int nx, ny;
    int px, py;
    long long plane;             // elements between population planes
    const void* src;             // post-collision populations of step t-1
    void* dst;                   // receives step t; holds step t-2 (stale slots)
    const uint8_t* flags;        // per-node BC flag, or null (all FLUID)
    const uint8_t* row_class;    // per row: 0 all FLUID, 1 non-FLUID only at x = 0 / nx-1, 2 other
    float omega;
    float S[9];
    float rho0;
    float fx, fy;                // uniform force
    int forced;                  // any non-zero force in play
    const float* force_field;    // AoS [N][2] or null
    const float* ibm_force;      // AoS [N][2] total force inside the IBM box, or null
    int ibm_x0, ibm_x1, ibm_y0, ibm_y1;
    int region;                  // one-step launches: 0 every node, 1 the nodes outside the IBM
                                 // box, 2 the box only (its x range is a multiple of 4 nodes)
    float wall_ux, wall_uy, u_lid;
    unsigned quirks;
    const float* avg;            // adapter averages (3) or null
    float* out_rho;              // macroscopic outputs (SoA) or null
    float* out_ux;
    float* out_uy;
    int ntx;                     // segments per row
    int nt;                      // non-temporal population stores
    // two-step launches: state t-1 (read) and t+1 (written) at the non-periodic edge nodes,
    // each [9][n_edge] in the storage format (edge_slot order)
    const void* edge_prev;
    void* edge_next;
    int n_edge;
    void* mid;                   // two-step launches: also receives state t+1, or null
};

// Row classes for the fast-path test: 0 all FLUID, 1 non-FLUID only at x = 0 or x = nx-1,
// 2 anything else (the segment's flags are then read).
inline void row_classes(const uint8_t* flags, int nx, int ny, uint8_t* out) {
    for (int y = 0; y < ny; y++) {
        uint8_t c = 0;
        for (int x = 0; x < nx; x++) {
            if (flags[(long long)y * nx + x] == LBM_FLUID) continue;
            uint8_t k = (x == 0 || x == nx - 1) ? 1 : 2;
            if (k > c) c = k;
        }
        out[y] = c;
    }
}

// ---- storage traits -------------------------------------------------------------------

LBM_HD float half_bits_to_float(uint16_t h) {
#if defined(__HIP_DEVICE_COMPILE__)
    _Float16 v;
    __builtin_memcpy(&v, &h, 2);
    return float(v);
#else
    uint32_t sign = uint32_t(h & 0x8000u) << 16, ex = (h >> 10) & 0x1fu, mant = h & 0x3ffu, u;
    if (ex == 0) {
        if (mant == 0) {
            u = sign;
        } else {
            int e = -1;
            do { mant <<= 1; e++; } while (!(mant & 0x400u));
            u = sign | (uint32_t(127 - 15 - e) << 23) | ((mant & 0x3ffu) << 13);
        }
    } else if (ex == 31) {
        u = sign | 0x7f800000u | (mant << 13);
    } else {
        u = sign | ((ex - 15 + 127) << 23) | (mant << 13);
    }
    float f;
    memcpy(&f, &u, 4);
    return f;
#endif
}

// Round to nearest even (v_cvt_f16_f32 on the GPU).
LBM_HD uint16_t float_to_half_bits(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    _Float16 v = _Float16(x);
    uint16_t h;
    __builtin_memcpy(&h, &v, 2);
    return h;
#else
    uint32_t u;
    memcpy(&u, &x, 4);
    uint32_t sign = (u >> 16) & 0x8000u, ex = (u >> 23) & 0xffu, mant = u & 0x7fffffu;
    if (ex == 0xffu) return uint16_t(sign | 0x7c00u | (mant ? 0x200u : 0u));
    int e = int(ex) - 127 + 15;
    if (e >= 31) return uint16_t(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return uint16_t(sign);
        mant |= 0x800000u;
        int shift = 14 - e;
        uint32_t hm = mant >> shift, rem = mant & ((1u << shift) - 1u), halfway = 1u << (shift - 1);
        if (rem > halfway || (rem == halfway && (hm & 1u))) hm++;
        return uint16_t(sign | hm);
    }
    uint32_t hm = mant >> 13, rem = mant & 0x1fffu, h = sign | (uint32_t(e) << 10) | hm;
    if (rem > 0x1000u || (rem == 0x1000u && (hm & 1u))) h++;
    return uint16_t(h);
#endif
}

LBM_HD float bf16_bits_to_float(uint16_t h) {
    uint32_t u = uint32_t(h) << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}
// Round to nearest even; finite inputs only (populations of a running solver).
LBM_HD uint16_t float_to_bf16_bits(float x) {
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

// Population storage of one plane row, seen as segments of V consecutive nodes (one
// segment per thread). STORE picks the element format, V the segment width: narrow
// segments give a small grid enough wavefronts to hide latency, wide ones move more bytes
// per instruction on HBM-resident grids.
template <int STORE, int V>
struct Storage;

// Stores of one 16-byte vector: plain, or non-temporal (streaming past the caches; used
// when the population buffers do not fit in the 256 MB Infinity Cache).
LBM_HD void store16(void* p, const uint32_t* d, bool nt) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (nt) {
        uint32_t* q = static_cast<uint32_t*>(p);
#pragma unroll
        for (int k = 0; k < 4; k++) __builtin_nontemporal_store(d[k], q + k);
        return;
    }
#endif
    (void)nt;
    __builtin_memcpy(p, d, 16);
}

// Stores of W dwords (W = 1, 2 or 4), plain or non-temporal.
template <int W>
LBM_HD void store_dwords(void* p, const uint32_t* d, bool nt) {
    if constexpr (W == 4) {
        store16(p, d, nt);
    } else {
#if defined(__HIP_DEVICE_COMPILE__)
        if (nt) {
#pragma unroll
            for (int k = 0; k < W; k++) __builtin_nontemporal_store(d[k], static_cast<uint32_t*>(p) + k);
            return;
        }
#endif
        (void)nt;
        __builtin_memcpy(p, d, 4 * W);
    }
}

// FP32: V nodes (1, 2 or 4) per vector. The x-shifted pull is one dword-aligned vector
// load (a single global_load_dwordx4 on gfx950 for 4 nodes).
template <int V_>
struct Storage<LBM_F32, V_> {
    using T = float;
    static constexpr int V = V_;
    static_assert(V == 1 || V == 2 || V == 4, "FP32 segments hold 1, 2 or 4 nodes");
    static constexpr int kRightHalo = 1;  // elements read past the segment on the right
    struct Vec {
        float v[V];
    };
    LBM_HD static Vec load_shifted(const T* row, int x0, int cx) {
        Vec r;
        __builtin_memcpy(r.v, row + (x0 - cx), 4 * V);
        return r;
    }
    LBM_HD static float get(const Vec& r, int j, int, float) { return r.v[j]; }
    LBM_HD static void put(Vec& r, int j, int, float, float v) { r.v[j] = v; }
    LBM_HD static void store_vec(T* p, const Vec& r, bool nt) {
        uint32_t d[V];
        __builtin_memcpy(d, r.v, 4 * V);
        store_dwords<V>(p, d, nt);
    }
    LBM_HD static float dec1(T raw, int, float) { return raw; }
    LBM_HD static T enc1(float v, int, float) { return v; }
    LBM_HD static float load1(const T* p, int, float) { return *p; }
    LBM_HD static void store1(T* p, int, float, float v) { *p = v; }
};

// 16-bit shifted storage: V nodes (2, 4 or 8) per 4-, 8- or 16-byte vector. A pull shifted
// by one node is rebuilt from the aligned vector and the neighbouring dword with a 16-bit
// funnel shift (v_alignbit_b32), keeping every load dword-aligned. At 8 nodes the MRT and CM
// bodies need over 200 VGPRs and the 8192^2 step runs ~2x slower than at 4 (MRT FP16
// 1721 us vs 868 us on MI355X), so 8 is not instantiated.
template <int STORE, int V_>
struct Shifted16 {
    using T = uint16_t;
    static constexpr int V = V_;
    static_assert(V == 2 || V == 4 || V == 8, "16-bit segments hold 2, 4 or 8 nodes");
    static constexpr int W = V / 2;       // dwords per vector
    static constexpr int kRightHalo = 2;
    struct Vec {
        uint32_t d[W];
    };
    LBM_HD static float dec(uint16_t h) { return STORE == LBM_F16_SHIFTED ? half_bits_to_float(h) : bf16_bits_to_float(h); }
    LBM_HD static uint16_t enc(float v) { return STORE == LBM_F16_SHIFTED ? float_to_half_bits(v) : float_to_bf16_bits(v); }
    LBM_HD static uint32_t funnel16(uint32_t hi, uint32_t lo) { return uint32_t(((uint64_t(hi) << 32) | lo) >> 16); }
    LBM_HD static Vec load_shifted(const T* row, int x0, int cx) {
        Vec a;
        __builtin_memcpy(a.d, row + x0, 4 * W);
        if (cx == 0) return a;
        Vec r;
        if (cx > 0) {  // nodes x0-1 .. x0+V-2
            uint32_t pv;
            __builtin_memcpy(&pv, row + x0 - 2, 4);
            r.d[0] = funnel16(a.d[0], pv);
#pragma unroll
            for (int k = 1; k < W; k++) r.d[k] = funnel16(a.d[k], a.d[k - 1]);
        } else {  // nodes x0+1 .. x0+V
            uint32_t nv;
            __builtin_memcpy(&nv, row + x0 + V, 4);
#pragma unroll
            for (int k = 0; k < W - 1; k++) r.d[k] = funnel16(a.d[k + 1], a.d[k]);
            r.d[W - 1] = funnel16(nv, a.d[W - 1]);
        }
        return r;
    }
    LBM_HD static float get(const Vec& r, int j, int q, float rho0) {
        uint32_t w = r.d[j >> 1];
        uint16_t h = uint16_t((j & 1) ? (w >> 16) : (w & 0xffffu));
        return dec(h) + wgt(q) * rho0;
    }
    LBM_HD static void put(Vec& r, int j, int q, float rho0, float v) {
        uint32_t h = enc(v - wgt(q) * rho0);
        uint32_t& w = r.d[j >> 1];
        w = (j & 1) ? ((w & 0xffffu) | (h << 16)) : ((w & 0xffff0000u) | h);
    }
    LBM_HD static void store_vec(T* p, const Vec& r, bool nt) { store_dwords<W>(p, r.d, nt); }
    LBM_HD static float dec1(T raw, int q, float rho0) { return dec(raw) + wgt(q) * rho0; }
    LBM_HD static T enc1(float v, int q, float rho0) { return enc(v - wgt(q) * rho0); }
    LBM_HD static float load1(const T* p, int q, float rho0) { return dec1(*p, q, rho0); }
    LBM_HD static void store1(T* p, int q, float rho0, float v) { *p = enc1(v, q, rho0); }
};
template <int V>
struct Storage<LBM_F16_SHIFTED, V> : Shifted16<LBM_F16_SHIFTED, V> {};
template <int V>
struct Storage<LBM_BF16_SHIFTED, V> : Shifted16<LBM_BF16_SHIFTED, V> {};

// Segment widths instantiated: wide (4 nodes, for HBM-resident grids) and narrow (FP32 1,
// 16-bit 2 nodes, for grids that stay in the Infinity Cache).
constexpr int kSegWide = 4;
template <int STORE>
constexpr int kSegNarrow = STORE == LBM_F32 ? 1 : 2;
// Single-element access (element type, load1/store1) does not depend on the width.
template <int STORE>
using Elem = Storage<STORE, kSegNarrow<STORE>>;

// ---- population arrays addressed as plane offset + node offset --------------------------
//
// The two-step kernels address a population array ([9][plane], or an edge store
// [9][n_edge]) as element `node` of plane q, with the plane offset (plus, for a pull, the
// uniform shift to the upstream node) a wavefront-uniform value. With BUF = true the GPU
// build makes the array a raw buffer resource: every access is one buffer_load/store with
// the node's byte offset in a VGPR and the plane offset in an SGPR (soffset), so there is
// no 64-bit address arithmetic per plane and no per-plane base pointer held in SGPRs.
// Buffer offsets are 32-bit, so BUF needs every byte offset below 4 GiB (buf_fits);
// BUF = false keeps 64-bit pointers for larger arrays. The host build (the CPU emulation)
// uses a plain pointer with the same offsets.
template <class T>
LBM_HD bool buf_fits(long long planes, long long plane) {
    return (unsigned long long)(planes * plane) * sizeof(T) < 0xffffffffull;
}
// Word 3 of a gfx9 raw buffer resource: DATA_FORMAT = 32 (untyped dword access).
constexpr int kBufWord3 = 0x00020000;

template <class T, bool BUF>
struct PopBuf {
    static_assert(sizeof(T) == 2 || sizeof(T) == 4, "16- or 32-bit populations");
    char* b;
#if defined(__HIP_DEVICE_COMPILE__)
    __amdgpu_buffer_rsrc_t r;
#endif
    LBM_HD explicit PopBuf(const void* ptr) : b(static_cast<char*>(const_cast<void*>(ptr))) {
#if defined(__HIP_DEVICE_COMPILE__)
        if constexpr (BUF) r = __builtin_amdgcn_make_buffer_rsrc(b, 0, -1, kBufWord3);
#endif
    }
    // element `node` of the array at element offset `off` (a plane offset, possibly plus a
    // uniform pull shift; off + node >= 0 always)
    LBM_HD T ld(uint32_t node, long long off) const {
#if defined(__HIP_DEVICE_COMPILE__)
        if constexpr (BUF) {
            const int vo = int(node * uint32_t(sizeof(T))), so = int(uint32_t(off) * uint32_t(sizeof(T)));
            if constexpr (sizeof(T) == 4)
                return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
            else
                return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b16(r, vo, so, 0));
        }
#endif
        T v;
        __builtin_memcpy(&v, b + (off + (long long)node) * (long long)sizeof(T), sizeof(T));
        return v;
    }
    LBM_HD void st(uint32_t node, long long off, T v) const {
#if defined(__HIP_DEVICE_COMPILE__)
        if constexpr (BUF) {
            const int vo = int(node * uint32_t(sizeof(T))), so = int(uint32_t(off) * uint32_t(sizeof(T)));
            if constexpr (sizeof(T) == 4)
                __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, vo, so, 0);
            else
                __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, v), r, vo, so, 0);
            return;
        }
#endif
        __builtin_memcpy(b + (off + (long long)node) * (long long)sizeof(T), &v, sizeof(T));
    }
    // st() with the non-temporal cache policy (aux bit 1: the line is not kept in L2 after
    // the write), for populations that no later access of the same launch reads
    LBM_HD void st_nt(uint32_t node, long long off, T v) const {
#if defined(__HIP_DEVICE_COMPILE__)
        if constexpr (BUF) {
            const int vo = int(node * uint32_t(sizeof(T))), so = int(uint32_t(off) * uint32_t(sizeof(T)));
            if constexpr (sizeof(T) == 4)
                __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, vo, so, 2);
            else
                __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, v), r, vo, so, 2);
            return;
        }
        __builtin_nontemporal_store(v, reinterpret_cast<T*>(b) + (off + (long long)node));
        return;
#endif
        __builtin_memcpy(b + (off + (long long)node) * (long long)sizeof(T), &v, sizeof(T));
    }
};

// Moment partial sums for the OptimalAdapter averages (FP64).
struct MomAcc {
    double rho, jmag, pimag;
};

// ---- general-path gather ----------------------------------------------------------------

// What the per-node general path needs, passed by value (kernel arguments must not be
// addressed through a generic pointer).
// `stale` is where a pull from outside a non-periodic domain reads (the slot the
// reference's push stream leaves untouched): the node's own slot in the buffer being
// written (stale_edge = 0, index = node, stride = plane), or, inside a two-step launch,
// the edge store (stale_edge = 1, index = edge_slot, stride = n_edge).
struct NodeArgs {
    const void* src;
    const void* dst;
    const void* stale;
    const uint8_t* flags;
    long long plane, stale_stride;
    int nx, ny, px, py, stale_edge;
    float rho0, u_lid, wall_ux, wall_uy;
    unsigned quirks;
};

// Slot of a node on a non-periodic domain edge in the edge store: rows y = 0 and NY-1
// first (when y is not periodic), then columns x = 0 and NX-1 (when x is not periodic).
// A corner has a row slot and a column slot; readers use the first, writers fill both.
// Returns -1 for nodes off the non-periodic edges.
LBM_HD int edge_slot(int x, int y, int nx, int ny, int px, int py) {
    if (!py) {
        if (y == 0) return x;
        if (y == ny - 1) return nx + x;
    }
    const int base = py ? 0 : 2 * nx;
    if (!px) {
        if (x == 0) return base + y;
        if (x == nx - 1) return base + ny + y;
    }
    return -1;
}
LBM_HD int edge_count(int nx, int ny, int px, int py) { return (py ? 0 : 2 * nx) + (px ? 0 : 2 * ny); }
// The node of edge slot e (inverse of edge_slot; a corner is reached from both its slots).
LBM_HD void edge_node(int e, int nx, int ny, int py, int& x, int& y) {
    const int rows = py ? 0 : 2 * nx;
    if (e < rows) {
        x = e < nx ? e : e - nx;
        y = e < nx ? 0 : ny - 1;
    } else {
        const int c = e - rows;
        x = c < ny ? 0 : nx - 1;
        y = c < ny ? c : c - ny;
    }
}

LBM_HD NodeArgs node_args(const StepParams& p) {
    NodeArgs a;
    a.src = p.src;
    a.dst = p.dst;
    a.stale = p.dst;
    a.stale_stride = p.plane;
    a.stale_edge = 0;
    a.flags = p.flags;
    a.plane = p.plane;
    a.nx = p.nx;
    a.ny = p.ny;
    a.px = p.px;
    a.py = p.py;
    a.rho0 = p.rho0;
    a.u_lid = p.u_lid;
    a.wall_ux = p.wall_ux;
    a.wall_uy = p.wall_uy;
    a.quirks = p.quirks;
    return a;
}

struct Pop9 {
    float v[Q];
};

// Post-stream value of population q at node (x, y).
template <int STORE>
LBM_HD float gather(const NodeArgs& p, int q, int x, int y) {
    using St = Elem<STORE>;
    const typename St::T* src = static_cast<const typename St::T*>(p.src);
    const typename St::T* dst = static_cast<const typename St::T*>(p.dst);
    long long node = (long long)y * p.nx + x;
    if (q == 0 && (p.quirks & LBM_QUIRK_STALE_REST)) return St::load1(dst + node, 0, p.rho0);
    int xs = x - CX[q], ys = y - CY[q];
    bool out = false;
    if (xs < 0 || xs >= p.nx) {
        if (p.px) xs = xs < 0 ? xs + p.nx : xs - p.nx;
        else out = true;
    }
    if (ys < 0 || ys >= p.ny) {
        if (p.py) ys = ys < 0 ? ys + p.ny : ys - p.ny;
        else out = true;
    }
    if (out) {
        const typename St::T* st = static_cast<const typename St::T*>(p.stale);
        long long idx = p.stale_edge ? (long long)edge_slot(x, y, p.nx, p.ny, p.px, p.py) : node;
        return St::load1(st + q * p.stale_stride + idx, q, p.rho0);
    }
    return St::load1(src + q * p.plane + (long long)ys * p.nx + xs, q, p.rho0);
}

// Flags whose functor also reads a neighbour's post-stream populations: their nodes take the
// one-step launch's per-node path (two-step launches do not cover them).
LBM_HD bool flag_reads_neighbour(uint8_t f) {
    return f == LBM_ZG_OUTFLOW || f == LBM_REGULARIZED_BOUNCE_BACK_CORNER || f == LBM_PRESSURE_OUTLET;
}

// Boundary functors that need only the node's own post-stream populations. Returns false
// for the flags that also read a neighbour (flag_reads_neighbour).
LBM_HD bool apply_local_bc(float* g, uint8_t fl, int x, int y, const NodeArgs& p) {
    Geom G{p.nx, p.ny, p.px, p.py};
    switch (fl) {
        case LBM_FLUID: return true;
        case LBM_BOUNCE_BACK: bc_bounce_back(g, x, y, G, false, 0.f, 0.f, p.rho0); return true;
        case LBM_MOVING_WALL_BOUNCE_BACK: bc_bounce_back(g, x, y, G, true, p.wall_ux, p.wall_uy, p.rho0); return true;
        case LBM_ZOU_HE_TOP: bc_zou_he_top(g, p.u_lid); return true;
        case LBM_ZOU_HE_LEFT: bc_zou_he_left(g, p.u_lid, (p.quirks & LBM_QUIRK_ZOUHE_RHO) != 0); return true;
        case LBM_REGULARIZED_INLET_TOP: bc_regularized_inlet_top(g, p.u_lid); return true;
        case LBM_REGULARIZED_BOUNCE_BACK: bc_regularized_bounce_back(g, x, y, G); return true;
        case LBM_CYLINDER: bc_cylinder(g); return true;
        default: return false;
    }
}

// General path for one node: gather, then the boundary functor of its flag. gather_node is
// the out-of-line form the step kernels call on their rare general path; gather_node_inl is
// the same body inlined, for kernels where it is the main path (the IBM band gather), whose
// call would pass the arguments and the result through scratch memory.
template <int STORE>
LBM_HD Pop9 gather_node_inl(const NodeArgs& p, int x, int y) {
    Pop9 r;
    float* g = r.v;
#pragma unroll
    for (int q = 0; q < Q; q++) g[q] = gather<STORE>(p, q, x, y);
    uint8_t fl = p.flags ? p.flags[(long long)y * p.nx + x] : uint8_t(LBM_FLUID);
    if (fl == LBM_FLUID) return r;
    if (apply_local_bc(g, fl, x, y, p)) return r;
    Geom G{p.nx, p.ny, p.px, p.py};
    switch (fl) {
        case LBM_ZG_OUTFLOW: {
            // zeroGradientOutflow.cuh:9-59: unknowns (c.n > 0) from the interior neighbour.
            int n0 = 0, n1 = 0, ix = x, iy = y;
            if (x == 0) { n0 = 1; ix = 1; }
            else if (x == p.nx - 1) { n0 = -1; ix = p.nx - 2; }
            else if (y == 0) { n1 = 1; iy = 1; }
            else if (y == p.ny - 1) { n1 = -1; iy = p.ny - 2; }
            else break;
#pragma unroll
            for (int i = 0; i < Q; i++)
                if (CX[i] * n0 + CY[i] * n1 > 0) g[i] = gather<STORE>(p, i, ix, iy);
            break;
        }
        case LBM_PRESSURE_OUTLET: {
            const int ix = x == 0 ? p.nx - 1 : x - 1;  // x = 0 only on periodic-x grids
            float gi[Q];
#pragma unroll
            for (int i = 0; i < Q; i++) gi[i] = gather<STORE>(p, i, ix, y);
            bc_pressure_outlet(g, gi);
            break;
        }
        case LBM_REGULARIZED_BOUNCE_BACK_CORNER: {
            int dx = x == 0 ? x + 1 : x - 1, dy = y == 0 ? y + 1 : y - 1;
            dx = dx < 1 ? 1 : (dx > p.nx - 2 ? p.nx - 2 : dx);
            dy = dy < 1 ? 1 : (dy > p.ny - 2 ? p.ny - 2 : dy);
            float rd = 0.f;
#pragma unroll
            for (int i = 0; i < Q; i++) rd += gather<STORE>(p, i, dx, dy);
            bc_regularized_corner(g, x, y, G, rd);
            break;
        }
        default: break;
    }
    return r;
}

template <int STORE>
__host__ __device__ __noinline__ Pop9 gather_node(NodeArgs p, int x, int y) {
    return gather_node_inl<STORE>(p, x, y);
}

// Node force: uniform, else the force field, else (inside the IBM box) the IBM total.
// Written as loads into zero-initialised values followed by value selects: "x = uniform
// field; if (c) x = load" is otherwise folded into one load through a selected address,
// which pins the kernel argument block in scratch memory and turns the load into a flat load.
struct ForceArgs {
    float fx, fy;
    const float* field;
    const float* ibm;
    int bx0, bx1, by0, by1;
};
LBM_HD ForceArgs force_args(const StepParams& p) {
    return ForceArgs{p.fx, p.fy, p.force_field, p.ibm_force, p.ibm_x0, p.ibm_x1, p.ibm_y0, p.ibm_y1};
}
LBM_HD void node_force(ForceArgs a, int x, int y, long long node, float& fx, float& fy) {
    const bool use_field = a.field != nullptr;
    const bool use_ibm = a.ibm != nullptr && x >= a.bx0 && x < a.bx1 && y >= a.by0 && y < a.by1;
    float vx = 0.0f, vy = 0.0f;
    if (use_ibm) {
        vx = a.ibm[2 * node];
        vy = a.ibm[2 * node + 1];
    } else if (use_field) {
        vx = a.field[2 * node];
        vy = a.field[2 * node + 1];
    }
    fx = (use_ibm || use_field) ? vx : a.fx;
    fy = (use_ibm || use_field) ? vy : a.fy;
}
LBM_HD void node_force(const StepParams& p, int x, int y, long long node, float& fx, float& fy) {
    node_force(force_args(p), x, y, node, fx, fy);
}

// Collide one node in place. MODE: 0 step, 1 moments only, 2 step + moments.
template <int COLL, int MODE>
LBM_HD void node_update(const StepParams& p, float* g, float fx, float fy, Macro& m, MomAcc& acc) {
    m = macroscopics(g, fx, fy);
    if (MODE != 0) {
        acc.rho += double(m.rho);
        acc.jmag += double(m.rho * sqrtf(m.ux * m.ux + m.uy * m.uy));
        acc.pimag += double(pi_magnitude(g));
    }
    if (MODE == 1) return;
    if (COLL == LBM_BGK) collide_bgk(g, m, fx, fy, p.omega, p.forced != 0);
    else if (COLL == LBM_MRT) collide_mrt(g, m, fx, fy, p.S, (p.quirks & LBM_QUIRK_MRT_FORCE_ROWS) != 0, p.forced != 0);
    else collide_cm(g, m, fx, fy, p.S, p.avg);
}

// The thread body. `seg` indexes segments row-major; out-of-range segments do nothing.
template <int COLL, int STORE, int MODE, int SEG>
LBM_HD void step_segment(const StepParams& p, long long seg, MomAcc& acc) {
    using St = Storage<STORE, SEG>;
    using T = typename St::T;
    using Vec = typename St::Vec;
    constexpr int V = St::V;
    int y, x0;
    if (p.region == 2) {  // segments of the IBM box, row by row
        const int bs = (p.ibm_x1 - p.ibm_x0 + V - 1) / V;
        if (seg >= (long long)bs * (p.ibm_y1 - p.ibm_y0)) return;
        const int r = int(seg / bs);
        y = p.ibm_y0 + r;
        x0 = p.ibm_x0 + int(seg - (long long)r * bs) * V;
    } else {
        const long long nseg = (long long)p.ntx * p.ny;
        if (seg >= nseg) return;
        y = int(seg / p.ntx);
        x0 = int(seg - (long long)y * p.ntx) * V;
        if (p.region == 1 && y >= p.ibm_y0 && y < p.ibm_y1 && x0 >= p.ibm_x0 && x0 < p.ibm_x1) return;
    }
    const long long row0 = (long long)y * p.nx;
    const T* src = static_cast<const T*>(p.src);
    T* dst = static_cast<T*>(p.dst);

    bool fast = x0 >= 1 && x0 + V + St::kRightHalo <= p.nx;
    if (fast && !p.py) fast = y >= 1 && y <= p.ny - 2;
    if (fast && p.row_class) {
        uint8_t rc = p.row_class[y];
        if (rc == 2) {
#pragma unroll
            for (int j = 0; j < V; j++) fast = fast && p.flags[row0 + x0 + j] == LBM_FLUID;
        }
    }
    if (fast && (p.quirks & LBM_QUIRK_STALE_REST)) fast = false;

    if (fast) {
        float pop[Q][V];
#pragma unroll
        for (int q = 0; q < Q; q++) {
            int ys = y - CY[q];
            if (p.py) ys = ys < 0 ? ys + p.ny : (ys >= p.ny ? ys - p.ny : ys);
            Vec raw = St::load_shifted(src + q * p.plane + (long long)ys * p.nx, x0, CX[q]);
#pragma unroll
            for (int j = 0; j < V; j++) pop[q][j] = St::get(raw, j, q, p.rho0);
        }
        float orho[V], oux[V], ouy[V];
#pragma unroll
        for (int j = 0; j < V; j++) {
            float g[Q];
#pragma unroll
            for (int q = 0; q < Q; q++) g[q] = pop[q][j];
            float fx, fy;
            node_force(p, x0 + j, y, row0 + x0 + j, fx, fy);
            Macro m;
            node_update<COLL, MODE>(p, g, fx, fy, m, acc);
#pragma unroll
            for (int q = 0; q < Q; q++) pop[q][j] = g[q];
            orho[j] = m.rho;
            oux[j] = m.ux;
            ouy[j] = m.uy;
        }
        if (MODE == 1) return;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            Vec raw;
#pragma unroll
            for (int j = 0; j < V; j++) St::put(raw, j, q, p.rho0, pop[q][j]);
            St::store_vec(dst + q * p.plane + row0 + x0, raw, p.nt != 0);
        }
        if (p.out_rho) {
            __builtin_memcpy(p.out_rho + row0 + x0, orho, sizeof(orho));
            __builtin_memcpy(p.out_ux + row0 + x0, oux, sizeof(oux));
            __builtin_memcpy(p.out_uy + row0 + x0, ouy, sizeof(ouy));
        }
        return;
    }

    // general path, one node at a time
    const int nvalid = p.nx - x0 < V ? p.nx - x0 : V;
    const NodeArgs na = node_args(p);
    for (int j = 0; j < nvalid; j++) {
        const int x = x0 + j;
        Pop9 r = gather_node<STORE>(na, x, y);
        float fx, fy;
        node_force(p, x, y, row0 + x, fx, fy);
        Macro m;
        node_update<COLL, MODE>(p, r.v, fx, fy, m, acc);
        if (MODE == 1) continue;
#pragma unroll
        for (int q = 0; q < Q; q++) St::store1(dst + q * p.plane + row0 + x, q, p.rho0, r.v[q]);
        if (p.out_rho) {
            p.out_rho[row0 + x] = m.rho;
            p.out_ux[row0 + x] = m.ux;
            p.out_uy[row0 + x] = m.uy;
        }
    }
}

// ---- two steps per launch (temporal blocking) --------------------------------------------
//
// A workgroup owns a tile of kT2W x TY nodes and runs TY + 2 wavefronts: wavefront w, lane
// l holds node (x0 - 1 + l, y0 - 1 + w), so the tile plus a one-node halo is exactly
// 64 x (TY + 2) threads. Phase A computes every thread's post-collision state t+1 from the
// state t in HBM and keeps it in LDS in the storage format (every value is what a one-step
// launch would have stored); phase B computes state t+2 of the tile's inner kT2W x TY nodes
// from LDS and writes it to HBM. The populations cross HBM once per two steps; the halo
// costs 64 (TY + 2) / (kT2W TY) phase-A node updates per node.
//
// Unstreamed slots: a one-step launch reads them from the buffer it writes (state t-1 at
// step t+1, state t at step t+2). Here phase A reads state t-1 from a small edge store,
// because the destination buffer is being overwritten by other workgroups during the
// launch, and saves state t+1 of the tile's own edge nodes into the next edge store. For
// phase B, the halo lanes outside a non-periodic domain copy the edge nodes' state-t slots
// from the source buffer into LDS (step2_fill_halo), so phase B pulls only from LDS. Boundary functors that read a neighbour's
// populations (ZG outflow, regularized corner) and the stale-rest quirk need the one-step
// launch.
constexpr int kT2L = 64;          // lanes of a tile row (one wavefront)
constexpr int kT2W = kT2L - 2;    // output columns per tile

template <int TY>
struct Tile2 {
    static constexpr int ROWS = TY + 2, NH = kT2L * ROWS, THREADS = NH;
};

// True when every pull source of columns x0 - 1 .. x0 + kT2W lies in the grid without a
// periodic wrap (columns x0 - 2 .. x0 + kT2W + 1), so none of these columns is an x edge.
LBM_HD bool interior_cols(const StepParams& p, int x0) { return x0 >= 2 && x0 + kT2W + 1 < p.nx; }
// The same for an unwrapped row y: rows y - 1 .. y + 1 in the grid.
LBM_HD bool interior_row(const StepParams& p, int y) { return y >= 1 && y < p.ny - 1; }

// Pull sources of phase-A node (gx, gy) (already wrapped and inside the domain); defined
// with the strip form below.
template <int STORE, bool BUF>
LBM_HD void step2_gather(const StepParams& p, int gx, int gy, float* g, bool affine);

// Phase-A slot of a halo position (xh, yh) outside a non-periodic domain (unwrapped
// coordinates; lds points at its direction-0 entry, planes `stride` apart). The only node
// that pulls direction q from here is (xh + c_x, yh + c_y); its upstream node is outside
// the domain, so it keeps its unstreamed slot, its own state-t population src[q][node]
// (the buffer a one-step launch would write at step t+2 holds state t). Storing that value
// (raw, in the storage format) lets phase B pull every direction from LDS, edge nodes
// included. Entries whose puller lies outside the domain are never read.
template <int STORE, bool BUF>
LBM_HD void step2_fill_halo(const StepParams& p, typename Elem<STORE>::T* lds, int stride, int xh, int yh) {
    using T = typename Elem<STORE>::T;
    const PopBuf<T, BUF> src(p.src);
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int tx = xh + CX[q], ty = yh + CY[q];
        if (tx < 0 || tx >= p.nx || ty < 0 || ty >= p.ny) continue;
        lds[q * stride] = src.ld(uint32_t(ty) * uint32_t(p.nx) + uint32_t(tx), q * p.plane);
    }
}

// Phase A of a tile, thread t: node (x0 - 1 + t % 64, y0 - 1 + t / 64) at state t+1 into
// the tile. interior_x = interior_cols(p, x0).
template <int COLL, int STORE, int TY, bool BUF>
LBM_HD void step2_phase_a(const StepParams& p, typename Elem<STORE>::T* tile, int t, int x0, int y0, bool interior_x) {
    using St = Elem<STORE>;
    using T = typename St::T;
    constexpr int NH = Tile2<TY>::NH;
    const int l = t % kT2L, w = t / kT2L;
    int gx = x0 - 1 + l, gy = y0 - 1 + w;
    const bool own = l >= 1 && l <= kT2W && w >= 1 && w <= TY && gx < p.nx && gy < p.ny;
    const bool affine = interior_x && interior_row(p, gy);  // uniform per wavefront
    if ((!p.px && (gx < 0 || gx >= p.nx)) || (!p.py && (gy < 0 || gy >= p.ny))) {
        step2_fill_halo<STORE, BUF>(p, tile + t, NH, gx, gy);
        return;
    }
    if (gx < 0 || gx >= p.nx) gx = ((gx % p.nx) + p.nx) % p.nx;
    if (gy < 0 || gy >= p.ny) gy = ((gy % p.ny) + p.ny) % p.ny;
    const uint32_t node = uint32_t(gy) * uint32_t(p.nx) + uint32_t(gx);
    const uint8_t fl = p.flags ? p.flags[node] : uint8_t(LBM_FLUID);
    const bool ex = !p.px && (gx == 0 || gx == p.nx - 1), ey = !p.py && (gy == 0 || gy == p.ny - 1);
    // pull sources; a source outside a non-periodic domain reads the node's unstreamed slot,
    // state t-1, from the edge store
    float g[Q];
    step2_gather<STORE, BUF>(p, gx, gy, g, affine);
    if (fl != LBM_FLUID) apply_local_bc(g, fl, gx, gy, node_args(p));
    float fx, fy;
    node_force(p, gx, gy, node, fx, fy);
    Macro m;
    MomAcc acc{0.0, 0.0, 0.0};
    node_update<COLL, 0>(p, g, fx, fy, m, acc);
#pragma unroll
    for (int q = 0; q < Q; q++) St::store1(tile + q * NH + t, q, p.rho0, g[q]);
    if (own && p.mid) {
        const PopBuf<T, BUF> md(p.mid);
#pragma unroll
        for (int q = 0; q < Q; q++) md.st(node, q * p.plane, St::enc1(g[q], q, p.rho0));
    }
    if (own && (ex || ey)) {
        const PopBuf<T, BUF> en(p.edge_next);
        const int base = p.py ? 0 : 2 * p.nx;
        int slot[2] = {-1, -1};
        if (ey) slot[0] = gy == 0 ? gx : p.nx + gx;
        if (ex) slot[1] = base + (gx == 0 ? gy : p.ny + gy);
#pragma unroll
        for (int k = 0; k < 2; k++) {
            if (slot[k] < 0) continue;
#pragma unroll
            for (int q = 0; q < Q; q++) en.st(uint32_t(slot[k]), q * (long long)p.n_edge, St::enc1(g[q], q, p.rho0));
        }
    }
}

// Phase B of a tile, thread t: state t+2 of the inner node from the tile to HBM. Every
// pull comes from the tile: unstreamed slots were placed there by step2_fill_halo.
template <int COLL, int STORE, int TY, bool BUF>
LBM_HD void step2_phase_b(const StepParams& p, const typename Elem<STORE>::T* tile, int t, int x0, int y0) {
    using St = Elem<STORE>;
    using T = typename St::T;
    constexpr int NH = Tile2<TY>::NH;
    const int l = t % kT2L, w = t / kT2L;
    if (l < 1 || l > kT2W || w < 1 || w > TY) return;
    const int x = x0 - 1 + l, y = y0 - 1 + w;
    if (x >= p.nx || y >= p.ny) return;
    const uint32_t node = uint32_t(y) * uint32_t(p.nx) + uint32_t(x);
    float g[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) g[q] = St::load1(tile + q * NH + t - CY[q] * kT2L - CX[q], q, p.rho0);
    const uint8_t fl = p.flags ? p.flags[node] : uint8_t(LBM_FLUID);
    if (fl != LBM_FLUID) apply_local_bc(g, fl, x, y, node_args(p));
    float fx, fy;
    node_force(p, x, y, node, fx, fy);
    Macro m;
    MomAcc acc{0.0, 0.0, 0.0};
    node_update<COLL, 0>(p, g, fx, fy, m, acc);
    const PopBuf<T, BUF> dst(p.dst);
#pragma unroll
    for (int q = 0; q < Q; q++) dst.st(node, q * p.plane, St::enc1(g[q], q, p.rho0));
    if (p.out_rho) {
        p.out_rho[node] = m.rho;
        p.out_ux[node] = m.ux;
        p.out_uy[node] = m.uy;
    }
}

// ---- two steps per launch, marching strips ------------------------------------------------
//
// A workgroup of W wavefronts owns a strip of kT2W columns and H rows and marches up it in
// chunks of W rows: iteration k computes phase-A rows kW .. kW+W-1 of the strip (strip row i
// = grid row y0 - 1 + i, one row per wavefront, lane l = column x0 - 1 + l) into a ring of
// W + 2 LDS rows, then phase-B rows kW - 2 .. kW + W - 3 (grid row y0 + j) from the three
// ring rows each needs. A phase-A row is computed once per strip (the tile form recomputes
// the two halo rows of every tile). Pull sources are loaded when their chunk starts: holding
// the next chunk's sources in registers through phase B measured slower (5-6 % on 8192^2),
// because the registers it costs lose more wavefronts per SIMD than the overlap gains.
// Unstreamed slots are handled as in the tile form (edge store for phase A, halo slots
// filled from the source buffer for phase B).

// Pull sources of phase-A node (gx, gy) (already wrapped and inside the domain). `affine`,
// the same for a whole wavefront, promises that every pull source of the wavefront is in
// the grid without a wrap (interior_cols and interior_row): the source of direction q is
// then node - (c_x + c_y nx), one uniform shift per plane, and the load needs no per-lane
// address arithmetic. Otherwise every node takes the general path (periodic wrap, and the
// edge store for sources outside a non-periodic domain), which loads the same values for
// interior nodes; a wavefront runs one of the two, never both.
template <int STORE, bool BUF>
LBM_HD void step2_gather(const StepParams& p, int gx, int gy, float* g, bool affine) {
    using St = Elem<STORE>;
    using T = typename St::T;
    const PopBuf<T, BUF> src(p.src);
    if (affine) {
        const uint32_t node = uint32_t(gy) * uint32_t(p.nx) + uint32_t(gx);
#pragma unroll
        for (int q = 0; q < Q; q++)
            g[q] = St::dec1(src.ld(node, q * p.plane - CX[q] - CY[q] * (long long)p.nx), q, p.rho0);
        return;
    }
    const bool ex = !p.px && (gx == 0 || gx == p.nx - 1), ey = !p.py && (gy == 0 || gy == p.ny - 1);
    const int xm = gx == 0 ? (p.px ? p.nx - 1 : -1) : gx - 1, xp = gx == p.nx - 1 ? (p.px ? 0 : -1) : gx + 1;
    const int ym = gy == 0 ? (p.py ? p.ny - 1 : -1) : gy - 1, yp = gy == p.ny - 1 ? (p.py ? 0 : -1) : gy + 1;
    const uint32_t rm = uint32_t(ym) * uint32_t(p.nx), r0 = uint32_t(gy) * uint32_t(p.nx), rp = uint32_t(yp) * uint32_t(p.nx);
    const PopBuf<T, BUF> prev(p.edge_prev);
    const uint32_t eslot = (ex || ey) ? uint32_t(edge_slot(gx, gy, p.nx, p.ny, p.px, p.py)) : 0u;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int xs = CX[q] > 0 ? xm : (CX[q] < 0 ? xp : gx);
        const int ys = CY[q] > 0 ? ym : (CY[q] < 0 ? yp : gy);
        const uint32_t row = CY[q] > 0 ? rm : (CY[q] < 0 ? rp : r0);
        if (xs < 0 || ys < 0)
            g[q] = St::dec1(prev.ld(eslot, q * (long long)p.n_edge), q, p.rho0);
        else
            g[q] = St::dec1(src.ld(row + uint32_t(xs), q * p.plane), q, p.rho0);
    }
}

// Grid coordinates of strip node (lane l, strip row i); false outside a non-periodic domain.
LBM_HD bool strip_node(const StepParams& p, int x0, int y0, int l, int i, int& gx, int& gy) {
    gx = x0 - 1 + l;
    gy = y0 - 1 + i;
    if (gx < 0 || gx >= p.nx) {
        if (!p.px) return false;
        gx = ((gx % p.nx) + p.nx) % p.nx;
    }
    if (gy < 0 || gy >= p.ny) {
        if (!p.py) return false;
        gy = ((gy % p.ny) + p.ny) % p.ny;
    }
    return true;
}

// Phase A of strip node (l, i) from its gathered sources g: collide, store into the ring
// row `slot` and, for the strip's own edge nodes, into the next edge store.
// AHEAD (one step per launch, the look-ahead form below): the launch's output is phase A
// itself, so the host points p.mid at the destination; own nodes also store the step's
// fields, and the next edge store receives their current (source) state, which is what
// the next launch's unstreamed slots hold.
template <int COLL, int STORE, bool BUF, bool AHEAD = false>
LBM_HD void step2s_phase_a(const StepParams& p, typename Elem<STORE>::T* ring, int ring_rows, int slot, int l, int i,
                           int x0, int y0, int hs, int gx, int gy, float* g) {
    using St = Elem<STORE>;
    using T = typename St::T;
    const uint32_t node = uint32_t(gy) * uint32_t(p.nx) + uint32_t(gx);
    const uint8_t fl = p.flags ? p.flags[node] : uint8_t(LBM_FLUID);
    if (fl != LBM_FLUID) apply_local_bc(g, fl, gx, gy, node_args(p));
    float fx, fy;
    node_force(p, gx, gy, node, fx, fy);
    Macro m;
    MomAcc acc{0.0, 0.0, 0.0};
    node_update<COLL, 0>(p, g, fx, fy, m, acc);
    const int stride = ring_rows * kT2L;
#pragma unroll
    for (int q = 0; q < Q; q++) St::store1(ring + q * stride + slot * kT2L + l, q, p.rho0, g[q]);
    const bool ex = !p.px && (gx == 0 || gx == p.nx - 1), ey = !p.py && (gy == 0 || gy == p.ny - 1);
    const bool own = l >= 1 && l <= kT2W && i >= 1 && i <= hs && x0 - 1 + l < p.nx;
    if (own && p.mid) {
        const PopBuf<T, BUF> md(p.mid);
#pragma unroll
        for (int q = 0; q < Q; q++) md.st(node, q * p.plane, St::enc1(g[q], q, p.rho0));
    }
    if (AHEAD && own && p.out_rho) {
        p.out_rho[node] = m.rho;
        p.out_ux[node] = m.ux;
        p.out_uy[node] = m.uy;
    }
    if (own && (ex || ey)) {
        const PopBuf<T, BUF> en(p.edge_next);
        const PopBuf<T, BUF> src(p.src);
        const int base = p.py ? 0 : 2 * p.nx;
        int es[2] = {-1, -1};
        if (ey) es[0] = gy == 0 ? gx : p.nx + gx;
        if (ex) es[1] = base + (gx == 0 ? gy : p.ny + gy);
#pragma unroll
        for (int k = 0; k < 2; k++) {
            if (es[k] < 0) continue;
#pragma unroll
            for (int q = 0; q < Q; q++)
                en.st(uint32_t(es[k]), q * (long long)p.n_edge,
                      AHEAD ? src.ld(node, q * p.plane) : St::enc1(g[q], q, p.rho0));
        }
    }
}

// Look-ahead phase B of strip row j, lane l (AHEAD launches): the post-stream, boundary-
// applied state of the next step from the ring, reduced to the OptimalAdapter's moment sums
// (macroscopics.cuh:50-120) in acc; nothing is collided or stored. The adapter's averages
// of step t+2 are thus ready when the launch of step t+2 starts, without a separate pass
// over the populations.
template <int STORE, bool BUF>
LBM_HD void step2s_phase_b_moments(const StepParams& p, const typename Elem<STORE>::T* ring, int ring_rows, int l,
                                   int j, int x0, int y0, MomAcc& acc) {
    using St = Elem<STORE>;
    if (l < 1 || l > kT2W) return;
    const int x = x0 - 1 + l, y = y0 + j;
    if (x >= p.nx) return;
    const uint32_t node = uint32_t(y) * uint32_t(p.nx) + uint32_t(x);
    const int stride = ring_rows * kT2L;
    const int s_lo = j % ring_rows, s_mid = (j + 1) % ring_rows, s_hi = (j + 2) % ring_rows;
    float g[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int srow = CY[q] > 0 ? s_lo : (CY[q] < 0 ? s_hi : s_mid);
        g[q] = St::load1(ring + q * stride + srow * kT2L + (l - CX[q]), q, p.rho0);
    }
    const uint8_t fl = p.flags ? p.flags[node] : uint8_t(LBM_FLUID);
    if (fl != LBM_FLUID) apply_local_bc(g, fl, x, y, node_args(p));
    float fx, fy;
    node_force(p, x, y, node, fx, fy);
    Macro m;
    node_update<LBM_BGK, 1>(p, g, fx, fy, m, acc);
}

// Phase B of strip row j (grid row y0 + j), lane l: pulls from ring rows j, j+1, j+2 (the
// unstreamed slots of edge nodes are there too, from step2_fill_halo).
template <int COLL, int STORE, bool BUF>
LBM_HD void step2s_phase_b(const StepParams& p, const typename Elem<STORE>::T* ring, int ring_rows, int l, int j,
                           int x0, int y0) {
    using St = Elem<STORE>;
    using T = typename St::T;
    if (l < 1 || l > kT2W) return;
    const int x = x0 - 1 + l, y = y0 + j;
    if (x >= p.nx) return;
    const uint32_t node = uint32_t(y) * uint32_t(p.nx) + uint32_t(x);
    const int stride = ring_rows * kT2L;
    // ring rows of strip rows j (below), j + 1 (own), j + 2 (above)
    const int s_lo = j % ring_rows, s_mid = (j + 1) % ring_rows, s_hi = (j + 2) % ring_rows;
    float g[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int srow = CY[q] > 0 ? s_lo : (CY[q] < 0 ? s_hi : s_mid);
        g[q] = St::load1(ring + q * stride + srow * kT2L + (l - CX[q]), q, p.rho0);
    }
    const uint8_t fl = p.flags ? p.flags[node] : uint8_t(LBM_FLUID);
    if (fl != LBM_FLUID) apply_local_bc(g, fl, x, y, node_args(p));
    float fx, fy;
    node_force(p, x, y, node, fx, fy);
    Macro m;
    MomAcc acc{0.0, 0.0, 0.0};
    node_update<COLL, 0>(p, g, fx, fy, m, acc);
    const PopBuf<T, BUF> dst(p.dst);
#pragma unroll
    for (int q = 0; q < Q; q++) dst.st(node, q * p.plane, St::enc1(g[q], q, p.rho0));
    if (p.out_rho) {
        p.out_rho[node] = m.rho;
        p.out_ux[node] = m.ux;
        p.out_uy[node] = m.uy;
    }
}

}  // namespace lbm
