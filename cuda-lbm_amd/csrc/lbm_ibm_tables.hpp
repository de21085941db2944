// lbm_ibm_tables.hpp — host-side construction of the IBM band tables (no HIP calls).
//
// IBMManager<2> (IBMManager.cuh:31-51, 222-252) touches, per marker, the nodes of the
// kernel stencil around it: 4x4 nodes at floor(p) - 1 .. floor(p) + 2 with the Peskin
// 4-point kernel (IBMUtils.cuh:23-43), or the reference's 2x2 stencil at floor(p) .. floor(p)+1
// (IBM_impl.cu:5-51, quirk bit LBM_QUIRK_IBM_2X2). The band is the union of those stencils.
// The tables built here drive the single-workgroup IBM pass:
//   - band_node[b]: global node index of band entry b (sorted by node index);
//   - mk_ptr / mk_band / mk_w: CSR marker -> (band entry, phi(dx) phi(dy)) in the reference's
//     loop order (x offset outer, y inner), out-of-domain nodes skipped;
//   - bn_ptr / bn_mk / bn_w: CSR band entry -> (marker, phi phi ds) in marker order, the
//     spreading gather whose sum order equals the reference's scatter order;
//   - box_map: band entry or -1 per cell of the band's bounding box, whose x range is widened
//     to whole 4-node segments (so one-step launches can split at it);
//   - mkf_band / mkf_w and bnf_mk / bnf_w: the same two lists at a fixed width, for a pass that
//     keeps them in registers: kIbmSlots slots per marker and `kspread` (8 or 16) per band entry,
//     stored slot-major ([slot][marker], [slot][band entry]); each list's entries in the CSR
//     order, then padding slots (index 0, weight 0) that add nothing to the sums. kspread is 0
//     when a band entry is touched by more than 16 markers (then only the CSR lists exist).
// Every index is checked before the function returns success; the device code forms no
// index that does not come from these tables.
#pragma once

#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

namespace lbm {

// delta4 (IBMUtils.cuh:23-33), FP32 as in the reference.
inline float delta4f(float r) {
    float rabs = std::fabs(r);
    if (rabs < 1.0f) return 0.125f * (3.0f - 2.0f * rabs + std::sqrt(1.0f + 4.0f * rabs - 4.0f * r * r));
    if (rabs < 2.0f) return 0.125f * (5.0f - 2.0f * rabs - std::sqrt(-7.0f + 12.0f * rabs - 4.0f * r * r));
    return 0.0f;
}

constexpr int kIbmSlots = 16;  // stencil slots per marker in the fixed-width lists (4x4)

struct IbmTables {
    std::vector<int> band_node, mk_ptr, mk_band, bn_ptr, bn_mk, box_map;
    std::vector<float> mk_w, bn_w;
    int x0 = 0, x1 = 0, y0 = 0, y1 = 0;  // bounding box of the band, exclusive ends
    int kspread = 0;                      // spread slots per band entry in the fixed-width lists
    std::vector<int> mkf_band, bnf_mk;
    std::vector<float> mkf_w, bnf_w;
};

// Builds the tables for n markers at xy (AoS [n][2]) with arc lengths ds (or null: 1) on an
// nx x ny grid. Returns null on success, else the reason the markers are rejected.
inline const char* ibm_build_tables(int nx, int ny, bool two_by_two, bool no_ds, const float* xy, const float* ds,
                                    int n, IbmTables& T) {
    T = IbmTables();
    if (nx < 1 || ny < 1 || n < 1 || !xy) return "IBM tables need a grid and at least one marker";
    for (int l = 0; l < n; l++) {
        if (!std::isfinite(xy[2 * l]) || !std::isfinite(xy[2 * l + 1])) return "marker coordinates must be finite";
        if (ds && !(ds[l] > 0.0f)) return "marker arc lengths must be positive";
    }
    const int W = two_by_two ? 2 : 4, off = two_by_two ? 0 : -1;
    struct Ent {
        int node;
        float k;
    };
    std::vector<std::vector<Ent>> st(static_cast<size_t>(n));
    std::map<int, int> band;
    for (int l = 0; l < n; l++) {
        const float px = xy[2 * l], py = xy[2 * l + 1];
        // a marker more than a stencil width outside the grid touches no node (and its floor
        // need not fit an int)
        if (px < -4.0f || py < -4.0f || px > float(nx) + 4.0f || py > float(ny) + 4.0f) continue;
        const int gx = int(std::floor(px)), gy = int(std::floor(py));
        for (int i = 0; i < W; i++)
            for (int j = 0; j < W; j++) {
                const int xx = gx + off + i, yy = gy + off + j;
                if (xx >= nx || xx < 0 || yy >= ny || yy < 0) continue;
                const float k = delta4f(px - float(xx)) * delta4f(py - float(yy));
                const int node = yy * nx + xx;
                st[size_t(l)].push_back({node, k});
                band.emplace(node, 0);
            }
    }
    if (band.empty()) return "no marker stencil touches the grid";
    int xmin = nx, xmax = -1, ymin = ny, ymax = -1;
    for (auto& kv : band) {
        kv.second = int(T.band_node.size());
        T.band_node.push_back(kv.first);
        const int y = kv.first / nx, x = kv.first % nx;
        xmin = std::min(xmin, x);
        xmax = std::max(xmax, x);
        ymin = std::min(ymin, y);
        ymax = std::max(ymax, y);
    }
    const int nb = int(T.band_node.size());
    T.mk_ptr.assign(size_t(n) + 1, 0);
    std::vector<std::vector<std::pair<int, float>>> per_node(static_cast<size_t>(nb));
    for (int l = 0; l < n; l++) {
        const float d = (ds && !no_ds) ? ds[l] : 1.0f;
        for (const Ent& e : st[size_t(l)]) {
            const int b = band.at(e.node);
            T.mk_band.push_back(b);
            T.mk_w.push_back(e.k);
            per_node[size_t(b)].push_back({l, e.k * d});
        }
        T.mk_ptr[size_t(l) + 1] = int(T.mk_band.size());
    }
    T.bn_ptr.assign(size_t(nb) + 1, 0);
    for (int b = 0; b < nb; b++) {
        for (const auto& pr : per_node[size_t(b)]) {
            T.bn_mk.push_back(pr.first);
            T.bn_w.push_back(pr.second);
        }
        T.bn_ptr[size_t(b) + 1] = int(T.bn_mk.size());
    }
    // fixed-width copies of both lists
    int longest = 0;
    for (int b = 0; b < nb; b++) longest = std::max(longest, T.bn_ptr[size_t(b) + 1] - T.bn_ptr[size_t(b)]);
    T.kspread = longest <= 8 ? 8 : (longest <= 16 ? 16 : 0);
    T.mkf_band.assign(size_t(kIbmSlots) * size_t(n), 0);
    T.mkf_w.assign(size_t(kIbmSlots) * size_t(n), 0.0f);
    for (int l = 0; l < n; l++) {
        if (T.mk_ptr[size_t(l) + 1] - T.mk_ptr[size_t(l)] > kIbmSlots) return "IBM stencil longer than its slots";
        for (int e = T.mk_ptr[size_t(l)]; e < T.mk_ptr[size_t(l) + 1]; e++) {
            const size_t slot = size_t(e - T.mk_ptr[size_t(l)]) * size_t(n) + size_t(l);
            T.mkf_band[slot] = T.mk_band[size_t(e)];
            T.mkf_w[slot] = T.mk_w[size_t(e)];
        }
    }
    if (T.kspread) {
        T.bnf_mk.assign(size_t(T.kspread) * size_t(nb), 0);
        T.bnf_w.assign(size_t(T.kspread) * size_t(nb), 0.0f);
        for (int b = 0; b < nb; b++)
            for (int e = T.bn_ptr[size_t(b)]; e < T.bn_ptr[size_t(b) + 1]; e++) {
                const size_t slot = size_t(e - T.bn_ptr[size_t(b)]) * size_t(nb) + size_t(b);
                T.bnf_mk[slot] = T.bn_mk[size_t(e)];
                T.bnf_w[slot] = T.bn_w[size_t(e)];
            }
    }
    xmin = xmin / 4 * 4;
    xmax = std::min(nx, (xmax + 1 + 3) / 4 * 4) - 1;
    const int bw = xmax - xmin + 1, bh = ymax - ymin + 1;
    T.box_map.assign(size_t(bw) * size_t(bh), -1);
    for (int b = 0; b < nb; b++) {
        const int y = T.band_node[size_t(b)] / nx, x = T.band_node[size_t(b)] % nx;
        T.box_map[size_t(y - ymin) * size_t(bw) + size_t(x - xmin)] = b;
    }
    T.x0 = xmin;
    T.x1 = xmax + 1;
    T.y0 = ymin;
    T.y1 = ymax + 1;
    // Every index the kernel forms comes from these tables: check them all.
    const long long N = (long long)nx * ny;
    for (int v : T.band_node)
        if (v < 0 || v >= N) return "IBM band node out of range";
    for (int v : T.mk_band)
        if (v < 0 || v >= nb) return "IBM stencil entry out of range";
    for (int v : T.bn_mk)
        if (v < 0 || v >= n) return "IBM spread entry out of range";
    if (T.mk_ptr.back() != int(T.mk_band.size()) || T.bn_ptr.back() != int(T.bn_mk.size()) ||
        T.mk_band.size() != T.mk_w.size() || T.bn_mk.size() != T.bn_w.size() || T.mk_band.size() != T.bn_mk.size())
        return "IBM table sizes inconsistent";
    for (int v : T.box_map)
        if (v < -1 || v >= nb) return "IBM box map out of range";
    for (int v : T.mkf_band)
        if (v < 0 || v >= nb) return "IBM fixed-width stencil entry out of range";
    for (int v : T.bnf_mk)
        if (v < 0 || v >= n) return "IBM fixed-width spread entry out of range";
    if (T.mkf_band.size() != size_t(kIbmSlots) * size_t(n) || T.mkf_w.size() != T.mkf_band.size() ||
        T.bnf_mk.size() != size_t(T.kspread) * size_t(nb) || T.bnf_w.size() != T.bnf_mk.size())
        return "IBM fixed-width table sizes inconsistent";
    if (T.x0 < 0 || T.y0 < 0 || T.x1 > nx || T.y1 > ny || T.x0 >= T.x1 || T.y0 >= T.y1)
        return "IBM box outside the grid";
    return nullptr;
}

}  // namespace lbm

namespace lbm {

// The 3-D tables (IBMManager<3>; IBM_impl.cu:53-118, 156-196): the 4x4x4 stencil of each
// marker at floor(p) - 1 .. floor(p) + 2 on every axis, in the reference's loop order (x outer,
// then y, then z), nodes outside the grid skipped; weights phi(dx) phi(dy) phi(dz); the spread
// weights carry the marker's area ds unless no_ds. Band nodes sorted by node index
// ((z NY + y) NX + x).
struct IbmTables3 {
    std::vector<int> band_node, mk_ptr, mk_band, bn_ptr, bn_mk;
    std::vector<float> mk_w, bn_w;
};

inline const char* ibm_build_tables3(int nx, int ny, int nz, bool no_ds, const float* xyz, const float* ds, int n,
                                     IbmTables3& T) {
    T = IbmTables3();
    if (nx < 1 || ny < 1 || nz < 1 || n < 1 || !xyz) return "IBM tables need a grid and at least one marker";
    for (int l = 0; l < n; l++) {
        for (int d = 0; d < 3; d++)
            if (!std::isfinite(xyz[3 * l + d])) return "marker coordinates must be finite";
        if (ds && !(ds[l] > 0.0f)) return "marker areas must be positive";
    }
    struct Ent {
        long long node;
        float k;
    };
    std::vector<std::vector<Ent>> st(static_cast<size_t>(n));
    std::map<long long, int> band;
    const long long N = (long long)nx * ny * nz;
    for (int l = 0; l < n; l++) {
        const float px = xyz[3 * l], py = xyz[3 * l + 1], pz = xyz[3 * l + 2];
        if (px < -4.0f || py < -4.0f || pz < -4.0f || px > float(nx) + 4.0f || py > float(ny) + 4.0f ||
            pz > float(nz) + 4.0f)
            continue;  // touches no node
        const int gx = int(std::floor(px)), gy = int(std::floor(py)), gz = int(std::floor(pz));
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++)
                for (int k = 0; k < 4; k++) {
                    const int x = gx + i - 1, y = gy + j - 1, z = gz + k - 1;
                    if (x >= nx || x < 0 || y >= ny || y < 0 || z >= nz || z < 0) continue;
                    const float w = delta4f(px - float(x)) * delta4f(py - float(y)) * delta4f(pz - float(z));
                    const long long node = ((long long)z * ny + y) * nx + x;
                    st[size_t(l)].push_back({node, w});
                    band.emplace(node, 0);
                }
    }
    if (band.empty()) return "no marker stencil touches the grid";
    for (auto& kv : band) {
        kv.second = int(T.band_node.size());
        T.band_node.push_back(int(kv.first));
    }
    const int nb = int(T.band_node.size());
    T.mk_ptr.assign(size_t(n) + 1, 0);
    std::vector<std::vector<std::pair<int, float>>> per_node(static_cast<size_t>(nb));
    for (int l = 0; l < n; l++) {
        const float d = (ds && !no_ds) ? ds[l] : 1.0f;
        for (const Ent& e : st[size_t(l)]) {
            const int b = band.at(e.node);
            T.mk_band.push_back(b);
            T.mk_w.push_back(e.k);
            per_node[size_t(b)].push_back({l, e.k * d});
        }
        T.mk_ptr[size_t(l) + 1] = int(T.mk_band.size());
    }
    T.bn_ptr.assign(size_t(nb) + 1, 0);
    for (int b = 0; b < nb; b++) {
        for (const auto& pr : per_node[size_t(b)]) {
            T.bn_mk.push_back(pr.first);
            T.bn_w.push_back(pr.second);
        }
        T.bn_ptr[size_t(b) + 1] = int(T.bn_mk.size());
    }
    for (int v : T.band_node)
        if (v < 0 || v >= N) return "IBM band node out of range";
    for (int v : T.mk_band)
        if (v < 0 || v >= nb) return "IBM stencil entry out of range";
    for (int v : T.bn_mk)
        if (v < 0 || v >= n) return "IBM spread entry out of range";
    if (T.mk_ptr.back() != int(T.mk_band.size()) || T.bn_ptr.back() != int(T.bn_mk.size()) ||
        T.mk_band.size() != T.bn_mk.size())
        return "IBM table sizes inconsistent";
    return nullptr;
}

}  // namespace lbm
