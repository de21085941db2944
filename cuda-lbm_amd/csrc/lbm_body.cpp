// lbm_body.cpp — 3-D immersed-body ingestion of include/lbm_body.h (host code, no HIP calls).
//
// The reference prepares a 3-D body on the host once per run: OBJ mesh -> affine placement in
// lattice units -> weighted sample elimination of the vertices -> block-Morton order of the
// markers. This file does the same in plain C++ with three differences of method:
//   - neighbour queries use a uniform grid of cells no smaller than the search diameter instead
//     of a k-d tree (the reference's nanoflann); the neighbour set of a query is the same (every
//     point with squared distance below the radius, in the same float arithmetic) and is visited
//     in ascending distance, ties by index (nanoflann leaves ties in tree order);
//   - errors return codes instead of exiting the process;
//   - nothing is written besides what the caller asks for.
// Floating-point steps that decide which points survive or where they go (the affine maps,
// distances, weights, volume, radius, Morton quantisation) follow the reference's operation
// order, so the same input gives the same markers. Compiled with FP contraction off.
#include "../../include/lbm_body.h"

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <queue>
#include <sstream>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct V3 {
    float x, y, z;
};

int fail_global(int code, const char* fmt, ...) {
    char b[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(b, sizeof b, fmt, ap);
    va_end(ap);
    g_err = b;
    return code;
}

// ---- 4x4 affine maps (math::Mat4, src/IBM/math/mat4.hpp) ------------------------------------
struct Affine {
    float m[4][4];
};
Affine identity() {
    Affine a{};
    for (int i = 0; i < 4; i++) a.m[i][i] = 1.0f;
    return a;
}
Affine translation(float tx, float ty, float tz) {
    Affine a = identity();
    a.m[0][3] = tx;
    a.m[1][3] = ty;
    a.m[2][3] = tz;
    return a;
}
Affine scaling(float sx, float sy, float sz) {
    Affine a = identity();
    a.m[0][0] = sx;
    a.m[1][1] = sy;
    a.m[2][2] = sz;
    return a;
}
// rotation about x (axis 0), y (1) or z (2), right-handed
Affine rotation(int axis, float angle) {
    Affine a = identity();
    const float c = cosf(angle), s = sinf(angle);
    const int i = axis == 0 ? 1 : 0, j = axis == 2 ? 1 : 2;  // the rotated pair
    a.m[i][i] = c;
    a.m[j][j] = c;
    if (axis == 1) {
        a.m[i][j] = s;
        a.m[j][i] = -s;
    } else {
        a.m[i][j] = -s;
        a.m[j][i] = s;
    }
    return a;
}
// row-by-column product, each entry summed from 0 over k = 0..3
Affine operator*(const Affine& a, const Affine& b) {
    Affine r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float acc = 0.0f;
            for (int k = 0; k < 4; k++) acc += a.m[i][k] * b.m[k][j];
            r.m[i][j] = acc;
        }
    return r;
}
// homogeneous point (w = 1), divided by w when w is neither 1 nor 0 (within 1e-6)
V3 apply(const Affine& a, const V3& p) {
    const float w1 = 1.0f;
    float r[4];
    for (int i = 0; i < 4; i++) r[i] = a.m[i][0] * p.x + a.m[i][1] * p.y + a.m[i][2] * p.z + a.m[i][3] * w1;
    if (std::fabs(r[3] - 1.0f) > 1e-6f && std::fabs(r[3]) > 1e-6f) return {r[0] / r[3], r[1] / r[3], r[2] / r[3]};
    return {r[0], r[1], r[2]};
}

struct Box {
    V3 lo{FLT_MAX, FLT_MAX, FLT_MAX}, hi{-FLT_MAX, -FLT_MAX, -FLT_MAX};
    bool valid = false;
    void add(const V3& p) {
        lo = {std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
        hi = {std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
        valid = true;
    }
    V3 centre() const { return {(lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f}; }
    V3 size() const { return {hi.x - lo.x, hi.y - lo.y, hi.z - lo.z}; }
};
Box box_of(const std::vector<V3>& v) {
    Box b;
    for (const V3& p : v) b.add(p);
    return b;
}

bool finite3(const V3& p) { return std::isfinite(p.x) && std::isfinite(p.y) && std::isfinite(p.z); }

// ---- neighbour queries ------------------------------------------------------------------
// Points binned in cubic cells of side >= the search diameter: the points closer than the
// diameter to a query lie in its cell or the 26 around it.
class CellGrid {
public:
    CellGrid(const std::vector<V3>& p, float diameter) : pts_(p) {
        if (!(diameter > 0.0f) || p.empty()) return;
        const Box b = box_of(p);
        lo_ = b.lo;
        const V3 s = b.size();
        const float ext = std::max({s.x, s.y, s.z});
        cell_ = std::max(diameter * (1.0f + 1e-5f), ext / float(1 << 20));
        for (int a = 0; a < 3; a++) dim_[a] = 1;
        for (const V3& q : p) {
            const auto c = coords(q);
            for (int a = 0; a < 3; a++) dim_[a] = std::max(dim_[a], c[a] + 1);
        }
        order_.resize(p.size());
        keys_.resize(p.size());
        for (size_t i = 0; i < p.size(); i++) order_[i] = int(i);
        std::vector<int64_t> k(p.size());
        for (size_t i = 0; i < p.size(); i++) k[i] = key(coords(p[i]));
        std::stable_sort(order_.begin(), order_.end(), [&](int a, int b2) { return k[size_t(a)] < k[size_t(b2)]; });
        for (size_t i = 0; i < p.size(); i++) keys_[i] = k[size_t(order_[i])];
        active_ = true;
    }
    // Points j with squared distance (query - p_j summed x, y, z in float) below r2, ascending
    // by that distance, ties by index.
    void query(const V3& q, float r2, std::vector<std::pair<float, int>>& out) const {
        out.clear();
        if (!active_ || !(r2 > 0.0f)) return;
        const auto c = coords(q);
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    const std::array<int64_t, 3> n{c[0] + dx, c[1] + dy, c[2] + dz};
                    bool in = true;
                    for (int a = 0; a < 3; a++) in = in && n[a] >= 0 && n[a] < dim_[a];
                    if (!in) continue;
                    const int64_t k = key(n);
                    auto range = std::equal_range(keys_.begin(), keys_.end(), k);
                    for (auto it = range.first; it != range.second; ++it) {
                        const int j = order_[size_t(it - keys_.begin())];
                        const V3& p = pts_[size_t(j)];
                        float d2 = 0.0f;
                        const float ex = q.x - p.x, ey = q.y - p.y, ez = q.z - p.z;
                        d2 += ex * ex;
                        d2 += ey * ey;
                        d2 += ez * ez;
                        if (d2 < r2) out.emplace_back(d2, j);
                    }
                }
        std::sort(out.begin(), out.end());
    }

private:
    std::array<int64_t, 3> coords(const V3& q) const {
        const float v[3] = {q.x - lo_.x, q.y - lo_.y, q.z - lo_.z};
        std::array<int64_t, 3> c{};
        for (int a = 0; a < 3; a++) c[a] = std::max<int64_t>(0, int64_t(std::floor(v[a] / cell_)));
        return c;
    }
    int64_t key(const std::array<int64_t, 3>& c) const { return (c[2] * dim_[1] + c[1]) * dim_[0] + c[0]; }

    const std::vector<V3>& pts_;
    V3 lo_{0, 0, 0};
    float cell_ = 1.0f;
    int64_t dim_[3] = {1, 1, 1};
    std::vector<int> order_;
    std::vector<int64_t> keys_;
    bool active_ = false;
};

// ---- weighted sample elimination (sampler::Sampler, sampler.hpp:20-164) -----------------------
// Weight of sample j as seen from i: (1 - d / 2r)^8 in FP64 rounded to float, 0 from d = 2r on.
float elim_weight(const V3& from, const V3& to, float r) {
    const float dx = to.x - from.x, dy = to.y - from.y, dz = to.z - from.z;
    const float d = sqrtf(dx * dx + dy * dy + dz * dz);
    if (d >= 2.0f * r) return 0.0f;
    return float(std::pow(double(1.0f - d / (2.0f * r)), 8.0));
}

struct HeapEntry {
    float w;
    int idx;
    bool operator<(const HeapEntry& o) const { return w < o.w; }
};

// The survivors (indices ascending) after removing points until `target` remain.
int eliminate(const std::vector<V3>& p, int target, float r, std::vector<int>& keep, std::string& err) {
    const int n = int(p.size());
    keep.clear();
    if (target >= n) {
        for (int i = 0; i < n; i++) keep.push_back(i);
        return LBM_OK;
    }
    const CellGrid grid(p, 2.0f * r);
    std::vector<float> w(size_t(n), 0.0f);
    std::vector<char> gone(size_t(n), 0);
    std::vector<std::pair<float, int>> nb;
    const float r2_weights = 4.0f * r * r;
    for (int i = 0; i < n; i++) {
        float acc = 0.0f;
        grid.query(p[size_t(i)], r2_weights, nb);
        for (const auto& e : nb) acc += e.second == i ? 0.0f : elim_weight(p[size_t(i)], p[size_t(e.second)], r);
        w[size_t(i)] = acc;
    }
    std::priority_queue<HeapEntry> heap;
    for (int i = 0; i < n; i++) heap.push({w[size_t(i)], i});
    const float r2_elim = (2.0f * r) * (2.0f * r);
    int removed = 0;
    while (n - removed > target) {
        if (heap.empty()) {
            err = "sample elimination ran out of candidates";
            return LBM_E_INVALID;
        }
        const HeapEntry top = heap.top();
        heap.pop();
        const int t = top.idx;
        if (gone[size_t(t)]) continue;
        if (std::fabs(w[size_t(t)] - top.w) > 1e-6f) continue;  // a stale entry
        gone[size_t(t)] = 1;
        removed++;
        grid.query(p[size_t(t)], r2_elim, nb);
        for (const auto& e : nb) {
            const int j = e.second;
            if (gone[size_t(j)] || j == t) continue;
            w[size_t(j)] -= elim_weight(p[size_t(j)], p[size_t(t)], r);
            heap.push({w[size_t(j)], j});
        }
    }
    for (int i = 0; i < n; i++)
        if (!gone[size_t(i)]) keep.push_back(i);
    return LBM_OK;
}

float mesh_volume(const std::vector<V3>& v, const std::vector<std::array<int, 3>>& f) {
    if (v.empty() || f.empty()) return 0.0f;
    double acc = 0.0;
    const long long nv = (long long)v.size();
    for (const auto& t : f) {
        if (t[0] < 0 || t[0] >= nv || t[1] < 0 || t[1] >= nv || t[2] < 0 || t[2] >= nv) continue;
        const V3 &a = v[size_t(t[0])], &b = v[size_t(t[1])], &c = v[size_t(t[2])];
        // six times the signed volume of the tetrahedron (origin, a, b, c): a . (b x c)
        const double v321 = double(c.x) * b.y * a.z, v231 = double(b.x) * c.y * a.z;
        const double v312 = double(c.x) * a.y * b.z, v132 = double(a.x) * c.y * b.z;
        const double v213 = double(b.x) * a.y * c.z, v123 = double(a.x) * b.y * c.z;
        acc += (-v321 + v231 + v312 - v132 - v213 + v123);
    }
    return float(std::abs(acc / 6.0));
}

float r_max_3d(float volume, int n) {
    if (volume <= 1e-6f || n <= 0) return 0.0f;
    return std::cbrt(volume / (4.0f * std::sqrt(2.0f) * float(n)));
}

// ---- block Morton order (block_morton_sort<dim>, IBM_sorting.cuh:41-159) -----------------------
uint32_t interleave(const uint32_t* c, int dim, int bits) {
    uint32_t m = 0;
    for (int i = 0; i < bits; i++)
        for (int a = 0; a < dim; a++) m |= ((c[a] >> i) & 1u) << (dim * i + a);
    return m;
}

int morton_order(float* xyz, int n, int dim, int block, int bits) {
    if (n <= 0) return LBM_OK;
    Box b;
    for (int i = 0; i < n; i++) b.add({xyz[dim * i], xyz[dim * i + 1], dim == 3 ? xyz[dim * i + 2] : 0.0f});
    const float bs = float(block);
    const float lo[3] = {b.lo.x, b.lo.y, b.lo.z}, hi[3] = {b.hi.x, b.hi.y, b.hi.z};
    float origin[3] = {0.0f, 0.0f, 0.0f};
    int nb[3] = {1, 1, 1};
    const float eps = bs * 1e-6f;
    for (int a = 0; a < dim; a++) {
        origin[a] = floorf(lo[a] / bs) * bs;
        nb[a] = std::max(1, int(ceilf((hi[a] - origin[a] + eps) / bs)));
    }
    const uint32_t qmax = (1u << bits) - 1u;
    struct Key {
        int block;
        uint32_t code;
        int src;
    };
    std::vector<Key> keys(static_cast<size_t>(n));
    for (int i = 0; i < n; i++) {
        int bi[3] = {0, 0, 0};
        uint32_t q[3] = {0, 0, 0};
        for (int a = 0; a < dim; a++) {
            const float p = xyz[dim * i + a];
            bi[a] = std::max(0, std::min(int(floorf((p - origin[a]) / bs)), nb[a] - 1));
            const float local = p - (origin[a] + float(bi[a]) * bs);
            q[a] = uint32_t(fmaxf(0.0f, fminf(local / bs, 1.0f)) * float(qmax));
        }
        keys[size_t(i)] = {(bi[2] * nb[1] + bi[1]) * nb[0] + bi[0], interleave(q, dim, bits), i};
    }
    std::sort(keys.begin(), keys.end(), [](const Key& a, const Key& c) {
        if (a.block != c.block) return a.block < c.block;
        return a.code < c.code;
    });
    std::vector<float> out(size_t(n) * size_t(dim));
    for (int i = 0; i < n; i++)
        for (int a = 0; a < dim; a++) out[size_t(dim * i + a)] = xyz[dim * keys[size_t(i)].src + a];
    std::copy(out.begin(), out.end(), xyz);
    return LBM_OK;
}

// ---- text files -------------------------------------------------------------------------------
std::string trim(const std::string& s) {
    const char* ws = " \t\n\r\f\v";
    const size_t a = s.find_first_not_of(ws);
    if (a == std::string::npos) return "";
    return s.substr(a, s.find_last_not_of(ws) - a + 1);
}
std::string lower(std::string s) {
    for (char& c : s) c = char(std::tolower(static_cast<unsigned char>(c)));
    return s;
}
std::string resolve(const std::string& path, const char* base_dir) {
    if (!base_dir || path.empty()) return path;
    const std::filesystem::path p(path);
    return p.is_absolute() ? path : (std::filesystem::path(base_dir) / p).string();
}

float deg2rad(float deg) {
    constexpr float kPi = 3.14159265358979323846f;
    return deg * (kPi / 180.0f);
}

}  // namespace

struct lbm_body {
    std::vector<V3> verts;
    std::vector<std::array<int, 3>> faces;
    V3 anchor{0, 0, 0};
    std::vector<V3> markers;
    std::string err;

    int fail(int code, const char* fmt, ...) {
        char b[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(b, sizeof b, fmt, ap);
        va_end(ap);
        err = b;
        g_err = b;
        return code;
    }
    void anchor_to_centroid() { anchor = box_of(verts).centre(); }
    void map_vertices(const Affine& a) {
        for (V3& v : verts) v = apply(a, v);
    }
    // translation to the anchor, `a`, translation back: (T(+anchor) * a) * T(-anchor)
    void about_anchor(const Affine& a) {
        map_vertices(translation(anchor.x, anchor.y, anchor.z) * a * translation(-anchor.x, -anchor.y, -anchor.z));
    }
    void translate(float tx, float ty, float tz) {
        const Affine t = translation(tx, ty, tz);
        map_vertices(t);
        anchor = apply(t, anchor);
    }
};

namespace {

int make_body(std::vector<V3> v, std::vector<std::array<int, 3>> f, lbm_body** out) {
    if (v.empty()) return fail_global(LBM_E_INVALID, "the mesh has no vertices");
    for (const V3& p : v)
        if (!finite3(p)) return fail_global(LBM_E_INVALID, "vertex coordinates must be finite");
    auto* b = new lbm_body();
    b->verts = std::move(v);
    b->faces = std::move(f);
    b->anchor_to_centroid();
    *out = b;
    return LBM_OK;
}

int load_obj_file(const std::string& path, lbm_body** out) {
    std::ifstream in(path);
    if (!in.is_open()) return fail_global(LBM_E_INVALID, "cannot open OBJ file \"%s\"", path.c_str());
    std::vector<V3> v;
    std::vector<std::array<int, 3>> f;
    std::string line;
    while (std::getline(in, line)) {
        std::stringstream ss(line);
        std::string kw;
        ss >> kw;
        if (kw == "v") {
            float x, y, z;
            if (ss >> x >> y >> z) v.push_back({x, y, z});
        } else if (kw == "f") {
            std::array<int, 3> t{};
            int count = 0;
            bool ok = true;
            for (int i = 0; i < 3; i++) {
                std::string tok;
                if (!(ss >> tok)) {
                    ok = false;
                    break;
                }
                count++;
                try {
                    t[size_t(i)] = std::stoi(tok.substr(0, tok.find('/'))) - 1;
                } catch (const std::exception&) {
                    ok = false;
                    break;
                }
            }
            if (ok && count == 3) f.push_back(t);
        }
    }
    if (v.empty()) return fail_global(LBM_E_INVALID, "no vertices in OBJ file \"%s\"", path.c_str());
    return make_body(std::move(v), std::move(f), out);
}

int parse_number(const std::string& s, float& out) {
    try {
        out = std::stof(s);
    } catch (const std::exception&) {
        return LBM_E_INVALID;
    }
    return LBM_OK;
}

void copy_str(char* dst, size_t cap, const std::string& s) {
    std::strncpy(dst, s.c_str(), cap - 1);
    dst[cap - 1] = '\0';
}

}  // namespace

extern "C" {

int lbm_body_load_obj(const char* path, lbm_body** out) {
    if (!path || !out) return fail_global(LBM_E_INVALID, "null argument");
    *out = nullptr;
    return load_obj_file(path, out);
}

int lbm_body_from_mesh(const float* xyz, int32_t nv, const int32_t* faces, int32_t nf, lbm_body** out) {
    if (!out || nv <= 0 || !xyz || nf < 0 || (nf > 0 && !faces)) return fail_global(LBM_E_INVALID, "invalid mesh arrays");
    *out = nullptr;
    std::vector<V3> v(static_cast<size_t>(nv));
    for (int i = 0; i < nv; i++) v[size_t(i)] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    std::vector<std::array<int, 3>> f(static_cast<size_t>(nf));
    for (int i = 0; i < nf; i++) f[size_t(i)] = {faces[3 * i], faces[3 * i + 1], faces[3 * i + 2]};
    return make_body(std::move(v), std::move(f), out);
}

int lbm_body_destroy(lbm_body* body) {
    delete body;
    return LBM_OK;
}

int lbm_body_info(const lbm_body* b, int32_t* nv, int32_t* nf, int32_t* nm) {
    if (!b) return fail_global(LBM_E_INVALID, "null body");
    if (nv) *nv = int32_t(b->verts.size());
    if (nf) *nf = int32_t(b->faces.size());
    if (nm) *nm = int32_t(b->markers.size());
    return LBM_OK;
}

int lbm_body_vertices(const lbm_body* b, float* xyz) {
    if (!b || !xyz) return fail_global(LBM_E_INVALID, "null argument");
    for (size_t i = 0; i < b->verts.size(); i++) {
        xyz[3 * i] = b->verts[i].x;
        xyz[3 * i + 1] = b->verts[i].y;
        xyz[3 * i + 2] = b->verts[i].z;
    }
    return LBM_OK;
}

int lbm_body_faces(const lbm_body* b, int32_t* faces) {
    if (!b || (!faces && !b->faces.empty())) return fail_global(LBM_E_INVALID, "null argument");
    for (size_t i = 0; i < b->faces.size(); i++)
        for (int k = 0; k < 3; k++) faces[3 * i + size_t(k)] = b->faces[i][size_t(k)];
    return LBM_OK;
}

int lbm_body_markers(const lbm_body* b, float* xyz) {
    if (!b || (!xyz && !b->markers.empty())) return fail_global(LBM_E_INVALID, "null argument");
    for (size_t i = 0; i < b->markers.size(); i++) {
        xyz[3 * i] = b->markers[i].x;
        xyz[3 * i + 1] = b->markers[i].y;
        xyz[3 * i + 2] = b->markers[i].z;
    }
    return LBM_OK;
}

int lbm_body_translate(lbm_body* b, float tx, float ty, float tz) {
    if (!b) return fail_global(LBM_E_INVALID, "null body");
    b->translate(tx, ty, tz);
    return LBM_OK;
}

int lbm_body_scale(lbm_body* b, float sx, float sy, float sz) {
    if (!b) return fail_global(LBM_E_INVALID, "null body");
    if (!std::isfinite(sx) || !std::isfinite(sy) || !std::isfinite(sz)) return b->fail(LBM_E_INVALID, "scale factors must be finite");
    b->about_anchor(scaling(sx, sy, sz));
    return LBM_OK;
}

int lbm_body_scale_to_overall_size(lbm_body* b, float size) {
    if (!b) return fail_global(LBM_E_INVALID, "null body");
    const V3 s = box_of(b->verts).size();
    const float m = std::max({std::abs(s.x), std::abs(s.y), std::abs(s.z)});
    if (!(m > 0.0f) || !std::isfinite(size)) return b->fail(LBM_E_INVALID, "a body of zero extent cannot be scaled to a size");
    const float f = size / m;
    return lbm_body_scale(b, f, f, f);
}

int lbm_body_rotate(lbm_body* b, int32_t axis, float angle) {
    if (!b) return fail_global(LBM_E_INVALID, "null body");
    if (axis < 0 || axis > 2 || !std::isfinite(angle)) return b->fail(LBM_E_INVALID, "axis must be 0, 1 or 2 and the angle finite");
    b->about_anchor(rotation(axis, angle));
    return LBM_OK;
}

int lbm_body_set_anchor(lbm_body* b, float x, float y, float z) {
    if (!b) return fail_global(LBM_E_INVALID, "null body");
    b->anchor = {x, y, z};
    return LBM_OK;
}

int lbm_body_set_anchor_to_centroid(lbm_body* b) {
    if (!b) return fail_global(LBM_E_INVALID, "null body");
    b->anchor_to_centroid();
    return LBM_OK;
}

int lbm_body_move_anchor_to_world(lbm_body* b, float x, float y, float z) {
    if (!b) return fail_global(LBM_E_INVALID, "null body");
    b->translate(x - b->anchor.x, y - b->anchor.y, z - b->anchor.z);
    return LBM_OK;
}

int lbm_body_move_centroid_to_world(lbm_body* b, float x, float y, float z) {
    if (!b) return fail_global(LBM_E_INVALID, "null body");
    const V3 c = box_of(b->verts).centre();
    b->translate(x - c.x, y - c.y, z - c.z);
    return LBM_OK;
}

int lbm_body_aabb(const lbm_body* b, float* mn, float* mx) {
    if (!b) return fail_global(LBM_E_INVALID, "null body");
    const Box x = box_of(b->verts);
    if (mn) mn[0] = x.lo.x, mn[1] = x.lo.y, mn[2] = x.lo.z;
    if (mx) mx[0] = x.hi.x, mx[1] = x.hi.y, mx[2] = x.hi.z;
    return LBM_OK;
}

int lbm_body_anchor(const lbm_body* b, float* a) {
    if (!b || !a) return fail_global(LBM_E_INVALID, "null argument");
    a[0] = b->anchor.x;
    a[1] = b->anchor.y;
    a[2] = b->anchor.z;
    return LBM_OK;
}

int lbm_body_fits_in_domain(const lbm_body* b, float nx, float ny, float nz, int32_t* fits) {
    if (!b || !fits) return fail_global(LBM_E_INVALID, "null argument");
    const Box x = box_of(b->verts);
    *fits = x.lo.x >= 0.0f && x.hi.x <= nx && x.lo.y >= 0.0f && x.hi.y <= ny && x.lo.z >= 0.0f && x.hi.z <= nz;
    return LBM_OK;
}

int lbm_body_save_obj(const lbm_body* b, const char* path) {
    if (!b || !path) return fail_global(LBM_E_INVALID, "null argument");
    std::ofstream o(path);
    if (!o.is_open()) return fail_global(LBM_E_INVALID, "cannot write \"%s\"", path);
    o << std::fixed << std::setprecision(6);
    o << "# immersed body mesh\n# Vertices: " << b->verts.size() << "\n";
    if (!b->faces.empty()) o << "# Faces: " << b->faces.size() << "\n";
    o << "\n";
    for (const V3& v : b->verts) o << "v " << v.x << " " << v.y << " " << v.z << "\n";
    o << "\n";
    for (const auto& f : b->faces) o << "f " << (f[0] + 1) << " " << (f[1] + 1) << " " << (f[2] + 1) << "\n";
    return o.good() ? LBM_OK : fail_global(LBM_E_INVALID, "write to \"%s\" failed", path);
}

int lbm_body_mesh_volume(const lbm_body* b, float* volume) {
    if (!b || !volume) return fail_global(LBM_E_INVALID, "null argument");
    *volume = mesh_volume(b->verts, b->faces);
    return LBM_OK;
}

float lbm_body_r_max(float volume, int32_t n) { return r_max_3d(volume, n); }

int lbm_body_sample(lbm_body* b, int32_t target) {
    if (!b) return fail_global(LBM_E_INVALID, "null body");
    if (target <= 0) return b->fail(LBM_E_INVALID, "target must be positive");
    const float r = r_max_3d(mesh_volume(b->verts, b->faces), target);
    std::vector<int> keep;
    std::string err;
    int rc = eliminate(b->verts, target, r, keep, err);
    if (rc) return b->fail(rc, "%s", err.c_str());
    b->markers.clear();
    for (int i : keep) b->markers.push_back(b->verts[size_t(i)]);
    return LBM_OK;
}

int lbm_body_markers_from_vertices(lbm_body* b) {
    if (!b) return fail_global(LBM_E_INVALID, "null body");
    b->markers = b->verts;
    return LBM_OK;
}

int lbm_points_morton_sort(float* xyz, int32_t n, int32_t dim, int32_t block, int32_t bits) {
    if (n < 0 || (n > 0 && !xyz) || (dim != 2 && dim != 3) || block < 1 || bits < 1 || dim * bits > 32)
        return fail_global(LBM_E_INVALID, "invalid Morton sort arguments (dim 2 or 3, block >= 1, 1 <= bits, dim * bits <= 32)");
    for (int i = 0; i < dim * n; i++)
        if (!std::isfinite(xyz[i])) return fail_global(LBM_E_INVALID, "marker coordinates must be finite");
    return morton_order(xyz, n, dim, block, bits);
}

int lbm_body_morton_sort(lbm_body* b, int32_t block, int32_t bits) {
    if (!b) return fail_global(LBM_E_INVALID, "null body");
    std::vector<float> x(3 * b->markers.size());
    lbm_body_markers(b, x.data());
    int rc = lbm_points_morton_sort(x.data(), int32_t(b->markers.size()), 3, block, bits);
    if (rc) return b->fail(rc, "%s", g_err.c_str());
    for (size_t i = 0; i < b->markers.size(); i++) b->markers[i] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    return LBM_OK;
}

int lbm_points_save_csv(const char* path, const float* xyz, int32_t n) {
    if (!path || n < 0 || (n > 0 && !xyz)) return fail_global(LBM_E_INVALID, "invalid arguments");
    std::ofstream o(path);
    if (!o.is_open()) return fail_global(LBM_E_INVALID, "cannot write \"%s\"", path);
    o << "x,y,z\n";
    for (int i = 0; i < n; i++) o << xyz[3 * i] << "," << xyz[3 * i + 1] << "," << xyz[3 * i + 2] << "\n";
    return o.good() ? LBM_OK : fail_global(LBM_E_INVALID, "write to \"%s\" failed", path);
}

int lbm_points_load_csv(const char* path, float* xyz, int32_t cap, int32_t* n) {
    if (!path || !n || cap < 0) return fail_global(LBM_E_INVALID, "invalid arguments");
    std::ifstream in(path);
    if (!in.is_open()) return fail_global(LBM_E_INVALID, "cannot open CSV file \"%s\"", path);
    std::string line;
    int count = 0, line_no = 0;
    while (std::getline(in, line)) {
        line_no++;
        if (line_no == 1 && (line.find('x') != std::string::npos || line.find('X') != std::string::npos)) continue;
        if (line.empty() || line[0] == '#') continue;
        std::stringstream ss(line);
        std::string seg;
        float v[3];
        bool ok = true;
        for (int k = 0; k < 3 && ok; k++) {
            if (!std::getline(ss, seg, ',')) ok = false;
            else if (parse_number(seg, v[k]) != LBM_OK) ok = false;
        }
        if (!ok) continue;
        if (xyz && count < cap)
            for (int k = 0; k < 3; k++) xyz[3 * count + k] = v[k];
        count++;
    }
    *n = count;
    return LBM_OK;
}

int lbm_body_parse_config(const char* ini_path, lbm_body_config* out) {
    if (!ini_path || !out) return fail_global(LBM_E_INVALID, "null argument");
    std::ifstream in(ini_path);
    if (!in.is_open()) return fail_global(LBM_E_INVALID, "cannot open body configuration \"%s\"", ini_path);
    lbm_body_config c;
    std::memset(&c, 0, sizeof c);
    c.sampling_type = LBM_BODY_SAMPLING_TARGET_POINTS;
    struct Field {
        const char* key;
        float* dst;
        uint32_t bit;
    };
    const Field fields[] = {
        {"scale_overall_size", &c.scale_overall_size, LBM_BODY_HAS_SCALE_OVERALL},
        {"scale_x", &c.scale_x, LBM_BODY_HAS_SCALE_X},
        {"scale_y", &c.scale_y, LBM_BODY_HAS_SCALE_Y},
        {"scale_z", &c.scale_z, LBM_BODY_HAS_SCALE_Z},
        {"rotate_x_deg", &c.rotate_x_deg, LBM_BODY_HAS_ROTATE_X},
        {"rotate_y_deg", &c.rotate_y_deg, LBM_BODY_HAS_ROTATE_Y},
        {"rotate_z_deg", &c.rotate_z_deg, LBM_BODY_HAS_ROTATE_Z},
        {"set_anchor_local_x", &c.set_anchor_local_x, LBM_BODY_HAS_ANCHOR_LOCAL_X},
        {"set_anchor_local_y", &c.set_anchor_local_y, LBM_BODY_HAS_ANCHOR_LOCAL_Y},
        {"set_anchor_local_z", &c.set_anchor_local_z, LBM_BODY_HAS_ANCHOR_LOCAL_Z},
        {"position_anchor_x", &c.position_anchor_x, LBM_BODY_HAS_POSITION_X},
        {"position_anchor_y", &c.position_anchor_y, LBM_BODY_HAS_POSITION_Y},
        {"position_anchor_z", &c.position_anchor_z, LBM_BODY_HAS_POSITION_Z},
    };
    std::string line, section;
    int line_no = 0;
    while (std::getline(in, line)) {
        line_no++;
        line = trim(line);
        if (line.empty() || line[0] == '#') continue;
        if (line[0] == '[' && line.back() == ']') {
            section = lower(trim(line.substr(1, line.size() - 2)));
            continue;
        }
        const size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        const std::string key = lower(trim(line.substr(0, eq)));
        const std::string val = trim(line.substr(eq + 1));
        if (val.empty()) continue;
        if (section == "general") {
            if (key == "obj_filepath") copy_str(c.obj_filepath, sizeof c.obj_filepath, val);
            else if (key == "save_transformed_obj_to") copy_str(c.save_transformed_obj_to, sizeof c.save_transformed_obj_to, val);
        } else if (section == "transformations") {
            for (const Field& f : fields)
                if (key == f.key) {
                    if (parse_number(val, *f.dst) != LBM_OK)
                        return fail_global(LBM_E_INVALID, "%s:%d: '%s' is not a number", ini_path, line_no, val.c_str());
                    c.has |= f.bit;
                }
        } else if (section == "sampling") {
            if (key == "type") {
                const std::string t = lower(val);
                if (t == "direct") c.sampling_type = LBM_BODY_SAMPLING_DIRECT;
                else if (t == "target_points") c.sampling_type = LBM_BODY_SAMPLING_TARGET_POINTS;
            } else if (key == "target_points") {
                try {
                    c.target_points = std::stoi(val);
                } catch (const std::exception&) {
                    return fail_global(LBM_E_INVALID, "%s:%d: '%s' is not an integer", ini_path, line_no, val.c_str());
                }
                c.has |= LBM_BODY_HAS_TARGET_POINTS;
            }
        }
    }
    if (!c.obj_filepath[0]) return fail_global(LBM_E_INVALID, "%s: obj_filepath not given in [General]", ini_path);
    if (c.sampling_type == LBM_BODY_SAMPLING_TARGET_POINTS &&
        (!(c.has & LBM_BODY_HAS_TARGET_POINTS) || c.target_points <= 0))
        return fail_global(LBM_E_INVALID, "%s: sampling type target_points needs a positive target_points", ini_path);
    *out = c;
    return LBM_OK;
}

int lbm_body_from_config(const char* ini_path, const char* base_dir, int32_t nx, int32_t ny, int32_t nz, lbm_body** out) {
    if (!ini_path || !out || nx < 1 || ny < 1 || nz < 1) return fail_global(LBM_E_INVALID, "invalid arguments");
    *out = nullptr;
    lbm_body_config c;
    int rc = lbm_body_parse_config(ini_path, &c);
    if (rc) return rc;
    lbm_body* b = nullptr;
    rc = load_obj_file(resolve(c.obj_filepath, base_dir), &b);
    if (rc) return rc;
    auto done = [&](int code) {
        if (code != LBM_OK) {
            g_err = b->err.empty() ? g_err : b->err;
            delete b;
            return code;
        }
        *out = b;
        return LBM_OK;
    };
    const uint32_t h = c.has;
    if (h & LBM_BODY_HAS_SCALE_OVERALL) {
        rc = lbm_body_scale_to_overall_size(b, c.scale_overall_size);
    } else if (h & (LBM_BODY_HAS_SCALE_X | LBM_BODY_HAS_SCALE_Y | LBM_BODY_HAS_SCALE_Z)) {
        rc = lbm_body_scale(b, (h & LBM_BODY_HAS_SCALE_X) ? c.scale_x : 1.0f, (h & LBM_BODY_HAS_SCALE_Y) ? c.scale_y : 1.0f,
                            (h & LBM_BODY_HAS_SCALE_Z) ? c.scale_z : 1.0f);
    }
    if (!rc && (h & LBM_BODY_HAS_ROTATE_Y)) rc = lbm_body_rotate(b, 1, deg2rad(c.rotate_y_deg));
    if (!rc && (h & LBM_BODY_HAS_ROTATE_X)) rc = lbm_body_rotate(b, 0, deg2rad(c.rotate_x_deg));
    if (!rc && (h & LBM_BODY_HAS_ROTATE_Z)) rc = lbm_body_rotate(b, 2, deg2rad(c.rotate_z_deg));
    const uint32_t anchor_all = LBM_BODY_HAS_ANCHOR_LOCAL_X | LBM_BODY_HAS_ANCHOR_LOCAL_Y | LBM_BODY_HAS_ANCHOR_LOCAL_Z;
    const uint32_t pos_all = LBM_BODY_HAS_POSITION_X | LBM_BODY_HAS_POSITION_Y | LBM_BODY_HAS_POSITION_Z;
    if (!rc && (h & anchor_all) == anchor_all) b->anchor = {c.set_anchor_local_x, c.set_anchor_local_y, c.set_anchor_local_z};
    if (!rc && (h & pos_all) == pos_all) rc = lbm_body_move_anchor_to_world(b, c.position_anchor_x, c.position_anchor_y, c.position_anchor_z);
    if (!rc && c.save_transformed_obj_to[0]) rc = lbm_body_save_obj(b, resolve(c.save_transformed_obj_to, base_dir).c_str());
    if (rc) return done(rc);
    int32_t fits = 0;
    lbm_body_fits_in_domain(b, float(nx), float(ny), float(nz), &fits);
    if (!fits) {
        const Box x = box_of(b->verts);
        return done(b->fail(LBM_E_INVALID, "the transformed body (%g..%g, %g..%g, %g..%g) does not fit the %dx%dx%d domain", x.lo.x,
                            x.hi.x, x.lo.y, x.hi.y, x.lo.z, x.hi.z, nx, ny, nz));
    }
    rc = c.sampling_type == LBM_BODY_SAMPLING_DIRECT ? lbm_body_markers_from_vertices(b) : lbm_body_sample(b, c.target_points);
    if (!rc && b->markers.empty()) rc = b->fail(LBM_E_INVALID, "the body has no markers");
    if (!rc) rc = lbm_body_morton_sort(b, 4, 10);
    return done(rc);
}

const char* lbm_body_last_error(const lbm_body* b) { return b && !b->err.empty() ? b->err.c_str() : g_err.c_str(); }

}  // extern "C"
