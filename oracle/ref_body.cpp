// ref_body.cpp — C entry points over the reference's own OBJ loader, mesh volume, sample
// elimination and 4x4 matrices (test infrastructure: the checker of lbm_body.cpp).
//
// Built by `make -C oracle ref` from the reference's sources where they lie
// (src/IBM/mesh/mesh_io.cpp, src/IBM/model_sampling/sampler.cu + sampler.hpp, the vendored
// third_party/nanoflann, src/IBM/math/mat4.hpp) into oracle/_ref/libref_body.so. The transformer
// (mesh_transformer.cpp) and the Morton sort (IBM_sorting.cuh) include CUDA headers and are not
// built; ref_transform composes the reference's Mat4 in MeshTransformer's order
// (mesh_transformer.cpp:45-137) instead.
#include <cstdio>

#include "IBM/math/mat4.hpp"
#include "IBM/mesh/mesh_io.hpp"
#include "IBM/model_sampling/sampler.hpp"

extern "C" {

// mesh::load_obj: counts, then up to cap vertices [n][3] and faces [n][3].
int ref_load_obj(const char* path, float* xyz, int cap_v, int* nv, int* faces, int cap_f, int* nf) {
    mesh::MeshData m = mesh::load_obj(path);
    *nv = int(m.vertices.size());
    *nf = int(m.faces.size());
    for (int i = 0; i < *nv && i < cap_v; i++)
        for (int k = 0; k < 3; k++) xyz[3 * i + k] = m.vertices[size_t(i)][size_t(k)];
    for (int i = 0; i < *nf && i < cap_f; i++)
        for (int k = 0; k < 3; k++) faces[3 * i + k] = m.faces[size_t(i)].v_indices[k];
    return 0;
}

// calculate_mesh_volume + calculate_r_max_3d + Sampler::eliminate_samples, as
// MeshTransformer::collect_ibm_body(target) runs them.
int ref_sample(const float* xyz, int nv, const int* faces, int nf, int target, float* out, int cap, int* n, float* volume,
               float* r_max) {
    std::vector<geom::Point3D> v(static_cast<size_t>(nv));
    for (int i = 0; i < nv; i++) v[size_t(i)] = geom::Point3D(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    std::vector<mesh::Face> f(static_cast<size_t>(nf));
    for (int i = 0; i < nf; i++)
        for (int k = 0; k < 3; k++) f[size_t(i)].v_indices[k] = faces[3 * i + k];
    *volume = sampler::calculate_mesh_volume(v, f);
    *r_max = sampler::calculate_r_max_3d(*volume, target);
    std::vector<sampler::Sample3> s = sampler::Sample3::from_points(v);
    sampler::Sampler smp(target, *r_max, s);
    std::vector<geom::Point3D> kept = smp.eliminate_samples();
    fflush(stdout);
    *n = int(kept.size());
    for (int i = 0; i < *n && i < cap; i++)
        for (int k = 0; k < 3; k++) out[3 * i + k] = kept[size_t(i)][size_t(k)];
    return 0;
}

// A sequence of MeshTransformer operations on vertices [nv][3] and the anchor a[3]:
// op 0 translate (p0..p2), 1 scale about the anchor (p0..p2), 2 rotate x, 3 rotate y, 4 rotate z
// about the anchor (p0 radians). Each map is composed as the transformer composes it.
int ref_transform(float* xyz, int nv, float* a, const int* ops, const float* params, int nops) {
    geom::Point3D anchor(a[0], a[1], a[2]);
    for (int o = 0; o < nops; o++) {
        const float* q = params + 3 * o;
        math::Mat4 m;
        if (ops[o] == 0) {
            m = math::Mat4::translation(q[0], q[1], q[2]);
        } else {
            math::Mat4 inner = ops[o] == 1 ? math::Mat4::scaling(q[0], q[1], q[2])
                               : ops[o] == 2 ? math::Mat4::rotationX(q[0])
                               : ops[o] == 3 ? math::Mat4::rotationY(q[0])
                                             : math::Mat4::rotationZ(q[0]);
            m = math::Mat4::translation(anchor.x(), anchor.y(), anchor.z()) * inner *
                math::Mat4::translation(-anchor.x(), -anchor.y(), -anchor.z());
        }
        for (int i = 0; i < nv; i++) {
            geom::Point3D p(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
            p = m.transform_point(p);
            for (int k = 0; k < 3; k++) xyz[3 * i + k] = p[size_t(k)];
        }
        if (ops[o] == 0) anchor = m.transform_point(anchor);
    }
    for (int k = 0; k < 3; k++) a[k] = anchor[size_t(k)];
    return 0;
}

}  // extern "C"
