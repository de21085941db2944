// Host check of the body ingestion (cuda-lbm_amd/csrc/lbm_body.cpp) built with AddressSanitizer
// and UndefinedBehaviorSanitizer: every entry point of include/lbm_body.h on generated meshes,
// the INI pipeline, edge cases (one vertex, flat meshes, zero targets, bad arguments, faces with
// out-of-range indices) and the error paths. Prints "body check ok" when every call returned
// what it should.
//
// usage: body_check <scratch directory>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/lbm_body.h"

#define CHECK(c)                                                            \
    do {                                                                    \
        if (!(c)) {                                                         \
            std::printf("check failed at line %d: %s\n", __LINE__, #c);    \
            return 1;                                                       \
        }                                                                   \
    } while (0)

static void sphere(int nt, int np, std::vector<float>& v, std::vector<int>& f) {
    v = {0.0f, 0.0f, 1.0f};
    for (int i = 1; i < nt; i++)
        for (int j = 0; j < np; j++) {
            const double th = M_PI * i / nt, ph = 2 * M_PI * j / np;
            v.push_back(float(std::sin(th) * std::cos(ph)));
            v.push_back(float(std::sin(th) * std::sin(ph)));
            v.push_back(float(std::cos(th)));
        }
    v.insert(v.end(), {0.0f, 0.0f, -1.0f});
    f.clear();
    for (int j = 0; j < np; j++) f.insert(f.end(), {0, 1 + j, 1 + (j + 1) % np});
    for (int i = 0; i < nt - 2; i++)
        for (int j = 0; j < np; j++) {
            const int a = 1 + i * np + j, b = 1 + i * np + (j + 1) % np;
            f.insert(f.end(), {a, a + np, b, b, a + np, b + np});
        }
    const int last = int(v.size() / 3) - 1, base = 1 + (nt - 2) * np;
    for (int j = 0; j < np; j++) f.insert(f.end(), {base + j, last, base + (j + 1) % np});
}

int main(int argc, char** argv) {
    if (argc != 2) return 2;
    const std::string dir = argv[1];
    std::vector<float> v;
    std::vector<int> f;
    sphere(12, 16, v, f);
    const int nv = int(v.size() / 3), nf = int(f.size() / 3);
    // an OBJ file with odd lines, and body files
    FILE* o = std::fopen((dir + "/s.obj").c_str(), "w");
    CHECK(o);
    for (int i = 0; i < nv; i++) std::fprintf(o, "v %.8g %.8g %.8g\n", v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    for (int i = 0; i < nf; i++) std::fprintf(o, "f %d/1 %d//2 %d\n", f[3 * i] + 1, f[3 * i + 1] + 1, f[3 * i + 2] + 1);
    std::fprintf(o, "f 1 2\nf\nv\nv 1\nf 99999 1 2\nf -3 1 2\nf 2147483648 1 2\n# end");
    std::fclose(o);
    lbm_body* b = nullptr;
    CHECK(lbm_body_load_obj((dir + "/s.obj").c_str(), &b) == LBM_OK);
    int32_t cv = 0, cf = 0, cm = 0;
    CHECK(lbm_body_info(b, &cv, &cf, &cm) == LBM_OK && cv == nv && cf == nf + 2 && cm == 0);
    float vol = 0.0f;
    CHECK(lbm_body_mesh_volume(b, &vol) == LBM_OK && vol > 3.0f && vol < 4.2f);  // bad faces skipped
    CHECK(lbm_body_scale_to_overall_size(b, 10.0f) == LBM_OK);
    CHECK(lbm_body_rotate(b, 0, 0.3f) == LBM_OK && lbm_body_rotate(b, 1, -0.2f) == LBM_OK && lbm_body_rotate(b, 2, 1.0f) == LBM_OK);
    CHECK(lbm_body_rotate(b, 3, 1.0f) == LBM_E_INVALID);
    CHECK(lbm_body_move_centroid_to_world(b, 20.0f, 20.0f, 20.0f) == LBM_OK);
    for (int target : {1, 2, 17, nv - 1, nv, nv + 3}) {
        CHECK(lbm_body_sample(b, target) == LBM_OK);
        CHECK(lbm_body_info(b, nullptr, nullptr, &cm) == LBM_OK && cm == (target < nv ? target : nv));
        CHECK(lbm_body_morton_sort(b, 4, 10) == LBM_OK);
        CHECK(lbm_body_morton_sort(b, 1, 1) == LBM_OK);
    }
    CHECK(lbm_body_sample(b, 0) == LBM_E_INVALID && lbm_body_sample(b, -5) == LBM_E_INVALID);
    std::vector<float> m(size_t(nv) * 3), vv(size_t(nv) * 3);
    std::vector<int32_t> ff(size_t(cf) * 3);
    CHECK(lbm_body_markers(b, m.data()) == LBM_OK && lbm_body_vertices(b, vv.data()) == LBM_OK && lbm_body_faces(b, ff.data()) == LBM_OK);
    CHECK(lbm_body_save_obj(b, (dir + "/t.obj").c_str()) == LBM_OK);
    CHECK(lbm_points_save_csv((dir + "/p.csv").c_str(), m.data(), cm) == LBM_OK);
    int32_t n = 0;
    CHECK(lbm_points_load_csv((dir + "/p.csv").c_str(), nullptr, 0, &n) == LBM_OK && n == cm);
    CHECK(lbm_points_load_csv((dir + "/p.csv").c_str(), m.data(), 3, &n) == LBM_OK && n == cm);  // capped
    lbm_body_destroy(b);
    // degenerate meshes: one vertex, no faces, a flat square (zero volume: r = 0, no neighbours)
    const float one[3] = {1.0f, 2.0f, 3.0f};
    CHECK(lbm_body_from_mesh(one, 1, nullptr, 0, &b) == LBM_OK);
    CHECK(lbm_body_sample(b, 1) == LBM_OK && lbm_body_scale_to_overall_size(b, 2.0f) == LBM_E_INVALID);
    lbm_body_destroy(b);
    const float sq[12] = {0, 0, 0, 1, 0, 0, 1, 1, 0, 0, 1, 0};
    const int32_t sqf[6] = {0, 1, 2, 0, 2, 3};
    CHECK(lbm_body_from_mesh(sq, 4, sqf, 2, &b) == LBM_OK);
    CHECK(lbm_body_sample(b, 2) == LBM_OK && lbm_body_info(b, nullptr, nullptr, &cm) == LBM_OK && cm == 2);
    lbm_body_destroy(b);
    const float nan3[3] = {NAN, 0.0f, 0.0f};
    CHECK(lbm_body_from_mesh(nan3, 1, nullptr, 0, &b) == LBM_E_INVALID);
    CHECK(lbm_body_from_mesh(sq, 0, nullptr, 0, &b) == LBM_E_INVALID);
    CHECK(lbm_body_load_obj((dir + "/missing.obj").c_str(), &b) == LBM_E_INVALID);
    float pts[6] = {1.0f, 2.0f, NAN, 4.0f, 5.0f, 6.0f};
    CHECK(lbm_points_morton_sort(pts, 2, 3, 4, 10) == LBM_E_INVALID);
    CHECK(lbm_points_morton_sort(pts, 0, 3, 4, 10) == LBM_OK);
    CHECK(lbm_points_morton_sort(pts + 3, 1, 3, 4, 10) == LBM_OK);
    CHECK(lbm_points_morton_sort(pts, 3, 2, 0, 10) == LBM_E_INVALID);
    // the INI pipeline, direct and sampled, and a body that does not fit
    for (const char* sampling : {"type = direct", "type = target_points\ntarget_points = 40"}) {
        FILE* c = std::fopen((dir + "/b.ini").c_str(), "w");
        CHECK(c);
        std::fprintf(c, "[General]\nobj_filepath = s.obj\nsave_transformed_obj_to = u.obj\n[Transformations]\n"
                        "scale_overall_size = 8\nrotate_x_deg = 20\nposition_anchor_x = 16\nposition_anchor_y = 12\n"
                        "position_anchor_z = 10\n[Sampling]\n%s\n", sampling);
        std::fclose(c);
        CHECK(lbm_body_from_config((dir + "/b.ini").c_str(), dir.c_str(), 32, 24, 20, &b) == LBM_OK);
        lbm_body_destroy(b);
        CHECK(lbm_body_from_config((dir + "/b.ini").c_str(), dir.c_str(), 18, 24, 20, &b) == LBM_E_INVALID);
        CHECK(b == nullptr);
    }
    lbm_body_config cfg;
    CHECK(lbm_body_parse_config((dir + "/b.ini").c_str(), &cfg) == LBM_OK && cfg.target_points == 40);
    CHECK(lbm_body_parse_config((dir + "/none.ini").c_str(), &cfg) == LBM_E_INVALID);
    std::printf("body check ok\n");
    return 0;
}
