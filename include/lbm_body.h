/* lbm_body.h — 3-D immersed-body ingestion: OBJ meshes, their placement in the lattice, marker
 * sampling and ordering, and the body INI files of the reference's 3-D scenarios.
 *
 * Replaces the host pipeline the reference runs in Scenario::add_bodies() before
 * IBMManager<3>::init_and_dispatch (IBMManager.cuh:54-110):
 *   conf::create_body_from_config (src/IBM/config/config_parser.cpp:132-209):
 *     parse_body_config (:32-130) -> mesh::MeshTransformer (src/IBM/mesh/mesh_transformer.cpp)
 *     over mesh::load_obj (src/IBM/mesh/mesh_io.cpp:13-93) -> fits_in_domain(NX, NY, NZ)
 *     -> every vertex (DIRECT) or weighted sample elimination to a target count
 *        (src/IBM/model_sampling/sampler.hpp:20-164, volume and radius sampler.cu:5-49)
 *   -> block_morton_sort<3>(block 4, 10 bits) (src/IBM/IBM_sorting.cuh:41-159).
 * Host code only (no HIP calls); the markers it yields go to lbm3d_ibm_set_markers.
 *
 * Every function returns LBM_OK or a negative LBM_E_* code (lbm_hip.h); after a failure,
 * lbm_body_last_error(body) (or lbm_body_last_error(NULL) for calls without a body) says why.
 */
#ifndef LBM_BODY_H
#define LBM_BODY_H

#include <stdint.h>

#include "lbm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lbm_body lbm_body;

/* ---- mesh and transforms (mesh::MeshTransformer, mesh_transformer.hpp:16-55) ---------------
 * A body holds a triangle mesh (vertices in float, faces as 0-based vertex triples), an anchor
 * (about which scale and rotations act; initially the centre of the vertices' bounding box) and
 * its markers (set by lbm_body_sample / lbm_body_markers_from_vertices). */

/* mesh::load_obj (mesh_io.cpp:13-93) + MeshTransformer(filename) (mesh_transformer.cpp:17-24):
 * "v x y z" lines; the first three vertex references of each "f" line ("i", "i/t", "i/t/n";
 * 1-based), as the reference reads them (a polygon with more vertices contributes its first
 * triangle only); malformed lines are skipped. A file without vertices is LBM_E_INVALID. */
int lbm_body_load_obj(const char* path, lbm_body** out);
/* MeshTransformer(MeshData) (mesh_transformer.cpp:26-33) from arrays: xyz [nv][3], faces
 * [nf][3] (may be NULL when nf = 0). */
int lbm_body_from_mesh(const float* xyz, int32_t nv, const int32_t* faces, int32_t nf, lbm_body** out);
int lbm_body_destroy(lbm_body* body);
/* Counts of vertices, faces and markers (any pointer may be NULL). */
int lbm_body_info(const lbm_body* body, int32_t* n_vertices, int32_t* n_faces, int32_t* n_markers);
/* Copies: vertices [nv][3], faces [nf][3], markers [nm][3]. */
int lbm_body_vertices(const lbm_body* body, float* xyz);
int lbm_body_faces(const lbm_body* body, int32_t* faces);
int lbm_body_markers(const lbm_body* body, float* xyz);

/* translate (:45-52), scale about the anchor (:54-65), scale_to_overall_size (largest bounding
 * box side -> size, :67-78), rotate_x / y / z about the anchor (:80-117; axis 0, 1, 2, angle in
 * radians), set_anchor (:40-43), set_anchor_to_centroid (:35-38), move_anchor_to_world (:119-126),
 * move_centroid_to_world (:128-137). Each is a 4x4 float affine map composed and applied as the
 * reference composes it (translation to the anchor, the map, translation back). */
int lbm_body_translate(lbm_body* body, float tx, float ty, float tz);
int lbm_body_scale(lbm_body* body, float sx, float sy, float sz);
int lbm_body_scale_to_overall_size(lbm_body* body, float size);
int lbm_body_rotate(lbm_body* body, int32_t axis, float angle_rad);
int lbm_body_set_anchor(lbm_body* body, float x, float y, float z);
int lbm_body_set_anchor_to_centroid(lbm_body* body);
int lbm_body_move_anchor_to_world(lbm_body* body, float x, float y, float z);
int lbm_body_move_centroid_to_world(lbm_body* body, float x, float y, float z);
/* get_aabb (:147-149; min[3], max[3]) and get_anchor (:151-153; a[3]). */
int lbm_body_aabb(const lbm_body* body, float* min3, float* max3);
int lbm_body_anchor(const lbm_body* body, float* a3);
/* fits_in_domain (:159-181): the bounding box inside [0, nx] x [0, ny] x [0, nz]. */
int lbm_body_fits_in_domain(const lbm_body* body, float nx, float ny, float nz, int32_t* fits);
/* mesh::save_obj (mesh_io.cpp:95-131), the reference's collect_file (mesh_transformer.cpp:155-157). */
int lbm_body_save_obj(const lbm_body* body, const char* path);

/* ---- markers ------------------------------------------------------------------------------
 * sampler::calculate_mesh_volume (sampler.cu:5-40): |sum of signed tetrahedra (origin, face)|,
 * accumulated in FP64, returned in float as the reference does. */
int lbm_body_mesh_volume(const lbm_body* body, float* volume);
/* sampler::calculate_r_max_3d (sampler.cu:42-49): cbrt(V / (4 sqrt(2) n)), 0 for V <= 1e-6 or
 * n <= 0. Needs no body. */
float lbm_body_r_max(float volume, int32_t n);
/* MeshTransformer::collect_ibm_body(target) (mesh_transformer.cpp:191-217): markers = the
 * vertices reduced by weighted sample elimination (Yuksel 2015 as sampler.hpp:20-164 runs it:
 * weights sum (1 - d / 2r)^8 over neighbours closer than 2r, r = lbm_body_r_max(mesh volume,
 * target); the heaviest sample goes first and its neighbours' weights drop) to `target`
 * points, kept in vertex order. target >= the vertex count keeps every vertex. */
int lbm_body_sample(lbm_body* body, int32_t target);
/* MeshTransformer::collect_ibm_body() (:183-189): markers = the vertices. */
int lbm_body_markers_from_vertices(lbm_body* body);
/* block_morton_sort<dim> (IBM_sorting.cuh:41-159) on the markers: blocks of block_size lattice
 * units from the floor of the markers' bounding box, ordered by block index (x fastest), then
 * by the Morton code of the position inside the block quantised to bits_per_dim bits.
 * IBMManager<3>::init_and_dispatch uses block 4, 10 bits. */
int lbm_body_morton_sort(lbm_body* body, int32_t block_size, int32_t bits_per_dim);
/* The same order on a caller's array xyz [n][dim] (dim 2 or 3), in place. Needs no body. */
int lbm_points_morton_sort(float* xyz, int32_t n, int32_t dim, int32_t block_size, int32_t bits_per_dim);

/* sampler::save_points_to_csv (utils.hpp:80-92): "x,y,z" then one point per line. */
int lbm_points_save_csv(const char* path, const float* xyz, int32_t n);
/* load_from_csv (IBM_utils.cu:111-189): an "x"/"X" header line, empty and '#' lines and lines
 * without three numbers are skipped. With xyz = NULL only *n is set; otherwise at most cap
 * points are written and *n is the number in the file. */
int lbm_points_load_csv(const char* path, float* xyz, int32_t cap, int32_t* n);

/* ---- body INI files (conf::BodyConfig, src/IBM/config/body_config.hpp:12-51) ---------------- */
enum {
    LBM_BODY_SAMPLING_DIRECT = 0,      /* every vertex                                    */
    LBM_BODY_SAMPLING_TARGET_POINTS = 1 /* weighted sample elimination to target_points    */
};
/* bits of lbm_body_config.has: which optional transform fields the file set */
enum {
    LBM_BODY_HAS_SCALE_OVERALL = 1 << 0,
    LBM_BODY_HAS_SCALE_X = 1 << 1,
    LBM_BODY_HAS_SCALE_Y = 1 << 2,
    LBM_BODY_HAS_SCALE_Z = 1 << 3,
    LBM_BODY_HAS_ROTATE_X = 1 << 4,
    LBM_BODY_HAS_ROTATE_Y = 1 << 5,
    LBM_BODY_HAS_ROTATE_Z = 1 << 6,
    LBM_BODY_HAS_ANCHOR_LOCAL_X = 1 << 7,
    LBM_BODY_HAS_ANCHOR_LOCAL_Y = 1 << 8,
    LBM_BODY_HAS_ANCHOR_LOCAL_Z = 1 << 9,
    LBM_BODY_HAS_POSITION_X = 1 << 10,
    LBM_BODY_HAS_POSITION_Y = 1 << 11,
    LBM_BODY_HAS_POSITION_Z = 1 << 12,
    LBM_BODY_HAS_TARGET_POINTS = 1 << 13
};
typedef struct lbm_body_config {
    char     obj_filepath[1024];
    char     save_transformed_obj_to[1024]; /* "" = not requested */
    uint32_t has;                           /* LBM_BODY_HAS_* */
    float    scale_overall_size, scale_x, scale_y, scale_z;
    float    rotate_x_deg, rotate_y_deg, rotate_z_deg;
    float    set_anchor_local_x, set_anchor_local_y, set_anchor_local_z;
    float    position_anchor_x, position_anchor_y, position_anchor_z;
    int32_t  sampling_type;                 /* LBM_BODY_SAMPLING_*, TARGET_POINTS by default */
    int32_t  target_points;
} lbm_body_config;

/* conf::parse_body_config (config_parser.cpp:32-130): sections [General], [Transformations],
 * [Sampling] and keys are case-insensitive, whitespace is trimmed, '#' starts a comment line,
 * numbers are read from the start of the value (so "128.0 # NX / 4" is 128). Lines without
 * '=', empty values, unknown sections and sampling types are skipped (the reference warns).
 * Errors: no obj_filepath; a number that does not parse; TARGET_POINTS without a positive
 * target_points (the reference goes on to fail later). */
int lbm_body_parse_config(const char* ini_path, lbm_body_config* out);
/* conf::create_body_from_config (config_parser.cpp:132-209) for a nx x ny x nz lattice, then
 * the Morton order IBMManager<3>::init_and_dispatch gives (IBMManager.cuh:57-66): load the OBJ,
 * scale (overall size, or per axis with 1 for missing axes), rotate y, x, z (degrees), set the
 * anchor and move it when all three of its coordinates are given, save the transformed mesh when
 * asked, check the bounding box lies in [0, n] on each axis (LBM_E_INVALID otherwise), take the
 * markers (every vertex or the sampled target count), Morton-sort them (block 4, 10 bits).
 * Relative paths in the file resolve against base_dir (NULL: the working directory, as in the
 * reference). The reference also writes its samples to debug/body.csv; this does not. */
int lbm_body_from_config(const char* ini_path, const char* base_dir, int32_t nx, int32_t ny, int32_t nz, lbm_body** out);

const char* lbm_body_last_error(const lbm_body* body);

#ifdef __cplusplus
}
#endif

#endif /* LBM_BODY_H */
