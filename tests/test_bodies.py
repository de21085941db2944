"""3-D body ingestion (include/lbm_body.h, cuda-lbm_amd/csrc/lbm_body.cpp): OBJ loading, the
anchor transforms, mesh volume, weighted sample elimination, block-Morton order and the body INI
files, against the reference's own loader, sampler and matrices where they compile
(oracle/_ref/libref_body.so, built from the reference tree by `make -C oracle ref`) and against
closed forms. Host code only: no GPU."""
import ctypes
import math
from pathlib import Path

import numpy as np
import pytest

from lbm_amd import bodies as B
from lbm_amd.solver import LbmError

ROOT = Path(__file__).resolve().parent.parent
REF_LIB = ROOT / "oracle" / "_ref" / "libref_body.so"
REF_TREE = Path("/root/reference")
P = ctypes.c_void_p


def uv_sphere(n_theta=32, n_phi=32, radius=1.0):
    """A closed, outward-oriented UV sphere: poles plus (n_theta - 1) rings of n_phi vertices."""
    v = [(0.0, 0.0, radius)]
    for i in range(1, n_theta):
        th = math.pi * i / n_theta
        for j in range(n_phi):
            ph = 2 * math.pi * j / n_phi
            v.append((radius * math.sin(th) * math.cos(ph), radius * math.sin(th) * math.sin(ph), radius * math.cos(th)))
    v.append((0.0, 0.0, -radius))
    f = [(0, 1 + j, 1 + (j + 1) % n_phi) for j in range(n_phi)]
    for i in range(n_theta - 2):
        for j in range(n_phi):
            a, b = 1 + i * n_phi + j, 1 + i * n_phi + (j + 1) % n_phi
            f += [(a, a + n_phi, b), (b, a + n_phi, b + n_phi)]
    last, base = len(v) - 1, 1 + (n_theta - 2) * n_phi
    f += [(base + j, last, base + (j + 1) % n_phi) for j in range(n_phi)]
    return np.array(v, np.float32), np.array(f, np.int32)


def cube():
    """Unit cube [0, 1]^3, 12 outward triangles."""
    v = np.array([[x, y, z] for z in (0, 1) for y in (0, 1) for x in (0, 1)], np.float32)
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    f = []
    for a, b, c, d in quads:
        f += [(a, b, c), (a, c, d)]
    return v, np.array(f, np.int32)


def write_obj(path, v, f, extra=""):
    lines = ["# test mesh", ""] + [f"v {x:.7g} {y:.7g} {z:.7g}" for x, y, z in v]
    lines += [f"f {a + 1} {b + 1} {c + 1}" for a, b, c in f]
    Path(path).write_text("\n".join(lines) + "\n" + extra)


@pytest.fixture(scope="module")
def ref():
    if not REF_LIB.exists():
        pytest.skip("oracle/_ref/libref_body.so is built only where the reference tree is present")
    return ctypes.CDLL(str(REF_LIB))


def ref_sample(ref, v, f, target):
    v = np.ascontiguousarray(v, np.float32)
    f = np.ascontiguousarray(f, np.int32)
    out = np.zeros((len(v), 3), np.float32)
    n, vol, r = ctypes.c_int(), ctypes.c_float(), ctypes.c_float()
    ref.ref_sample(v.ctypes.data_as(P), len(v), f.ctypes.data_as(P), len(f), int(target), out.ctypes.data_as(P), len(v),
                   ctypes.byref(n), ctypes.byref(vol), ctypes.byref(r))
    return out[: n.value], vol.value, r.value


def jittered_sphere(seed, radius=8.0, centre=20.0, n=32):
    v, f = uv_sphere(n, n)
    v = v * (1 + 0.01 * np.random.default_rng(seed).standard_normal((len(v), 1)))
    return (v * radius + centre).astype(np.float32), f


# ---- against the reference's own code ------------------------------------------------------

def test_obj_loader_matches_the_reference_loader(ref, tmp_path):
    """mesh::load_obj: vertex lines, 'i/t/n' face references, a quad (its first triangle only),
    malformed lines skipped."""
    v, f = cube()
    extra = "\n".join(["vt 0.5 0.5", "vn 0 0 1", "f 1/1/1 2/2/2 4/4/4", "f 1 2 3 4", "f 1 2", "f a b c",
                       "v 1.5 2.5", "  v 0.25 -0.5 3e-1", "f 12abc 2 3", "# the end"])
    path = tmp_path / "m.obj"
    write_obj(path, v, f, extra)
    xyz = np.zeros((64, 3), np.float32)
    fc = np.zeros((64, 3), np.int32)
    nv, nf = ctypes.c_int(), ctypes.c_int()
    ref.ref_load_obj(str(path).encode(), xyz.ctypes.data_as(P), 64, ctypes.byref(nv), fc.ctypes.data_as(P), 64,
                     ctypes.byref(nf))
    with B.Body.load_obj(path) as b:
        assert np.array_equal(b.vertices, xyz[: nv.value])
        assert np.array_equal(b.faces, fc[: nf.value])
        assert len(b.faces) == 12 + 3  # the i/t/n face, the quad's first triangle, "12abc" read as 12


@pytest.mark.parametrize("seed,target", [(1, 50), (1, 200), (2, 600), (3, 993)])
def test_sample_elimination_matches_the_reference_bit_for_bit(ref, seed, target):
    """Weighted sample elimination on a jittered sphere (no two neighbour distances equal): the
    same volume, radius and surviving vertices, in the same order, as the reference's Sampler
    over nanoflann."""
    v, f = jittered_sphere(seed)
    want, vol, r = ref_sample(ref, v, f, target)
    with B.Body.from_mesh(v, f) as b:
        assert b.mesh_volume() == vol
        assert B.r_max_3d(vol, target) == r
        got = b.collect_ibm_body(target)
    assert got.shape == (target, 3)
    assert np.array_equal(got, want)


def test_sample_elimination_with_tied_distances_matches_in_count_and_spacing(ref):
    """On the exact UV sphere many neighbour distances are equal; which of two tied samples goes
    first follows nanoflann's tree order there and the index here, so the sets differ, but the
    count and the spacing they reach agree."""
    from scipy.spatial import cKDTree

    v, f = uv_sphere(32, 32)
    v = (v * 8 + 20).astype(np.float32)
    for target in (50, 200, 600):
        want, _, _ = ref_sample(ref, v, f, target)
        with B.Body.from_mesh(v, f) as b:
            got = b.collect_ibm_body(target)
        assert got.shape == want.shape == (target, 3)
        d_got = cKDTree(got).query(got, k=2)[0][:, 1]
        d_want = cKDTree(want).query(want, k=2)[0][:, 1]
        assert abs(d_got.mean() / d_want.mean() - 1) < 0.03
        assert d_got.min() >= 0.97 * d_want.min()


def test_transforms_match_the_reference_matrices(ref):
    """translate, scale and rotate about the anchor: the reference's Mat4 products applied in
    MeshTransformer's order give the same float vertices and anchor."""
    v, f = jittered_sphere(4, radius=1.0, centre=0.3)
    ops = [(1, (2.5, 0.75, 1.25)), (3, (0.4, 0, 0)), (2, (-1.1, 0, 0)), (0, (3.0, -2.0, 7.5)), (4, (2.2, 0, 0)),
           (1, (0.5, 0.5, 0.5))]
    with B.Body.from_mesh(v, f) as b:
        a = b.get_anchor().copy()
        for op, (p, q, r) in ops:
            if op == 0:
                b.translate(p, q, r)
            elif op == 1:
                b.scale(p, q, r)
            else:
                [b.rotate_x, b.rotate_y, b.rotate_z][op - 2](p)
        got, got_anchor = b.vertices, b.get_anchor()
    xyz = v.copy()
    codes = np.array([o for o, _ in ops], np.int32)
    params = np.array([p for _, p in ops], np.float32)
    ref.ref_transform(xyz.ctypes.data_as(P), len(xyz), a.ctypes.data_as(P), codes.ctypes.data_as(P),
                      params.ctypes.data_as(P), len(ops))
    assert np.array_equal(got, xyz)
    assert np.array_equal(got_anchor, a)


def test_config_pipeline_matches_the_reference_steps(ref, tmp_path):
    """create_body_from_config on a windTest-style file (scale to an overall size 16, anchor to
    (128, 64, 48), 256 target points): the reference's matrices and sampler, then the Morton
    order, give the same markers."""
    v, f = jittered_sphere(5, radius=0.5, centre=0.0, n=24)
    write_obj(tmp_path / "sphere.obj", v, f)
    (tmp_path / "body.ini").write_text(
        "[General]\nobj_filepath = sphere.obj\n[Transformations]\nscale_overall_size = 16.0\n"
        "position_anchor_x = 128.0 # NX / 4\nposition_anchor_y = 64.0\nposition_anchor_z = 48.0\n"
        "[Sampling]\ntype = target_points\ntarget_points = 256\n")
    with B.Body.from_config(tmp_path / "body.ini", (512, 128, 96), base_dir=tmp_path) as b:
        got = b.markers
    with B.Body.load_obj(tmp_path / "sphere.obj") as b0:
        xyz, a = b0.vertices, b0.get_anchor().copy()
        lo, hi = b0.get_aabb()
    s = np.float32(16.0) / np.float32(max(abs(hi - lo)))
    codes = np.array([1], np.int32)
    ref.ref_transform(xyz.ctypes.data_as(P), len(xyz), a.ctypes.data_as(P), codes.ctypes.data_as(P),
                      np.array([[s, s, s]], np.float32).ctypes.data_as(P), 1)
    t = np.array([[128.0 - a[0], 64.0 - a[1], 48.0 - a[2]]], np.float32)
    ref.ref_transform(xyz.ctypes.data_as(P), len(xyz), a.ctypes.data_as(P), np.array([0], np.int32).ctypes.data_as(P),
                      t.ctypes.data_as(P), 1)
    want, _, _ = ref_sample(ref, xyz, f, 256)
    assert np.array_equal(got, B.morton_sort(want, 4, 10))


# ---- closed forms and properties ---------------------------------------------------------

def test_mesh_volume_closed_forms():
    v, f = cube()
    with B.Body.from_mesh(v, f) as b:
        assert b.mesh_volume() == 1.0
        b.scale(2.0, 3.0, 0.5)  # about the centre: volume x 3
        assert abs(b.mesh_volume() - 3.0) < 1e-6
        b.translate(10.0, -4.0, 7.0)  # the signed sum does not depend on the origin
        assert abs(b.mesh_volume() - 3.0) < 1e-5
    tet = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    with B.Body.from_mesh(tet, [[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]]) as b:
        assert abs(b.mesh_volume() - 1 / 6) < 1e-7
    v, f = uv_sphere(64, 64)
    with B.Body.from_mesh(v, f) as b:  # inscribed polyhedron: just below 4/3 pi
        assert 0.99 * 4 / 3 * math.pi < b.mesh_volume() < 4 / 3 * math.pi
    with B.Body.from_mesh(v, np.zeros((0, 3), np.int32)) as b:
        assert b.mesh_volume() == 0.0


def test_r_max_is_the_radius_of_the_close_packing():
    """n spheres of radius r in close packing fill V = 4 sqrt(2) n r^3 (Yuksel's r_max)."""
    for n, r in ((100, 2.0), (256, 0.75), (7, 3.5)):
        assert abs(B.r_max_3d(4 * math.sqrt(2) * n * r ** 3, n) / r - 1) < 1e-6
    assert B.r_max_3d(1e-7, 10) == 0.0 and B.r_max_3d(5.0, 0) == 0.0


def test_sample_elimination_properties():
    v, f = jittered_sphere(6)
    with B.Body.from_mesh(v, f) as b:
        for target in (10, 123, 500):
            m = b.collect_ibm_body(target)
            assert m.shape == (target, 3)
            idx = [int(np.flatnonzero((v == p).all(1))[0]) for p in m]
            assert idx == sorted(idx)  # a subset of the vertices, in vertex order
        assert np.array_equal(b.collect_ibm_body(len(v) + 5), v)  # nothing to eliminate
        assert np.array_equal(b.collect_ibm_body(), v)
        with pytest.raises(LbmError):
            b.collect_ibm_body(0)
    # fewer samples keep the survivors further apart
    from scipy.spatial import cKDTree

    with B.Body.from_mesh(v, f) as b:
        gaps = [cKDTree(m).query(m, k=2)[0][:, 1].mean() for m in (b.collect_ibm_body(t) for t in (600, 200, 50))]
    assert gaps[0] < gaps[1] < gaps[2]


def test_morton_codes_within_a_block_and_block_order():
    """Block 4, 2 bits (quantiser max 3): local coordinates 0 and 1.5 quantise to 0 and 1, so 8
    points of one block sort x fastest, then y, then z; a second block (x in [4, 8)) follows."""
    corners = [(x, y, z) for z in (0.0, 1.5) for y in (0.0, 1.5) for x in (0.0, 1.5)]
    pts = np.array(corners + [(4.0, 0.0, 0.0), (5.5, 1.5, 1.5)], np.float32)
    order = np.random.default_rng(3).permutation(len(pts))
    got = B.morton_sort(pts[order], block_size=4, bits_per_dim=2)
    assert np.array_equal(got, pts)
    # 2-D: x bit 0 -> bit 0, y bit 0 -> bit 1, x bit 1 -> bit 2, ...; local 3.0 quantises to 2
    p2 = np.array([[0, 0], [1.5, 0], [0, 1.5], [1.5, 1.5], [3.0, 0], [0, 3.0]], np.float32)
    got2 = B.morton_sort(p2[::-1], block_size=4, bits_per_dim=2)
    assert np.array_equal(got2, p2)


def test_morton_sort_properties():
    rng = np.random.default_rng(7)
    pts = (rng.random((3000, 3)) * [37.0, 21.0, 13.0] + [5.3, -2.2, 1.1]).astype(np.float32)
    got = B.morton_sort(pts, 4, 10)
    assert sorted(map(tuple, got)) == sorted(map(tuple, pts))  # a permutation
    origin = np.floor(pts.min(0) / 4) * 4
    nb = np.ceil((pts.max(0) - origin + 4e-6) / 4).astype(int)
    bi = np.clip(np.floor((got - origin) / 4).astype(int), 0, nb - 1)
    block = (bi[:, 2] * nb[1] + bi[:, 1]) * nb[0] + bi[:, 0]
    assert np.all(np.diff(block) >= 0)
    with pytest.raises(LbmError):
        B.morton_sort(pts, 4, 11)  # 33 bits of code


def test_body_config_parser(tmp_path):
    ini = tmp_path / "b.ini"
    ini.write_text(
        "# a body\n\n[ General ]\n  OBJ_FilePath = meshes/My Body.obj  \nsave_transformed_obj_to = out/t.obj\n"
        "[TRANSFORMATIONS]\nscale_x = 2\nrotate_y_deg = 90 # degrees\nposition_anchor_x = 128.0 # NX / 4\n"
        "no equals sign here\nscale_y =\n[Sampling]\nType = DIRECT\n[Other]\nkey = 1\n")
    c = B.parse_body_config(ini)
    assert c["obj_filepath"] == "meshes/My Body.obj"
    assert c["save_transformed_obj_to"] == "out/t.obj"
    assert c["scale_x"] == 2.0 and c["rotate_y_deg"] == 90.0 and c["position_anchor_x"] == 128.0
    assert "scale_y" not in c and "scale_overall_size" not in c
    assert c["sampling_type"] == "direct"
    ini.write_text("[General]\nobj_filepath = a.obj\n[Sampling]\ntarget_points = 12\n")
    c = B.parse_body_config(ini)
    assert c["sampling_type"] == "target_points" and c["target_points"] == 12
    for bad in ("[Sampling]\ntype = direct\n",                                 # no obj_filepath
                "[General]\nobj_filepath = a.obj\n[Sampling]\ntype = target_points\n",  # no target
                "[General]\nobj_filepath = a.obj\n[Transformations]\nscale_x = big\n"):
        ini.write_text(bad)
        with pytest.raises(LbmError):
            B.parse_body_config(ini)


@pytest.mark.parametrize("name", ["windTest/sphere_config.ini", "zeroFlow/sphere_config.ini",
                                  "turbulentChannel/sphere_body_config.ini",
                                  "flowPastSquareCylinder/sphere_config.ini",
                                  "flowPastSquareCylinder/square_cylinder_config.ini"])
def test_reference_body_files_parse(name):
    """The reference scenarios' own body files, read where they lie."""
    path = REF_TREE / "src" / "scenarios" / name
    if not path.exists():
        pytest.skip("the reference tree is not present")
    c = B.parse_body_config(path)
    assert c["obj_filepath"].startswith("assets/") and c["obj_filepath"].endswith(".obj")
    if name.startswith("windTest"):
        assert c["scale_overall_size"] == 16.0 and c["target_points"] == 256
        assert (c["position_anchor_x"], c["position_anchor_y"], c["position_anchor_z"]) == (128.0, 64.0, 48.0)


def test_config_pipeline_order_and_errors(tmp_path):
    """The file's transforms run as create_body_from_config runs them (scale, rotate y, x, z,
    anchor, move), the transformed mesh is saved when asked, a body outside the lattice and a
    missing mesh are errors."""
    v, f = uv_sphere(16, 16)
    write_obj(tmp_path / "s.obj", v, f)
    (tmp_path / "b.ini").write_text(
        "[General]\nobj_filepath = s.obj\nsave_transformed_obj_to = t.obj\n[Transformations]\n"
        "scale_x = 4\nscale_z = 2\nrotate_x_deg = 30\nrotate_y_deg = -45\nrotate_z_deg = 10\n"
        "set_anchor_local_x = 0\nset_anchor_local_y = 0\nset_anchor_local_z = 0\n"
        "position_anchor_x = 20\nposition_anchor_y = 12\nposition_anchor_z = 9\n[Sampling]\ntype = direct\n")
    with B.Body.from_config(tmp_path / "b.ini", (40, 24, 18), base_dir=tmp_path) as b:
        got = b.markers
    rad = np.float32(math.pi) / np.float32(180.0)
    with B.Body.load_obj(tmp_path / "s.obj") as m:
        m.scale(4, 1, 2).rotate_y(np.float32(-45) * rad).rotate_x(np.float32(30) * rad).rotate_z(np.float32(10) * rad)
        m.set_anchor(0, 0, 0).move_anchor_to_world(20, 12, 9)
        want = B.morton_sort(m.vertices, 4, 10)
        assert np.array_equal(got, want)
    with B.Body.load_obj(tmp_path / "t.obj") as t:
        assert np.allclose(np.sort(t.vertices, 0), np.sort(want, 0), atol=1e-5)
    with pytest.raises(LbmError) as e:
        B.Body.from_config(tmp_path / "b.ini", (22, 24, 18), base_dir=tmp_path)  # x reaches about 24.5
    assert "does not fit" in str(e.value)
    (tmp_path / "c.ini").write_text("[General]\nobj_filepath = missing.obj\n[Sampling]\ntype = direct\n")
    with pytest.raises(LbmError):
        B.Body.from_config(tmp_path / "c.ini", (40, 24, 18), base_dir=tmp_path)


def test_points_csv_round_trip_and_reader_rules(tmp_path):
    pts = np.random.default_rng(9).random((17, 3)).astype(np.float32) * 50
    B.save_points_csv(tmp_path / "p.csv", pts)
    back = B.load_points_csv(tmp_path / "p.csv")
    assert back.shape == pts.shape and np.allclose(back, pts, rtol=1e-5)  # six significant digits
    (tmp_path / "q.csv").write_text("X,Y,Z\n1,2,3\n\n# comment\n4,5\n6,7,8\nbad,1,2\n")
    assert np.array_equal(B.load_points_csv(tmp_path / "q.csv"), np.array([[1, 2, 3], [6, 7, 8]], np.float32))


def test_body_ingestion_under_address_and_ub_sanitizers(tmp_path):
    """Every entry point of include/lbm_body.h, degenerate meshes, malformed files and the INI
    pipeline, run by tests/emu/body_check (lbm_body.cpp built with ASan + UBSan)."""
    import subprocess

    exe = ROOT / "tests" / "emu" / "body_check"
    if not exe.exists():
        pytest.skip("tests/emu/body_check not built")
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "body check ok" in r.stdout
