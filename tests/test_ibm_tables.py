"""The IBM band tables (cuda-lbm_amd/csrc/lbm_ibm_tables.hpp) built on the host under
AddressSanitizer + UndefinedBehaviorSanitizer (tests/emu/ibm_tables_check), against an
independent numpy construction of the stencils (IBM_impl.cu:5-51, 120-154; IBMUtils.cuh:23-43).

Covers the marker set of the BGK + IBM-quirk GPU case (2x2 stencil, no ds), the default 4x4
stencil, markers at and beyond the grid edges, and seeded random sets."""
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from lbm_amd import scenarios as sc

ROOT = Path(__file__).resolve().parent.parent
CHECK = ROOT / "tests" / "emu" / "ibm_tables_check"


def delta4(r):
    """Peskin's 4-point kernel (IBMUtils.cuh:23-33) in FP64."""
    a = abs(r)
    if a < 1.0:
        return 0.125 * (3.0 - 2.0 * a + math.sqrt(1.0 + 4.0 * a - 4.0 * r * r))
    if a < 2.0:
        return 0.125 * (5.0 - 2.0 * a - math.sqrt(-7.0 + 12.0 * a - 4.0 * r * r))
    return 0.0


def run_check(tmp_path, nx, ny, xy, ds, two, no_ds):
    if not CHECK.exists():
        pytest.fail("tests/emu/ibm_tables_check is not built (run __graft_entry__.build())")
    xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
    n = xy.shape[0]
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as fh:
        fh.write(np.array([nx, ny, int(two), int(no_ds), n, int(ds is not None)], np.int32).tobytes())
        fh.write(xy.tobytes())
        if ds is not None:
            fh.write(np.ascontiguousarray(ds, np.float32).tobytes())
    r = subprocess.run([str(CHECK), str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr  # a sanitizer report aborts with a non-zero status
    b = np.fromfile(fout, np.uint8)
    i32 = lambda o, k: (np.frombuffer(b, np.int32, k, o), o + 4 * k)
    f32 = lambda o, k: (np.frombuffer(b, np.float32, k, o), o + 4 * k)
    (status,), o = i32(0, 1)
    if status:
        return None
    (nb, ne, x0, x1, y0, y1), o = i32(o, 6)
    t = {"box": (x0, x1, y0, y1)}
    t["band_node"], o = i32(o, nb)
    t["mk_ptr"], o = i32(o, n + 1)
    t["mk_band"], o = i32(o, ne)
    t["mk_w"], o = f32(o, ne)
    t["bn_ptr"], o = i32(o, nb + 1)
    t["bn_mk"], o = i32(o, ne)
    t["bn_w"], o = f32(o, ne)
    t["box_map"], o = i32(o, (x1 - x0) * (y1 - y0))
    (ks,), o = i32(o, 1)
    t["kspread"] = ks
    t["mkf_band"], o = i32(o, 16 * n)
    t["mkf_w"], o = f32(o, 16 * n)
    t["bnf_mk"], o = i32(o, ks * nb)
    t["bnf_w"], o = f32(o, ks * nb)
    assert o == len(b)
    return t


def expected(nx, ny, xy, ds, two, no_ds):
    """Per marker, the (node, weight) list in the reference's loop order, in FP64."""
    W, off = (2, 0) if two else (4, -1)
    out = []
    for l, (px, py) in enumerate(np.asarray(xy, np.float32)):
        ent = []
        if -4 <= px <= nx + 4 and -4 <= py <= ny + 4:
            gx, gy = math.floor(px), math.floor(py)
            for i in range(W):
                for j in range(W):
                    x, y = gx + off + i, gy + off + j
                    if 0 <= x < nx and 0 <= y < ny:
                        ent.append((y * nx + x, delta4(float(px) - x) * delta4(float(py) - y)))
        out.append(ent)
    return out


def verify(t, nx, ny, xy, ds, two, no_ds):
    n = len(xy)
    exp = expected(nx, ny, xy, ds, two, no_ds)
    band = sorted({node for ent in exp for node, _ in ent})
    assert list(t["band_node"]) == band
    pos = {node: b for b, node in enumerate(band)}
    # marker -> band CSR in the reference's loop order
    for l in range(n):
        a, e = t["mk_ptr"][l], t["mk_ptr"][l + 1]
        assert [band[b] for b in t["mk_band"][a:e]] == [node for node, _ in exp[l]]
        np.testing.assert_allclose(t["mk_w"][a:e], [w for _, w in exp[l]], rtol=2e-6, atol=1e-7)
    # band -> marker CSR: marker order, weight phi phi ds
    dsv = np.ones(n) if (ds is None or no_ds) else np.asarray(ds, np.float64)
    for node, b in pos.items():
        a, e = t["bn_ptr"][b], t["bn_ptr"][b + 1]
        want = [(l, w * dsv[l]) for l in range(n) for nd, w in exp[l] if nd == node]
        assert list(t["bn_mk"][a:e]) == [l for l, _ in want]
        np.testing.assert_allclose(t["bn_w"][a:e], [w for _, w in want], rtol=2e-6, atol=1e-7)
    # box: covers the band, x range in whole 4-node segments, inverse of band_node
    x0, x1, y0, y1 = t["box"]
    xs, ys = [v % nx for v in band], [v // nx for v in band]
    assert x0 % 4 == 0 and (x1 % 4 == 0 or x1 == nx) and x0 <= min(xs) and x1 > max(xs)
    assert y0 == min(ys) and y1 == max(ys) + 1
    bm = t["box_map"].reshape(y1 - y0, x1 - x0)
    for b, node in enumerate(band):
        assert bm[node // nx - y0, node % nx - x0] == b
    assert int(np.sum(bm >= 0)) == len(band)
    # fixed-width lists: slot-major, each list's CSR entries in order, then (0, 0.0) padding
    longest = int(np.max(np.diff(t["bn_ptr"])))
    assert t["kspread"] == (8 if longest <= 8 else 16 if longest <= 16 else 0)
    mb, mw = t["mkf_band"].reshape(16, n), t["mkf_w"].reshape(16, n)
    for l in range(n):
        a, e = t["mk_ptr"][l], t["mk_ptr"][l + 1]
        assert list(mb[: e - a, l]) == list(t["mk_band"][a:e]) and not mb[e - a:, l].any()
        assert np.array_equal(mw[: e - a, l], t["mk_w"][a:e]) and not mw[e - a:, l].any()
    ks, nb = t["kspread"], len(band)
    if ks:
        km, kw = t["bnf_mk"].reshape(ks, nb), t["bnf_w"].reshape(ks, nb)
        for b in range(nb):
            a, e = t["bn_ptr"][b], t["bn_ptr"][b + 1]
            assert list(km[: e - a, b]) == list(t["bn_mk"][a:e]) and not km[e - a:, b].any()
            assert np.array_equal(kw[: e - a, b], t["bn_w"][a:e]) and not kw[e - a:, b].any()


def test_tables_of_the_bgk_ibm_quirk_case(tmp_path):
    """The marker set of test_gpu_ibm_matches_oracle with the 2x2 | clamp | no-ds quirk bits."""
    case = sc.ibm_cylinder(nx=96, ny=48, u_in=0.05, D=12.0, n_markers=38, Re=20.0, cx=30.0)
    for two, no_ds in ((True, True), (False, False), (True, False), (False, True)):
        t = run_check(tmp_path, 96, 48, case.markers, case.ds, two, no_ds)
        verify(t, 96, 48, case.markers, case.ds, two, no_ds)


def test_tables_of_the_bench_cylinder(tmp_path):
    case = sc.ibm_cylinder()
    t = run_check(tmp_path, 2048, 512, case.markers, case.ds, False, False)
    verify(t, 2048, 512, case.markers, case.ds, False, False)


@pytest.mark.parametrize("seed", range(12))
def test_tables_of_random_marker_sets(tmp_path, seed):
    rng = np.random.default_rng(seed)
    nx, ny = int(rng.integers(5, 40)), int(rng.integers(5, 40))
    n = int(rng.integers(1, 60))
    # markers inside, on the edges and up to 6 cells outside the grid
    xy = np.stack([rng.uniform(-6, nx + 6, n), rng.uniform(-6, ny + 6, n)], -1).astype(np.float32)
    xy[: n // 4, 0] = np.floor(xy[: n // 4, 0])  # some on grid lines
    ds = rng.uniform(0.1, 2.0, n).astype(np.float32) if seed % 2 else None
    two, no_ds = bool(seed % 3 == 0), bool(seed % 4 == 1)
    t = run_check(tmp_path, nx, ny, xy, ds, two, no_ds)
    if t is None:  # rejected: only when no stencil touches the grid
        assert not any(expected(nx, ny, xy, ds, two, no_ds))
        return
    verify(t, nx, ny, xy, ds, two, no_ds)


def test_crowded_band_has_no_fixed_spread_lists(tmp_path):
    # 40 markers within a quarter cell: each band node of their common 4x4 stencil is touched
    # by all of them (more than 16), so only the CSR spread lists exist
    rng = np.random.default_rng(5)
    xy = (np.array([6.3, 7.6]) + rng.uniform(0, 0.25, (40, 2))).astype(np.float32)
    t = run_check(tmp_path, 16, 16, xy, None, False, False)
    verify(t, 16, 16, xy, None, False, False)
    assert t["kspread"] == 0 and len(t["bnf_mk"]) == 0
    xy12 = xy[:12]
    t = run_check(tmp_path, 16, 16, xy12, None, False, False)
    verify(t, 16, 16, xy12, None, False, False)
    assert t["kspread"] == 16


def test_far_away_and_invalid_markers(tmp_path):
    # a marker far outside touches nothing (no float -> int overflow), the others still count
    xy = np.array([[1e30, -1e30], [5.5, 5.5]], np.float32)
    t = run_check(tmp_path, 12, 12, xy, None, False, False)
    verify(t, 12, 12, xy, None, False, False)
    assert t["mk_ptr"][1] == 0
    assert run_check(tmp_path, 12, 12, np.array([[np.nan, 1.0]], np.float32), None, False, False) is None
    assert run_check(tmp_path, 12, 12, np.array([[3.0, 1.0]], np.float32), np.array([0.0], np.float32), False,
                     False) is None
    assert run_check(tmp_path, 12, 12, np.array([[300.0, 1.0]], np.float32), None, False, False) is None
