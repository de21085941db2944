"""The reference's output files (src/core/lbm.cuh:173-343) written by liblbm_hip.so's host
writers, read back the way their consumers read them: a VTK XML appended-raw reader (arrays at
their declared offsets after '_', UInt64 byte counts), graphics/visualize.py's
np.fromfile(...).reshape((ny, nx, 2)) for the .bin dumps and graphics/plot_midplane.py's
reshape((NY, NX, 3)) for mid-plane slices. CPU only (host writers, no device work)."""
import re

import numpy as np
import pytest

from lbm_amd import _abi as A
from lbm_amd import output


def read_vti(path):
    """{name: float32 array} of a VTK ImageData file with raw appended data and UInt64 headers,
    plus the header attributes under '_attrs'."""
    data = path.read_bytes()
    mark = data.index(b"<AppendedData")
    head = data[:mark].decode()
    start = data.index(b"_", mark) + 1
    attrs = dict(re.findall(r'(\w+)="([^"]*)"', head.split("<ImageData", 1)[0]))
    assert attrs.get("header_type") == "UInt64"
    out = {"_attrs": attrs, "_extent": re.search(r'WholeExtent="([^"]*)"', head).group(1)}
    for m in re.finditer(r"<DataArray ([^>]*)/>", head):
        a = dict(re.findall(r'(\w+)="([^"]*)"', m.group(1)))
        assert a["type"] == "Float32" and a["format"] == "appended"
        off = start + int(a["offset"])
        nbytes = int(np.frombuffer(data[off:off + 8], np.uint64)[0])
        arr = np.frombuffer(data[off + 8:off + 8 + nbytes], np.float32)
        out[a["Name"]] = arr.reshape(-1, int(a.get("NumberOfComponents", 1))).squeeze()
    assert data.rstrip().endswith(b"</VTKFile>")
    return out


def fields(shape, comps, seed):
    rng = np.random.default_rng(seed)
    return (1 + 0.01 * rng.standard_normal(shape)).astype(np.float32), \
        (0.05 * rng.standard_normal(shape + (comps,))).astype(np.float32)


def test_vti_2d_as_the_reference_writes_it(tmp_path):
    ny, nx = 37, 53
    rho, u = fields((ny, nx), 2, 1)
    path = output.save_vtk(tmp_path, 42, rho, u)
    assert path.endswith("vtk/sim_data_000042.vti")
    r = read_vti(tmp_path / "vtk" / "sim_data_000042.vti")
    assert r["_attrs"]["version"] == "1.0" and r["_attrs"]["byte_order"] == "LittleEndian"
    assert r["_extent"] == f"0 {nx - 1} 0 {ny - 1} 0 0"
    # node order x fastest, then y (lbm.cuh:317-330); u_z padded with zeros for 2-D
    assert np.array_equal(r["Density"], rho.reshape(-1))
    assert np.array_equal(r["Velocity"][:, :2], u.reshape(-1, 2)) and not r["Velocity"][:, 2].any()


def test_vti_3d(tmp_path):
    nz, ny, nx = 6, 5, 7
    rho, u = fields((nz, ny, nx), 3, 2)
    output.save_vtk(tmp_path, 7, rho, u)
    r = read_vti(tmp_path / "vtk" / "sim_data_000007.vti")
    assert r["_extent"] == f"0 {nx - 1} 0 {ny - 1} 0 {nz - 1}"
    assert np.array_equal(r["Density"], rho.reshape(-1))
    assert np.array_equal(r["Velocity"], u.reshape(-1, 3))


def test_vti_reference_offset_quirk(tmp_path):
    """The reference declares the Velocity offset as the density bytes alone (lbm.cuh:298-299),
    8 bytes short of the array's length word; LBM_VTI_REFERENCE_OFFSET reproduces it."""
    ny, nx = 4, 6
    rho, u = fields((ny, nx), 2, 3)
    output.save_vtk(tmp_path, 1, rho, u, reference_offset=True)
    text = (tmp_path / "vtk" / "sim_data_000001.vti").read_bytes()
    offs = [int(v) for v in re.findall(rb'offset="(\d+)"', text)]
    assert offs == [0, 4 * nx * ny]
    output.save_vtk(tmp_path / "fixed", 1, rho, u)
    text = (tmp_path / "fixed" / "vtk" / "sim_data_000001.vti").read_bytes()
    assert [int(v) for v in re.findall(rb'offset="(\d+)"', text)] == [0, 4 * nx * ny + 8]


@pytest.mark.parametrize("dim", [2, 3])
def test_raw_dumps_read_as_the_reference_plotter_does(tmp_path, dim):
    shape = (9, 11) if dim == 2 else (4, 9, 11)
    rho, u = fields(shape, dim, 4)
    output.save_macroscopics(tmp_path, 300, rho, u)
    r = np.fromfile(tmp_path / "density" / "density_300.bin", dtype=np.float32)
    v = np.fromfile(tmp_path / "velocity" / "velocity_300.bin", dtype=np.float32)
    assert np.array_equal(r.reshape(shape), rho)
    assert np.array_equal(v.reshape(shape + (dim,)), u)  # the "aos" layout of visualize.py


def test_midplane_slice(tmp_path):
    nz, ny, nx = 7, 5, 8
    rho, u = fields((nz, ny, nx), 3, 5)
    output.save_midplane_slice(tmp_path, 100, rho, u)
    r = np.fromfile(tmp_path / "midplane_density" / "density_100.bin", dtype=np.float32).reshape(ny, nx)
    v = np.fromfile(tmp_path / "midplane_velocity" / "velocity_100.bin", dtype=np.float32).reshape((ny, nx, 3))
    assert np.array_equal(r, rho[nz // 2]) and np.array_equal(v, u[nz // 2])


def test_writers_reject_bad_arguments(tmp_path):
    lib = A.load_library()
    x = np.zeros(4, np.float32)
    assert lib.lbm_write_vti(bytes(tmp_path), 0, 2, 2, 2, 2, x.ctypes.data, x.ctypes.data, 0) == A.LBM_E_INVALID
    assert lib.lbm_write_vti(bytes(tmp_path), 0, 0, 2, 1, 2, x.ctypes.data, x.ctypes.data, 0) == A.LBM_E_INVALID
    assert lib.lbm_write_macros(bytes(tmp_path), -1, 4, 2, x.ctypes.data, x.ctypes.data) == A.LBM_E_INVALID
    assert lib.lbm_write_midplane(None, 0, 2, 2, 1, x.ctypes.data, x.ctypes.data) == A.LBM_E_INVALID
    (tmp_path / "file").write_text("x")  # a file where a directory should go
    with pytest.raises(OSError):
        output.save_macroscopics(tmp_path / "file", 0, np.ones((2, 2)), np.zeros((2, 2, 2)))
