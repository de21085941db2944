"""The C++ drop-in route on the GPU: the lbm_run drivers (a scenario header + LBM<2> from
cuda-lbm_amd/compat over liblbm_hip.so, the role of src/main.cu) run, validate and write the
reference's output formats (save_macroscopics .bin dumps, save_vtk .vti)."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from lbm_amd import scenarios as sc

pytestmark = pytest.mark.gpu

APPS = Path(__file__).resolve().parents[1] / "cuda-lbm_amd" / "apps"


def run_app(name, steps, out):
    exe = APPS / name
    if not exe.exists():
        pytest.fail(f"{exe} not built (make -C cuda-lbm_amd)")
    r = subprocess.run([str(exe), str(steps), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    print(r.stdout, flush=True)
    return r.stdout


def read_vti_velocity(path, n):
    """(Density, Velocity [n][3]) of the appended-raw .vti, read as VTK does: each DataArray
    at its declared offset after the '_' marker, a UInt64 byte count, then the values."""
    from test_output_files import read_vti

    arrays = read_vti(path)
    return arrays["Density"], arrays["Velocity"].reshape(n, 3)


def test_taylor_green_driver_validates_and_writes_outputs(gpu_available, tmp_path):
    out = run_app("lbm_run_tg128", 1000, tmp_path)
    err = float(re.search(r"validation error ([0-9.eE+-]+) %", out).group(1))
    n = 128 * 128
    rho = np.fromfile(tmp_path / "density" / "density_1000.bin", np.float32)
    u = np.fromfile(tmp_path / "velocity" / "velocity_1000.bin", np.float32)
    assert rho.size == n and u.size == 2 * n
    u = u.reshape(128, 128, 2)  # AoS, the reference's "aos" reader layout
    # the driver's TaylorGreenScenario::compute_error and the same formula on the dump agree
    assert abs(sc.tg_error(u, 0.01, 0.04, 1000) - err) <= 1e-4 * max(err, 1e-3)
    assert err < 1.0
    vrho, vel = read_vti_velocity(tmp_path / "vtk" / "sim_data_001000.vti", n)
    assert np.array_equal(vrho, rho) and np.array_equal(vel[:, :2], u.reshape(n, 2)) and not vel[:, 2].any()


def test_lid_driven_driver_runs_and_writes_outputs(gpu_available, tmp_path):
    out = run_app("lbm_run_ldc256", 400, tmp_path)
    assert "centre u_x(y)" in out
    u = np.fromfile(tmp_path / "velocity" / "velocity_400.bin", np.float32).reshape(256, 256, 2)
    rho = np.fromfile(tmp_path / "density" / "density_400.bin", np.float32).reshape(256, 256)
    assert np.all(np.isfinite(u)) and np.all(np.isfinite(rho))
    # the lid drags the fluid below it in +x; the bulk stays near rest early on
    assert u[-2, 128, 0] > 0.01 and abs(u[128, 128, 0]) < 0.05
