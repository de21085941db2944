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


def test_unchanged_reference_taylor_green_header_against_the_oracle(gpu_available, tmp_path):
    """The reference's own src/scenarios/taylorGreen/taylorGreenScenario.cuh, included unchanged
    from where it lies (built here into apps/lbm_run_ref_tg128 next to the library), run through
    compat/ on the GPU: BGK, nu = 1/6 (omega = 1), u_max = 0.04, FLUID everywhere on a
    non-periodic 128^2 grid (PERIODIC_X/Y are never defined there). Its initial fields (the dump
    at step 0) seed an FP32 oracle run of the same configuration, whose fields after 100 steps
    the driver's dumps match, and TaylorGreenScenario::compute_error as printed by the driver
    equals the validator's formula on the dumped field."""
    import math

    import oracle as orc
    from lbm_amd import _abi as A
    from lbm_amd.solver import Config

    exe = APPS / "lbm_run_ref_tg128"
    if not exe.exists():
        pytest.skip("apps/lbm_run_ref_tg128 is built only where the reference tree is present")
    n = 128
    out0 = run_app("lbm_run_ref_tg128", 0, tmp_path / "t0")
    assert "TaylorGreen" in out0
    rho0 = np.fromfile(tmp_path / "t0" / "density" / "density_0.bin", np.float32).reshape(n, n)
    u0 = np.fromfile(tmp_path / "t0" / "velocity" / "velocity_0.bin", np.float32).reshape(n, n, 2)
    # the reference's init: rho = 1 - rho0 u0^2 / (4 cs2) (cos 2kx + cos 2ky), u = u0 (cos kx sin ky, -sin kx cos ky)
    assert abs(float(np.abs(u0).max()) - 0.04) < 1e-3
    steps = 100
    out = run_app("lbm_run_ref_tg128", steps, tmp_path / "t1")
    err = float(re.search(r"validation error ([0-9.eE+-]+) %", out).group(1))
    rho = np.fromfile(tmp_path / "t1" / "density" / f"density_{steps}.bin", np.float32).reshape(n, n)
    u = np.fromfile(tmp_path / "t1" / "velocity" / f"velocity_{steps}.bin", np.float32).reshape(n, n, 2)
    tau = 3.0 * np.float32(1.0 / 6.0) + 0.5
    cfg = Config(nx=n, ny=n, collision=A.LBM_BGK, omega=float(np.float32(1.0) / np.float32(tau)))
    o = orc.Oracle(cfg.to_c(), precision=32)
    o.init_from_macros(rho0, u0)
    o.step(steps)
    r_o, u_o = o.get_macros()
    assert float(np.max(np.abs(rho - r_o))) < 1e-5
    assert float(np.max(np.abs(u - u_o))) < 1e-6
    # compute_error (taylorGreenScenario.cuh:59-88) on the dumped field, from the validator's analytic
    # decay at nu = 1/6 over the same steps
    want = sc.tg_error(u, 1.0 / 6.0, 0.04, steps)
    assert abs(want - err) <= 1e-3 * max(err, 1e-3), (want, err)
    assert math.isfinite(err)
