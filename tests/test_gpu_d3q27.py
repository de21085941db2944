"""D3Q27 on the GPU through the C ABI of include/lbm3d_hip.h, against the FP32 oracle
(oracle/lbm_oracle3d.hpp) and size-independent properties."""
import numpy as np
import pytest

import oracle as O
from lbm_amd import _abi as A
from lbm_amd.solver import LbmError
from lbm_amd.solver3d import Config3D, Solver3D
from test_emu3d_parity import (CASES, CM_CASES, GUO_CASES, MRT_CASES, REG_CASES, ZG_CASES, box_flags, cm_case, guo_case,
                               initial_state, make_cfg, mrt_case, oracle_run, reg_case, zg_case)

pytestmark = pytest.mark.gpu
W = O.tables3d()[0]


def gpu_config(c):
    return Config3D(c.nx, c.ny, c.nz, collision=c.collision, adapter=c.adapter, omega=c.omega, S=tuple(c.S),
                    storage=c.storage,
                    periodic_x=bool(c.periodic_x), periodic_y=bool(c.periodic_y), periodic_z=bool(c.periodic_z),
                    force=tuple(c.force), u_inlet=c.u_inlet, quirks=c.quirks)


@pytest.mark.parametrize("case", range(len(GUO_CASES)))
def test_gpu_guo_inlet_outlet_matches_oracle(gpu_available, case):
    c, f0, fl = guo_case(case)
    f0 = f0.astype(np.float32)
    o = oracle_run(c, f0, fl, None, 8)
    with Solver3D(gpu_config(c)) as s:
        s.set_pops(f0)
        s.set_flags(fl)
        s.step(8)
        g = s.get_pops()
        rho, u = s.get_macros()
    assert np.abs(g - o.get_pops()).max() < 1e-5
    r_rho, r_u = o.get_macros()
    assert np.abs(rho - r_rho).max() < 1e-5
    assert np.abs(u - r_u).max() < 1e-6


def test_gpu_wind_tunnel_runs_and_carries_the_inlet_flux(gpu_available):
    """The reference's default 3-D case (windTestScenario.cuh: 256x128x96, CM<3, NoAdapter>, Guo
    velocity inlet u = 0.05 at x = 0, pressure outlet at x = NX-1, bounce-back on y and z; no IBM
    body): 400 steps from the plug-flow initial state stay finite, the inlet plane keeps
    u_x = 0.05 on average and the flow stays a plug flow within 10 % away from the walls."""
    nx, ny, nz = 256, 128, 96
    nu = 0.05 * ny / 1000.0
    omega = 1.0 / (3.0 * nu + 0.5)
    fl = np.zeros((nz, ny, nx), np.uint8)
    fl[:, 0, :] = fl[:, -1, :] = A.BOUNCE_BACK
    fl[0], fl[-1] = A.BOUNCE_BACK, A.BOUNCE_BACK
    fl[:, :, 0] = A.GUO_VELOCITY_INLET
    fl[:, :, -1] = A.GUO_PRESSURE_OUTLET
    u0 = np.zeros((nz, ny, nx, 3), np.float32)
    u0[..., 0] = 0.05
    with Solver3D(Config3D(nx, ny, nz, collision=A.LBM_CM, omega=omega, u_inlet=0.05)) as s:
        s.set_flags(fl)
        s.init_from_macros(np.ones((nz, ny, nx), np.float32), u0)
        s.step(400)
        rho, u = s.get_macros()
    assert np.isfinite(u).all() and np.isfinite(rho).all()
    core = u[8:-8, 8:-8, 1:-1, 0]
    assert abs(core.mean() / 0.05 - 1) < 0.1
    assert abs(rho.mean() - 1) < 0.05


@pytest.mark.parametrize("case", range(len(CM_CASES)))
def test_gpu_cm3_matches_dense_oracle(gpu_available, case):
    c, f0, flags, force = cm_case(case)
    f0 = f0.astype(np.float32)
    o = oracle_run(c, f0, flags, force, 8)
    with Solver3D(gpu_config(c)) as s:
        s.set_pops(f0)
        s.set_flags(flags)
        s.set_force_field(None if force is None else force.astype(np.float32))
        s.step(8)
        g = s.get_pops()
        rho, u = s.get_macros()
    assert np.abs(g - o.get_pops()).max() < 1e-5
    r_rho, r_u = o.get_macros()
    assert np.abs(rho - r_rho).max() < 1e-5
    assert np.abs(u - r_u).max() < 1e-6


@pytest.mark.parametrize("case", range(len(MRT_CASES)))
def test_gpu_mrt3_matches_dense_oracle(gpu_available, case):
    c, f0, flags, force = mrt_case(case)
    f0 = f0.astype(np.float32)
    o = oracle_run(c, f0, flags, force, 8)
    with Solver3D(gpu_config(c)) as s:
        s.set_pops(f0)
        s.set_flags(flags)
        s.set_force_field(None if force is None else force.astype(np.float32))
        s.step(8)
        g = s.get_pops()
        rho, u = s.get_macros()
    assert np.abs(g - o.get_pops()).max() < 1e-5
    r_rho, r_u = o.get_macros()
    assert np.abs(rho - r_rho).max() < 1e-5
    assert np.abs(u - r_u).max() < 1e-6


@pytest.mark.parametrize("case", range(len(REG_CASES)))
def test_gpu_regularized_boundaries_match_oracle(gpu_available, case):
    c, f0, fl = reg_case(case)
    f0 = f0.astype(np.float32)
    o = oracle_run(c, f0, fl, None, 8)
    with Solver3D(gpu_config(c)) as s:
        s.set_pops(f0)
        s.set_flags(fl)
        s.step(8)
        g = s.get_pops()
        rho, u = s.get_macros()
    assert np.abs(g - o.get_pops()).max() < 1e-5
    r_rho, r_u = o.get_macros()
    assert np.abs(rho - r_rho).max() < 1e-5
    assert np.abs(u - r_u).max() < 1e-6


@pytest.mark.parametrize("case", range(len(ZG_CASES)))
def test_gpu_zg_outflow_matches_oracle(gpu_available, case):
    c, f0, fl = zg_case(case)
    f0 = f0.astype(np.float32)
    o = oracle_run(c, f0, fl, None, 8)
    with Solver3D(gpu_config(c)) as s:
        s.set_pops(f0)
        s.set_flags(fl)
        s.step(8)
        g = s.get_pops()
        rho, u = s.get_macros()
    assert np.abs(g - o.get_pops()).max() < 1e-5
    r_rho, r_u = o.get_macros()
    assert np.abs(rho - r_rho).max() < 1e-5
    assert np.abs(u - r_u).max() < 1e-6


@pytest.mark.parametrize("case", range(len(CASES)))
def test_gpu_d3q27_matches_fp32_oracle(gpu_available, case):
    shape, periodic, omega, faces, field, quirks = CASES[case]
    nx, ny, nz = shape
    c = make_cfg(shape, periodic, omega, force=(1e-5, -2e-5, 5e-6), quirks=quirks)
    f0 = initial_state(shape, case).astype(np.float32)
    flags = box_flags(shape, faces) if faces else None
    force = None
    if field:
        force = (1e-5 * np.random.default_rng(100 + case).standard_normal((nz, ny, nx, 3))).astype(np.float32)
    o = O.Oracle3D(c, 32)
    o.set_pops(f0)
    if flags is not None:
        o.set_flags(flags)
    if force is not None:
        o.set_force_field(force)
    o.step(8)
    with Solver3D(gpu_config(c)) as s:
        s.set_pops(f0)
        s.set_flags(flags)
        s.set_force_field(force)
        s.step(8)
        g = s.get_pops()
        rho, u = s.get_macros()
    assert np.abs(g - o.get_pops()).max() < 1e-5
    r_rho, r_u = o.get_macros()
    assert np.abs(rho - r_rho).max() < 1e-5
    assert np.abs(u - r_u).max() < 1e-6


@pytest.mark.parametrize("storage,eps", [(A.LBM_F16_SHIFTED, 2.0 ** -11), (A.LBM_BF16_SHIFTED, 2.0 ** -8)])
def test_gpu_d3q27_reduced_storage(gpu_available, storage, eps):
    shape = (70, 9, 6)
    c = make_cfg(shape, (1, 0, 1), 1.3, storage=storage, force=(1e-5, 0, 0))
    f0 = initial_state(shape, 7).astype(np.float32)
    flags = box_flags(shape, "y")
    o = O.Oracle3D(c, 32)
    o.set_pops(f0)
    o.set_flags(flags)
    o.step(6)
    with Solver3D(gpu_config(c)) as s:
        s.set_pops(f0)
        s.set_flags(flags)
        s.step(6)
        g = s.get_pops()
    ref = o.get_pops()
    dev = np.abs(ref - W[:, None, None, None]).max()
    assert np.abs(g - ref).max() <= 4 * eps * dev


def test_gpu_d3q27_uniform_force_equals_the_same_force_field(gpu_available):
    shape = (66, 7, 5)
    F = (3e-6, -1e-6, 2e-6)
    f0 = initial_state(shape, 3).astype(np.float32)
    cfg = Config3D(*shape, omega=1.6, periodic_x=True, periodic_z=True, force=F)
    flags = box_flags(shape, "y")
    with Solver3D(cfg) as a, Solver3D(Config3D(*shape, omega=1.6, periodic_x=True, periodic_z=True)) as b:
        for s in (a, b):
            s.set_pops(f0)
            s.set_flags(flags)
        b.set_force_field(np.broadcast_to(np.array(F, np.float32), (*shape[::-1], 3)).copy())
        a.step(5)
        b.step(5)
        assert np.array_equal(a.get_pops(), b.get_pops())


@pytest.mark.parametrize("storage", [A.LBM_F32, A.LBM_F16_SHIFTED])
def test_gpu_d3q27_large_grid_tiles_a_small_periodic_one(gpu_available, storage):
    """A fully periodic field with period 8 on a 344^3 grid (27 FP32 planes pass 4 GiB, so the
    step addresses them with 64-bit pointers) evolves exactly as its 8^3 period does (buffer
    resources): every node computes the same operations on the same values."""
    n_small, n_big = 8, 344
    k = 2 * np.pi / n_small
    z, y, x = np.meshgrid(np.arange(n_small), np.arange(n_small), np.arange(n_small), indexing="ij")
    u = np.stack([0.02 * np.sin(k * x) * np.cos(k * y), -0.02 * np.cos(k * x) * np.sin(k * y), 0.01 * np.sin(k * z)], -1)
    rho = 1 + 0.01 * np.cos(k * (x + 2 * y + 3 * z))
    kw = dict(omega=1.7, periodic_x=True, periodic_y=True, periodic_z=True, storage=storage, force=(1e-6, 0, 0))
    with Solver3D(Config3D(n_small, n_small, n_small, **kw)) as s:
        s.init_from_macros(rho.astype(np.float32), u.astype(np.float32))
        s.step(6)
        r_s, u_s = s.get_macros()
    reps = n_big // n_small
    with Solver3D(Config3D(n_big, n_big, n_big, **kw)) as b:
        b.init_from_macros(np.tile(rho, (reps,) * 3).astype(np.float32), np.tile(u, (reps, reps, reps, 1)).astype(np.float32))
        b.step(6)
        r_b, u_b = b.get_macros()
    assert np.array_equal(r_b, np.tile(r_s, (reps,) * 3))
    assert np.array_equal(u_b, np.tile(u_s, (reps, reps, reps, 1)))


def test_gpu_taylor_green_3d_kinetic_energy(gpu_available):
    """TaylorGreen3DInit (taylorGreen3DFunctors.cuh:21-46) on 64^3 at small amplitude: the GPU's
    kinetic energy follows the FP32 oracle and decays at the viscous rate exp(-6 nu k^2 t)."""
    n = 64
    omega = 1.5
    nu = (1 / omega - 0.5) / 3
    k = 2 * np.pi / n
    U = 0.01
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    u = np.stack([U * np.sin(k * x) * np.cos(k * y) * np.cos(k * z), -U * np.cos(k * x) * np.sin(k * y) * np.cos(k * z),
                  np.zeros_like(x, dtype=float)], -1).astype(np.float32)
    rho = np.ones((n, n, n), np.float32)
    cfg = Config3D(n, n, n, omega=omega, periodic_x=True, periodic_y=True, periodic_z=True)
    o = O.Oracle3D(cfg.to_c(), 32)
    o.init_from_macros(rho, u)
    with Solver3D(cfg) as s:
        s.init_from_macros(rho, u)
        s.step(40)
        o.step(40)
        e0 = s.kinetic_energy(None)
        r_o, u_o = o.get_macros()
        e0_o = 0.5 * (r_o * (u_o ** 2).sum(-1)).sum()
        assert abs(e0 / e0_o - 1) < 1e-5
        s.step(200)
        e1 = s.kinetic_energy(None)
        rho_g, u_g = s.get_macros()
        assert abs(e1 / (0.5 * (rho_g.astype(np.float64) * (u_g.astype(np.float64) ** 2).sum(-1)).sum()) - 1) < 1e-9
    rate = -np.log(e1 / e0) / 200
    assert abs(rate / (6 * nu * k * k) - 1) < 0.01


def test_gpu_mrt3_and_cm3_taylor_green_decay_and_f16_storage(gpu_available):
    """MRT<3> with S[4..8] = 1.5 and CM<3> with omega = 1.5 on 64^3: the kinetic energy decays at
    exp(-6 nu k^2 t), nu = (1/1.5 - 1/2)/3, in FP32 and in FP16 storage."""
    n, s48 = 64, 1.5
    nu = (1 / s48 - 0.5) / 3
    k = 2 * np.pi / n
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    U = 0.01
    u = np.stack([U * np.sin(k * x) * np.cos(k * y) * np.cos(k * z), -U * np.cos(k * x) * np.sin(k * y) * np.cos(k * z),
                  np.zeros_like(x, dtype=float)], -1).astype(np.float32)
    S = [1.0] * 27
    S[4:9] = [s48] * 5
    for coll, storage in ((A.LBM_MRT, A.LBM_F32), (A.LBM_MRT, A.LBM_F16_SHIFTED), (A.LBM_CM, A.LBM_F32),
                          (A.LBM_CM, A.LBM_F16_SHIFTED)):
        cfg = Config3D(n, n, n, collision=coll, omega=s48, S=tuple(S), storage=storage, periodic_x=True,
                       periodic_y=True, periodic_z=True)
        with Solver3D(cfg) as s:
            s.init_from_macros(np.ones((n, n, n), np.float32), u)
            s.step(40)
            e0 = s.kinetic_energy(None)
            s.step(200)
            e1 = s.kinetic_energy(None)
        assert abs(-np.log(e1 / e0) / 200 / (6 * nu * k * k) - 1) < 0.01


def test_gpu_d3q27_reductions_and_conservation(gpu_available):
    shape = (67, 10, 9)
    f0 = initial_state(shape, 11).astype(np.float32)
    with Solver3D(Config3D(*shape, omega=1.2, periodic_x=True, periodic_y=True, periodic_z=True)) as s:
        s.set_pops(f0)
        s.step(10)
        rho, u = s.get_macros()
        m = s.reduce(A.LBM3D_REDUCE_MASS)
        p = s.reduce(A.LBM3D_REDUCE_MOMENTUM)
        r64, u64 = rho.astype(np.float64), u.astype(np.float64)
        assert abs(m - r64.sum()) < 1e-9 * m
        assert np.allclose(p, (r64[..., None] * u64).sum((0, 1, 2)), rtol=1e-9, atol=1e-12)
        assert abs(m - f0.astype(np.float64).sum()) < 1e-6 * m


def test_gpu_d3q27_errors_are_reported(gpu_available):
    with pytest.raises(LbmError) as e:
        Solver3D(Config3D(8, 8, 8, collision=A.LBM_CM, adapter=A.LBM_OPTIMAL_ADAPTER))
    assert e.value.code == A.LBM_E_UNSUPPORTED
    with Solver3D(Config3D(8, 8, 8)) as s:
        fl = np.zeros((8, 8, 8), np.uint8)
        fl[:, :, 3] = A.GUO_VELOCITY_INLET  # inlets are built at x = 0 only
        with pytest.raises(LbmError) as e:
            s.set_flags(fl)
        assert e.value.code == A.LBM_E_UNSUPPORTED
    with pytest.raises(LbmError) as e:
        Solver3D(Config3D(2, 8, 8))
    assert e.value.code == A.LBM_E_INVALID
    with Solver3D(Config3D(8, 8, 8)) as s:
        with pytest.raises(LbmError) as e:
            s.step(1)
        assert e.value.code == A.LBM_E_STATE
        fl = np.zeros((8, 8, 8), np.uint8)
        fl[:, 3, 0] = A.ZG_OUTFLOW
        fl[:, 0, 4] = A.ZG_OUTFLOW  # ZG is built on the x faces only
        with pytest.raises(LbmError) as e:
            s.set_flags(fl)
        assert e.value.code == A.LBM_E_UNSUPPORTED
        fl = np.zeros((8, 8, 8), np.uint8)
        fl[0, 0, 0] = A.ZOU_HE_LEFT
        with pytest.raises(LbmError) as e:
            s.set_flags(fl)
        assert e.value.code == A.LBM_E_UNSUPPORTED
        with pytest.raises(LbmError) as e:
            s.reduce(A.LBM3D_REDUCE_MASS)
        assert e.value.code == A.LBM_E_STATE
