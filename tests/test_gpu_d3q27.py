"""D3Q27 on the GPU through the C ABI of include/lbm3d_hip.h, against the FP32 oracle
(oracle/lbm_oracle3d.hpp) and size-independent properties."""
import numpy as np
import pytest

import oracle as O
from lbm_amd import _abi as A
from lbm_amd.solver import LbmError
from lbm_amd.solver3d import Config3D, Solver3D
from test_emu3d_parity import (CASES, CM_CASES, GUO_CASES, IBM3_CASES, MRT_CASES, NEW3_CASES, REG_CASES, ZG_CASES, box_flags,
                               cm_case, guo_case, ibm3_case, initial_state, make_cfg, mrt_case, new3_case, oracle_run,
                               oracle_run_ibm, reg_case, zg_case)

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
        fl[:, 3, 4] = A.REGULARIZED_OUTLET  # regularized outlets are built at x = NX-1 only
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


@pytest.mark.parametrize("case", range(len(NEW3_CASES)))
def test_gpu_outlet_extrapolation_and_zg_faces_match_oracle(gpu_available, case):
    """REGULARIZED_OUTLET, EXTRAPOLATED_CORNER_EDGE (with and without its velocity quirk) and
    ZG_OUTFLOW on every face (nodes whose neighbour's stale slots another wavefront rewrites run
    in pre-launches) against the oracle's functors on the post-stream snapshot."""
    c, f0, fl = new3_case(case)
    print(f"case d3q27 new flags {case} {NEW3_CASES[case][0]} {NEW3_CASES[case][3]}", flush=True)
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


def test_gpu_flags_reading_each_other_in_a_cycle_are_rejected(gpu_available):
    """Every node of a 3x3x3 box ZG: the centre plane's neighbours form no cycle, so the set-up
    succeeds; a Guo outlet whose neighbour is a Guo inlet on a 3-wide axis reads and is read
    within one wavefront, also fine. Unsupported flags still fail loudly."""
    c = make_cfg((3, 3, 3), (0, 0, 0), 1.0)
    with Solver3D(gpu_config(c)) as s:
        s.set_flags(np.full((3, 3, 3), A.ZG_OUTFLOW, np.uint8))
        with pytest.raises(LbmError):
            s.set_flags(np.full((3, 3, 3), A.ZOU_HE_LEFT, np.uint8))


def run_ibm3_gpu(c, f0, fl, force, xyz, ds, iters, steps, force_after=0):
    """Solver3D with markers; with force_after > 0 the force field is set after that many steps."""
    with Solver3D(gpu_config(c)) as s:
        s.set_pops(f0)
        if fl is not None:
            s.set_flags(fl)
        if force is not None and not force_after:
            s.set_force_field(force.astype(np.float32))
        s.ibm_set_markers(xyz, ds, iters)
        if force_after:
            s.step(force_after)
            s.set_force_field(force.astype(np.float32))
        s.step(steps - force_after)
        g = s.get_pops()
        rho, u = s.get_macros()
        lag, tot = s.ibm_get_forces()
    return g, rho, u, lag, tot


@pytest.mark.parametrize("case", range(len(IBM3_CASES)))
def test_gpu_ibm3_matches_oracle(gpu_available, case):
    """3-D multi-direct forcing (IBMManager<3>::multi_direct) on the GPU: the band gather, the
    interpolation and spreading passes and the force array read by the step, against the
    oracle's whole-grid forcing; Lagrangian forces and their total as well."""
    c, f0, fl, force, xyz, ds, iters = ibm3_case(case)
    print(f"case d3q27 ibm {case} {IBM3_CASES[case][0]} {IBM3_CASES[case][3]}", flush=True)
    f0 = f0.astype(np.float32)
    steps = 6
    o = oracle_run_ibm(c, f0, fl, force, xyz, ds, iters, steps)
    g, rho, u, lag, tot = run_ibm3_gpu(c, f0, fl, force, xyz, ds, iters, steps)
    assert np.abs(g - o.get_pops()).max() < 1e-5
    r_rho, r_u = o.get_macros()
    assert np.abs(rho - r_rho).max() < 1e-5
    assert np.abs(u - r_u).max() < 1e-6
    o_lag, o_tot = o.ibm_get_forces()
    assert np.abs(o_lag).max() > 1e-3
    assert np.abs(lag - o_lag).max() < 1e-5 * max(1.0, float(np.abs(o_lag).max()) * 10)
    assert np.abs(tot - o_tot).max() < 1e-5 * max(1.0, float(np.abs(o_tot).max()))


def test_gpu_ibm3_force_field_set_after_the_markers(gpu_available):
    """A force field set after the markers and two steps replaces the uniform force outside the
    band (the step's force array is refilled) and under it."""
    c, f0, fl, force, xyz, ds, iters = ibm3_case(2)
    f0 = f0.astype(np.float32)
    o = O.Oracle3D(c, 32)
    o.set_pops(f0)
    o.ibm_set_markers(xyz.astype(np.float64), ds.astype(np.float64), iters)
    o.step(2)
    o.set_force_field(force.astype(np.float32))
    o.step(4)
    g, rho, u, lag, tot = run_ibm3_gpu(c, f0, fl, force, xyz, ds, iters, 6, force_after=2)
    assert np.abs(g - o.get_pops()).max() < 1e-5
    assert np.abs(u - o.get_macros()[1]).max() < 1e-6


@pytest.mark.parametrize("storage,eps", [(A.LBM_F16_SHIFTED, 2.0 ** -11), (A.LBM_BF16_SHIFTED, 2.0 ** -8)])
def test_gpu_ibm3_reduced_storage(gpu_available, storage, eps):
    c, f0, fl, force, xyz, ds, iters = ibm3_case(1)
    c.storage = storage
    f0 = f0.astype(np.float32)
    o = oracle_run_ibm(c, f0, fl, force, xyz, ds, iters, 6)
    g, rho, u, lag, tot = run_ibm3_gpu(c, f0, fl, force, xyz, ds, iters, 6)
    ref = o.get_pops()
    dev = np.abs(ref - W[:, None, None, None]).max()
    assert np.abs(g - ref).max() <= 4 * eps * dev
    assert np.abs(lag - o.ibm_get_forces()[0]).max() <= 8 * eps * float(np.abs(o.ibm_get_forces()[0]).max())


def test_gpu_ibm3_markers_removed_and_errors(gpu_available):
    """n = 0 removes the body: the following steps equal a run that never had one. Forces
    without markers, non-finite coordinates and a body no stencil of which touches the grid are
    reported."""
    c, f0, fl, force, xyz, ds, iters = ibm3_case(0)
    f0 = f0.astype(np.float32)
    with Solver3D(gpu_config(c)) as s, Solver3D(gpu_config(c)) as t:
        with pytest.raises(LbmError) as e:
            s.ibm_get_forces()
        assert e.value.code == A.LBM_E_STATE
        s.set_pops(f0)
        t.set_pops(f0)
        s.ibm_set_markers(xyz, ds, iters)
        s.step(1)
        s.ibm_set_markers(np.zeros((0, 3), np.float32))
        g1 = s.get_pops()
        t.set_pops(g1)
        s.step(3)
        t.step(3)
        assert np.array_equal(s.get_pops(), t.get_pops())
        bad = xyz.copy()
        bad[3, 1] = np.nan
        with pytest.raises(LbmError) as e:
            s.ibm_set_markers(bad, ds)
        assert e.value.code == A.LBM_E_INVALID
        with pytest.raises(LbmError) as e:
            s.ibm_set_markers(xyz + 100.0, ds)
        assert e.value.code == A.LBM_E_INVALID


@pytest.mark.parametrize("case", range(len(IBM3_CASES)))
def test_gpu_ibm3_split_step_equals_serial_step(gpu_available, case):
    """An IBM step whose band box is clear of non-periodic faces and neighbour-reading flags
    steps the nodes outside the box on a second stream beside the IBM pass, then the box: the
    same node updates as the serial step (LBM_IBM3_SPLIT=0), bit for bit, in every case (the
    cases whose box touches a face or a Guo flag run serially either way)."""
    import os

    c, f0, fl, force, xyz, ds, iters = ibm3_case(case)
    print(f"case d3q27 ibm split vs serial {case} {IBM3_CASES[case][0]}", flush=True)
    out = {}
    for mode in ("split", "serial"):
        if mode == "serial":
            os.environ["LBM_IBM3_SPLIT"] = "0"
        try:
            out[mode] = run_ibm3_gpu(c, f0, fl, force, xyz, ds, iters, 7)
        finally:
            os.environ.pop("LBM_IBM3_SPLIT", None)
    for a, b in zip(out["split"], out["serial"]):
        assert np.array_equal(a, b)


def test_gpu_ibm3_split_step_in_the_wind_tunnel(gpu_available):
    """A sphere of markers in a CM<3> channel with Guo inlet/outlet and bounce-back walls (the
    band box clear of them, so the split applies): split and serial steps agree bit for bit
    over 40 steps, with the fields stored midway."""
    import os

    nx, ny, nz = 48, 24, 24
    k = np.arange(120)
    th, ph = np.arccos(1 - 2 * (k + 0.5) / 120), np.pi * (1 + 5 ** 0.5) * k
    xyz = np.stack([16 + 4 * np.sin(th) * np.cos(ph), 12 + 4 * np.sin(th) * np.sin(ph), 12 + 4 * np.cos(th)],
                   -1).astype(np.float32)
    fl = np.zeros((nz, ny, nx), np.uint8)
    fl[:, 0, :] = fl[:, -1, :] = A.BOUNCE_BACK
    fl[0], fl[-1] = A.BOUNCE_BACK, A.BOUNCE_BACK
    fl[:, :, 0] = A.GUO_VELOCITY_INLET
    fl[:, :, -1] = A.GUO_PRESSURE_OUTLET
    c = make_cfg((nx, ny, nz), (0, 0, 0), 1.0 / (3 * 0.05 * ny / 100.0 + 0.5), collision=A.LBM_CM)
    c.u_inlet = 0.05
    f0 = np.empty((27, nz, ny, nx), np.float32)
    f0[:] = O.equilibrium3d(1.0, 0.05, 0.0, 0.0)[:, None, None, None]
    ds = np.full(120, 4 * np.pi * 16.0 / 120, np.float32)
    print("case d3q27 ibm split vs serial wind tunnel 48x24x24", flush=True)
    out = {}
    for mode in ("split", "serial"):
        if mode == "serial":
            os.environ["LBM_IBM3_SPLIT"] = "0"
        try:
            with Solver3D(gpu_config(c)) as s:
                s.set_pops(f0)
                s.set_flags(fl)
                s.ibm_set_markers(xyz, ds, 3)
                snaps = []
                for n in (17, 23):
                    s.step(n)
                    rho, u = s.get_macros()
                    snaps += [s.get_pops(), rho, u, s.ibm_get_forces()[0]]
            out[mode] = snaps
        finally:
            os.environ.pop("LBM_IBM3_SPLIT", None)
    for a, b in zip(out["split"], out["serial"]):
        assert np.array_equal(a, b)


def test_gpu_sphere_body_from_a_config_file_in_the_wind_tunnel(gpu_available, tmp_path):
    """The reference's wind-tunnel body path at reduced size: a UV-sphere OBJ scaled to D = 8
    and placed at (16, 12, 12) by a body file with 96 target points (Morton order), in a CM<3>
    channel with Guo inlet/outlet and walls. The GPU matches the oracle over 6 steps; after 300
    steps the force the markers apply to the fluid points upstream (drag)."""
    from lbm_amd import bodies as Bd

    v, f = Bd.uv_sphere(24, 24)
    Bd.Body.from_mesh(v, f).collect_file(tmp_path / "sphere.obj")
    (tmp_path / "sphere.ini").write_text(
        "[General]\nobj_filepath = sphere.obj\n[Transformations]\nscale_overall_size = 8.0\n"
        "position_anchor_x = 16.0\nposition_anchor_y = 12.0\nposition_anchor_z = 12.0\n"
        "[Sampling]\ntype = target_points\ntarget_points = 96\n")
    nx, ny, nz = 48, 24, 24
    with Bd.Body.from_config(tmp_path / "sphere.ini", (nx, ny, nz), base_dir=tmp_path) as b:
        xyz = b.markers
    assert xyz.shape == (96, 3)
    fl = np.zeros((nz, ny, nx), np.uint8)
    fl[:, 0, :] = fl[:, -1, :] = A.BOUNCE_BACK
    fl[0], fl[-1] = A.BOUNCE_BACK, A.BOUNCE_BACK
    fl[:, :, 0] = A.GUO_VELOCITY_INLET
    fl[:, :, -1] = A.GUO_PRESSURE_OUTLET
    c = make_cfg((nx, ny, nz), (0, 0, 0), 1.0 / (3 * 0.05 * ny / 100.0 + 0.5), collision=A.LBM_CM)
    c.u_inlet = 0.05
    f0 = np.empty((27, nz, ny, nx), np.float32)
    f0[:] = O.equilibrium3d(1.0, 0.05, 0.0, 0.0)[:, None, None, None]
    ds = np.full(96, np.pi * 8.0 ** 2 / 96, np.float32)
    o = oracle_run_ibm(c, f0, fl, None, xyz, ds, 3, 6)
    with Solver3D(gpu_config(c)) as s:
        s.set_pops(f0)
        s.set_flags(fl)
        s.ibm_set_markers(xyz, ds, 3)
        s.step(6)
        assert np.abs(s.get_pops() - o.get_pops()).max() < 1e-5
        lag, tot = s.ibm_get_forces()
        assert np.abs(lag - o.ibm_get_forces()[0]).max() < 1e-5
        s.step(294)
        lag, tot = s.ibm_get_forces()
        rho, u = s.get_macros()
    assert np.isfinite(u).all()
    assert tot[0] < 0 and abs(tot[0]) > 5 * max(abs(tot[1]), abs(tot[2]))
    # the wake: slower flow just behind the sphere than beside it
    assert u[12, 12, 22, 0] < u[12, 5, 22, 0]
