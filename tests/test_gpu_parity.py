"""GPU parity: liblbm_hip.so on the MI355X against the oracle, through the C ABI.

Bar: FP32 populations within 1e-5 relative of the FP32 oracle after N steps (the
tolerance the build targets); reduced storages within one storage quantum."""
import numpy as np
import pytest

from helpers import layout_case, max_rel, run_gpu, run_oracle
from lbm_amd import _abi as A
from lbm_amd import scenarios as sc

pytestmark = pytest.mark.gpu

LAYOUTS = ["periodic", "channel", "cavity_mw", "cavity_reg", "inout", "zouhe_top", "open", "cyl_outlet"]
OPS = [(A.LBM_BGK, 0), (A.LBM_MRT, 0), (A.LBM_CM, 0), (A.LBM_CM, 1)]


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("op", OPS, ids=["bgk", "mrt", "cm", "cm_optimal"])
def test_gpu_matches_oracle_fp32(gpu_available, layout, op):
    coll, adapter = op
    print(f"case layout={layout} op={op} 37x21 x 12 steps", flush=True)
    case = layout_case(layout, 37, 21, coll, adapter=adapter)
    fg, rg, ug, s = run_gpu(case, 12)
    fo, ro, uo, _ = run_oracle(case, 12, 32)
    ef, er, eu = max_rel(fg, fo), max_rel(rg, ro), float(np.max(np.abs(ug - uo)))
    assert ef <= 1e-5, ef
    assert er <= 1e-5, er
    assert eu <= 1e-6, eu
    s.close()


@pytest.mark.parametrize("storage", [A.LBM_F16_SHIFTED, A.LBM_BF16_SHIFTED])
@pytest.mark.parametrize("coll", [A.LBM_BGK, A.LBM_MRT, A.LBM_CM])
def test_gpu_reduced_storage(gpu_available, storage, coll):
    print(f"case storage={storage} coll={coll}", flush=True)
    case = layout_case("cavity_mw", 40, 24, coll, storage=storage)
    fg, *_ = run_gpu(case, 10)
    fo, *_ = run_oracle(case, 10, 32)
    quantum = 2.0 ** -11 if storage == A.LBM_F16_SHIFTED else 2.0 ** -8
    # FP32 rounding differences (FMA contraction on the GPU) occasionally flip the rounding
    # of a 16-bit deviation (|d| < 0.05): bound = a few storage quanta.
    err = float(np.max(np.abs(fg - fo)))
    assert err <= 4 * 0.05 * quantum, err


@pytest.mark.parametrize("storage,shape", [(A.LBM_F32, (2048, 1400)), (A.LBM_F32, (4096, 2100)),
                                           (A.LBM_F16_SHIFTED, (2560, 2304))],
                         ids=["f32_2048x1400", "f32_4096x2100", "f16_2560x2304"])
def test_gpu_hbm_sized_grid_launch_shapes(gpu_available, storage, shape):
    # Population buffers above 192 MB take the HBM launch shapes (non-temporal stores; FP32
    # grids above 2^23 nodes also 4 nodes per thread); the small grids above take the
    # cache-resident shape (1 node per thread FP32, 2 for 16-bit storage).
    nx, ny = shape
    assert 2 * 9 * nx * ny * (4 if storage == A.LBM_F32 else 2) > 192e6
    print(f"case hbm-sized cavity {nx}x{ny} storage={storage} x 3 steps", flush=True)
    case = layout_case("cavity_mw", nx, ny, A.LBM_MRT, storage=storage)
    fg, rg, ug, s = run_gpu(case, 3)
    s.close()
    fo, ro, uo, _ = run_oracle(case, 3, 32)
    if storage == A.LBM_F32:
        assert max_rel(fg, fo) <= 1e-5
        assert float(np.max(np.abs(ug - uo))) <= 1e-6
    else:
        assert float(np.max(np.abs(fg - fo))) <= 4 * 0.05 * 2.0 ** -11


@pytest.mark.parametrize("quirk", [A.LBM_QUIRK_STALE_REST, A.LBM_QUIRK_ZOUHE_RHO, A.LBM_QUIRK_MRT_FORCE_ROWS])
def test_gpu_quirk_modes(gpu_available, quirk):
    coll = A.LBM_MRT if quirk == A.LBM_QUIRK_MRT_FORCE_ROWS else A.LBM_BGK
    case = layout_case("inout", 33, 17, coll, quirks=quirk)
    fg, *_ = run_gpu(case, 8)
    fo, *_ = run_oracle(case, 8, 32)
    assert max_rel(fg, fo) <= 1e-5


@pytest.mark.parametrize("shape", [(3, 3), (5, 4), (9, 5), (31, 6), (67, 3), (130, 66)])
def test_gpu_ragged_and_tiny_grids(gpu_available, shape):
    for layout in ["periodic", "cavity_mw", "open", "inout"]:
        print(f"case {layout} {shape}", flush=True)
        case = layout_case(layout, *shape, A.LBM_MRT)
        fg, *_ = run_gpu(case, 6)
        fo, *_ = run_oracle(case, 6, 32)
        assert max_rel(fg, fo) <= 1e-5, (layout, shape)


def test_gpu_taylor_green_128_against_oracle_and_analytic(gpu_available):
    """Taylor-Green 128^2, BGK, nu=0.01, u0=0.04, for 1000 steps."""
    case = sc.taylor_green(128, nu=0.01, u0=0.04)
    fg, rg, ug, s = run_gpu(case, 1000)
    fo, ro, uo, _ = run_oracle(case, 1000, 32)
    f64, _, u64, _ = run_oracle(case, 1000, 64)
    # Over 1000 steps FP32 rounding alone moves the FP32 oracle 1.9e-5 (max relative) from
    # the FP64 oracle, so 1e-5 is a per-step bar (the layout tests above, 6 steps). Here the
    # GPU must sit inside that rounding envelope and be as close to FP64 as the FP32 oracle.
    envelope = max_rel(fo, f64)
    assert max_rel(fg, fo) <= max(1e-5, envelope), (max_rel(fg, fo), envelope)
    assert max_rel(fg, f64) <= 1.25 * envelope, (max_rel(fg, f64), envelope)
    e_gpu = sc.tg_error(ug, 0.01, 0.04, 1000)
    e_64 = sc.tg_error(u64, 0.01, 0.04, 1000)
    assert abs(e_gpu - e_64) <= 1e-3 * max(e_64, 1.0)
    assert e_gpu < 1.0  # percent; analytic decay exp(-2 nu k^2 t)


def test_gpu_mass_conservation_periodic(gpu_available):
    case = layout_case("periodic", 256, 128, A.LBM_MRT, forced=True)
    from lbm_amd import Solver

    s = Solver(case.config)
    case.apply(s)
    s.step(1)
    m0 = s.reduce(A.LBM_REDUCE_MASS)[0]
    s.step(200)
    m1 = s.reduce(A.LBM_REDUCE_MASS)[0]
    assert abs(m1 - m0) / m0 < 1e-6
    # momentum grows by exactly F per node per step (Guo forcing)
    p0 = s.reduce(A.LBM_REDUCE_MOMENTUM)
    s.step(100)
    p1 = s.reduce(A.LBM_REDUCE_MOMENTUM)
    N = case.config.N
    assert abs((p1[0] - p0[0]) - 100 * N * 2e-5) < 1e-3 * 100 * N * 2e-5
    s.close()


def test_gpu_errors_are_reported(gpu_available):
    from lbm_amd import Config, LbmError, Solver

    with pytest.raises(LbmError):
        Solver(Config(nx=2, ny=8))
    s = Solver(Config(nx=8, ny=8))
    with pytest.raises(LbmError):
        s.step(1)  # not initialised
    flags = np.full((8, 8), A.GUO_PRESSURE_OUTLET, np.uint8)
    with pytest.raises(LbmError):
        s.set_flags(flags)
    s.close()


@pytest.mark.parametrize("storage", [A.LBM_F32, A.LBM_F16_SHIFTED])
@pytest.mark.parametrize("layout", ["periodic", "channel", "cavity_mw", "open", "inout"])
@pytest.mark.parametrize("coll", [A.LBM_BGK, A.LBM_MRT, A.LBM_CM])
def test_gpu_two_step_launches_match_one_step(gpu_available, layout, coll, storage):
    # steps_per_launch = 0 runs pairs of steps in one launch (state t+1 in LDS) where the
    # configuration allows; 1 forces one step per launch. 150x70 spans several tiles in x
    # and y with partial tiles; 9 steps = four pairs and one single step.
    print(f"case two-step {layout} coll={coll} storage={storage}", flush=True)
    case = layout_case(layout, 150, 70, coll, storage=storage)
    case.config.steps_per_launch = 1
    f1, r1, u1, s1 = run_gpu(case, 9)
    s1.close()
    case.config.steps_per_launch = 0
    f2, r2, u2, s2 = run_gpu(case, 9)
    s2.close()
    # The host emulation shows the two launch forms are the same algorithm bit for bit; on
    # the GPU the compiler contracts multiply-adds differently in the two kernels, so they
    # agree to FP32 rounding (FP16: a rounding flip of a few deviations), and both meet the
    # oracle bar.
    fo, *_ = run_oracle(case, 9, 32)
    if storage == A.LBM_F32:
        assert max_rel(f2, f1) <= 1e-5 and float(np.max(np.abs(u2 - u1))) <= 1e-6
        assert max_rel(f2, fo) <= 1e-5
    else:
        assert float(np.max(np.abs(f2 - f1))) <= 4 * 0.05 * 2.0 ** -11
        assert float(np.max(np.abs(f2 - fo))) <= 4 * 0.05 * 2.0 ** -11


@pytest.mark.parametrize("layout", ["periodic", "open", "channel"])
def test_gpu_two_step_strips_on_hbm_sized_grid(gpu_available, layout):
    # Above 192 MB of populations two-step launches march 62-column strips (phase-A rows in
    # an LDS ring, next rows prefetched); 2600x2100 FP32 = 393 MB. 5 steps = two strip
    # launches and a single step.
    print(f"case strips {layout} 2600x2100", flush=True)
    case = layout_case(layout, 2600, 2100, A.LBM_MRT)
    case.config.steps_per_launch = 1
    f1, _, u1, s1 = run_gpu(case, 5)
    s1.close()
    case.config.steps_per_launch = 0
    f2, _, u2, s2 = run_gpu(case, 5)
    assert s2.steps_per_launch == 2
    s2.close()
    assert max_rel(f2, f1) <= 1e-5 and float(np.max(np.abs(u2 - u1))) <= 1e-6
    fo, *_ = run_oracle(case, 5, 32)
    assert max_rel(f2, fo) <= 1e-5


@pytest.mark.parametrize("storage", [A.LBM_F32, A.LBM_F16_SHIFTED])
@pytest.mark.parametrize("mode", [A.LBM_ADAPTER_SAME_STEP, A.LBM_ADAPTER_LAGGED])
def test_gpu_optimal_adapter_on_a_uniform_equilibrium(gpu_available, mode, storage):
    """The adapter path on the GPU against its closed form on a uniform equilibrium state
    (periodic, no force): LBM_REDUCE_MOMENT_AVG = (rho0, rho0 |u0|, |Pi(rho0, u0)|), and the
    step relaxes moments 6..8 with s = 1 / (3 x 0.00141 + 0.5) (adapters.cuh:48-111)."""
    import math

    from helpers import THETA, cm_collide_uniform
    from lbm_amd.solver import Config, Solver

    rho0, ux, uy = 1.03, 0.04, -0.025
    S = (0.0, 0.0, 0.0, 1.0, 1.7, 1.7, 1.0, 1.0, 1.0)
    ny, nx = 40, 70
    cfg = Config(nx=nx, ny=ny, collision=A.LBM_CM, adapter=A.LBM_OPTIMAL_ADAPTER, adapter_mode=mode, omega=1.7, S=S,
                 periodic_x=True, periodic_y=True, storage=storage)
    print(f"case adapter uniform mode={mode} storage={storage}", flush=True)
    s = Solver(cfg)
    s.init_from_macros(np.full((ny, nx), rho0, np.float32), np.tile(np.float32([ux, uy]), (ny, nx, 1)))
    s.step(1)
    pi_cf = math.sqrt((rho0 / 3 + rho0 * ux * ux) ** 2 + 2 * (rho0 * ux * uy) ** 2 + (rho0 / 3 + rho0 * uy * uy) ** 2)
    avg = s.reduce(A.LBM_REDUCE_MOMENT_AVG)
    q = 1e-6 if storage == A.LBM_F32 else 4 * 2.0 ** -11 * 0.05  # FP16 deviations: a few quanta
    np.testing.assert_allclose(avg, [rho0, rho0 * math.hypot(ux, uy), pi_cf], rtol=1e-6 if storage == A.LBM_F32 else 1e-3)
    want = cm_collide_uniform(rho0, ux, uy, S, 1.0 / (3.0 * sum(THETA) + 0.5))
    f = s.get_pops()
    assert float(np.max(np.abs(f - want[:, None, None]))) <= q
    s.close()


def test_gpu_reductions_against_host_sums(gpu_available):
    """LBM_REDUCE_MASS / MOMENTUM / L2_VS_REF / ROW_MEAN_UX (FP64 device reductions of the
    stored fields) against FP64 sums of the same fields read back to the host."""
    print("case reductions 300x77 channel", flush=True)
    case = layout_case("channel", 300, 77, A.LBM_MRT)
    _, rho, u, s = run_gpu(case, 9)
    r64, u64 = rho.astype(np.float64), u.astype(np.float64)
    np.testing.assert_allclose(s.reduce(A.LBM_REDUCE_MASS)[0], r64.sum(), rtol=1e-12)
    mom = s.reduce(A.LBM_REDUCE_MOMENTUM)[:2]
    np.testing.assert_allclose(mom, [(r64 * u64[..., 0]).sum(), (r64 * u64[..., 1]).sum()], rtol=1e-9, atol=1e-12)
    rng = np.random.default_rng(11)
    ref = (u + 0.01 * rng.standard_normal(u.shape)).astype(np.float32)
    l2 = s.reduce(A.LBM_REDUCE_L2_VS_REF, ref=ref)[:2]
    ref64 = ref.astype(np.float64)
    np.testing.assert_allclose(l2, [((u64 - ref64) ** 2).sum(), (ref64 ** 2).sum()], rtol=1e-9)
    rows = s.reduce(A.LBM_REDUCE_ROW_MEAN_UX)
    np.testing.assert_allclose(rows, u64[..., 0].mean(axis=1), rtol=1e-9, atol=1e-15)
    s.close()
