"""lbm_reduce_noalloc (the reductions with their partial sums in a fixed per-GPU array, for
callers that cannot allocate) against lbm_reduce through the C ABI: every kind, identical
results, on two calls in a row with the second on a smaller grid."""
import numpy as np
import pytest

from helpers import layout_case, run_gpu
from lbm_amd import _abi as A

pytestmark = pytest.mark.gpu

KINDS = [A.LBM_REDUCE_MASS, A.LBM_REDUCE_MOMENTUM, A.LBM_REDUCE_L2_VS_REF, A.LBM_REDUCE_ROW_MEAN_UX]


def check_grid(nx, ny, seed):
    import torch

    print(f"case noalloc {nx}x{ny}", flush=True)
    case = layout_case("channel", nx, ny, A.LBM_CM, adapter=A.LBM_OPTIMAL_ADAPTER, seed=seed)
    _, _, u, s = run_gpu(case, 5)
    ref = (u + 0.01 * np.random.default_rng(seed).standard_normal(u.shape)).astype(np.float32)
    ref_dev = torch.as_tensor(ref, device="cuda").contiguous()
    for kind in KINDS + [A.LBM_REDUCE_MOMENT_AVG]:
        want = s.reduce(kind, ref=ref if kind == A.LBM_REDUCE_L2_VS_REF else None)
        got = s.reduce_noalloc(kind, ref=ref_dev if kind == A.LBM_REDUCE_L2_VS_REF else None)
        used = 1 if kind == A.LBM_REDUCE_MASS else (2 if kind in (A.LBM_REDUCE_MOMENTUM, A.LBM_REDUCE_L2_VS_REF)
                                                    else len(want))
        # the same kernels over the same fields in the same order: bit-identical
        assert np.array_equal(got[:used], want[:used]), (kind, got, want)
        # the device-output form, written on the solver's stream
        dev = torch.full((len(want),), -1.0, dtype=torch.float64, device="cuda")
        s.reduce_noalloc(kind, ref=ref_dev if kind == A.LBM_REDUCE_L2_VS_REF else None, out=dev)
        s.sync()
        assert np.array_equal(dev.cpu().numpy()[:used], want[:used]), kind
    s.close()


def test_noalloc_reductions_match_lbm_reduce_on_two_grids(gpu_available):
    check_grid(700, 450, 1)
    check_grid(90, 37, 2)  # second call: a smaller grid reusing the same per-GPU array


def test_noalloc_rejects_what_it_cannot_hold(gpu_available):
    from lbm_amd.solver import LbmError

    case = layout_case("periodic", 16, 12, A.LBM_BGK)
    _, _, _, s = run_gpu(case, 1)
    with pytest.raises(LbmError):
        s.reduce_noalloc(A.LBM_REDUCE_L2_VS_REF)  # no device reference
    with pytest.raises(LbmError):
        s.reduce_noalloc(A.LBM_REDUCE_MOMENT_AVG)  # no adapter averages in a BGK run
    s.close()


def test_noalloc_rejects_rows_beyond_its_scratch(gpu_available):
    from lbm_amd.solver import Config, LbmError, Solver

    ny = (1 << 20) // 8 + 1  # one row more than the per-GPU array holds
    s = Solver(Config(nx=3, ny=ny, periodic_x=True, periodic_y=True))
    s.init_from_macros(np.ones((ny, 3), np.float32), np.zeros((ny, 3, 2), np.float32))
    with pytest.raises(LbmError):
        s.reduce_noalloc(A.LBM_REDUCE_ROW_MEAN_UX)
    assert len(s.reduce(A.LBM_REDUCE_ROW_MEAN_UX)) == ny  # the allocating form still serves it
    s.close()
