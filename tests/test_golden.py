"""Golden vectors (tests/golden/d2q9_small.npz, written by tests/golden/make_golden.py from
the FP64 oracle): small seeded cases of every boundary layout x collision operator, and the
reduced storages, 5 steps each. The CPU test guards the oracle against regressions; the GPU
test checks the library against the stored FP64 populations without running the oracle."""
from pathlib import Path

import numpy as np
import pytest

import oracle as orc
from helpers import max_rel
from lbm_amd import _abi as A
from lbm_amd.scenarios import Case
from lbm_amd.solver import Config

GOLDEN = Path(__file__).resolve().parent / "golden" / "d2q9_small.npz"
DATA = np.load(GOLDEN, allow_pickle=False)
KEYS = sorted({k.split("/")[0] for k in DATA.files if not k.startswith("meta/")})
STEPS = int(DATA["meta/steps"])


def load_case(key):
    fields = [str(x) for x in DATA["meta/fields"]]
    v = dict(zip(fields, DATA[f"{key}/cfg"]))
    cfg = Config(nx=int(v["nx"]), ny=int(v["ny"]), collision=int(v["collision"]), adapter=int(v["adapter"]),
                 omega=float(v["omega"]), S=tuple(float(x) for x in DATA[f"{key}/S"]), rho0=float(v["rho0"]),
                 storage=int(v["storage"]), periodic_x=bool(v["periodic_x"]), periodic_y=bool(v["periodic_y"]),
                 force=tuple(float(x) for x in DATA[f"{key}/force"]),
                 wall_u=tuple(float(x) for x in DATA[f"{key}/wall_u"]), u_lid=float(v["u_lid"]),
                 quirks=int(v["quirks"]))
    flags = DATA[f"{key}/flags"]
    return Case(key, cfg, DATA[f"{key}/rho"], DATA[f"{key}/u"], flags=flags if flags.any() else None)


def test_golden_file_covers_every_layout_and_operator():
    assert len(KEYS) == 8 * 4 + 2
    for key in KEYS:
        assert DATA[f"{key}/f"].shape == (9, 7, 10)


@pytest.mark.parametrize("key", KEYS)
def test_oracle_reproduces_golden(key):
    case = load_case(key)
    o = orc.Oracle(case.config.to_c(), precision=64)
    case.apply_oracle(o)
    o.step(STEPS)
    assert np.max(np.abs(o.get_pops() - DATA[f"{key}/f"])) <= 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("key", KEYS)
def test_gpu_matches_golden(gpu_available, key):
    from lbm_amd import Solver

    case = load_case(key)
    s = Solver(case.config)
    case.apply(s)
    s.step(STEPS)
    f = s.get_pops()
    s.close()
    ref = DATA[f"{key}/f"]
    if case.config.storage == A.LBM_F32:
        assert max_rel(f, ref) <= 1e-5
    else:
        quantum = 2.0 ** -11 if case.config.storage == A.LBM_F16_SHIFTED else 2.0 ** -8
        assert float(np.max(np.abs(f - ref))) <= 4 * 0.05 * quantum
