"""Write tests/golden/d2q9_small.npz: inputs and FP64-oracle outputs of small seeded cases.

Regenerate with `python tests/golden/make_golden.py` (from the repository root, after
`make -C oracle`). Every case is a 10 x 7 grid run for 5 steps from f = f_eq(rho, u): each
boundary layout x collision operator (BGK, MRT, CM, CM with the OptimalAdapter), FP32
storage, plus FP16 and BF16 shifted storage for the cavity. Stored per case: the config
fields, flags, rho, u (AoS) and the populations after 5 steps ([9][ny][nx], float64)."""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path[:0] = [str(ROOT / "tests"), str(ROOT / "oracle"), str(ROOT / "cuda-lbm_amd")]

from helpers import layout_case  # noqa: E402
from lbm_amd import _abi as A  # noqa: E402
import oracle as orc  # noqa: E402

NX, NY, STEPS = 10, 7, 5
LAYOUTS = ["periodic", "channel", "cavity_mw", "cavity_reg", "inout", "zouhe_top", "open", "cyl_outlet"]
OPS = {"bgk": (A.LBM_BGK, 0), "mrt": (A.LBM_MRT, 0), "cm": (A.LBM_CM, 0), "cm_optimal": (A.LBM_CM, 1)}
FIELDS = ["nx", "ny", "collision", "adapter", "omega", "rho0", "storage", "periodic_x", "periodic_y", "u_lid",
          "quirks"]


def cases():
    for layout in LAYOUTS:
        for name, (coll, adapter) in OPS.items():
            yield f"{layout}-{name}-f32", layout_case(layout, NX, NY, coll, adapter=adapter)
    for st, tag in ((A.LBM_F16_SHIFTED, "f16"), (A.LBM_BF16_SHIFTED, "bf16")):
        yield f"cavity_mw-mrt-{tag}", layout_case("cavity_mw", NX, NY, A.LBM_MRT, storage=st)


def main():
    out = {}
    for key, case in cases():
        c = case.config
        o = orc.Oracle(c.to_c(), precision=64)
        case.apply_oracle(o)
        o.step(STEPS)
        out[f"{key}/cfg"] = np.array([float(getattr(c, k)) for k in FIELDS])
        out[f"{key}/S"] = np.array(c.S, np.float64)
        out[f"{key}/force"] = np.array(c.force, np.float64)
        out[f"{key}/wall_u"] = np.array(c.wall_u, np.float64)
        out[f"{key}/flags"] = (np.zeros((NY, NX), np.uint8) if case.flags is None else case.flags.astype(np.uint8))
        out[f"{key}/rho"] = case.rho.astype(np.float32)
        out[f"{key}/u"] = case.u.astype(np.float32)
        out[f"{key}/f"] = o.get_pops()
    out["meta/fields"] = np.array(FIELDS)
    out["meta/steps"] = np.array(STEPS)
    np.savez_compressed(HERE / "d2q9_small.npz", **out)
    print(f"wrote {len(out)} arrays")


if __name__ == "__main__":
    main()
