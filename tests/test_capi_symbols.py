"""The C-ABI library loads and exports every entry point include/lbm_hip.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "lbm_hip.h"
LIB = ROOT / "cuda-lbm_amd" / "lbm_amd" / "liblbm_hip.so"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(lbm_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ["lbm_create", "lbm_destroy", "lbm_set_flags", "lbm_init_from_macros", "lbm_step",
                 "lbm_get_macros", "lbm_get_pops", "lbm_set_pops", "lbm_reduce", "lbm_ibm_set_markers",
                 "lbm_adjoint_forward", "lbm_adjoint_backward", "lbm_last_error"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    if not LIB.exists():
        pytest.fail("liblbm_hip.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_prototypes_cover_the_header():
    from lbm_amd import _abi

    assert sorted(n for n, _, _ in _abi.PROTOTYPES) == [n for n in declared_functions() if n != "lbm_config_default"] or \
        set(n for n, _, _ in _abi.PROTOTYPES) == set(declared_functions())


def test_config_struct_layout_matches_header():
    from lbm_amd import _abi

    # 5 int32 + omega + S[9] + rho0 + storage + 2 periodic + force[2] + wall_u[2] + u_lid + quirks
    # + steps_per_launch + reserved[7]
    assert ctypes.sizeof(_abi.LbmConfig) == 4 * (5 + 1 + 9 + 1 + 1 + 2 + 2 + 2 + 1 + 1 + 8)


def test_version_and_defaults_without_gpu():
    from lbm_amd import _abi

    lib = _abi.load_library()
    assert b"gfx950" in lib.lbm_version()
    c = _abi.LbmConfig()
    assert lib.lbm_config_default(ctypes.byref(c), 16, 8) == 0
    assert c.nx == 16 and c.ny == 8 and abs(c.omega - 1.0) < 1e-7 and c.rho0 == 1.0
    assert list(c.S) == [0.0, 1.0, 1.0, 0.0, 1.0, 0.0, 1.0, 1.0, 1.0]
