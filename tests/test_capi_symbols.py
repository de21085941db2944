"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADERS = sorted((ROOT / "include").glob("*.h"))
LIB = ROOT / "cuda-lbm_amd" / "lbm_amd" / "liblbm_hip.so"


def declared_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(lbm\w*_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ["lbm_create", "lbm_destroy", "lbm_set_flags", "lbm_init_from_macros", "lbm_step",
                 "lbm_get_macros", "lbm_get_pops", "lbm_set_pops", "lbm_reduce", "lbm_ibm_set_markers",
                 "lbm_adjoint_forward", "lbm_adjoint_backward", "lbm_last_error",
                 "lbm3d_create", "lbm3d_step", "lbm3d_get_pops", "lbm3d_reduce", "lbm3d_last_error"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    if not LIB.exists():
        pytest.fail("liblbm_hip.so is not built (run __graft_entry__.build())")
    from lbm_amd import _abi

    lib = _abi.load_library()  # binds to PyTorch's HIP runtime, as every user of the library does
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_prototypes_cover_the_header():
    from lbm_amd import _abi

    assert set(n for n, _, _ in _abi.PROTOTYPES) == set(declared_functions())


def test_config_struct_layout_matches_header():
    from lbm_amd import _abi

    # 5 int32 + omega + S[9] + rho0 + storage + 2 periodic + force[2] + wall_u[2] + u_lid + quirks
    # + steps_per_launch + reserved[7]
    assert ctypes.sizeof(_abi.LbmConfig) == 4 * (5 + 1 + 9 + 1 + 1 + 2 + 2 + 2 + 1 + 1 + 8)
    # nx, ny, nz, collision, adapter, omega, S[27], rho0, storage, 3 periodic, force[3], u_inlet, quirks,
    # reserved[8]
    assert ctypes.sizeof(_abi.Lbm3dConfig) == 4 * (5 + 1 + 27 + 1 + 1 + 3 + 3 + 1 + 1 + 8)


def test_version_and_defaults_without_gpu():
    from lbm_amd import _abi

    lib = _abi.load_library()
    assert b"gfx950" in lib.lbm_version()
    c = _abi.LbmConfig()
    assert lib.lbm_config_default(ctypes.byref(c), 16, 8) == 0
    assert c.nx == 16 and c.ny == 8 and abs(c.omega - 1.0) < 1e-7 and c.rho0 == 1.0
    assert list(c.S) == [0.0, 1.0, 1.0, 0.0, 1.0, 0.0, 1.0, 1.0, 1.0]


def test_3d_defaults_without_gpu():
    from lbm_amd import _abi

    lib = _abi.load_library()
    c = _abi.Lbm3dConfig()
    assert lib.lbm3d_config_default(ctypes.byref(c), 16, 8, 4) == 0
    assert (c.nx, c.ny, c.nz) == (16, 8, 4) and c.omega == 1.0 and c.rho0 == 1.0
    assert c.collision == _abi.LBM_BGK and c.storage == _abi.LBM_F32 and c.quirks == 0
    assert list(c.S) == [1.0] * 27


def test_one_hip_runtime_per_process():
    """The test-infrastructure libraries and liblbm_hip.so, loaded in any order with torch, leave
    exactly one HIP runtime mapped (PyTorch's): the library binds to the runtime PyTorch loads."""
    import subprocess
    import sys

    code = (
        "import ctypes, sys\n"
        f"sys.path.insert(0, {str(ROOT / 'cuda-lbm_amd')!r})\n"
        f"ctypes.CDLL({str(ROOT / 'tests' / 'emu' / 'liblbm_emu.so')!r})\n"
        f"ctypes.CDLL({str(ROOT / 'tests' / 'emu' / 'liblbm3d_emu.so')!r})\n"
        "from lbm_amd import _abi\n"
        "_abi.load_library()\n"
        "import torch\n"
        "libs = _abi.hip_runtimes_loaded()\n"
        "assert len(libs) == 1 and 'torch' in libs[0], libs\n"
        "print('ok')\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
