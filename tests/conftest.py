import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
# Under pytest-xdist every worker would start an OpenMP team of all cores for the oracle, and
# spinning teams starve each other; one thread per worker (set before libgomp loads).
if os.environ.get("PYTEST_XDIST_WORKER"):
    os.environ.setdefault("OMP_NUM_THREADS", "1")
for p in (ROOT / "cuda-lbm_amd", ROOT / "oracle", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through liblbm_hip.so on the GPU)")


def _ensure_built():
    # Build the CPU-side test infrastructure if it is missing (never the GPU library here:
    # GPU tests use the in-tree build and fail loudly without it).
    if not (ROOT / "oracle" / "liblbm_oracle.so").exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    emu = ROOT / "tests" / "emu"
    if not (emu / "liblbm_emu.so").exists() and os.path.exists("/opt/rocm/bin/hipcc"):
        subprocess.run(["make", "-s", "-C", str(emu)], check=True)


_ensure_built()


@pytest.fixture(scope="session")
def gpu_available():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return True
