"""Shared test helpers: run a case through the oracle, the host emulation of the kernel body,
or liblbm_hip.so on the GPU, and compare."""
from __future__ import annotations

import ctypes
import math
from pathlib import Path

import numpy as np

import oracle as orc
from lbm_amd import _abi as A
from lbm_amd import scenarios as sc
from lbm_amd.solver import Config

ROOT = Path(__file__).resolve().parent.parent
EMU = ROOT / "tests" / "emu" / "liblbm_emu.so"
_emu = None


def emu_lib():
    global _emu
    if _emu is None:
        L = ctypes.CDLL(str(EMU))
        L.emu_run.restype = ctypes.c_int
        L.emu_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 4 + \
            [ctypes.c_int] + [ctypes.c_void_p] * 4
        _emu = L
    return _emu


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def run_emu(case: sc.Case, nsteps: int, seg: int = 0, two_ty: int = 0, reverse: bool = False):
    """Host emulation of the fused kernels (same bodies, same grids). seg: nodes per thread of
    one-step launches (4, or the storage's narrow width: FP32 1, 16-bit 2; 0 = 4). two_ty: 0
    one step per launch, 6 or 14 two-step tile launches with that tile height, (W, H) two-step
    marching strips of W wavefronts and H rows; both where the library would use them; reverse
    runs tiles or strips in reverse order. Returns (f, rho, u, avg)."""
    if isinstance(two_ty, tuple):
        two_ty = -(two_ty[0] * 1000 + two_ty[1])
    c = case.config.to_c()
    N = case.config.N
    fl = None if case.flags is None else np.ascontiguousarray(case.flags, np.uint8).reshape(-1)
    rho = np.ascontiguousarray(case.rho, np.float32).reshape(-1)
    u = np.ascontiguousarray(case.u, np.float32).reshape(-1)
    F = None if case.force_field is None else np.ascontiguousarray(case.force_field, np.float32).reshape(-1)
    f = np.zeros(9 * N, np.float32)
    r = np.zeros(N, np.float32)
    uu = np.zeros(2 * N, np.float32)
    avg = np.zeros(3, np.float64)
    rc = emu_lib().emu_run(ctypes.byref(c), seg, two_ty, int(reverse), _p(fl), _p(rho), _p(u), _p(F), nsteps, _p(f),
                           _p(r), _p(uu), _p(avg))
    assert rc == 0
    ny, nx = case.config.ny, case.config.nx
    return f.reshape(9, ny, nx), r.reshape(ny, nx), uu.reshape(ny, nx, 2), avg


def run_oracle(case: sc.Case, nsteps: int, precision: int = 32):
    o = orc.Oracle(case.config.to_c(), precision=precision)
    case.apply_oracle(o)
    o.step(nsteps)
    r, u = o.get_macros()
    return o.get_pops(), r, u, o


def run_gpu(case: sc.Case, nsteps: int):
    from lbm_amd import Solver

    s = Solver(case.config)
    case.apply(s)
    s.step(nsteps)
    f = s.get_pops()
    r, u = s.get_macros()
    return f, r, u, s


def max_rel(a, b, floor=1e-3):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


def layout_case(layout: str, nx: int, ny: int, collision: int, storage: int = A.LBM_F32, adapter: int = 0,
                forced: bool = True, quirks: int = 0, seed: int = 7) -> sc.Case:
    """Small seeded cases covering every supported boundary layout.

    layouts: periodic, channel (periodic x, bounce-back walls), cavity_mw (moving-wall lid),
    cavity_reg (the reference's regularized cavity), inout (Zou-He inlet, ZG outlet, walls,
    one solid node), open (no flags, non-periodic: the reference's unstreamed edge slots),
    cyl_outlet (Zou-He inlet, pressure outlet, walls, a block of full-way bounce-back nodes).
    """
    rng = np.random.default_rng(seed)
    om = 1.25
    if collision == A.LBM_MRT:
        S = (0.0, 1.1, 1.2, 0.0, 1.4, 0.0, 1.4, om, om)
    elif collision == A.LBM_CM:
        S = (0.0, 0.0, 0.0, 1.0, om, om, 1.0, 1.0, 1.0)
    else:
        S = (0.0, om, om, 0.0, om, 0.0, om, om, om)
    flags = np.zeros((ny, nx), np.uint8)
    px = py = False
    if layout == "periodic":
        px = py = True
    elif layout == "channel":
        px = True
        flags[0, :] = flags[-1, :] = A.BOUNCE_BACK
    elif layout == "cavity_mw":
        flags[:, 0] = flags[:, -1] = flags[0, :] = flags[-1, :] = A.BOUNCE_BACK
        flags[-1, 1:-1] = A.MOVING_WALL_BOUNCE_BACK
    elif layout == "cavity_reg":
        flags[:, 0] = flags[:, -1] = flags[0, :] = A.REGULARIZED_BOUNCE_BACK
        flags[-1, :] = A.REGULARIZED_INLET_TOP
        for yy, xx in ((0, 0), (0, nx - 1), (ny - 1, 0), (ny - 1, nx - 1)):
            flags[yy, xx] = A.REGULARIZED_BOUNCE_BACK_CORNER
    elif layout == "inout":
        flags[:, 0] = A.ZOU_HE_LEFT
        flags[:, -1] = A.ZG_OUTFLOW
        flags[0, :] = flags[-1, :] = A.BOUNCE_BACK
        flags[ny // 2, nx // 2] = A.BOUNCE_BACK
    elif layout == "zouhe_top":
        flags[:, 0] = flags[:, -1] = flags[0, :] = A.BOUNCE_BACK
        flags[-1, 1:-1] = A.ZOU_HE_TOP
    elif layout == "open":
        pass
    elif layout == "cyl_outlet":
        # Zou-He inlet, pressure outlet, bounce-back walls, a full-way bounce-back block
        flags[:, 0] = A.ZOU_HE_LEFT
        flags[:, -1] = A.PRESSURE_OUTLET
        flags[0, :] = flags[-1, :] = A.BOUNCE_BACK
        cy, cx = ny // 2, nx // 3
        flags[max(cy - 2, 1):cy + 2, max(cx - 2, 1):cx + 2] = A.CYLINDER
    else:
        raise ValueError(layout)
    cfg = Config(nx=nx, ny=ny, collision=collision, adapter=adapter, omega=om, S=S, storage=storage,
                 periodic_x=px, periodic_y=py, force=(2e-5, -1e-5) if forced else (0.0, 0.0),
                 wall_u=(0.05, 0.0), u_lid=0.05, quirks=quirks)
    rho = (1.0 + 0.02 * rng.uniform(-1, 1, size=(ny, nx))).astype(np.float32)
    u = (0.04 * rng.uniform(-1, 1, size=(ny, nx, 2))).astype(np.float32)
    return sc.Case(layout, cfg, rho, u, flags=None if not flags.any() else flags)


# ---- closed forms from the lattice definition (no oracle code) -----------------------------

D2Q9_C = np.array([(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1)], np.float64)
D2Q9_W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
THETA = (0.0003, -0.00775, 0.00016, 0.0087)  # OptimalAdapter (adapters.cuh:63)


def feq2(rho, ux, uy):
    """Second-order equilibrium w rho [1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u^2] (equilibrium.cu:5-39)."""
    cu = D2Q9_C[:, 0] * ux + D2Q9_C[:, 1] * uy
    return D2Q9_W * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * (ux * ux + uy * uy))


def central_moment_matrix(ux, uy):
    """T(u): rows 1, x, y, x^2+y^2, x^2-y^2, xy, x^2 y, x y^2, x^2 y^2 of (c - u) (CM.cuh:53-68)."""
    x = D2Q9_C[:, 0] - ux
    y = D2Q9_C[:, 1] - uy
    return np.stack([np.ones(9), x, y, x * x + y * y, x * x - y * y, x * y, x * x * y, x * y * y, x * x * y * y])


def momentum_flux_magnitude(f):
    """|Pi| = sqrt(Pxx^2 + 2 Pxy^2 + Pyy^2), Pi = sum f c c (macroscopics.cu:18-33)."""
    pxx = np.sum(f * D2Q9_C[:, 0] ** 2)
    pxy = np.sum(f * D2Q9_C[:, 0] * D2Q9_C[:, 1])
    pyy = np.sum(f * D2Q9_C[:, 1] ** 2)
    return math.sqrt(pxx * pxx + 2 * pxy * pxy + pyy * pyy)


def cm_collide_uniform(rho, ux, uy, S, high_rate):
    """Post-collision populations of the uniform equilibrium state f = feq2(rho, u) under CM
    with rates S[0..5] and `high_rate` for moments 6..8 (CM.cuh:104-115, no force), inverted
    by solving T(u) f = k* (independent of the closed-form T^-1)."""
    T = central_moment_matrix(ux, uy)
    k = T @ feq2(rho, ux, uy)
    cs2 = 1.0 / 3.0
    keq = np.array([rho, 0, 0, 2 * rho * cs2, 0, 0, 0, 0, rho * cs2 * cs2])
    s = np.array(list(S[:6]) + [high_rate] * 3, np.float64)
    return np.linalg.solve(T, k - s * (k - keq))
