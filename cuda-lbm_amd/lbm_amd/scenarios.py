"""The benchmark configurations and the reference's validators, on the C ABI.

Each builder returns a `Case` (config, initial rho/u, flags, optional force field and IBM
markers) equivalent to a reference `ScenarioTrait` (src/scenarios/scenario.cuh:22-78) with
its Init and Boundary functors evaluated on the host. The error functions restate the
reference's Validation formulas:
  - Taylor-Green relative L2 %         taylorGreenScenario.cuh:59-88, taylorGreenFunctors.cuh:57-98
  - Poiseuille RMS % of u_max          poiseuilleScenario.cuh:55-77, poiseuilleFunctors.cuh:63-87
  - lid-driven cavity Ghia NRMSE %     lidDrivenCavityScenario.cuh:89-157, lidDrivenCavityFunctors.cuh:59-222
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import _abi as A
from .solver import Config


@dataclass
class Case:
    name: str
    config: Config
    rho: np.ndarray            # [ny, nx] float32
    u: np.ndarray              # [ny, nx, 2] float32
    flags: np.ndarray | None = None      # [ny, nx] uint8
    force_field: np.ndarray | None = None  # [ny, nx, 2] float32
    markers: np.ndarray | None = None    # [n, 2] float32
    ds: np.ndarray | None = None         # [n] float32
    steps: int = 0
    extra: dict = field(default_factory=dict)

    def apply(self, solver):
        """Load this case into a Solver (flags, force, markers, initial state)."""
        if self.flags is not None:
            solver.set_flags(self.flags)
        if self.force_field is not None:
            solver.set_force_field(self.force_field)
        if self.markers is not None:
            solver.ibm_set_markers(self.markers, self.ds, self.extra.get("ibm_iters", 3))
        solver.init_from_macros(self.rho, self.u)

    def apply_oracle(self, orc):
        if self.flags is not None:
            orc.set_flags(self.flags)
        if self.force_field is not None:
            orc.set_force_field(self.force_field)
        if self.markers is not None:
            orc.ibm_set_markers(self.markers, self.ds, self.extra.get("ibm_iters", 3))
        orc.init_from_macros(self.rho, self.u)


def viscosity_to_omega(nu: float) -> float:
    """omega = 1/tau, tau = 3 nu + 0.5 in float32 (lbm_constants.cuh:365-367, scenario.cuh:37-38)."""
    tau = np.float32(3.0) * np.float32(nu) + np.float32(0.5)
    return float(np.float32(1.0) / tau)


# ---- Taylor-Green -----------------------------------------------------------------------

def tg_field(nx: int, ny: int, nu: float, u0: float, t: float):
    """TaylorGreenInit / TaylorGreenValidation fields (taylorGreenFunctors.cuh:24-47, 67-93),
    cell-centred coordinates x+0.5, decay exp(-t nu (kx^2+ky^2))."""
    x = np.arange(nx, dtype=np.float64) + 0.5
    y = np.arange(ny, dtype=np.float64) + 0.5
    X, Y = np.meshgrid(x, y)
    kx, ky = 2 * math.pi / nx, 2 * math.pi / ny
    td = 1.0 / (nu * (kx * kx + ky * ky))
    ux = -u0 * math.sqrt(ky / kx) * np.cos(kx * X) * np.sin(ky * Y) * math.exp(-t / td)
    uy = u0 * math.sqrt(kx / ky) * np.sin(kx * X) * np.cos(ky * Y) * math.exp(-t / td)
    P = -0.25 * u0 * u0 * ((ky / kx) * np.cos(2 * kx * X) + (kx / ky) * np.cos(2 * ky * Y)) * math.exp(-2 * t / td)
    rho = 1.0 + 3.0 * P
    return rho, np.stack([ux, uy], axis=-1)


def taylor_green(n: int = 128, nu: float = 0.01, u0: float = 0.04, steps: int = 1000,
                 collision: int = A.LBM_BGK, storage: int = A.LBM_F32, ny: int | None = None) -> Case:
    ny = ny or n
    om = viscosity_to_omega(nu)
    cfg = Config(nx=n, ny=ny, collision=collision, omega=om,
                 S=(0.0, 1.0, 1.4, 0.0, 1.2, 0.0, 1.9, om, om) if collision == A.LBM_MRT else
                 (0.0, 0.0, 0.0, 1.0, om, om, 1.0, 1.0, 1.0) if collision == A.LBM_CM else
                 (0.0, om, om, 0.0, om, 0.0, om, om, om),
                 storage=storage, periodic_x=True, periodic_y=True, u_lid=u0)
    rho, u = tg_field(n, ny, nu, u0, 0.0)
    return Case("taylor_green", cfg, rho.astype(np.float32), u.astype(np.float32), steps=steps,
                extra={"nu": nu, "u0": u0})


def tg_error(u_sim: np.ndarray, nu: float, u0: float, t: float) -> float:
    """TaylorGreenScenario::compute_error: 100 sqrt(sum|u-u_a|^2 / sum|u_a|^2)."""
    ny, nx = u_sim.shape[:2]
    _, ua = tg_field(nx, ny, nu, u0, t)
    d = u_sim.astype(np.float64) - ua
    return 100.0 * math.sqrt(float(np.sum(d * d)) / float(np.sum(ua * ua)))


# ---- lid-driven cavity ------------------------------------------------------------------

def lid_driven_cavity(n: int = 1024, u_lid: float = 0.1, Re: float = 1000.0, collision: int = A.LBM_MRT,
                      storage: int = A.LBM_F32, steps: int = 50000, rates=(1.1, 1.2, 1.4),
                      lid: str = "moving_wall") -> Case:
    """D2Q9 cavity, lid on the top row. lid = "moving_wall" (Ladd bounce-back, walls plain
    bounce-back) or "regularized" (the reference's LidDrivenBoundary, lidDrivenCavityFunctors.cuh:37-56)."""
    nu = u_lid * n / Re
    om = viscosity_to_omega(nu)
    se, seps, sq = rates
    if collision == A.LBM_MRT:
        S = (0.0, se, seps, 0.0, sq, 0.0, sq, om, om)  # [rho,e,eps,jx,qx,jy,qy,pxx,pxy]
    elif collision == A.LBM_CM:
        S = (0.0, 0.0, 0.0, 1.0, om, om, 1.0, 1.0, 1.0)
    else:
        S = (0.0, om, om, 0.0, om, 0.0, om, om, om)
    flags = np.full((n, n), A.BOUNCE_BACK, np.uint8)
    flags[1:-1, 1:-1] = A.FLUID
    if lid == "moving_wall":
        flags[n - 1, 1:-1] = A.MOVING_WALL_BOUNCE_BACK
    else:
        flags[:, 0] = flags[:, -1] = flags[0, :] = A.REGULARIZED_BOUNCE_BACK
        flags[n - 1, :] = A.REGULARIZED_INLET_TOP
        for (yy, xx) in ((0, 0), (0, n - 1), (n - 1, 0), (n - 1, n - 1)):
            flags[yy, xx] = A.REGULARIZED_BOUNCE_BACK_CORNER
    cfg = Config(nx=n, ny=n, collision=collision, omega=om, S=S, storage=storage,
                 wall_u=(u_lid, 0.0), u_lid=u_lid)
    rho = np.ones((n, n), np.float32)
    u = np.zeros((n, n, 2), np.float32)
    return Case("lid_driven_cavity", cfg, rho, u, flags=flags, steps=steps, extra={"nu": nu, "Re": Re, "u_lid": u_lid})


# Ghia, Ghia & Shin (1982), Tables I and II: u along the vertical centre line, v along the
# horizontal one (the data lidDrivenCavityFunctors.cuh:59-170 holds). The Re=3200 entry at
# y = 0.4531 is -0.08636 (the reference's -0.86636 breaks the column's monotone trend).
GHIA_Y = [1.0000, 0.9766, 0.9688, 0.9609, 0.9531, 0.8516, 0.7344, 0.6172, 0.5000, 0.4531, 0.2813, 0.1719,
          0.1016, 0.0703, 0.0625, 0.0547, 0.0000]
GHIA_X = [1.0000, 0.9688, 0.9609, 0.9531, 0.9453, 0.9063, 0.8594, 0.8047, 0.5000, 0.2344, 0.2266, 0.1563,
          0.0938, 0.0781, 0.0703, 0.0625, 0.0000]
GHIA_U = {
    100: [1.0000, 0.84123, 0.78871, 0.73722, 0.68717, 0.23151, 0.00332, -0.13641, -0.20581, -0.21090, -0.15662,
          -0.10150, -0.06434, -0.04775, -0.04192, -0.03717, 0.00000],
    400: [1.00000, 0.75837, 0.68439, 0.61756, 0.55892, 0.29093, 0.16256, 0.02135, -0.11477, -0.17119, -0.32726,
          -0.24299, -0.14612, -0.10338, -0.09266, -0.08186, 0.00000],
    1000: [1.0000, 0.65928, 0.57492, 0.51117, 0.46604, 0.33304, 0.18719, 0.05702, -0.06080, -0.10648, -0.27805,
           -0.38289, -0.29730, -0.22220, -0.20196, -0.18109, 0.00000],
    3200: [1.0000, 0.53236, 0.48296, 0.46547, 0.46101, 0.34682, 0.19791, 0.07156, -0.04272, -0.08636, -0.24427,
           -0.34323, -0.41933, -0.37827, -0.35344, -0.32407, 0.00000],
    5000: [1.0000, 0.48223, 0.46120, 0.45992, 0.46036, 0.33556, 0.20087, 0.08183, -0.03039, -0.07404, -0.22855,
           -0.33050, -0.40435, -0.43643, -0.42901, -0.41165, 0.00000],
    7500: [1.0000, 0.47244, 0.47048, 0.47323, 0.47167, 0.34228, 0.20591, 0.08342, -0.03800, -0.07503, -0.23176,
           -0.32393, -0.38324, -0.43025, -0.43590, -0.43154, 0.00000],
    10000: [1.0000, 0.47221, 0.47783, 0.48070, 0.47804, 0.34635, 0.20673, 0.08344, 0.03111, -0.07540, -0.23186,
            -0.32709, -0.38000, -0.41657, -0.42537, -0.42735, 0.00000],
}
GHIA_V = {
    100: [0.00000, -0.05906, -0.07391, -0.08864, -0.10313, -0.16914, -0.22445, -0.24533, 0.05454, 0.17527, 0.17507,
          0.16077, 0.12317, 0.10890, 0.10091, 0.09233, 0.00000],
    400: [0.00000, -0.12146, -0.15663, -0.19254, -0.22847, -0.23827, -0.44993, -0.38598, 0.05188, 0.30174, 0.30203,
          0.28124, 0.22965, 0.20920, 0.19713, 0.18360, 0.00000],
    1000: [0.00000, -0.21388, -0.27669, -0.33714, -0.39188, -0.51550, -0.42665, -0.31966, 0.02526, 0.32235, 0.33075,
           0.37095, 0.32627, 0.30353, 0.29012, 0.27485, 0.00000],
    3200: [0.00000, -0.39017, -0.47425, -0.52357, -0.54053, -0.44307, -0.37401, -0.31184, 0.00999, 0.28188, 0.29030,
           0.37119, 0.42768, 0.41906, 0.40917, 0.39560, 0.00000],
    5000: [0.00000, -0.49774, -0.55069, -0.55408, -0.52876, -0.41442, -0.36214, -0.30018, 0.00945, 0.27280, 0.28066,
           0.35368, 0.42951, 0.43648, 0.43329, 0.42447, 0.00000],
    7500: [0.00000, -0.53858, -0.55216, -0.52347, -0.48590, -0.41050, -0.36213, -0.30448, 0.00824, 0.27348, 0.28117,
           0.35060, 0.41824, 0.43564, 0.44030, 0.43979, 0.00000],
    10000: [0.00000, -0.54302, -0.52987, -0.49099, -0.45863, -0.41496, -0.36737, -0.30719, 0.00831, 0.27224, 0.28003,
            0.35070, 0.41487, 0.43124, 0.43733, 0.43983, 0.00000],
}


def ghia_closest(re: int) -> int:
    """LidDrivenValidation::get_closest_ref_data (lidDrivenCavityFunctors.cuh:199-219)."""
    avail = [100, 400, 1000, 3200, 5000, 7500, 10000]
    best = avail[0]
    for a in avail[1:]:
        if abs(re - a) < abs(re - best):
            best = a
    return best


def ghia_error(u_sim: np.ndarray, u_lid: float, re: float) -> float:
    """LidDrivenScenario::compute_error: mean of the u and v NRMSE along the centre lines, %."""
    ny, nx = u_sim.shape[:2]
    ref = ghia_closest(int(re))
    cx, cy = nx // 2, ny // 2
    eu = ru = ev = rv = 0.0
    for i in range(len(GHIA_Y)):
        y = min(max(int(math.floor(GHIA_Y[i] * (ny - 1) + 0.5)), 0), ny - 1)
        ul = float(u_sim[y, cx, 0]) / u_lid
        eu += (ul - GHIA_U[ref][i]) ** 2
        ru += GHIA_U[ref][i] ** 2
        x = min(max(int(math.floor(GHIA_X[i] * (nx - 1) + 0.5)), 0), nx - 1)
        vl = float(u_sim[cy, x, 1]) / u_lid
        ev += (vl - GHIA_V[ref][i]) ** 2
        rv += GHIA_V[ref][i] ** 2
    return 100.0 * (math.sqrt(eu / ru) + math.sqrt(ev / rv)) / 2.0


# ---- Poiseuille ---------------------------------------------------------------------------

def poiseuille(nx: int = 4096, ny: int = 1024, nu: float = 0.05, u_max: float = 0.05,
               collision: int = A.LBM_CM, adapter: int = A.LBM_OPTIMAL_ADAPTER, storage: int = A.LBM_F32,
               steps: int = 20000, adapter_mode: int = A.LBM_ADAPTER_SAME_STEP) -> Case:
    """Force-driven channel: periodic x, bounce-back rows y = 0 and ny-1, Guo force
    F = 8 nu u_max / ny^2 (PoiseuilleInit::apply_forces, poiseuilleFunctors.cuh:21-39)."""
    om = viscosity_to_omega(nu)
    F = float(np.float32(8.0) * np.float32(nu) * np.float32(u_max) / np.float32(ny * ny))
    if collision == A.LBM_CM:
        S = (0.0, 0.0, 0.0, 1.0, om, om, 1.0, 1.0, 1.0)
    else:
        S = (0.0, om, om, 0.0, om, 0.0, om, om, om)
    flags = np.zeros((ny, nx), np.uint8)
    flags[0, :] = flags[ny - 1, :] = A.BOUNCE_BACK
    cfg = Config(nx=nx, ny=ny, collision=collision, adapter=adapter if collision == A.LBM_CM else 0,
                 adapter_mode=adapter_mode, omega=om, S=S, storage=storage, periodic_x=True, force=(F, 0.0))
    return Case("poiseuille", cfg, np.ones((ny, nx), np.float32), np.zeros((ny, nx, 2), np.float32),
                flags=flags, steps=steps, extra={"nu": nu, "u_max": u_max, "F": F})


def poiseuille_profile(ny: int, nu: float, u_max: float) -> np.ndarray:
    """PoiseuilleValidation: u(y) = F/(2 nu) y (ny - y), F = 8 nu u_max / ny^2."""
    F = 8.0 * nu * u_max / (ny * ny)
    y = np.arange(ny, dtype=np.float64)
    return (F / (2.0 * nu)) * y * (ny - y)


def poiseuille_startup_profile(ny: int, nu: float, u_max: float, t: float, terms: int = 4001) -> np.ndarray:
    """The channel flow from rest under the same force, with the validator's walls (u = 0 at
    y = 0 and y = ny): u_t = nu u_yy + F, u(y, 0) = 0, whose solution is the parabola minus its
    decaying sine series,
        u(y, t) = F/(2 nu) y (ny - y) - sum_{n odd} 4 F ny^2 / (nu n^3 pi^3) sin(n pi y / ny) exp(-nu (n pi / ny)^2 t).
    It tends to poiseuille_profile as t grows, so it measures the start-up at any step count."""
    F = 8.0 * nu * u_max / (ny * ny)
    y = np.arange(ny, dtype=np.float64)
    u = (F / (2.0 * nu)) * y * (ny - y)
    for n in range(1, terms + 1, 2):
        a = 4.0 * F * ny * ny / (nu * n ** 3 * math.pi ** 3) * math.exp(-nu * (n * math.pi / ny) ** 2 * t)
        if a < 1e-30:
            break
        u -= a * np.sin(n * math.pi * y / ny)
    return u


def poiseuille_error(row_mean_ux: np.ndarray, nu: float, u_max: float, profile: np.ndarray | None = None) -> float:
    """PoiseuilleScenario::compute_error: RMS over y of (x-mean u_x - profile), % of u_max."""
    prof = poiseuille_profile(len(row_mean_ux), nu, u_max) if profile is None else profile
    d = np.asarray(row_mean_ux, np.float64) - prof
    return math.sqrt(float(np.mean(d * d))) * 100.0 / u_max


# ---- IBM cylinder -------------------------------------------------------------------------

def cylinder_markers(cx: float, cy: float, r: float, n: int):
    """create_cylinder (IBM_generators.cu:5-25): n equispaced markers; ds = 2 pi r / n."""
    ang = 2.0 * math.pi / n
    i = np.arange(n)
    xy = np.stack([cx + r * np.cos(i * ang), cy + r * np.sin(i * ang)], axis=-1).astype(np.float32)
    ds = np.full(n, 2.0 * math.pi * r / n, np.float32)
    return xy, ds


def ibm_cylinder(nx: int = 2048, ny: int = 512, u_in: float = 0.05, D: float = 64.0, n_markers: int = 201,
                 Re: float = 100.0, seed: int = 0x1B3, steps: int = 100000, storage: int = A.LBM_F32,
                 collision: int = A.LBM_CM, cx: float = 512.0, iters: int = 3) -> Case:
    """Channel with a Zou-He velocity inlet (x = 0), zero-gradient outlet (x = nx-1), bounce-back
    side walls and an IBM cylinder (FlowPastCylinderBoundary, flowPastCylinderFunctors.cuh:48-77)."""
    rng = np.random.default_rng(seed)
    dy = float(rng.uniform(-2.0, 2.0))
    nu = u_in * D / Re
    om = viscosity_to_omega(nu)
    S = (0.0, 0.0, 0.0, 1.0, om, om, 1.0, 1.0, 1.0) if collision == A.LBM_CM else (0.0, om, om, 0.0, om, 0.0, om, om, om)
    flags = np.zeros((ny, nx), np.uint8)
    flags[:, 0] = A.ZOU_HE_LEFT
    flags[:, nx - 1] = A.ZG_OUTFLOW
    flags[0, :] = flags[ny - 1, :] = A.BOUNCE_BACK
    cfg = Config(nx=nx, ny=ny, collision=collision, omega=om, S=S, storage=storage, u_lid=u_in)
    xy, ds = cylinder_markers(cx, ny / 2.0 + dy, D / 2.0, n_markers)
    u = np.zeros((ny, nx, 2), np.float32)
    u[..., 0] = u_in
    return Case("ibm_cylinder", cfg, np.ones((ny, nx), np.float32), u, flags=flags, markers=xy, ds=ds,
                steps=steps, extra={"nu": nu, "u_in": u_in, "D": D, "ibm_iters": iters, "cy_offset": dy})


# ---- inverse problem / bandwidth grid ----------------------------------------------------

def fourier_force(nx: int, ny: int, modes: int = 16, amp: float = 1e-5, seed: int = 0xF0F0) -> np.ndarray:
    """Smooth body-force field: sum of `modes` random Fourier modes, each of amplitude `amp`."""
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny))
    F = np.zeros((ny, nx, 2))
    for _ in range(modes):
        kx, ky = rng.integers(1, 5, size=2)
        ph = rng.uniform(0, 2 * math.pi, size=2)
        d = rng.normal(size=2)
        d /= np.linalg.norm(d)
        s = np.sin(2 * math.pi * (kx * X / nx + ky * Y / ny) + ph[0])
        F[..., 0] += amp * d[0] * s
        F[..., 1] += amp * d[1] * s
    return F.astype(np.float32)


def inverse_problem(n: int = 512, nu: float = 0.02, steps: int = 2000, seed: int = 0xF0F0,
                    collision: int = A.LBM_BGK) -> Case:
    om = viscosity_to_omega(nu)
    S = (0.0, om, om, 0.0, om, 0.0, om, om, om)
    cfg = Config(nx=n, ny=n, collision=collision, omega=om, S=S, periodic_x=True, periodic_y=True)
    F = fourier_force(n, n, seed=seed)
    return Case("inverse_problem", cfg, np.ones((n, n), np.float32), np.zeros((n, n, 2), np.float32),
                force_field=F, steps=steps, extra={"nu": nu})


def bandwidth_grid(n: int = 8192, storage: int = A.LBM_F32, collision: int = A.LBM_BGK, nu: float = 0.02,
                   seed: int = 0x5EED) -> Case:
    """Periodic BGK grid large enough that the populations live in HBM, not the 256 MB cache."""
    rng = np.random.default_rng(seed)
    om = viscosity_to_omega(nu)
    cfg = Config(nx=n, ny=n, collision=collision, omega=om,
                 S=(0.0, 1.1, 1.2, 0.0, 1.4, 0.0, 1.4, om, om), storage=storage, periodic_x=True, periodic_y=True)
    rho = (1.0 + 0.01 * rng.uniform(-1, 1, size=(n, n))).astype(np.float32)
    u = (0.05 * rng.uniform(-1, 1, size=(n, n, 2))).astype(np.float32)
    return Case("bandwidth_grid", cfg, rho, u)
