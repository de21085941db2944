"""Python handle on one D3Q27 context of liblbm_hip.so (include/lbm3d_hip.h).

Mirrors the host-side surface of the reference's `LBM<3>` (src/core/lbm.cuh:33-382) for
BGK<3>, MRT<3> and CM<3>: allocate/init (`Solver3D(...)`, `init_from_macros`), the time loop (`step`),
`update_macroscopics` (`get_macros`) and the kinetic-energy validator of the 3-D Taylor-Green
scenario (`kinetic_energy`). Arrays are indexed [z, y, x]; host arrays are numpy, device arrays
torch CUDA tensors (passed by pointer, stream-ordered on the context's stream).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _abi as A
from .solver import LbmError, _ptr


@dataclass
class Config3D:
    """Python form of `lbm3d_config`. Defaults follow lbm3d_config_default()."""

    nx: int
    ny: int
    nz: int
    collision: int = A.LBM_BGK  # or A.LBM_MRT (rates S[4..8]) or A.LBM_CM (omega)
    adapter: int = A.LBM_NO_ADAPTER
    omega: float = 1.0
    S: tuple = (1.0,) * 27
    rho0: float = 1.0
    storage: int = A.LBM_F32
    periodic_x: bool = False
    periodic_y: bool = False
    periodic_z: bool = False
    force: tuple = (0.0, 0.0, 0.0)
    u_inlet: float = 0.0
    quirks: int = 0

    def to_c(self) -> A.Lbm3dConfig:
        c = A.Lbm3dConfig()
        c.nx, c.ny, c.nz = self.nx, self.ny, self.nz
        c.collision = self.collision
        c.adapter = self.adapter
        c.omega = self.omega
        for k in range(27):
            c.S[k] = self.S[k]
        c.rho0 = self.rho0
        c.storage = self.storage
        c.periodic_x, c.periodic_y, c.periodic_z = int(self.periodic_x), int(self.periodic_y), int(self.periodic_z)
        for k in range(3):
            c.force[k] = self.force[k]
        c.u_inlet = self.u_inlet
        c.quirks = self.quirks
        return c

    @property
    def N(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)


class Solver3D:
    """One D3Q27 solver context on the current HIP device."""

    def __init__(self, cfg: Config3D, lib=None):
        self.lib = lib or A.load_library()
        self.cfg = cfg
        self._c = cfg.to_c()
        h = ctypes.c_void_p()
        rc = self.lib.lbm3d_create(ctypes.byref(self._c), ctypes.byref(h))
        if rc != 0:
            raise LbmError(rc, self.lib.lbm3d_last_error(None).decode())
        self.h = h

    def _check(self, rc):
        if rc != 0:
            raise LbmError(rc, self.lib.lbm3d_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.lbm3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def N(self) -> int:
        return self.cfg.N

    def set_stream(self, stream_handle: int | None):
        self._check(self.lib.lbm3d_set_stream(self.h, ctypes.c_void_p(stream_handle or 0)))

    def set_flags(self, flags):
        if flags is None:
            self._check(self.lib.lbm3d_set_flags(self.h, None, 0))
            return
        p, dev, keep = _ptr(flags, np.uint8, self.N)
        self._check(self.lib.lbm3d_set_flags(self.h, p, dev))

    def init_from_macros(self, rho, u):
        pr, d1, k1 = _ptr(rho, np.float32, self.N)
        pu, d2, k2 = _ptr(u, np.float32, 3 * self.N)
        if d1 != d2:
            raise ValueError("rho and u must both be host or both be device arrays")
        self._check(self.lib.lbm3d_init_from_macros(self.h, pr, pu, d1))

    def set_force_field(self, force):
        if force is None:
            self._check(self.lib.lbm3d_set_force_field(self.h, None, 0))
            return
        p, dev, keep = _ptr(force, np.float32, 3 * self.N)
        self._check(self.lib.lbm3d_set_force_field(self.h, p, dev))

    def step(self, n: int = 1):
        self._check(self.lib.lbm3d_step(self.h, int(n)))

    def sync(self):
        self._check(self.lib.lbm3d_sync(self.h))

    @property
    def timestep(self) -> int:
        t = ctypes.c_int64()
        self._check(self.lib.lbm3d_get_timestep(self.h, ctypes.byref(t)))
        return t.value

    def get_macros(self):
        """(rho [nz, ny, nx], u [nz, ny, nx, 3]) as float32 numpy arrays."""
        rho = np.empty(self.N, np.float32)
        u = np.empty(3 * self.N, np.float32)
        self._check(self.lib.lbm3d_get_macros(self.h, rho.ctypes.data_as(ctypes.c_void_p),
                                              u.ctypes.data_as(ctypes.c_void_p), 0))
        return rho.reshape(self.cfg.shape), u.reshape(*self.cfg.shape, 3)

    def get_pops(self):
        f = np.empty(27 * self.N, np.float32)
        self._check(self.lib.lbm3d_get_pops(self.h, f.ctypes.data_as(ctypes.c_void_p), 0))
        return f.reshape(27, *self.cfg.shape)

    def set_pops(self, f):
        p, dev, keep = _ptr(f, np.float32, 27 * self.N)
        self._check(self.lib.lbm3d_set_pops(self.h, p, dev))

    def reduce(self, kind: int):
        out = np.zeros(3, np.float64)
        self._check(self.lib.lbm3d_reduce(self.h, kind, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out if kind == A.LBM3D_REDUCE_MOMENTUM else out[0]

    # -- immersed boundary (IBMManager<3>, IBMManager.cuh:54-110, 222-252) ----------------
    def ibm_set_markers(self, xyz, ds=None, iters: int = 3):
        """Static markers AoS [n][3] in lattice units, per-marker area ds (None = 1)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        n = xyz.shape[0]
        dsa = None if ds is None else np.ascontiguousarray(ds, dtype=np.float32).reshape(n)
        self._check(self.lib.lbm3d_ibm_set_markers(
            self.h, xyz.ctypes.data_as(ctypes.c_void_p),
            None if dsa is None else dsa.ctypes.data_as(ctypes.c_void_p), n, int(iters)))
        self._ibm_n = n

    def ibm_get_forces(self):
        """(Lagrangian forces of the last step [n, 3], total force on the fluid [3] in FP64)."""
        n = getattr(self, "_ibm_n", 0)
        lag = np.empty(3 * max(n, 1), np.float32)
        tot = np.zeros(3, np.float64)
        self._check(self.lib.lbm3d_ibm_get_forces(self.h, lag.ctypes.data_as(ctypes.c_void_p),
                                                  tot.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return lag[: 3 * n].reshape(n, 3), tot

    # -- output files (src/core/lbm.cuh:173-343) -------------------------------------------
    def save_vtk(self, directory, reference_offset: bool = False):
        from . import output

        rho, u = self.get_macros()
        return output.save_vtk(directory, self.timestep, rho, u, reference_offset)

    def save_macroscopics(self, directory):
        """Full 3-D dumps (N densities, 3N velocities). The reference's LBM<3> writes only the
        first NX*NY densities and 2*NX*NY velocity floats here (lbm.cuh:174, 200)."""
        from . import output

        rho, u = self.get_macros()
        output.save_macroscopics(directory, self.timestep, rho, u)

    def save_midplane_slice(self, directory):
        from . import output

        rho, u = self.get_macros()
        output.save_midplane_slice(directory, self.timestep, rho, u)

    def kinetic_energy(self, normalise_u: float | None = 0.05) -> float:
        """calculate_kinetic_energy (kineticEnergy.cuh:44-67): sum 0.5 rho |u|^2 over the last
        stored fields, divided by N * normalise_u^2 (the reference divides by N * 0.05^2);
        None returns the raw sum."""
        e = float(self.reduce(A.LBM3D_REDUCE_KINETIC_ENERGY))
        return e if normalise_u is None else e / (self.N * normalise_u * normalise_u)
