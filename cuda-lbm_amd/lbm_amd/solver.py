"""Python handle on one liblbm_hip.so context.

Mirrors the host-side surface of the reference's `LBM<2>` (src/core/lbm.cuh:33-382):
allocate/init (`Solver(...)`, `init_from_macros`), the time loop (`step`),
`update_macroscopics` (`get_macros`) and the validators' inputs, over the C ABI of
include/lbm_hip.h. Host arrays are numpy; device arrays are torch CUDA tensors (passed by
pointer, stream-ordered on the context's stream).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _abi as A


class LbmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


@dataclass
class Config:
    """Python form of `lbm_config`. Defaults follow lbm_config_default()."""

    nx: int
    ny: int
    collision: int = A.LBM_BGK
    adapter: int = A.LBM_NO_ADAPTER
    adapter_mode: int = A.LBM_ADAPTER_SAME_STEP
    omega: float = 1.0
    S: tuple = (0.0, 1.0, 1.0, 0.0, 1.0, 0.0, 1.0, 1.0, 1.0)
    rho0: float = 1.0
    storage: int = A.LBM_F32
    periodic_x: bool = False
    periodic_y: bool = False
    force: tuple = (0.0, 0.0)
    wall_u: tuple = (0.0, 0.0)
    u_lid: float = 0.1
    quirks: int = 0
    steps_per_launch: int = 0  # 0: two steps per launch where supported, 1: one

    def to_c(self) -> A.LbmConfig:
        c = A.LbmConfig()
        c.nx, c.ny = self.nx, self.ny
        c.collision, c.adapter, c.adapter_mode = self.collision, self.adapter, self.adapter_mode
        c.omega = self.omega
        for k in range(9):
            c.S[k] = self.S[k]
        c.rho0 = self.rho0
        c.storage = self.storage
        c.periodic_x, c.periodic_y = int(self.periodic_x), int(self.periodic_y)
        c.force[0], c.force[1] = self.force
        c.wall_u[0], c.wall_u[1] = self.wall_u
        c.u_lid = self.u_lid
        c.quirks = self.quirks
        c.steps_per_launch = self.steps_per_launch
        return c

    @property
    def N(self) -> int:
        return self.nx * self.ny


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _ptr(x, dtype, n):
    """(pointer, on_device, keepalive) for a numpy array or a torch tensor."""
    if x is None:
        return None, 0, None
    if _is_torch(x):
        if not x.is_contiguous():
            raise ValueError("torch tensors must be contiguous")
        if x.numel() != n:
            raise ValueError(f"expected {n} elements, got {x.numel()}")
        return ctypes.c_void_p(x.data_ptr()), int(x.is_cuda), x
    a = np.ascontiguousarray(x, dtype=dtype)
    if a.size != n:
        raise ValueError(f"expected {n} elements, got {a.size}")
    return a.ctypes.data_as(ctypes.c_void_p), 0, a


class Solver:
    """One D2Q9 solver context on the current HIP device."""

    def __init__(self, cfg: Config, lib=None):
        self.lib = lib or A.load_library()
        self.cfg = cfg
        self._c = cfg.to_c()
        h = ctypes.c_void_p()
        rc = self.lib.lbm_create(ctypes.byref(self._c), ctypes.byref(h))
        if rc != 0:
            raise LbmError(rc, self.lib.lbm_last_error(None).decode())
        self.h = h

    # -- plumbing ---------------------------------------------------------------------
    def _check(self, rc):
        if rc != 0:
            raise LbmError(rc, self.lib.lbm_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.lbm_destroy(self.h)
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
        self._check(self.lib.lbm_set_stream(self.h, ctypes.c_void_p(stream_handle or 0)))

    # -- state ------------------------------------------------------------------------
    def set_flags(self, flags):
        p, dev, keep = _ptr(flags, np.uint8, self.N)
        self._check(self.lib.lbm_set_flags(self.h, p, dev))

    def init_from_macros(self, rho, u):
        pr, d1, k1 = _ptr(rho, np.float32, self.N)
        pu, d2, k2 = _ptr(u, np.float32, 2 * self.N)
        if d1 != d2:
            raise ValueError("rho and u must both be host or both be device arrays")
        self._check(self.lib.lbm_init_from_macros(self.h, pr, pu, d1))

    def set_force_field(self, force):
        p, dev, keep = _ptr(force, np.float32, 2 * self.N)
        self._check(self.lib.lbm_set_force_field(self.h, p, dev))

    def step(self, n: int = 1):
        self._check(self.lib.lbm_step(self.h, int(n)))

    def sync(self):
        self._check(self.lib.lbm_sync(self.h))

    @property
    def timestep(self) -> int:
        t = ctypes.c_int64()
        self._check(self.lib.lbm_get_timestep(self.h, ctypes.byref(t)))
        return t.value

    @property
    def steps_per_launch(self) -> int:
        """Time steps per kernel launch lbm_step uses now (2 or 1)."""
        n = ctypes.c_int32()
        self._check(self.lib.lbm_get_steps_per_launch(self.h, ctypes.byref(n)))
        return n.value

    def get_macros(self):
        """(rho [ny, nx], u [ny, nx, 2]) as float32 numpy arrays."""
        rho = np.empty(self.N, np.float32)
        u = np.empty(2 * self.N, np.float32)
        self._check(self.lib.lbm_get_macros(self.h, rho.ctypes.data_as(ctypes.c_void_p),
                                            u.ctypes.data_as(ctypes.c_void_p), 0))
        return rho.reshape(self.cfg.ny, self.cfg.nx), u.reshape(self.cfg.ny, self.cfg.nx, 2)

    def get_macros_device(self, rho=None, u=None):
        """Stream-ordered copy into torch CUDA tensors (rho [N], u [N*2])."""
        pr = ctypes.c_void_p(rho.data_ptr()) if rho is not None else None
        pu = ctypes.c_void_p(u.data_ptr()) if u is not None else None
        self._check(self.lib.lbm_get_macros(self.h, pr, pu, 1))

    def get_pops(self):
        f = np.empty(9 * self.N, np.float32)
        self._check(self.lib.lbm_get_pops(self.h, f.ctypes.data_as(ctypes.c_void_p), 0))
        return f.reshape(9, self.cfg.ny, self.cfg.nx)

    def set_pops(self, f):
        p, dev, keep = _ptr(f, np.float32, 9 * self.N)
        self._check(self.lib.lbm_set_pops(self.h, p, dev))

    def reduce(self, kind: int, ref=None):
        n = self.cfg.ny if kind == A.LBM_REDUCE_ROW_MEAN_UX else 3
        out = np.zeros(n, np.float64)
        p = None
        keep = None
        if ref is not None:
            keep = np.ascontiguousarray(ref, dtype=np.float32).reshape(-1)
            p = keep.ctypes.data_as(ctypes.c_void_p)
        self._check(self.lib.lbm_reduce(self.h, kind, p, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def reduce_noalloc(self, kind: int, ref=None, out=None):
        """lbm_reduce_noalloc: the same reductions without any device allocation. `ref` (for
        LBM_REDUCE_L2_VS_REF) must be a contiguous device tensor, AoS [N][2]. With `out` a
        float64 device tensor the result is written there asynchronously; else returned."""
        n = self.cfg.ny if kind == A.LBM_REDUCE_ROW_MEAN_UX else 3
        pr = None
        if ref is not None:
            if not _is_torch(ref) or not ref.is_cuda or not ref.is_contiguous():
                raise ValueError("reduce_noalloc needs the reference field as a contiguous device tensor")
            pr = ctypes.c_void_p(ref.data_ptr())
        if out is not None:
            self._check(self.lib.lbm_reduce_noalloc(self.h, kind, pr, ctypes.cast(ctypes.c_void_p(out.data_ptr()),
                                                                                 ctypes.POINTER(ctypes.c_double)), 1))
            return out
        h = np.zeros(n, np.float64)
        self._check(self.lib.lbm_reduce_noalloc(self.h, kind, pr, h.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 0))
        return h

    # -- output files (src/core/lbm.cuh:173-204, 269-343) -----------------------------------
    def save_vtk(self, directory, reference_offset: bool = False):
        from . import output

        rho, u = self.get_macros()
        return output.save_vtk(directory, self.timestep, rho, u, reference_offset)

    def save_macroscopics(self, directory):
        from . import output

        rho, u = self.get_macros()
        output.save_macroscopics(directory, self.timestep, rho, u)

    # -- immersed boundary --------------------------------------------------------------
    def ibm_set_markers(self, xy, ds=None, iters: int = 3):
        xy = np.ascontiguousarray(xy, dtype=np.float32).reshape(-1, 2)
        n = xy.shape[0]
        dsa = None if ds is None else np.ascontiguousarray(ds, dtype=np.float32).reshape(n)
        self._check(self.lib.lbm_ibm_set_markers(
            self.h, xy.ctypes.data_as(ctypes.c_void_p),
            None if dsa is None else dsa.ctypes.data_as(ctypes.c_void_p), n, iters))
        self._ibm_n = n
        self._ibm_keep = (xy, dsa)

    def ibm_get_forces(self):
        n = getattr(self, "_ibm_n", 0)
        lag = np.empty(2 * max(n, 1), np.float32)
        tot = np.zeros(2, np.float64)
        self._check(self.lib.lbm_ibm_get_forces(self.h, lag.ctypes.data_as(ctypes.c_void_p),
                                                tot.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return lag[: 2 * n].reshape(n, 2), tot

    # -- adjoint ------------------------------------------------------------------------
    def adjoint_forward(self, nsteps: int, checkpoint_every: int):
        self._check(self.lib.lbm_adjoint_forward(self.h, int(nsteps), int(checkpoint_every)))

    def adjoint_backward(self, dL_du, dL_dforce=None, dL_du0=None):
        """Host numpy in/out, or torch CUDA tensors (all of one kind)."""
        if _is_torch(dL_du):
            ptr = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
            self._check(self.lib.lbm_adjoint_backward(self.h, ptr(dL_du), ptr(dL_dforce), ptr(dL_du0), 1))
            return dL_dforce, dL_du0
        g = np.ascontiguousarray(dL_du, dtype=np.float32).reshape(-1)
        gf = np.zeros(2 * self.N, np.float32)
        gu = np.zeros(2 * self.N, np.float32)
        self._check(self.lib.lbm_adjoint_backward(self.h, g.ctypes.data_as(ctypes.c_void_p),
                                                  gf.ctypes.data_as(ctypes.c_void_p),
                                                  gu.ctypes.data_as(ctypes.c_void_p), 0))
        shp = (self.cfg.ny, self.cfg.nx, 2)
        return gf.reshape(shp), gu.reshape(shp)
