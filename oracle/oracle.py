"""ctypes wrapper of the CPU oracle (liblbm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker. The product never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liblbm_oracle.so"
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_P = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        sig = {
            "oracle_create": (_P, [_P, ctypes.c_int]),
            "oracle_destroy": (None, [_P]),
            "oracle_set_eq_literal": (ctypes.c_int, [_P, ctypes.c_int]),
            "oracle_set_flags": (ctypes.c_int, [_P, _P]),
            "oracle_init_from_macros": (ctypes.c_int, [_P, _D, _D]),
            "oracle_set_force_field": (ctypes.c_int, [_P, _D]),
            "oracle_step": (ctypes.c_int, [_P, ctypes.c_int]),
            "oracle_get_macros": (ctypes.c_int, [_P, _D, _D]),
            "oracle_get_pops": (ctypes.c_int, [_P, _D]),
            "oracle_get_pops_back": (ctypes.c_int, [_P, _D]),
            "oracle_set_pops": (ctypes.c_int, [_P, _D, _D]),
            "oracle_get_avg": (ctypes.c_int, [_P, _D]),
            "oracle_ibm_set_markers": (ctypes.c_int, [_P, _D, _D, ctypes.c_int, ctypes.c_int]),
            "oracle_ibm_get_forces": (ctypes.c_int, [_P, _D, _D]),
            "oracle_get_force": (ctypes.c_int, [_P, _D]),
            "oracle_ibm_operators": (ctypes.c_int, [_P, _D, _D, _D, _D, ctypes.c_int, _D, _D, _D, _D, _D]),
            "oracle_last_error": (ctypes.c_char_p, [_P]),
            "oracle_adjoint": (ctypes.c_int, [_P, ctypes.c_int, _D, _D, _D]),
            "oracle_node_jacobian": (ctypes.c_int, [_P, _D, _D, _D]),
            "oracle_num_threads": (ctypes.c_int, []),
            "oracle_set_num_threads": (None, [ctypes.c_int]),
            "oracle_tables": (None, [_D, _P, _P, _P, _D, _D]),
            "oracle_equilibrium": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double, _D]),
            "oracle_cm_matrix_inverse": (None, [ctypes.c_double, ctypes.c_double, _D]),
            "oracle_delta4": (ctypes.c_double, [ctypes.c_double]),
            "oracle_adapter_rate": (ctypes.c_int, [ctypes.c_double, ctypes.c_double, ctypes.c_double, _D, _D]),
            "oracle_collide_node": (ctypes.c_int, [ctypes.c_int, _P, _D, _D, _D, _D]),
            "oracle_f32_to_f16": (ctypes.c_uint16, [ctypes.c_float]),
            "oracle_f16_to_f32": (ctypes.c_float, [ctypes.c_uint16]),
            "oracle_f32_to_bf16": (ctypes.c_uint16, [ctypes.c_float]),
            "oracle_bf16_to_f32": (ctypes.c_float, [ctypes.c_uint16]),
            # D3Q27 (lbm_oracle3d.cpp)
            "oracle3d_create": (_P, [_P, ctypes.c_int]),
            "oracle3d_destroy": (None, [_P]),
            "oracle3d_set_flags": (ctypes.c_int, [_P, _P]),
            "oracle3d_init_from_macros": (ctypes.c_int, [_P, _D, _D]),
            "oracle3d_set_force_field": (ctypes.c_int, [_P, _D]),
            "oracle3d_step": (ctypes.c_int, [_P, ctypes.c_int]),
            "oracle3d_get_macros": (ctypes.c_int, [_P, _D, _D]),
            "oracle3d_get_pops": (ctypes.c_int, [_P, _D]),
            "oracle3d_set_pops": (ctypes.c_int, [_P, _D]),
            "oracle3d_set_pops_back": (ctypes.c_int, [_P, _D]),
            "oracle3d_ibm_set_markers": (ctypes.c_int, [_P, _D, _D, ctypes.c_int, ctypes.c_int]),
            "oracle3d_ibm_get_forces": (ctypes.c_int, [_P, _D, _D]),
            "oracle3d_stream_only": (ctypes.c_int, [_P]),
            "oracle3d_stream_boundaries": (ctypes.c_int, [_P]),
            "oracle3d_last_error": (ctypes.c_char_p, [_P]),
            "oracle3d_tables": (None, [_D, _P, _P, _P, _P]),
            "oracle3d_equilibrium": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_double, _D]),
            "oracle3d_guo_source": (None, [ctypes.c_double] * 7 + [_D]),
            "oracle3d_mrt_basis": (None, [_D, _D]),
            "oracle3d_collide_cm": (None, [_D, _D, _D, ctypes.c_double, _D]),
            "oracle3d_collide_mrt": (None, [_D, _D, _D, ctypes.POINTER(ctypes.c_float), ctypes.c_int, _D]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_D)


def tables():
    L = lib()
    w = np.zeros(9)
    cx = np.zeros(9, np.int32)
    cy = np.zeros(9, np.int32)
    opp = np.zeros(9, np.int32)
    M = np.zeros(81)
    Mi = np.zeros(81)
    L.oracle_tables(_dp(w), cx.ctypes.data_as(_P), cy.ctypes.data_as(_P), opp.ctypes.data_as(_P), _dp(M), _dp(Mi))
    return w, cx, cy, opp, M.reshape(9, 9), Mi.reshape(9, 9)


def equilibrium(rho, ux, uy, prec=64, literal=False):
    out = np.zeros(9)
    lib().oracle_equilibrium(prec, int(literal), rho, ux, uy, _dp(out))
    return out


def cm_matrix_inverse(ux, uy):
    out = np.zeros(81)
    lib().oracle_cm_matrix_inverse(ux, uy, _dp(out))
    return out.reshape(9, 9)


def delta4(r):
    return lib().oracle_delta4(float(r))


def ibm_operators(cfg_c, rho, u, xy, ds, F_lag):
    """One IBM iteration's operators on given fields (FP64): returns (u_lag [2, n],
    rho_lag [n], F_lag_out [2, n] = Lagrangian force of u_lag, F_eul [2, ny, nx] = spread of
    F_lag [2, n]). rho [ny, nx], u [2, ny, nx], xy [n, 2]."""
    n = len(xy)
    N = cfg_c.nx * cfg_c.ny
    r = np.ascontiguousarray(rho, np.float64).reshape(-1)
    uu = np.ascontiguousarray(u, np.float64).reshape(-1)
    p = np.ascontiguousarray(xy, np.float64).reshape(-1)
    d = np.ascontiguousarray(ds, np.float64).reshape(-1)
    fl = np.ascontiguousarray(F_lag, np.float64).reshape(-1)
    ul, rl, flo, fe = np.zeros(2 * n), np.zeros(n), np.zeros(2 * n), np.zeros(2 * N)
    rc = lib().oracle_ibm_operators(ctypes.byref(cfg_c), _dp(r), _dp(uu), _dp(p), _dp(d), n, _dp(fl), _dp(ul),
                                    _dp(rl), _dp(flo), _dp(fe))
    if rc:
        raise RuntimeError("oracle_ibm_operators failed")
    return ul.reshape(2, n), rl, flo.reshape(2, n), fe.reshape(2, cfg_c.ny, cfg_c.nx)


def adapter_rate(rho, j, pi, avg):
    out = np.zeros(1)
    a = np.asarray(avg, np.float64)
    lib().oracle_adapter_rate(rho, j, pi, _dp(a), _dp(out))
    return out[0]


def collide_node(cfg_c, g, F=(0.0, 0.0), avg=None, prec=64):
    out = np.zeros(9)
    g = np.asarray(g, np.float64)
    Fa = np.asarray(F, np.float64)
    av = None if avg is None else np.asarray(avg, np.float64)
    lib().oracle_collide_node(prec, ctypes.byref(cfg_c), _dp(g), _dp(Fa), None if av is None else _dp(av), _dp(out))
    return out


class Oracle:
    """One oracle solver (precision 32 or 64) following the reference's stage order."""

    def __init__(self, cfg_c, precision=64, eq_literal=False):
        self.L = lib()
        self.cfg = cfg_c
        self.nx, self.ny = cfg_c.nx, cfg_c.ny
        self.N = self.nx * self.ny
        self.h = self.L.oracle_create(ctypes.byref(cfg_c), precision)
        if not self.h:
            raise RuntimeError("oracle_create failed")
        if eq_literal:
            self.L.oracle_set_eq_literal(self.h, 1)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.oracle_destroy(self.h)
            self.h = None

    def _ck(self, rc):
        if rc != 0:
            raise RuntimeError(self.L.oracle_last_error(self.h).decode())

    def set_flags(self, flags):
        f = np.ascontiguousarray(flags, np.uint8).reshape(-1)
        self._ck(self.L.oracle_set_flags(self.h, f.ctypes.data_as(_P)))

    def init_from_macros(self, rho, u):
        r = np.ascontiguousarray(rho, np.float64).reshape(-1)
        uu = np.ascontiguousarray(u, np.float64).reshape(-1)
        self._ck(self.L.oracle_init_from_macros(self.h, _dp(r), _dp(uu)))

    def set_force_field(self, F):
        if F is None:
            self._ck(self.L.oracle_set_force_field(self.h, None))
        else:
            a = np.ascontiguousarray(F, np.float64).reshape(-1)
            self._ck(self.L.oracle_set_force_field(self.h, _dp(a)))

    def step(self, n=1):
        self._ck(self.L.oracle_step(self.h, int(n)))

    def get_macros(self):
        r = np.zeros(self.N)
        u = np.zeros(2 * self.N)
        self._ck(self.L.oracle_get_macros(self.h, _dp(r), _dp(u)))
        return r.reshape(self.ny, self.nx), u.reshape(self.ny, self.nx, 2)

    def get_pops(self):
        f = np.zeros(9 * self.N)
        self._ck(self.L.oracle_get_pops(self.h, _dp(f)))
        return f.reshape(9, self.ny, self.nx)

    def get_pops_back(self):
        f = np.zeros(9 * self.N)
        self._ck(self.L.oracle_get_pops_back(self.h, _dp(f)))
        return f.reshape(9, self.ny, self.nx)

    def set_pops(self, f, f_back=None):
        a = np.ascontiguousarray(f, np.float64).reshape(-1)
        b = None if f_back is None else np.ascontiguousarray(f_back, np.float64).reshape(-1)
        self._ck(self.L.oracle_set_pops(self.h, _dp(a), None if b is None else _dp(b)))

    def get_avg(self):
        a = np.zeros(3)
        self._ck(self.L.oracle_get_avg(self.h, _dp(a)))
        return a

    def get_force(self):
        F = np.zeros(2 * self.N)
        self._ck(self.L.oracle_get_force(self.h, _dp(F)))
        return F.reshape(self.ny, self.nx, 2)

    def adjoint(self, T, dLdu):
        """FP64 adjoint over T steps from the freshly initialised state: (dL/dF, dL/du0)."""
        g = np.ascontiguousarray(dLdu, np.float64).reshape(-1)
        gF = np.zeros(2 * self.N)
        gu = np.zeros(2 * self.N)
        self._ck(self.L.oracle_adjoint(self.h, int(T), _dp(g), _dp(gF), _dp(gu)))
        return gF.reshape(self.ny, self.nx, 2), gu.reshape(self.ny, self.nx, 2)

    def ibm_set_markers(self, xy, ds=None, iters=3):
        a = np.ascontiguousarray(xy, np.float64).reshape(-1)
        n = a.size // 2
        d = None if ds is None else np.ascontiguousarray(ds, np.float64).reshape(-1)
        self._ck(self.L.oracle_ibm_set_markers(self.h, _dp(a), None if d is None else _dp(d), n, iters))
        self._n = n

    def ibm_get_forces(self):
        lag = np.zeros(2 * self._n)
        tot = np.zeros(2)
        self._ck(self.L.oracle_ibm_get_forces(self.h, _dp(lag), _dp(tot)))
        return lag.reshape(self._n, 2), tot


def node_jacobian(cfg_c, g, F=(0.0, 0.0)):
    J = np.zeros(99)
    gg = np.asarray(g, np.float64)
    FF = np.asarray(F, np.float64)
    lib().oracle_node_jacobian(ctypes.byref(cfg_c), _dp(gg), _dp(FF), _dp(J))
    return J.reshape(9, 11)


def num_threads():
    return lib().oracle_num_threads()


def set_num_threads(n):
    lib().oracle_set_num_threads(int(n))


# ---- D3Q27 -------------------------------------------------------------------------------

def tables3d():
    """(w, cx, cy, cz, opp) of the 27-velocity set, reference order."""
    w = np.zeros(27)
    c = [np.zeros(27, np.int32) for _ in range(4)]
    lib().oracle3d_tables(_dp(w), *[a.ctypes.data_as(_P) for a in c])
    return (w, *c)


def equilibrium3d(rho, ux, uy, uz, prec=64, literal=False):
    out = np.zeros(27)
    lib().oracle3d_equilibrium(prec, int(literal), rho, ux, uy, uz, _dp(out))
    return out


def guo_source3d(u, F, omega):
    out = np.zeros(27)
    lib().oracle3d_guo_source(*[float(v) for v in u], *[float(v) for v in F], float(omega), _dp(out))
    return out


def mrt3_basis():
    M = np.zeros(27 * 27)
    Mi = np.zeros(27 * 27)
    lib().oracle3d_mrt_basis(_dp(M), _dp(Mi))
    return M.reshape(27, 27), Mi.reshape(27, 27)


def collide_mrt3(g, u, F, S, literal=False):
    """One node through MRT<3> (FP64): g post-stream populations, u corrected velocity;
    literal = the reference's momentum and moment-19 forms (LBM_QUIRK_MRT3_LITERAL)."""
    gg = np.ascontiguousarray(g, np.float64)
    uu = np.ascontiguousarray(u, np.float64)
    FF = np.ascontiguousarray(F, np.float64)
    SS = (ctypes.c_float * 27)(*[float(v) for v in S])
    out = np.zeros(27)
    lib().oracle3d_collide_mrt(_dp(gg), _dp(uu), _dp(FF), SS, int(literal), _dp(out))
    return out


def collide_cm3(g, u, F, omega):
    """One node through CM<3, NoAdapter> (FP64): g post-stream populations, u corrected velocity."""
    gg = np.ascontiguousarray(g, np.float64)
    uu = np.ascontiguousarray(u, np.float64)
    FF = np.ascontiguousarray(F, np.float64)
    out = np.zeros(27)
    lib().oracle3d_collide_cm(_dp(gg), _dp(uu), _dp(FF), float(omega), _dp(out))
    return out


class Oracle3D:
    """One D3Q27 oracle solver (precision 32 or 64) in the reference's stage order.

    Arrays are indexed [z, y, x] (node = (z*NY + y)*NX + x); populations [27, z, y, x]."""

    def __init__(self, cfg_c, precision=64):
        self.L = lib()
        self.cfg = cfg_c
        self.nx, self.ny, self.nz = cfg_c.nx, cfg_c.ny, cfg_c.nz
        self.N = self.nx * self.ny * self.nz
        self.shape = (self.nz, self.ny, self.nx)
        self.h = self.L.oracle3d_create(ctypes.byref(cfg_c), precision)
        if not self.h:
            raise RuntimeError("oracle3d_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.oracle3d_destroy(self.h)
            self.h = None

    def _ck(self, rc):
        if rc != 0:
            raise RuntimeError(self.L.oracle3d_last_error(self.h).decode())

    def set_flags(self, flags):
        f = np.ascontiguousarray(flags, np.uint8).reshape(-1)
        self._ck(self.L.oracle3d_set_flags(self.h, f.ctypes.data_as(_P)))

    def init_from_macros(self, rho, u):
        r = np.ascontiguousarray(rho, np.float64).reshape(-1)
        uu = np.ascontiguousarray(u, np.float64).reshape(-1)
        self._ck(self.L.oracle3d_init_from_macros(self.h, _dp(r), _dp(uu)))

    def set_force_field(self, F):
        if F is None:
            self._ck(self.L.oracle3d_set_force_field(self.h, None))
        else:
            a = np.ascontiguousarray(F, np.float64).reshape(-1)
            self._ck(self.L.oracle3d_set_force_field(self.h, _dp(a)))

    def step(self, n=1):
        self._ck(self.L.oracle3d_step(self.h, int(n)))

    def stream_only(self):
        self._ck(self.L.oracle3d_stream_only(self.h))

    def stream_boundaries(self):
        """Streaming and the boundary functors, no collision."""
        self._ck(self.L.oracle3d_stream_boundaries(self.h))

    def get_macros(self):
        r = np.zeros(self.N)
        u = np.zeros(3 * self.N)
        self._ck(self.L.oracle3d_get_macros(self.h, _dp(r), _dp(u)))
        return r.reshape(self.shape), u.reshape(*self.shape, 3)

    def get_pops(self):
        f = np.zeros(27 * self.N)
        self._ck(self.L.oracle3d_get_pops(self.h, _dp(f)))
        return f.reshape(27, *self.shape)

    def ibm_set_markers(self, xyz, ds=None, iters=3):
        a = np.ascontiguousarray(xyz, np.float64).reshape(-1)
        n = a.size // 3
        d = None if ds is None else np.ascontiguousarray(ds, np.float64).reshape(n)
        self._ck(self.L.oracle3d_ibm_set_markers(self.h, _dp(a), None if d is None else _dp(d), n, iters))
        self._n = n

    def ibm_get_forces(self):
        lag = np.zeros(3 * self._n)
        tot = np.zeros(3)
        self._ck(self.L.oracle3d_ibm_get_forces(self.h, _dp(lag), _dp(tot)))
        return lag.reshape(self._n, 3), tot

    def set_pops(self, f, f_back=None):
        """Both buffers = f; with f_back, the second buffer (the stale slots streaming leaves
        at non-periodic edges) takes f_back."""
        a = np.ascontiguousarray(f, np.float64).reshape(-1)
        self._ck(self.L.oracle3d_set_pops(self.h, _dp(a)))
        if f_back is not None:
            b = np.ascontiguousarray(f_back, np.float64).reshape(-1)
            self._ck(self.L.oracle3d_set_pops_back(self.h, _dp(b)))
