"""The reference's output files from numpy fields, through the writers in liblbm_hip.so
(lbm_write_vti / lbm_write_macros / lbm_write_midplane, include/lbm_hip.h):

- save_vtk            LBM::save_vtk (src/core/lbm.cuh:269-343): <dir>/vtk/sim_data_%06d.vti
- save_macroscopics   LBM::save_macroscopics (lbm.cuh:173-204): density_<t>.bin, velocity_<t>.bin
- save_midplane_slice LBM<3>::save_midplane_slice (lbm.cuh:206-267): the z = NZ/2 plane

Fields are host arrays: rho [ny, nx] or [nz, ny, nx], u [..., 2] or [..., 3] (AoS).
"""
from __future__ import annotations

import os

import numpy as np

from . import _abi as A


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _check(rc, what):
    if rc != 0:
        raise OSError(f"{what} failed ({rc})")


def save_vtk(directory, timestep: int, rho, u, reference_offset: bool = False):
    rho, u = _f32(rho), _f32(u)
    dim = u.shape[-1]
    shape = rho.shape if rho.ndim == 3 else (1,) + rho.shape
    nz, ny, nx = shape
    if u.shape[:-1] != rho.shape or dim not in (2, 3) or (dim == 2 and rho.ndim != 2):
        raise ValueError("rho [ny, nx] with u [ny, nx, 2], or rho [nz, ny, nx] with u [nz, ny, nx, 3]")
    lib = A.load_library()
    _check(lib.lbm_write_vti(os.fsencode(directory), int(timestep), nx, ny, nz, dim, rho.ctypes.data, u.ctypes.data,
                             A.LBM_VTI_REFERENCE_OFFSET if reference_offset else 0), "lbm_write_vti")
    return os.path.join(directory, "vtk", f"sim_data_{int(timestep):06d}.vti")


def save_macroscopics(directory, timestep: int, rho, u):
    rho, u = _f32(rho), _f32(u)
    lib = A.load_library()
    _check(lib.lbm_write_macros(os.fsencode(directory), int(timestep), rho.size, u.shape[-1], rho.ctypes.data,
                                u.ctypes.data), "lbm_write_macros")


def save_midplane_slice(directory, timestep: int, rho, u):
    rho, u = _f32(rho), _f32(u)
    if rho.ndim != 3 or u.shape != rho.shape + (3,):
        raise ValueError("rho [nz, ny, nx] and u [nz, ny, nx, 3]")
    nz, ny, nx = rho.shape
    lib = A.load_library()
    _check(lib.lbm_write_midplane(os.fsencode(directory), int(timestep), nx, ny, nz, rho.ctypes.data, u.ctypes.data),
           "lbm_write_midplane")
