"""3-D immersed bodies from OBJ meshes and body INI files, over include/lbm_body.h.

Mirrors the reference's host pipeline for a 3-D scenario's body: mesh::MeshTransformer
(src/IBM/mesh/mesh_transformer.hpp:16-55: translate / scale / rotate about an anchor, AABB,
fits_in_domain, collect_ibm_body), conf::parse_body_config and conf::create_body_from_config
(src/IBM/config/config_parser.cpp:32-209), and the Morton order IBMManager<3>::init_and_dispatch
gives the markers (IBMManager.cuh:57-66, IBM_sorting.cuh:41-159). The markers go to
Solver3D.ibm_set_markers.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _abi as A
from .solver import LbmError


def _lib():
    return A.load_library()


def _check(rc, body=None):
    if rc != 0:
        raise LbmError(rc, _lib().lbm_body_last_error(body).decode())


def _f3(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1, 3)
    if n is not None and a.shape[0] != n:
        raise ValueError(f"expected {n} points")
    return a


class Body:
    """A triangle mesh with an anchor and markers (mesh::MeshTransformer)."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle)

    # -- construction ---------------------------------------------------------------------
    @classmethod
    def load_obj(cls, path) -> "Body":
        """mesh::load_obj + MeshTransformer(filename): anchor at the bounding-box centre."""
        h = ctypes.c_void_p()
        _check(_lib().lbm_body_load_obj(os.fsencode(path), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def from_mesh(cls, vertices, faces=None) -> "Body":
        v = _f3(vertices)
        f = None if faces is None else np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 3)
        h = ctypes.c_void_p()
        _check(_lib().lbm_body_from_mesh(v.ctypes.data_as(ctypes.c_void_p), v.shape[0],
                                         None if f is None else f.ctypes.data_as(ctypes.c_void_p),
                                         0 if f is None else f.shape[0], ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def from_config(cls, ini_path, domain, base_dir=None) -> "Body":
        """conf::create_body_from_config for a lattice domain = (nx, ny, nz), markers in the
        order IBMManager<3>::init_and_dispatch gives them. Relative paths in the file resolve
        against base_dir (None: the working directory)."""
        nx, ny, nz = (int(d) for d in domain)
        h = ctypes.c_void_p()
        _check(_lib().lbm_body_from_config(os.fsencode(ini_path), None if base_dir is None else os.fsencode(base_dir),
                                           nx, ny, nz, ctypes.byref(h)))
        return cls(h.value)

    def close(self):
        if self._h and self._h.value:
            _lib().lbm_body_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- queries --------------------------------------------------------------------------
    def _counts(self):
        nv, nf, nm = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(_lib().lbm_body_info(self._h, ctypes.byref(nv), ctypes.byref(nf), ctypes.byref(nm)), self._h)
        return nv.value, nf.value, nm.value

    @property
    def vertices(self) -> np.ndarray:
        out = np.empty((self._counts()[0], 3), np.float32)
        _check(_lib().lbm_body_vertices(self._h, out.ctypes.data_as(ctypes.c_void_p)), self._h)
        return out

    @property
    def faces(self) -> np.ndarray:
        out = np.empty((self._counts()[1], 3), np.int32)
        _check(_lib().lbm_body_faces(self._h, out.ctypes.data_as(ctypes.c_void_p) if out.size else None), self._h)
        return out

    @property
    def markers(self) -> np.ndarray:
        out = np.empty((self._counts()[2], 3), np.float32)
        _check(_lib().lbm_body_markers(self._h, out.ctypes.data_as(ctypes.c_void_p) if out.size else None), self._h)
        return out

    def get_aabb(self):
        lo, hi = np.empty(3, np.float32), np.empty(3, np.float32)
        _check(_lib().lbm_body_aabb(self._h, lo.ctypes.data_as(ctypes.c_void_p), hi.ctypes.data_as(ctypes.c_void_p)),
               self._h)
        return lo, hi

    def get_anchor(self) -> np.ndarray:
        a = np.empty(3, np.float32)
        _check(_lib().lbm_body_anchor(self._h, a.ctypes.data_as(ctypes.c_void_p)), self._h)
        return a

    def fits_in_domain(self, nx, ny, nz) -> bool:
        f = ctypes.c_int32()
        _check(_lib().lbm_body_fits_in_domain(self._h, float(nx), float(ny), float(nz), ctypes.byref(f)), self._h)
        return bool(f.value)

    def mesh_volume(self) -> float:
        v = ctypes.c_float()
        _check(_lib().lbm_body_mesh_volume(self._h, ctypes.byref(v)), self._h)
        return v.value

    # -- transforms (chainable, as the reference's) ------------------------------------------
    def translate(self, tx, ty, tz):
        _check(_lib().lbm_body_translate(self._h, tx, ty, tz), self._h)
        return self

    def scale(self, sx, sy, sz):
        _check(_lib().lbm_body_scale(self._h, sx, sy, sz), self._h)
        return self

    def scale_to_overall_size(self, size):
        _check(_lib().lbm_body_scale_to_overall_size(self._h, size), self._h)
        return self

    def rotate_x(self, rad):
        _check(_lib().lbm_body_rotate(self._h, 0, rad), self._h)
        return self

    def rotate_y(self, rad):
        _check(_lib().lbm_body_rotate(self._h, 1, rad), self._h)
        return self

    def rotate_z(self, rad):
        _check(_lib().lbm_body_rotate(self._h, 2, rad), self._h)
        return self

    def set_anchor(self, x, y, z):
        _check(_lib().lbm_body_set_anchor(self._h, x, y, z), self._h)
        return self

    def set_anchor_to_centroid(self):
        _check(_lib().lbm_body_set_anchor_to_centroid(self._h), self._h)
        return self

    def move_anchor_to_world(self, x, y, z):
        _check(_lib().lbm_body_move_anchor_to_world(self._h, x, y, z), self._h)
        return self

    def move_centroid_to_world(self, x, y, z):
        _check(_lib().lbm_body_move_centroid_to_world(self._h, x, y, z), self._h)
        return self

    def collect_file(self, path):
        """mesh::save_obj of the transformed mesh."""
        _check(_lib().lbm_body_save_obj(self._h, os.fsencode(path)), self._h)

    # -- markers ---------------------------------------------------------------------------
    def collect_ibm_body(self, target: int | None = None) -> np.ndarray:
        """Markers: every vertex (None), or the vertices reduced to `target` by weighted
        sample elimination (in vertex order)."""
        if target is None:
            _check(_lib().lbm_body_markers_from_vertices(self._h), self._h)
        else:
            _check(_lib().lbm_body_sample(self._h, int(target)), self._h)
        return self.markers

    def morton_sort(self, block_size: int = 4, bits_per_dim: int = 10) -> np.ndarray:
        _check(_lib().lbm_body_morton_sort(self._h, int(block_size), int(bits_per_dim)), self._h)
        return self.markers


def uv_sphere(n_theta: int = 64, n_phi: int = 64, radius: float = 1.0):
    """(vertices [n, 3] float32, faces [m, 3] int32) of a closed outward-oriented UV sphere: the
    two poles and n_theta - 1 rings of n_phi vertices (the shape of the reference's
    assets/sphere_unit_64x64.obj, which its tree does not ship)."""
    th = np.pi * np.arange(1, n_theta) / n_theta
    ph = 2 * np.pi * np.arange(n_phi) / n_phi
    ring = np.stack([np.outer(np.sin(th), np.cos(ph)), np.outer(np.sin(th), np.sin(ph)),
                     np.repeat(np.cos(th)[:, None], n_phi, 1)], -1).reshape(-1, 3)
    v = np.concatenate([[[0.0, 0.0, 1.0]], ring, [[0.0, 0.0, -1.0]]]) * radius
    j = np.arange(n_phi)
    f = [np.stack([np.zeros(n_phi, int), 1 + j, 1 + (j + 1) % n_phi], 1)]
    for i in range(n_theta - 2):
        a, b = 1 + i * n_phi + j, 1 + i * n_phi + (j + 1) % n_phi
        f += [np.stack([a, a + n_phi, b], 1), np.stack([b, a + n_phi, b + n_phi], 1)]
    last, base = len(v) - 1, 1 + (n_theta - 2) * n_phi
    f.append(np.stack([base + j, np.full(n_phi, last), base + (j + 1) % n_phi], 1))
    return v.astype(np.float32), np.concatenate(f).astype(np.int32)


def r_max_3d(volume: float, n: int) -> float:
    return float(_lib().lbm_body_r_max(float(volume), int(n)))


def morton_sort(points, block_size: int = 4, bits_per_dim: int = 10) -> np.ndarray:
    """block_morton_sort<dim> of a [n, 2] or [n, 3] array (a sorted copy)."""
    p = np.array(points, dtype=np.float32, copy=True, order="C")
    dim = p.shape[1]
    _check(_lib().lbm_points_morton_sort(p.ctypes.data_as(ctypes.c_void_p), p.shape[0], dim, int(block_size),
                                         int(bits_per_dim)))
    return p


def save_points_csv(path, points):
    p = _f3(points)
    _check(_lib().lbm_points_save_csv(os.fsencode(path), p.ctypes.data_as(ctypes.c_void_p), p.shape[0]))


def load_points_csv(path) -> np.ndarray:
    n = ctypes.c_int32()
    _check(_lib().lbm_points_load_csv(os.fsencode(path), None, 0, ctypes.byref(n)))
    out = np.empty((n.value, 3), np.float32)
    _check(_lib().lbm_points_load_csv(os.fsencode(path), out.ctypes.data_as(ctypes.c_void_p), n.value, ctypes.byref(n)))
    return out


def parse_body_config(ini_path) -> dict:
    """conf::parse_body_config: the file's fields, the optional ones only when set."""
    c = A.LbmBodyConfig()
    _check(_lib().lbm_body_parse_config(os.fsencode(ini_path), ctypes.byref(c)))
    out = {"obj_filepath": c.obj_filepath.decode(),
           "save_transformed_obj_to": c.save_transformed_obj_to.decode() or None,
           "sampling_type": "direct" if c.sampling_type == A.LBM_BODY_SAMPLING_DIRECT else "target_points"}
    for k, bit in A.LBM_BODY_HAS.items():
        if c.has & bit:
            out[k] = getattr(c, k)
    return out
