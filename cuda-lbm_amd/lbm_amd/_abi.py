"""ctypes view of include/lbm_hip.h and the loader of liblbm_hip.so.

The library is the product: if it is missing this module raises, it never falls back to
anything else.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "liblbm_hip.so"

# status codes
LBM_OK = 0
LBM_E_INVALID = -1
LBM_E_UNSUPPORTED = -2
LBM_E_HIP = -3
LBM_E_NOMEM = -4
LBM_E_STATE = -5

# collision operators / adapters / storage
LBM_BGK, LBM_MRT, LBM_CM = 0, 1, 2
LBM_NO_ADAPTER, LBM_OPTIMAL_ADAPTER = 0, 1
LBM_ADAPTER_SAME_STEP, LBM_ADAPTER_LAGGED = 0, 1
LBM_F32, LBM_F16_SHIFTED, LBM_BF16_SHIFTED = 0, 1, 2
LBM_VTI_REFERENCE_OFFSET = 1

# boundary flags (src/core/lbm_constants.cuh:377-397 order, + MOVING_WALL_BOUNCE_BACK)
FLUID = 0
BOUNCE_BACK = 1
ZOU_HE_TOP = 2
ZOU_HE_LEFT = 3
ZOU_HE_TOP_LEFT_TOP_INFLOW = 4
ZOU_HE_TOP_RIGHT_TOP_INFLOW = 5
CYLINDER = 6
ZG_OUTFLOW = 7
PRESSURE_OUTLET = 8
REGULARIZED_INLET_TOP = 9
REGULARIZED_INLET_LEFT = 10
REGULARIZED_BOUNCE_BACK = 11
REGULARIZED_BOUNCE_BACK_CORNER = 12
EXTRAPOLATED_CORNER_EDGE = 13
CORNER_EDGE_BOUNCE_BACK = 14
PRESSURE_INLET_LEFT = 15
GUO_VELOCITY_INLET = 16
GUO_PRESSURE_OUTLET = 17
REGULARIZED_OUTLET = 18
MOVING_WALL_BOUNCE_BACK = 19

# quirks
LBM_QUIRK_STALE_REST = 1 << 0
LBM_QUIRK_ZOUHE_RHO = 1 << 1
LBM_QUIRK_MRT_FORCE_ROWS = 1 << 2
LBM_QUIRK_IBM_2X2 = 1 << 3
LBM_QUIRK_IBM_CLAMP = 1 << 4
LBM_QUIRK_IBM_NO_DS = 1 << 5
LBM_QUIRK_EQ3D_LINEAR = 1 << 6   # include/lbm3d_hip.h
LBM_QUIRK_MRT3_LITERAL = 1 << 7
LBM_QUIRK_GUO3D_OWN_MOMENTS = 1 << 8
LBM_QUIRK_CORNER3D_UZ = 1 << 9

# reductions
LBM_REDUCE_MASS = 0
LBM_REDUCE_MOMENTUM = 1
LBM_REDUCE_L2_VS_REF = 2
LBM_REDUCE_ROW_MEAN_UX = 3
LBM_REDUCE_MOMENT_AVG = 4


class LbmConfig(ctypes.Structure):
    """struct lbm_config (include/lbm_hip.h)."""

    _fields_ = [
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("collision", ctypes.c_int32),
        ("adapter", ctypes.c_int32),
        ("adapter_mode", ctypes.c_int32),
        ("omega", ctypes.c_float),
        ("S", ctypes.c_float * 9),
        ("rho0", ctypes.c_float),
        ("storage", ctypes.c_int32),
        ("periodic_x", ctypes.c_int32),
        ("periodic_y", ctypes.c_int32),
        ("force", ctypes.c_float * 2),
        ("wall_u", ctypes.c_float * 2),
        ("u_lid", ctypes.c_float),
        ("quirks", ctypes.c_uint32),
        ("steps_per_launch", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 7),
    ]


class Lbm3dConfig(ctypes.Structure):
    """struct lbm3d_config (include/lbm3d_hip.h)."""

    _fields_ = [
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32),
        ("collision", ctypes.c_int32),
        ("adapter", ctypes.c_int32),
        ("omega", ctypes.c_float),
        ("S", ctypes.c_float * 27),
        ("rho0", ctypes.c_float),
        ("storage", ctypes.c_int32),
        ("periodic_x", ctypes.c_int32),
        ("periodic_y", ctypes.c_int32),
        ("periodic_z", ctypes.c_int32),
        ("force", ctypes.c_float * 3),
        ("u_inlet", ctypes.c_float),
        ("quirks", ctypes.c_uint32),
        ("reserved", ctypes.c_int32 * 8),
    ]


LBM3D_EMPIRICAL_ADAPTER = 2  # CM<3> adapter ids (include/lbm3d_hip.h); only LBM_NO_ADAPTER is built

# 3-D reductions (include/lbm3d_hip.h)
LBM3D_REDUCE_MASS = 0
LBM3D_REDUCE_MOMENTUM = 1
LBM3D_REDUCE_KINETIC_ENERGY = 2


class LbmBodyConfig(ctypes.Structure):
    """struct lbm_body_config (include/lbm_body.h)."""
    _fields_ = [("obj_filepath", ctypes.c_char * 1024), ("save_transformed_obj_to", ctypes.c_char * 1024),
                ("has", ctypes.c_uint32)] + [(k, ctypes.c_float) for k in (
                    "scale_overall_size", "scale_x", "scale_y", "scale_z", "rotate_x_deg", "rotate_y_deg",
                    "rotate_z_deg", "set_anchor_local_x", "set_anchor_local_y", "set_anchor_local_z",
                    "position_anchor_x", "position_anchor_y", "position_anchor_z")] + [
                ("sampling_type", ctypes.c_int32), ("target_points", ctypes.c_int32)]


# body configuration (include/lbm_body.h)
LBM_BODY_SAMPLING_DIRECT = 0
LBM_BODY_SAMPLING_TARGET_POINTS = 1
LBM_BODY_HAS = {k: 1 << i for i, k in enumerate((
    "scale_overall_size", "scale_x", "scale_y", "scale_z", "rotate_x_deg", "rotate_y_deg", "rotate_z_deg",
    "set_anchor_local_x", "set_anchor_local_y", "set_anchor_local_z", "position_anchor_x", "position_anchor_y",
    "position_anchor_z", "target_points"))}

# (name, restype, argtypes) of every entry point include/lbm_hip.h, include/lbm3d_hip.h and
# include/lbm_body.h declare.
_P = ctypes.c_void_p
_F = ctypes.POINTER(ctypes.c_float)
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.c_int32
PROTOTYPES = [
    ("lbm_config_default", _I, [ctypes.POINTER(LbmConfig), _I, _I]),
    ("lbm_create", _I, [ctypes.POINTER(LbmConfig), ctypes.POINTER(_P)]),
    ("lbm_destroy", _I, [_P]),
    ("lbm_set_stream", _I, [_P, _P]),
    ("lbm_set_flags", _I, [_P, _P, _I]),
    ("lbm_init_from_macros", _I, [_P, _P, _P, _I]),
    ("lbm_set_force_field", _I, [_P, _P, _I]),
    ("lbm_step", _I, [_P, _I]),
    ("lbm_sync", _I, [_P]),
    ("lbm_get_timestep", _I, [_P, ctypes.POINTER(ctypes.c_int64)]),
    ("lbm_get_steps_per_launch", _I, [_P, ctypes.POINTER(ctypes.c_int32)]),
    ("lbm_get_macros", _I, [_P, _P, _P, _I]),
    ("lbm_get_pops", _I, [_P, _P, _I]),
    ("lbm_set_pops", _I, [_P, _P, _I]),
    ("lbm_reduce", _I, [_P, _I, _P, _D]),
    ("lbm_reduce_noalloc", _I, [_P, _I, _P, _D, _I]),
    ("lbm_ibm_set_markers", _I, [_P, _P, _P, _I, _I]),
    ("lbm_ibm_get_forces", _I, [_P, _P, _D]),
    ("lbm_adjoint_forward", _I, [_P, _I, _I]),
    ("lbm_adjoint_backward", _I, [_P, _P, _P, _P, _I]),
    ("lbm_write_vti", _I, [ctypes.c_char_p, ctypes.c_int64, _I, _I, _I, _I, _P, _P, ctypes.c_uint32]),
    ("lbm_write_macros", _I, [ctypes.c_char_p, ctypes.c_int64, ctypes.c_int64, _I, _P, _P]),
    ("lbm_write_midplane", _I, [ctypes.c_char_p, ctypes.c_int64, _I, _I, _I, _P, _P]),
    ("lbm_last_error", ctypes.c_char_p, [_P]),
    ("lbm_version", ctypes.c_char_p, []),
    # include/lbm3d_hip.h
    ("lbm3d_config_default", _I, [ctypes.POINTER(Lbm3dConfig), _I, _I, _I]),
    ("lbm3d_create", _I, [ctypes.POINTER(Lbm3dConfig), ctypes.POINTER(_P)]),
    ("lbm3d_destroy", _I, [_P]),
    ("lbm3d_set_stream", _I, [_P, _P]),
    ("lbm3d_set_flags", _I, [_P, _P, _I]),
    ("lbm3d_init_from_macros", _I, [_P, _P, _P, _I]),
    ("lbm3d_set_force_field", _I, [_P, _P, _I]),
    ("lbm3d_step", _I, [_P, _I]),
    ("lbm3d_sync", _I, [_P]),
    ("lbm3d_get_timestep", _I, [_P, ctypes.POINTER(ctypes.c_int64)]),
    ("lbm3d_get_macros", _I, [_P, _P, _P, _I]),
    ("lbm3d_get_pops", _I, [_P, _P, _I]),
    ("lbm3d_set_pops", _I, [_P, _P, _I]),
    ("lbm3d_reduce", _I, [_P, _I, _D]),
    ("lbm3d_ibm_set_markers", _I, [_P, _P, _P, _I, _I]),
    ("lbm3d_ibm_get_forces", _I, [_P, _P, _D]),
    # include/lbm_body.h
    ("lbm_body_load_obj", _I, [ctypes.c_char_p, ctypes.POINTER(_P)]),
    ("lbm_body_from_mesh", _I, [_P, _I, _P, _I, ctypes.POINTER(_P)]),
    ("lbm_body_destroy", _I, [_P]),
    ("lbm_body_info", _I, [_P, _P, _P, _P]),
    ("lbm_body_vertices", _I, [_P, _P]),
    ("lbm_body_faces", _I, [_P, _P]),
    ("lbm_body_markers", _I, [_P, _P]),
    ("lbm_body_translate", _I, [_P, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    ("lbm_body_scale", _I, [_P, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    ("lbm_body_scale_to_overall_size", _I, [_P, ctypes.c_float]),
    ("lbm_body_rotate", _I, [_P, _I, ctypes.c_float]),
    ("lbm_body_set_anchor", _I, [_P, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    ("lbm_body_set_anchor_to_centroid", _I, [_P]),
    ("lbm_body_move_anchor_to_world", _I, [_P, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    ("lbm_body_move_centroid_to_world", _I, [_P, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    ("lbm_body_aabb", _I, [_P, _P, _P]),
    ("lbm_body_anchor", _I, [_P, _P]),
    ("lbm_body_fits_in_domain", _I, [_P, ctypes.c_float, ctypes.c_float, ctypes.c_float, _P]),
    ("lbm_body_save_obj", _I, [_P, ctypes.c_char_p]),
    ("lbm_body_mesh_volume", _I, [_P, _P]),
    ("lbm_body_r_max", ctypes.c_float, [ctypes.c_float, _I]),
    ("lbm_body_sample", _I, [_P, _I]),
    ("lbm_body_markers_from_vertices", _I, [_P]),
    ("lbm_body_morton_sort", _I, [_P, _I, _I]),
    ("lbm_points_morton_sort", _I, [_P, _I, _I, _I, _I]),
    ("lbm_points_save_csv", _I, [ctypes.c_char_p, _P, _I]),
    ("lbm_points_load_csv", _I, [ctypes.c_char_p, _P, _I, _P]),
    ("lbm_body_parse_config", _I, [ctypes.c_char_p, ctypes.POINTER(LbmBodyConfig)]),
    ("lbm_body_from_config", _I, [ctypes.c_char_p, ctypes.c_char_p, _I, _I, _I, ctypes.POINTER(_P)]),
    ("lbm_body_last_error", ctypes.c_char_p, [_P]),
    ("lbm3d_last_error", ctypes.c_char_p, [_P]),
]

_lib = None


def _bind_torch_runtime():
    """Load PyTorch's HIP runtime first when PyTorch is installed.

    The library needs `libamdhip64.so.7`. PyTorch-ROCm ships its own copy under that SONAME,
    but its modules ask for it as `libamdhip64.so`, a name the dynamic loader does not match
    against a copy loaded from /opt/rocm. Loaded first, the library therefore pulls in
    /opt/rocm's HIP and HSA runtimes, and a later `import torch` adds PyTorch's: two runtimes
    that each drive the GPU from one process, and tensor pointers of one passed to the other.
    Importing torch first makes the library bind to the copy already loaded."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def hip_runtimes_loaded() -> list:
    """Paths of the HIP runtime libraries mapped into this process (Linux)."""
    try:
        with open("/proc/self/maps") as fh:
            return sorted({ln.split()[-1] for ln in fh if "libamdhip64" in ln and "/" in ln})
    except OSError:
        return []


def _check_one_hip_runtime():
    libs = hip_runtimes_loaded()
    if len(libs) > 1:
        raise RuntimeError("more than one HIP runtime is loaded into this process (" + ", ".join(libs) +
                           "); load liblbm_hip.so before any other HIP library, or import torch first")


def load_library(path: str | os.PathLike | None = None) -> ctypes.CDLL:
    """Load liblbm_hip.so (built by `make -C cuda-lbm_amd`) and declare its prototypes."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"liblbm_hip.so not found at {p}; build it with `make -C cuda-lbm_amd`")
    _bind_torch_runtime()
    lib = ctypes.CDLL(str(p))
    _check_one_hip_runtime()
    for name, res, args in PROTOTYPES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib
