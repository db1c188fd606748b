"""ctypes binding of the in-tree HIP library ``_jt.so`` (C ABI in
``include/jaxtrace_amd.h``).

``torch`` is imported before the library is loaded so that the HIP runtime
torch already holds is the one the library binds to (one runtime, one set of
device pointers and streams per process).  There is no CPU fallback: if the
library or a GPU is missing, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import torch  # noqa: F401  (must precede loading _jt.so)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("JAXTRACE_AMD_LIB") or os.path.join(_HERE, "_jt.so")

JT_OK, JT_EINVAL, JT_ENOMEM, JT_EHIP, JT_EUNSUPPORTED = 0, 1, 2, 3, 4

JT_EULER, JT_RK2, JT_RK4 = 0, 1, 2
JT_DTYPE_F32, JT_DTYPE_BF16, JT_DTYPE_F16 = 0, 1, 2
JT_BOUNDS = {"clamp": 0, "zero": 1, "nan": 2, "periodic": 3}
JT_TIME = {"linear": 0, "nearest": 1}
JT_EXTRAP = {"constant": 0, "linear": 1, "nan": 2, "zero": 3}
JT_BC_NONE, JT_BC_PERIODIC, JT_BC_REFLECTIVE, JT_BC_CLAMP, JT_BC_ABSORB = 0, 1, 2, 3, 4
JT_INLET_OFF, JT_INLET_GRID, JT_INLET_ZERO, JT_INLET_HOST = 0, 1, 2, 3

c_float3 = C.c_float * 3
c_int3 = C.c_int32 * 3


class GridField(C.Structure):
    _fields_ = [
        ("data", C.c_void_p), ("times", C.c_void_p), ("n_snap", C.c_int32),
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("dtype", C.c_int32), ("ncomp", C.c_int32), ("z_halo", C.c_int32),
        ("origin", c_float3), ("spacing", c_float3),
        ("bounds_mode", C.c_int32), ("time_interp", C.c_int32), ("time_extrap", C.c_int32),
        ("pair_data", C.c_void_p),
    ]


class Boundary(C.Structure):
    _fields_ = [
        ("axis_kind", c_int3), ("lo", c_float3), ("hi", c_float3),
        ("absorb_rows", C.c_int32), ("outside_value", C.c_float),
        ("inlet", C.c_int32), ("inlet_axis", C.c_int32), ("inlet_positive", C.c_int32),
        ("inlet_backstop", C.c_float),
        ("inlet_coords", C.c_void_p * 2), ("inlet_ncoords", C.c_int32 * 2),
    ]


class HashGridDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("key_mode", C.c_int32), ("origin", c_float3), ("inv", C.c_float),
        ("dims", c_int3), ("n_points", C.c_int64), ("order", C.c_void_p), ("sorted_pos", C.c_void_p),
        ("cell_start", C.c_void_p), ("cell_end", C.c_void_p),
    ]


JT_KEY_X32, JT_KEY_X64 = 0, 1
JT_BOUNDS_WS_BYTES = 24576          # include/jaxtrace_amd.h JT_BOUNDS_WS_BYTES
JT_SPH = {"cubic_spline": 0, "wendland": 1, "gaussian": 2}


class OctreeField(C.Structure):
    _fields_ = [
        ("points", C.c_void_p), ("conn", C.c_void_p), ("element_bounds", C.c_void_p),
        ("nodes_min", C.c_void_p), ("nodes_max", C.c_void_p), ("nodes_elements", C.c_void_p),
        ("nodes_elem_counts", C.c_void_p), ("nodes_children", C.c_void_p), ("nodes_is_leaf", C.c_void_p),
        ("n_points", C.c_int64), ("n_elems", C.c_int64), ("n_nodes", C.c_int64),
        ("max_elems", C.c_int32), ("max_depth", C.c_int32),
        ("data", C.c_void_p), ("times", C.c_void_p),
        ("n_snap", C.c_int32), ("time_interp", C.c_int32), ("time_extrap", C.c_int32),
        ("element_records", C.c_void_p),
    ]


class NodesField(C.Structure):
    _fields_ = [
        ("node_pos", C.c_void_p), ("n_nodes", C.c_int64), ("data", C.c_void_p), ("times", C.c_void_p),
        ("n_snap", C.c_int32), ("time_interp", C.c_int32), ("time_extrap", C.c_int32),
    ]


class TrackOpts(C.Structure):
    _fields_ = [("steps_per_launch", C.c_int32), ("sort_every", C.c_int32), ("block", C.c_int32),
                ("n_launch_events", C.c_int32), ("launch_events", C.POINTER(C.c_void_p))]


_lib = None

# name -> (restype, argtypes); every symbol the header declares
_SIGS = {
    "jt_last_error": (C.c_char_p, []),
    "jt_version": (C.c_int32, []),
    "jt_random_seeds": (C.c_int, [C.c_int64, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "jt_inlet_grid_particles": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_float, C.c_void_p, C.c_int32,
                                          C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "jt_octree_build_device_workspace_size": (C.c_size_t, [C.c_int64, C.c_int64]),
    "jt_octree_build_device": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int32, C.c_int32,
                                         C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                         C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                         C.c_void_p]),
    "jt_octree_export_device": (C.c_int, [C.c_void_p, C.c_size_t, C.c_int64, C.c_int64, C.c_int64, C.c_int32,
                                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p]),
    "jt_grid_seeds": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32,
                                C.c_void_p, C.c_void_p]),
    "jt_grid_sample": (C.c_int, [C.POINTER(GridField), C.c_void_p, C.c_int64, C.c_float,
                                 C.c_void_p, C.c_void_p]),
    "jt_grid_pair_records": (C.c_int, [C.POINTER(GridField), C.c_void_p, C.c_void_p]),
    "jt_octree_element_records": (C.c_int, [C.POINTER(OctreeField), C.c_void_p, C.c_void_p]),
    "jt_apply_boundary": (C.c_int, [C.POINTER(Boundary), C.c_void_p, C.c_int64, C.c_void_p]),
    "jt_inlet_exits_workspace_size": (C.c_size_t, [C.c_int64]),
    "jt_inlet_exits": (C.c_int, [C.POINTER(Boundary), C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_size_t, C.c_void_p]),
    "jt_inlet_place": (C.c_int, [C.POINTER(Boundary), C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                                 C.c_int64, C.c_void_p]),
    "jt_track_grid_workspace_size": (C.c_size_t, [C.c_int64, C.c_int32, C.POINTER(TrackOpts)]),
    "jt_track_grid": (C.c_int, [C.POINTER(GridField), C.POINTER(Boundary), C.c_int32, C.c_void_p,
                                C.c_int64, C.c_void_p, C.c_float, C.c_int32, C.c_int32, C.c_void_p,
                                C.POINTER(TrackOpts), C.c_void_p, C.c_size_t, C.c_void_p]),
    "jt_track_grid_noalloc": (C.c_int, [C.POINTER(GridField), C.POINTER(Boundary), C.c_int32, C.c_void_p,
                                        C.c_int64, C.c_void_p, C.c_float, C.c_int32, C.c_int32, C.c_void_p,
                                        C.POINTER(TrackOpts), C.c_void_p]),
    "jt_points_bounds": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p,
                                   C.c_size_t, C.c_void_p]),
    "jt_hashgrid_num_cells": (C.c_int64, [C.POINTER(HashGridDesc)]),
    "jt_hashgrid_workspace_size": (C.c_size_t, [C.c_int64]),
    "jt_hashgrid_build": (C.c_int, [C.POINTER(HashGridDesc), C.c_void_p, C.c_void_p, C.c_size_t,
                                    C.c_void_p]),
    "jt_hashgrid_query": (C.c_int, [C.POINTER(HashGridDesc), C.c_void_p, C.c_int64, C.c_float,
                                    C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "jt_sph_eval": (C.c_int, [C.POINTER(HashGridDesc), C.c_void_p, C.c_int64, C.c_float, C.c_int32,
                              C.c_int32, C.c_void_p, C.c_void_p, C.c_float, C.c_int32, C.c_void_p,
                              C.c_void_p]),
    "jt_sph_tiled_workspace_size": (C.c_size_t, [C.POINTER(HashGridDesc), C.c_int64]),
    "jt_sph_eval_tiled": (C.c_int, [C.POINTER(HashGridDesc), C.c_void_p, C.c_int64, C.c_float, C.c_int32,
                                    C.c_int32, C.c_float, C.c_int32, C.c_void_p, C.c_void_p, C.c_size_t,
                                    C.c_void_p]),
    "jt_kde_eval": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_int64, C.c_float,
                              C.c_void_p, C.c_void_p]),
    "jt_kde_eval_workspace_size": (C.c_size_t, [C.c_int64, C.c_int64]),
    "jt_kde_eval_ws": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_int64, C.c_float, C.c_void_p,
                                 C.c_void_p, C.c_size_t, C.c_void_p]),
    "jt_kde_grid_cutoff": (C.c_int, [C.POINTER(HashGridDesc), C.c_int64, C.c_void_p, C.c_int32,
                                     C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_float, C.c_float,
                                     C.c_void_p, C.c_void_p]),
    "jt_kde_grid_cutoff_workspace_size": (C.c_size_t, [C.c_int32, C.c_int32, C.c_int32]),
    "jt_kde_grid3d_workspace_size": (C.c_size_t, [C.c_int64, C.c_int32, C.c_int32, C.c_int32]),
    "jt_kde_grid3d": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p,
                                C.c_int32, C.c_float, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "jt_kde_grid2d_workspace_size": (C.c_size_t, [C.c_int64, C.c_int32, C.c_int32]),
    "jt_kde_grid2d": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_float,
                                C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "jt_kde_grid_cutoff_ws": (C.c_int, [C.POINTER(HashGridDesc), C.c_int64, C.c_void_p, C.c_int32,
                                        C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_float, C.c_float,
                                        C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "jt_octree_build": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int32, C.c_int32,
                                  C.POINTER(C.c_void_p)]),
    "jt_octree_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                 C.POINTER(C.c_int32)]),
    "jt_octree_export": (C.c_int, [C.c_void_p] + [C.c_void_p] * 8),
    "jt_octree_free": (None, [C.c_void_p]),
    "jt_octree_sample": (C.c_int, [C.POINTER(OctreeField), C.c_void_p, C.c_int64, C.c_float, C.c_void_p,
                                   C.c_void_p, C.c_void_p]),
    "jt_track_octree_workspace_size": (C.c_size_t, [C.c_int64, C.c_int32, C.POINTER(TrackOpts)]),
    "jt_hashgrid_query_stats": (C.c_int, [C.POINTER(HashGridDesc), C.c_void_p, C.c_int64, C.c_float, C.c_int32, C.c_void_p, C.c_void_p]),
    "jt_octree_locate_stats": (C.c_int, [C.POINTER(OctreeField), C.c_void_p, C.c_int64, C.c_float, C.c_void_p, C.c_void_p]),
    "jt_track_octree": (C.c_int, [C.POINTER(OctreeField), C.POINTER(Boundary), C.c_int32, C.c_void_p,
                                  C.c_int64, C.c_void_p, C.c_float, C.c_int32, C.c_int32, C.c_void_p,
                                  C.POINTER(TrackOpts), C.c_void_p, C.c_size_t, C.c_void_p]),
    "jt_nodes_sample": (C.c_int, [C.POINTER(NodesField), C.c_void_p, C.c_int64, C.c_float, C.c_void_p,
                                  C.c_void_p, C.c_void_p]),
    "jt_track_nodes_workspace_size": (C.c_size_t, [C.c_int64, C.c_int32]),
    "jt_track_nodes": (C.c_int, [C.POINTER(NodesField), C.POINTER(Boundary), C.c_int32, C.c_void_p,
                                 C.c_int64, C.c_void_p, C.c_float, C.c_int32, C.c_int32, C.c_void_p,
                                 C.POINTER(TrackOpts), C.c_void_p, C.c_size_t, C.c_void_p]),
    "jt_track_grid_adjoint_workspace_size": (C.c_size_t, [C.c_int64, C.c_int32, C.c_int32]),
    "jt_track_grid_adjoint": (C.c_int, [C.POINTER(GridField), C.POINTER(Boundary), C.c_void_p, C.c_int64,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_int32, C.c_int32,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
}


def load():
    """Load ``_jt.so`` (once) and declare every exported signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"jaxtrace_amd HIP library not built ({LIB_PATH} missing); "
            "run `python -m jaxtrace_amd.build`")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    return sorted(_SIGS)


_EXC = {JT_EINVAL: ValueError, JT_ENOMEM: MemoryError, JT_EHIP: RuntimeError,
        JT_EUNSUPPORTED: TypeError}


def check(rc: int, what: str):
    """Map a status code to the reference's exception types."""
    if rc == JT_OK:
        return
    msg = load().jt_last_error().decode(errors="replace")
    raise _EXC.get(rc, RuntimeError)(f"{what}: {msg}")


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("jaxtrace_amd requires a ROCm GPU (torch.cuda.is_available() is False)")


def device():
    require_gpu()
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)
