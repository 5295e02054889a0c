"""ctypes binding of libsvlg.so (the C ABI declared in include/svlg.h).

The library is loaded from this package directory only; if it is missing the import
of the device API raises, so no code path silently falls back to the CPU.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from .mesh import Model

LIB_PATH = Path(__file__).resolve().parent / "libsvlg.so"

i32p = C.POINTER(C.c_int32)
f64p = C.POINTER(C.c_double)


class MeshDesc(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32),
        ("elem_type", C.c_int32),
        ("nnodes", C.c_int32),
        ("nelem", C.c_int32),
        ("coords", f64p),
        ("free_dof", i32p),
        ("nfree", C.c_int32),
        ("conn", i32p),
        ("mat_type", i32p),
        ("E", f64p),
        ("nu", f64p),
        ("rho", f64p),
        ("j2", f64p),
        ("thickness", C.c_double),
        ("lumped", C.c_int32),
        ("ray_am", f64p),
        ("ray_ak", f64p),
        ("ndash", C.c_int32),
        ("dash_node", i32p),
        ("dash_dir", i32p),
        ("dash_eta", f64p),
        ("nload", C.c_int32),
        ("load_nt", C.c_int32),
        ("load_node", i32p),
        ("load_dir", f64p),
        ("load_factor", f64p),
        ("load_series", f64p),
        ("ndrm_elem", C.c_int32),
        ("drm_elem", i32p),
        ("ndrm_node", C.c_int32),
        ("drm_node", i32p),
        ("drm_cond", i32p),
        ("drm_nt", C.c_int32),
        ("drm_signal", f64p),
        ("drm_factor", C.c_double),
        ("nsm", C.c_int32),
        ("sm_node", i32p),
        ("sm_dir", i32p),
        ("sm_factor", f64p),
        ("sm_ptr", i32p),
        ("sm_values", f64p),
        ("sm_offset", C.c_int32),
    ]


class ExecOpts(C.Structure):
    _fields_ = [("tile_nodes", C.c_int32), ("block", C.c_int32), ("variant", C.c_int32),
                ("geo_table", C.c_int32), ("elem_precision", C.c_int32), ("cg_launch", C.c_int32),
                ("reserved", C.c_int32 * 2)]


def _p(a, t):
    if a is None:
        return C.cast(None, t)
    return a.ctypes.data_as(t)


def make_desc(m: Model):
    """Builds an svlg_mesh_desc that views the model's arrays (returns desc, keepalive)."""
    keep = []

    def arr(x, dt):
        a = np.ascontiguousarray(x, dtype=dt)
        keep.append(a)
        return a

    free = arr(m.free_dof().reshape(-1), np.int32)
    d = MeshDesc()
    d.ndim = m.ndim
    d.elem_type = m.elem_type
    d.nnodes = m.nnodes
    d.nelem = m.nelem
    d.coords = _p(arr(m.coords, np.float64), f64p)
    d.free_dof = _p(free, i32p)
    d.nfree = m.nfree
    d.conn = _p(arr(m.conn, np.int32), i32p)
    d.mat_type = _p(arr(m.mat_type, np.int32), i32p)
    d.E = _p(arr(m.E, np.float64), f64p)
    d.nu = _p(arr(m.nu, np.float64), f64p)
    d.rho = _p(arr(m.rho, np.float64), f64p)
    d.j2 = _p(arr(m.j2, np.float64), f64p)
    d.thickness = float(m.thickness)
    d.lumped = 1 if m.massform_lumped else 0
    d.ray_am = _p(arr(m.ray_am, np.float64), f64p)
    d.ray_ak = _p(arr(m.ray_ak, np.float64), f64p)
    d.ndash = int(m.dash_node.size)
    d.dash_node = _p(arr(m.dash_node, np.int32), i32p)
    d.dash_dir = _p(arr(m.dash_dir, np.int32), i32p)
    d.dash_eta = _p(arr(m.dash_eta, np.float64), f64p)
    d.nload = int(m.load_node.size)
    d.load_nt = int(m.load_series.shape[1]) if m.load_node.size else 0
    d.load_node = _p(arr(m.load_node, np.int32), i32p)
    d.load_dir = _p(arr(m.load_dir, np.float64), f64p)
    d.load_factor = _p(arr(m.load_factor, np.float64), f64p)
    d.load_series = _p(arr(m.load_series, np.float64), f64p)
    if m.drm_elem is not None and len(m.drm_elem):
        sig = arr(m.drm_signal, np.float64)
        d.ndrm_elem = int(len(m.drm_elem))
        d.drm_elem = _p(arr(m.drm_elem, np.int32), i32p)
        d.ndrm_node = int(len(m.drm_node))
        d.drm_node = _p(arr(m.drm_node, np.int32), i32p)
        d.drm_cond = _p(arr(m.drm_cond, np.int32), i32p)
        d.drm_nt = int(sig.shape[1])
        d.drm_signal = _p(sig, f64p)
        d.drm_factor = float(m.drm_factor)
    if len(m.sm_node):
        ptr = np.concatenate([[0], np.cumsum([len(g) for g in m.sm_series])]).astype(np.int32)
        d.nsm = int(len(m.sm_node))
        d.sm_node = _p(arr(m.sm_node, np.int32), i32p)
        d.sm_dir = _p(arr(m.sm_dir, np.int32), i32p)
        d.sm_factor = _p(arr(m.sm_factor, np.float64), f64p)
        d.sm_ptr = _p(arr(ptr, np.int32), i32p)
        d.sm_values = _p(arr(np.concatenate(m.sm_series), np.float64), f64p)
        d.sm_offset = int(m.sm_offset)
    return d, keep


# Every exported symbol with its signature; also used by the CPU test that checks the
# library exports exactly what the header declares.
SIGNATURES = {
    "svlg_last_error": (C.c_char_p, []),
    "svlg_version": (C.c_char_p, []),
    "svlg_device_count": (C.c_int, []),
    "svlg_model_create": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(C.c_void_p)]),
    "svlg_model_create_ex": (C.c_int, [C.POINTER(MeshDesc), C.c_void_p, C.POINTER(C.c_void_p)]),
    "svlg_plan_tiles": (C.c_int, [C.POINTER(MeshDesc), C.c_int32, C.POINTER(C.c_int64)]),
    "svlg_model_destroy": (C.c_int, [C.c_void_p]),
    "svlg_model_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "svlg_model_stream": (C.c_void_p, [C.c_void_p]),
    "svlg_model_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "svlg_element_colors": (C.c_int, [C.c_void_p, i32p]),
    "svlg_mass_diagonal": (C.c_int, [C.c_void_p, f64p]),
    "svlg_damping_diagonal": (C.c_int, [C.c_void_p, f64p]),
    "svlg_internal_force": (C.c_int, [C.c_void_p, f64p, C.c_double, f64p]),
    "svlg_cd_init": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "svlg_cd_steps": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, i32p, f64p]),
    "svlg_cd_steps_async": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "svlg_cd_get_state": (C.c_int, [C.c_void_p, f64p, f64p, f64p]),
    "svlg_sync": (C.c_int, [C.c_void_p]),
    "svlg_reactions": (C.c_int, [C.c_void_p, C.c_int32, i32p, C.c_int32, f64p]),
    "svlg_cd_init_batch": (C.c_int, [C.c_void_p, C.c_int32, i32p, f64p, C.c_double]),
    "svlg_cd_graph_steps": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    "svlg_autotune_tiles": (C.c_int, [C.POINTER(MeshDesc), C.c_void_p, i32p, C.c_int32, C.c_int32, i32p, f64p,
                                      C.c_int32, i32p, f64p]),
    "svlg_nm_init": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_int32, C.c_int32]),
    "svlg_nm_steps": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, i32p, f64p, i32p, f64p]),
    "svlg_nm_get_state": (C.c_int, [C.c_void_p, f64p, f64p, f64p]),
    "svlg_keff_apply": (C.c_int, [C.c_void_p, f64p, f64p]),
    "svlg_cd_effective_force": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, f64p]),
    "svlg_cd_effective_stiffness": (C.c_int, [C.c_void_p, f64p]),
    "svlg_cd_apply_increment": (C.c_int, [C.c_void_p, C.c_int32, f64p]),
    "svlg_external_force": (C.c_int, [C.c_void_p, C.c_int32, f64p]),
    "svlg_nm_effective_force": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, f64p]),
    "svlg_nm_apply_increment": (C.c_int, [C.c_void_p, C.c_int32, f64p]),
    "svlg_support_motion_vector": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, f64p]),
    "svlg_get_displacements": (C.c_int, [C.c_void_p, f64p]),
    "svlg_keff_apply_noalloc": (C.c_int, [C.c_void_p, f64p, f64p]),
    "svlg_solve": (C.c_int, [C.c_void_p, f64p, f64p, C.c_double, C.c_int32, i32p, f64p]),
    "svlg_nm_set_precond": (C.c_int, [C.c_void_p, C.c_int32]),
    "svlg_nl_init": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_int32, C.c_double, C.c_int32, C.c_int32]),
    "svlg_nl_steps": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, i32p, f64p, i32p, f64p, i32p]),
    "svlg_gp_stress": (C.c_int, [C.c_void_p, C.c_int32, f64p]),
    "svlg_gnn_init": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, f64p]),
    "svlg_gnn_info": (C.c_int, [C.c_void_p, i32p, f64p]),
    "svlg_gnn_get_params": (C.c_int, [C.c_void_p, f64p]),
    "svlg_gnn_set_params": (C.c_int, [C.c_void_p, f64p]),
    "svlg_gnn_apply": (C.c_int, [C.c_void_p, C.c_int32, f64p, f64p]),
    "svlg_gnn_loss_grad": (C.c_int, [C.c_void_p, C.c_int32, f64p, f64p, f64p, f64p]),
    "svlg_gnn_train": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double, f64p, i32p]),
    "svlg_stream_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "svlg_pml_create": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_int32, C.POINTER(C.c_void_p)]),
    "svlg_pml_destroy": (C.c_int, [C.c_void_p]),
    "svlg_pml_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "svlg_pml_steps": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, i32p, f64p, f64p, f64p, i32p]),
    "svlg_pml_get_state": (C.c_int, [C.c_void_p, f64p, f64p, f64p]),
    "svlg_pml_get_history": (C.c_int, [C.c_void_p, f64p]),
    "svlg_pml_set_graph": (C.c_int, [C.c_void_p, C.c_int32]),
}

_LIB = None


class SvlgError(RuntimeError):
    """A libsvlg entry point returned nonzero (SVL's `stop == true`)."""


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run `python -m svl_amd.build` (no CPU fallback exists)")
    L = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def check(rc: int, what: str):
    if rc != 0:
        msg = lib().svlg_last_error()
        raise SvlgError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def fptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(f64p)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(i32p)


def device_count() -> int:
    return int(lib().svlg_device_count())
