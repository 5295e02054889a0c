"""CPU oracle: FP64 restatement of SVL's dynamic-FEM hot path (test infrastructure).

Only the tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use this package, and only as the checker. The product library
(``svl_amd/libsvlg.so``) never links or calls it.

``oracle/svl_oracle.cpp`` is compiled with g++ (no GPU code) into
``oracle/libsvl_oracle.so`` by :func:`build`. SVL's own C++ cannot be compiled in this
environment (Eigen, PETSc and MUMPS are absent), so ``oracle/_ref`` stays empty.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SRC = HERE / "svl_oracle.cpp"
LIB = HERE / "libsvl_oracle.so"

# SVL's release flags are -O3 -m64 -std=c++14 (02-Run_Process/Makefile.mk:33); the
# restatement needs C++17 only for std headers it shares with the product header.
CXXFLAGS = ["-O3", "-m64", "-std=c++17", "-fPIC", "-shared"]


def build(force: bool = False) -> Path:
    if force or not LIB.exists() or LIB.stat().st_mtime < SRC.stat().st_mtime:
        tmp = LIB.with_suffix(".so.tmp")
        cmd = ["g++", *CXXFLAGS, str(SRC), "-o", str(tmp)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"oracle build failed:\n{r.stderr}")
        os.replace(tmp, LIB)
    return LIB


class OracleOpts(C.Structure):
    _fields_ = [
        ("integrator", C.c_int32),
        ("nt", C.c_int32),
        ("dt", C.c_double),
        ("mtol", C.c_double),
        ("ktol", C.c_double),
        ("ftol", C.c_double),
        ("algorithm", C.c_int32),
        ("nmax", C.c_int32),
        ("cnvgtest", C.c_int32),
        ("pad", C.c_int32),
        ("cnvgtol", C.c_double),
    ]


CENTRAL_DIFFERENCE = 0
EXTENDED_NEWMARK = 1
NEWMARK = 2

LINEAR = 0
NEWTON = 1
# Algorithm::ComputeConvergence flags (SVL's 'cnvgtest' names).
CNVG = {"UnbalanceForce": 1, "IncrementalDisplacement": 2, "InterpolatedEnergy": 3,
        "RelativeUnbalanceForce": 4, "RelativeIncrementalDisplacement": 5, "RelativeInterpolatedEnergy": 6,
        "TotalRelativeIncrementalDisplacement": 7, "MaxIterations": 8}

_L = None


def _lib():
    global _L
    if _L is None:
        build()
        from svl_amd._capi import MeshDesc
        L = C.CDLL(str(LIB))
        f64p = C.POINTER(C.c_double)
        i32p = C.POINTER(C.c_int32)
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_run.restype = C.c_int
        L.oracle_run.argtypes = [C.POINTER(MeshDesc), C.POINTER(OracleOpts), C.c_int32, i32p, f64p, f64p, f64p,
                                 f64p, f64p, i32p, f64p, C.c_int32, i32p, f64p]
        L.oracle_element_matrices.restype = C.c_int
        L.oracle_element_matrices.argtypes = [C.POINTER(MeshDesc), C.c_int32, f64p, f64p, f64p]
        L.oracle_internal_force.restype = C.c_int
        L.oracle_internal_force.argtypes = [C.POINTER(MeshDesc), f64p, C.c_double, f64p]
        L.oracle_diagonals.restype = C.c_int
        L.oracle_diagonals.argtypes = [C.POINTER(MeshDesc), f64p, f64p]
        L.oracle_j2_path.restype = C.c_int
        L.oracle_j2_path.argtypes = [f64p, C.c_int32, C.c_int32, f64p, f64p, f64p, f64p]
        L.oracle_newmark_keff_dense.restype = C.c_int
        L.oracle_newmark_keff_dense.argtypes = [C.POINTER(MeshDesc), C.c_double, f64p]
        L.oracle_newmark_keff_apply.restype = C.c_int
        L.oracle_newmark_keff_apply.argtypes = [C.POINTER(MeshDesc), C.c_double, f64p, f64p]
        _L = L
    return _L


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed ({rc}): {_lib().oracle_last_error().decode()}")


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def run(model, integrator: int, dt: float, nt: int, rec_dofs=(), ftol: float = 1e-12, want_va: bool = False,
        algorithm: int = LINEAR, nmax: int = 20, cnvgtest: int = 5, cnvgtol: float = 1e-6, reac_tot=()):
    """Runs k = 1..nt-1; returns dict with U (and V, A) histories [nt-1, nrec] (rec_dofs: free
    DOF indices, or -1 - t for total DOF t, e.g. a restrained DOF with support motion),
    the final free-DOF displacement, the time-loop wall time in seconds and, for Newton,
    the iterations of every step and the convergence-test value of every iteration
    ("residuals", [nt-1, nmax], NaN past the last iteration). reac_tot: total DOF indices
    (mesh node n, component c: n*ndim + c; the restrained partner node of the dashpots of mesh
    node n follows the mesh nodes in order of first appearance) whose reaction history lands
    in "R"."""
    from svl_amd._capi import make_desc
    desc, keep = make_desc(model)
    rec = np.ascontiguousarray(rec_dofs, dtype=np.int32)
    ns = max(nt - 1, 0)
    hu = np.zeros((ns, rec.size))
    hv = np.zeros((ns, rec.size)) if want_va else None
    ha = np.zeros((ns, rec.size)) if want_va else None
    secs = C.c_double(0.0)
    uf = np.zeros(model.nfree)
    o = OracleOpts(integrator, nt, dt, 1e-12, 1e-12, ftol, algorithm, nmax, cnvgtest, 0, cnvgtol)
    its = np.zeros(max(ns, 1), np.int32)
    res = np.full((max(ns, 1), max(nmax, 1)), np.nan) if algorithm == NEWTON else None
    rt = np.ascontiguousarray(reac_tot, dtype=np.int32)
    hr = np.zeros((max(ns, 1), rt.size)) if rt.size else None
    _chk(_lib().oracle_run(C.byref(desc), C.byref(o), rec.size, rec.ctypes.data_as(C.POINTER(C.c_int32)), _f(hu),
                           _f(hv), _f(ha), C.byref(secs), _f(uf), its.ctypes.data_as(C.POINTER(C.c_int32)),
                           _f(res), int(rt.size), rt.ctypes.data_as(C.POINTER(C.c_int32)), _f(hr)), "run")
    out = {"U": hu, "final_U": uf, "seconds": secs.value, "iters": its[:ns]}
    if res is not None:
        out["residuals"] = res[:ns]
    if hr is not None:
        out["R"] = hr[:ns]
    if want_va:
        out["V"] = hv
        out["A"] = ha
    return out


def element_matrices(model, e: int = 0):
    from svl_amd._capi import make_desc
    desc, keep = make_desc(model)
    nd = model.nen * model.ndim
    K, M, Cm = np.zeros((nd, nd)), np.zeros((nd, nd)), np.zeros((nd, nd))
    _chk(_lib().oracle_element_matrices(C.byref(desc), e, _f(K), _f(M), _f(Cm)), "element_matrices")
    return K, M, Cm


def internal_force(model, U, ftol: float = 1e-12):
    from svl_amd._capi import make_desc
    desc, keep = make_desc(model)
    U = np.ascontiguousarray(U, dtype=np.float64)
    F = np.zeros(model.nfree)
    _chk(_lib().oracle_internal_force(C.byref(desc), _f(U), ftol, _f(F)), "internal_force")
    return F


def keff_apply(model, dt: float, x):
    """y = Keff x (ExtendedNewmarkBeta effective stiffness, free DOFs)."""
    from svl_amd._capi import make_desc
    desc, keep = make_desc(model)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(model.nfree)
    _chk(_lib().oracle_newmark_keff_apply(C.byref(desc), dt, _f(x), _f(y)), "keff_apply")
    return y


def keff_dense(model, dt: float):
    """Dense Keff = T^T (K + 4/dt^2 M + 2/dt C) T over the free DOFs (small models only)."""
    from svl_amd._capi import make_desc
    desc, keep = make_desc(model)
    n = model.nfree
    if n > 6000:
        raise ValueError("keff_dense is meant for models of at most 6000 free DOFs")
    K = np.zeros((n, n))
    _chk(_lib().oracle_newmark_keff_dense(C.byref(desc), dt, _f(K)), "keff_dense")
    return K


def j2_path(params, strains, plane: bool = False):
    """One J2 Gauss point driven through engineering strains [n, 6] (or [n, 3] plane
    strain), one UpdateState per row: returns (stress [n, nv], tangent [n, nv, nv], alpha [n])."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    e = np.ascontiguousarray(strains, dtype=np.float64)
    n, nv = e.shape
    if nv != (3 if plane else 6) or p.size != 5:
        raise ValueError("j2_path: strains must have 3 columns in plane strain, 6 in 3D; 5 parameters")
    sig = np.zeros((n, nv))
    tan = np.zeros((n, nv, nv))
    al = np.zeros(n)
    _chk(_lib().oracle_j2_path(_f(p), int(plane), n, _f(e), _f(sig), _f(tan), _f(al)), "j2_path")
    return sig, tan, al


def diagonals(model):
    from svl_amd._capi import make_desc
    desc, keep = make_desc(model)
    M, Cm = np.zeros(model.nfree), np.zeros(model.nfree)
    _chk(_lib().oracle_diagonals(C.byref(desc), _f(M), _f(Cm)), "diagonals")
    return M, Cm
