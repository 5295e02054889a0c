"""Host-side mirror of SVL's Integrator / Assembler / DynamicAnalysis interfaces.

Method names, argument meaning, the step-index convention and the boolean stop
semantics follow SVL (RP = 02-Run_Process):
  * Assembler            RP/07-Assembler/Assembler.hpp:50-181
  * Integrator           RP/10-Integrators/Integrator.hpp:53-144
  * CentralDifference    RP/10-Integrators/02-CentralDifference/CentralDifference.hpp:63
  * ExtendedNewmarkBeta  RP/10-Integrators/03-Newmark/ExtendedNewmarkBeta.cpp (G = 0)
  * LinearSystem         RP/11-Solvers/LinearSystem.hpp:44-66
  * DynamicAnalysis      RP/08-Analysis/02-Dynamic/DynamicAnalysis.cpp:25-64
Every numerical call goes through libsvlg (HIP on gfx950); a failing call raises
:class:`SvlgError` or, where SVL returns ``bool stop``, returns ``True``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as capi
from .mesh import Model


class DeviceModel:
    """A model resident in HBM (Mesh::Initialize + Element::SetDomain)."""

    def __init__(self, model: Model, tile_nodes: int = 0, block: int = 0, variant: int = 0, geo_table: int = 0,
                 elem_precision: int = 0, cg_launch: int = 0):
        self.model = model
        self._desc, self._keep = capi.make_desc(model)
        h = C.c_void_p()
        opts = capi.ExecOpts(tile_nodes, block, variant, geo_table, elem_precision, cg_launch)
        capi.check(capi.lib().svlg_model_create_ex(C.byref(self._desc), C.byref(opts), C.byref(h)),
                   "svlg_model_create")
        self.h = h
        self.nfree = model.nfree
        self.ndim = model.ndim

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            capi.lib().svlg_model_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        a = (C.c_int64 * 10)()
        capi.check(capi.lib().svlg_model_info(self.h, a), "svlg_model_info")
        keys = ["ncolors", "nelem", "nnodes", "nfree", "variant", "ntiles", "elem_path", "lds_bytes",
                "elem_precision", "ngeo"]
        d = dict(zip(keys, [int(v) for v in a]))
        d["affine"] = int(d["elem_path"] > 0)
        return d

    def element_colors(self) -> np.ndarray:
        out = np.zeros(self.model.nelem, np.int32)
        capi.check(capi.lib().svlg_element_colors(self.h, capi.iptr(out)), "svlg_element_colors")
        return out

    def set_stream(self, stream_handle: int | None):
        capi.check(capi.lib().svlg_model_set_stream(self.h, C.c_void_p(stream_handle or 0)), "svlg_model_set_stream")

    def stream(self) -> int:
        return int(capi.lib().svlg_model_stream(self.h) or 0)

    # ---------------------------------------------------------------- Assembler
    def reactions(self, nodes, k: int) -> np.ndarray:
        """Reactions [len(nodes), ndim] at mesh nodes after step k for the integrator
        initialised last (Integrator::ComputeReactionForce restricted to these nodes)."""
        q = np.ascontiguousarray(nodes, dtype=np.int32)
        R = np.zeros((q.size, self.ndim))
        capi.check(capi.lib().svlg_reactions(self.h, int(q.size), capi.iptr(q), int(k), capi.fptr(R)),
                   "svlg_reactions")
        return R

    def ComputeMassMatrix(self) -> np.ndarray:
        """Lumped mass diagonal in free DOFs (Assembler::ComputeMassMatrix)."""
        out = np.zeros(self.nfree)
        capi.check(capi.lib().svlg_mass_diagonal(self.h, capi.fptr(out)), "svlg_mass_diagonal")
        return out

    def ComputeDampingMatrix(self) -> np.ndarray:
        """Diagonal damping (dashpots + Rayleigh am*M) in free DOFs."""
        out = np.zeros(self.nfree)
        capi.check(capi.lib().svlg_damping_diagonal(self.h, capi.fptr(out)), "svlg_damping_diagonal")
        return out

    def ComputeInternalForceVector(self, U: np.ndarray, ftol: float = 1e-12) -> np.ndarray:
        """Fint(U) in free DOFs (Assembler::ComputeInternalForceVector)."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == (self.nfree,)
        out = np.zeros(self.nfree)
        capi.check(capi.lib().svlg_internal_force(self.h, capi.fptr(U), float(ftol), capi.fptr(out)),
                   "svlg_internal_force")
        return out

    def sync(self):
        capi.check(capi.lib().svlg_sync(self.h), "svlg_sync")


def plan_tiles(model: Model, tile_nodes: int = 0) -> dict:
    """Host-only tile plan of the fused explicit kernel (no GPU needed)."""
    desc, keep = capi.make_desc(model)
    st = (C.c_int64 * 8)()
    capi.check(capi.lib().svlg_plan_tiles(C.byref(desc), int(tile_nodes), st), "svlg_plan_tiles")
    keys = ["ntiles", "max_own", "max_stage", "max_elem", "tile_elems", "halo_nodes", "lds_bytes", "ncolors"]
    out = dict(zip(keys, [int(v) for v in st]))
    out["redundancy"] = out["tile_elems"] / model.nelem
    return out


def external_force(mesh: DeviceModel, k: int) -> np.ndarray:
    """Fext(k) at the mesh nodes, [nnodes, ndim] (Integrator::ComputeProgressiveForce with Fbar = 0)."""
    info = mesh.info()
    out = np.zeros(info["nnodes"] * mesh.ndim)
    capi.check(capi.lib().svlg_external_force(mesh.h, int(k), capi.fptr(out)), "svlg_external_force")
    return out.reshape(info["nnodes"], mesh.ndim)


def support_motion_vector(mesh: DeviceModel, Feff: np.ndarray, k: int, factor: float = 1.0) -> np.ndarray:
    """Integrator::ComputeSupportMotionVector of the integrator initialised last: a copy of Feff
    (free DOFs) minus T^T K a_k."""
    out = np.array(Feff, dtype=np.float64, copy=True)
    capi.check(capi.lib().svlg_support_motion_vector(mesh.h, int(k), float(factor), capi.fptr(out)),
               "svlg_support_motion_vector")
    return out


def total_displacements(mesh: DeviceModel) -> np.ndarray:
    """The total displacement vector [nnodes, ndim] of the integrator initialised last
    (restrained DOFs: 0, or their support motion)."""
    out = np.zeros((mesh.model.nnodes, mesh.ndim))
    capi.check(capi.lib().svlg_get_displacements(mesh.h, capi.fptr(out)), "svlg_get_displacements")
    return out


class CentralDifference:
    """CentralDifference(mesh, TimeStep, mtol, ktol, ftol) on the device."""

    def __init__(self, mesh: DeviceModel, TimeStep: float, mtol: float = 1e-12, ktol: float = 1e-12,
                 ftol: float = 1e-12):
        self.mesh = mesh
        self.dt = float(TimeStep)
        self.ftol = float(ftol)

    def Initialize(self):
        capi.check(capi.lib().svlg_cd_init(self.mesh.h, self.dt, self.ftol), "svlg_cd_init")

    def ComputeNewStep(self, k: int) -> bool:
        """One step with load sample k; returns SVL's stop flag."""
        rc = capi.lib().svlg_cd_steps(self.mesh.h, int(k), 1, 0, None, None)
        return rc != 0

    def advance(self, k0: int, nsteps: int, rec_dofs=None) -> np.ndarray | None:
        """nsteps steps from step index k0; returns U history [nsteps, nrec] if recording."""
        if rec_dofs is None or len(rec_dofs) == 0:
            capi.check(capi.lib().svlg_cd_steps(self.mesh.h, int(k0), int(nsteps), 0, None, None), "svlg_cd_steps")
            return None
        rec = np.ascontiguousarray(rec_dofs, dtype=np.int32)
        hist = np.zeros((nsteps, rec.size))
        capi.check(capi.lib().svlg_cd_steps(self.mesh.h, int(k0), int(nsteps), int(rec.size), capi.iptr(rec),
                                            capi.fptr(hist)), "svlg_cd_steps")
        return hist

    def advance_async(self, k0: int, nsteps: int):
        capi.check(capi.lib().svlg_cd_steps_async(self.mesh.h, int(k0), int(nsteps)), "svlg_cd_steps_async")

    # The pieces of one step, for a caller that drives SVL's Linear algorithm itself
    # (Linear::ComputeNewIncrement, Linear.cpp:22-56): Feff, Keff, solve, then the update.
    def ComputeEffectiveForce(self, k: int, factor: float = 1.0) -> np.ndarray:
        out = np.zeros(self.mesh.nfree)
        capi.check(capi.lib().svlg_cd_effective_force(self.mesh.h, int(k), float(factor), capi.fptr(out)),
                   "svlg_cd_effective_force")
        return out

    def ComputeEffectiveStiffness(self) -> np.ndarray:
        """The diagonal of Keff = T^T (M/dt^2 + C/2dt) T over the free DOFs."""
        out = np.zeros(self.mesh.nfree)
        capi.check(capi.lib().svlg_cd_effective_stiffness(self.mesh.h, capi.fptr(out)), "svlg_cd_effective_stiffness")
        return out

    def ComputeSupportMotionVector(self, Feff: np.ndarray, k: int, factor: float = 1.0) -> np.ndarray:
        """Feff - T^T K a_k (a_k: the step's support increments); K = M/dt^2 + C/2dt is diagonal
        here, so Feff comes back unchanged."""
        return support_motion_vector(self.mesh, Feff, k, factor)

    def ApplyIncrement(self, dU: np.ndarray, k: int):
        """The update of ComputeNewStep for step k with the algorithm's increment (free DOFs);
        the support-motion DOFs move by their increment of step k."""
        d = np.ascontiguousarray(dU, dtype=np.float64)
        capi.check(capi.lib().svlg_cd_apply_increment(self.mesh.h, int(k), capi.fptr(d)), "svlg_cd_apply_increment")

    def ComputeProgressiveForce(self, k: int) -> np.ndarray:
        return external_force(self.mesh, k)

    def _state(self):
        n = self.mesh.nfree
        U, V, A = np.zeros(n), np.zeros(n), np.zeros(n)
        capi.check(capi.lib().svlg_cd_get_state(self.mesh.h, capi.fptr(U), capi.fptr(V), capi.fptr(A)),
                   "svlg_cd_get_state")
        return U, V, A

    def GetDisplacements(self) -> np.ndarray:
        return self._state()[0]

    def GetVelocities(self) -> np.ndarray:
        return self._state()[1]

    def GetAccelerations(self) -> np.ndarray:
        return self._state()[2]


PRECOND_JACOBI = 0
PRECOND_LEARNED = 1


class ExtendedNewmarkBeta:
    """ExtendedNewmarkBeta(mesh, TimeStep, mtol, ktol, ftol) with the Linear algorithm on the
    device: gamma = 1/2, beta = 1/4, Keff = K + 4/dt^2 M + 2/dt C applied matrix-free and
    solved by preconditioned CG to ``||b - Keff x|| <= rtol ||b||`` (SVL factors Keff
    directly; the tolerance stands in for the direct solve's round-off)."""

    def __init__(self, mesh: DeviceModel, TimeStep: float, mtol: float = 1e-12, ktol: float = 1e-12,
                 ftol: float = 1e-12, rtol: float = 1e-10, maxit: int = 5000):
        self.mesh = mesh
        self.dt = float(TimeStep)
        self.rtol = float(rtol)
        self.maxit = int(maxit)
        self.iters = np.zeros(0, np.int32)
        self.relres = np.zeros(0)

    def Initialize(self):
        capi.check(capi.lib().svlg_nm_init(self.mesh.h, self.dt, self.rtol, self.maxit, PRECOND_JACOBI),
                   "svlg_nm_init")

    def ComputeNewStep(self, k: int) -> bool:
        """One step with load sample k; returns SVL's stop flag."""
        it = np.zeros(1, np.int32)
        rr = np.zeros(1)
        rc = capi.lib().svlg_nm_steps(self.mesh.h, int(k), 1, 0, None, None, capi.iptr(it), capi.fptr(rr))
        self.iters, self.relres = it, rr
        return rc != 0

    def advance(self, k0: int, nsteps: int, rec_dofs=None) -> np.ndarray | None:
        """nsteps steps from step index k0; returns U history [nsteps, nrec] if recording.
        CG iteration counts and final relative residuals land in ``iters`` / ``relres``."""
        it = np.zeros(max(nsteps, 1), np.int32)
        rr = np.zeros(max(nsteps, 1))
        if rec_dofs is None or len(rec_dofs) == 0:
            rc = capi.lib().svlg_nm_steps(self.mesh.h, int(k0), int(nsteps), 0, None, None, capi.iptr(it),
                                          capi.fptr(rr))
            hist = None
        else:
            rec = np.ascontiguousarray(rec_dofs, dtype=np.int32)
            hist = np.zeros((nsteps, rec.size))
            rc = capi.lib().svlg_nm_steps(self.mesh.h, int(k0), int(nsteps), int(rec.size), capi.iptr(rec),
                                          capi.fptr(hist), capi.iptr(it), capi.fptr(rr))
        self.iters, self.relres = it[:nsteps], rr[:nsteps]
        capi.check(rc, "svlg_nm_steps")
        return hist

    def ComputeEffectiveForce(self, k: int, factor: float = 1.0) -> np.ndarray:
        out = np.zeros(self.mesh.nfree)
        capi.check(capi.lib().svlg_nm_effective_force(self.mesh.h, int(k), float(factor), capi.fptr(out)),
                   "svlg_nm_effective_force")
        return out

    def ComputeSupportMotionVector(self, Feff: np.ndarray, k: int, factor: float = 1.0) -> np.ndarray:
        """Feff - T^T Keff a_k (ExtendedNewmarkBeta::ComputeSupportMotionVector, :191-204)."""
        return support_motion_vector(self.mesh, Feff, k, factor)

    def ApplyIncrement(self, dU: np.ndarray, k: int):
        """The update of ComputeNewStep for step k with the algorithm's increment (free DOFs):
        U += T dU + SupportMotion; V and A from dU alone."""
        d = np.ascontiguousarray(dU, dtype=np.float64)
        capi.check(capi.lib().svlg_nm_apply_increment(self.mesh.h, int(k), capi.fptr(d)), "svlg_nm_apply_increment")

    def ComputeProgressiveForce(self, k: int) -> np.ndarray:
        return external_force(self.mesh, k)

    def _state(self):
        n = self.mesh.nfree
        U, V, A = np.zeros(n), np.zeros(n), np.zeros(n)
        capi.check(capi.lib().svlg_nm_get_state(self.mesh.h, capi.fptr(U), capi.fptr(V), capi.fptr(A)),
                   "svlg_nm_get_state")
        return U, V, A

    def GetDisplacements(self) -> np.ndarray:
        return self._state()[0]

    def GetVelocities(self) -> np.ndarray:
        return self._state()[1]

    def GetAccelerations(self) -> np.ndarray:
        return self._state()[2]


class NewmarkBeta(ExtendedNewmarkBeta):
    """NewmarkBeta(mesh, TimeStep) with the NewtonRaphson algorithm (nstep = nmax,
    cnvgtol, cnvgtest as SVL's convergence flags) on the device, for models with J2
    elements: each Newton solve is Jacobi-preconditioned CG on the tangent effective
    stiffness to ``rtol`` (SVL factors it directly)."""

    def __init__(self, mesh: DeviceModel, TimeStep: float, cnvgtol: float = 1e-6, nmax: int = 50,
                 cnvgtest: int = 5, rtol: float = 1e-12, maxit: int = 5000):
        super().__init__(mesh, TimeStep, rtol=rtol, maxit=maxit)
        self.cnvgtol = float(cnvgtol)
        self.nmax = int(nmax)
        self.cnvgtest = int(cnvgtest)

    def Initialize(self):
        capi.check(capi.lib().svlg_nl_init(self.mesh.h, self.dt, self.rtol, self.maxit, self.cnvgtol, self.nmax,
                                           self.cnvgtest), "svlg_nl_init")

    def ComputeNewStep(self, k: int) -> bool:
        it, cg = np.zeros(1, np.int32), np.zeros(1, np.int32)
        res = np.full((1, self.nmax), np.nan)
        rc = capi.lib().svlg_nl_steps(self.mesh.h, int(k), 1, 0, None, None, capi.iptr(it), capi.fptr(res),
                                      capi.iptr(cg))
        self.iters, self.residuals, self.cg_iters = it, res, cg
        return rc != 0

    def advance(self, k0: int, nsteps: int, rec_dofs=None) -> np.ndarray | None:
        """nsteps steps from step index k0; Newton iteration counts land in ``iters``, the
        convergence-test value of every iteration in ``residuals`` [nsteps, nmax] (NaN past
        the last iteration) and the CG iterations of each step in ``cg_iters``."""
        it = np.zeros(max(nsteps, 1), np.int32)
        cg = np.zeros(max(nsteps, 1), np.int32)
        res = np.full((max(nsteps, 1), self.nmax), np.nan)
        if rec_dofs is None or len(rec_dofs) == 0:
            rec, hist, nrec, hp = None, None, 0, None
        else:
            rec = np.ascontiguousarray(rec_dofs, dtype=np.int32)
            hist = np.zeros((nsteps, rec.size))
            nrec, hp = int(rec.size), capi.fptr(hist)
        rc = capi.lib().svlg_nl_steps(self.mesh.h, int(k0), int(nsteps), nrec, capi.iptr(rec) if nrec else None, hp,
                                      capi.iptr(it), capi.fptr(res), capi.iptr(cg))
        self.iters, self.residuals, self.cg_iters = it[:nsteps], res[:nsteps], cg[:nsteps]
        capi.check(rc, "svlg_nl_steps")
        return hist

    def GetStress(self, elem: int) -> np.ndarray:
        """Gauss-point stresses of J2 element ``elem``: [ngp, 3] (2D) or [ngp, 6] (3D)."""
        nd = self.mesh.ndim
        s = np.zeros((4, 3) if nd == 2 else (8, 6))
        capi.check(capi.lib().svlg_gp_stress(self.mesh.h, int(elem), capi.fptr(s)), "svlg_gp_stress")
        return s


class LearnedPreconditioner:
    """Graph-convolution preconditioner of an initialised :class:`ExtendedNewmarkBeta`.

    ``channels`` C in [1, 4], ``layers`` L in [2, 16]; ``mfma`` 1 / 0 / -1 (auto) selects the
    BF16 MFMA element step. ``theta`` None starts from the Chebyshev preconditioner."""

    def __init__(self, integrator: ExtendedNewmarkBeta, channels: int = 1, layers: int = 6, mfma: int = -1,
                 theta=None):
        self.integrator = integrator
        self.mesh = integrator.mesh
        t = None
        if theta is not None:
            t = np.ascontiguousarray(theta, dtype=np.float64)
        capi.check(capi.lib().svlg_gnn_init(self.mesh.h, int(channels), int(layers), int(mfma),
                                            capi.fptr(t) if t is not None else None), "svlg_gnn_init")
        self.loss_hist = np.zeros(0)

    def info(self) -> dict:
        a = np.zeros(6, np.int32)
        b = np.zeros(2)
        capi.check(capi.lib().svlg_gnn_info(self.mesh.h, capi.iptr(a), capi.fptr(b)), "svlg_gnn_info")
        return {"channels": int(a[0]), "layers": int(a[1]), "mfma": int(a[2]), "nparams": int(a[3]),
                "val_iters_before": int(a[4]), "val_iters_after": int(a[5]), "lmin": float(b[0]),
                "lmax": float(b[1])}

    @property
    def theta(self) -> np.ndarray:
        out = np.zeros(self.info()["nparams"])
        capi.check(capi.lib().svlg_gnn_get_params(self.mesh.h, capi.fptr(out)), "svlg_gnn_get_params")
        return out

    @theta.setter
    def theta(self, value):
        v = np.ascontiguousarray(value, dtype=np.float64)
        capi.check(capi.lib().svlg_gnn_set_params(self.mesh.h, capi.fptr(v)), "svlg_gnn_set_params")

    def apply(self, r: np.ndarray, mfma: bool = False) -> np.ndarray:
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros(self.mesh.nfree)
        capi.check(capi.lib().svlg_gnn_apply(self.mesh.h, int(bool(mfma)), capi.fptr(r), capi.fptr(z)),
                   "svlg_gnn_apply")
        return z

    def loss_grad(self, R: np.ndarray, E: np.ndarray):
        R = np.ascontiguousarray(R, dtype=np.float64)
        E = np.ascontiguousarray(E, dtype=np.float64)
        loss = np.zeros(1)
        g = np.zeros(self.info()["nparams"])
        capi.check(capi.lib().svlg_gnn_loss_grad(self.mesh.h, int(R.shape[0]), capi.fptr(R), capi.fptr(E),
                                                 capi.fptr(loss), capi.fptr(g)), "svlg_gnn_loss_grad")
        return float(loss[0]), g

    def train(self, nsteps: int = 40, every: int = 4, sweeps: int = 2, epochs: int = 60, lr: float = 2e-3):
        """Trains on pairs from the model's own solves; returns the loss history. ``accepted``
        tells whether the trained parameters beat the previous ones on the validation solve."""
        hist = np.zeros(max(epochs, 1))
        acc = np.zeros(1, np.int32)
        capi.check(capi.lib().svlg_gnn_train(self.mesh.h, int(nsteps), int(every), int(sweeps), int(epochs),
                                             float(lr), capi.fptr(hist), capi.iptr(acc)), "svlg_gnn_train")
        self.loss_hist = hist[:epochs]
        self.accepted = bool(acc[0])
        return self.loss_hist

    def use(self, on: bool = True):
        """Selects this preconditioner (flexible CG) or Jacobi-PCG for later steps."""
        capi.check(capi.lib().svlg_nm_set_precond(self.mesh.h, PRECOND_LEARNED if on else PRECOND_JACOBI),
                   "svlg_nm_set_precond")


class LinearSystem:
    """The effective-stiffness system of an initialised :class:`ExtendedNewmarkBeta`
    (LinearSystem::SetMatrix / SetVector / SolveSystem / GetSolution)."""

    def __init__(self, integrator: ExtendedNewmarkBeta):
        self.mesh = integrator.mesh
        self.rtol = integrator.rtol
        self.maxit = integrator.maxit
        self._b = None
        self._x = None
        self.iters = 0
        self.relres = 0.0

    def Apply(self, x: np.ndarray) -> np.ndarray:
        """y = Keff x (free DOFs)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.mesh.nfree)
        capi.check(capi.lib().svlg_keff_apply(self.mesh.h, capi.fptr(x), capi.fptr(y)), "svlg_keff_apply")
        return y

    def SetVector(self, b: np.ndarray):
        self._b = np.ascontiguousarray(b, dtype=np.float64)

    def SolveSystem(self) -> bool:
        x = np.zeros(self.mesh.nfree)
        it = np.zeros(1, np.int32)
        rr = np.zeros(1)
        rc = capi.lib().svlg_solve(self.mesh.h, capi.fptr(self._b), capi.fptr(x), self.rtol, self.maxit,
                                   capi.iptr(it), capi.fptr(rr))
        self.iters, self.relres = int(it[0]), float(rr[0])
        self._x = x
        return rc != 0

    def GetSolution(self) -> np.ndarray:
        return self._x


class DynamicAnalysis:
    """DynamicAnalysis(mesh, integrator, nt): k = 1..nt-1, recorder row k-1."""

    def __init__(self, integrator, nt: int):
        self.integrator = integrator
        self.nt = int(nt)

    def Analyze(self, rec_dofs=None):
        self.integrator.Initialize()
        return self.integrator.advance(1, self.nt - 1, rec_dofs)
