"""svl_amd -- MI355X-native SeismoVLAB dynamic-FEM hot path (HIP on gfx950 + C ABI).

The numerical work lives in ``libsvlg.so`` (``svl_amd/csrc``); this package builds
models in SVL's pre-processor conventions (:mod:`svl_amd.mesh`), binds the C ABI
(:mod:`svl_amd._capi`) and mirrors SVL's integrator interfaces
(:mod:`svl_amd.integrators`).
"""
from .mesh import Model  # noqa: F401

__all__ = ["Model"]
