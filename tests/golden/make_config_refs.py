"""Writes tests/golden/config_refs.npz: oracle reference histories of the benchmark workloads
at sizes too long for a live oracle run inside a GPU test (test infrastructure; the oracle is
the checker).

    python tests/golden/make_config_refs.py            # about 8 minutes on one core

Contents (every workload from svl_amd.workloads with its default seeds):
  c1_*   the 2D layer (lin2DQuad4 200 x 100, Lysmer, Ricker point load), CentralDifference,
         all 2000 steps: U history of 16 free DOFs and the final U of every free DOF;
  c2s_*  the layered, lognormal-Vs half-space (lin3DHexa8) at 24 x 24 x 12 elements,
         CentralDifference, 4000 steps: U of 16 free DOFs after every 4th step (steps 4, 8, ...,
         4000), and the U field after step 1000 (while the wave is inside the model: after 4000
         steps the field has left through the Lysmer boundaries to 1e-6 of its peak);
  c4_*   the 64 batched lin2DQuad4 realizations, each run alone by the oracle with its own dt,
         50 steps: U history of 4 free DOFs per realization (offsets into the part's own
         numbering);
  *_sig  input signatures (sums of coordinates, moduli, dashpots and load series), so that a
         change of a workload generator is told apart from a change of the device results.
"""
from __future__ import annotations

import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from svl_amd import workloads as W  # noqa: E402


def signature(m) -> np.ndarray:
    return np.array([m.coords.sum(), np.sum(m.E), np.sum(m.dash_eta), np.sum(m.load_series), m.nfree, m.nelem])


def rec_dofs(m, n: int, extra=()) -> list:
    base = [int(d) for d in extra]
    step = max(1, m.nfree // (n - len(base)))
    rest = [d for d in range(step // 2, m.nfree, step) if d not in base]
    return (base + rest)[:n]


def main():
    out = {}
    t0 = time.time()
    m, dt, nt = W.layer2d(200, 100, nt=2001)
    rec = rec_dofs(m, 16, W.dofs_of(m, int(m.load_node[0])))
    r = oracle.run(m, oracle.CENTRAL_DIFFERENCE, dt, nt, rec)
    out.update(c1_rec=np.array(rec, np.int32), c1_U=r["U"], c1_final=r["final_U"], c1_sig=signature(m))
    print(f"config 1 layer: {time.time() - t0:.0f} s", flush=True)
    t0 = time.time()
    m, dt, nt = W.halfspace3d(24, 24, 12, nt=4001)
    rec = rec_dofs(m, 16, W.dofs_of(m, int(m.load_node[0])))
    r = oracle.run(m, oracle.CENTRAL_DIFFERENCE, dt, nt, rec)
    mid = oracle.run(m, oracle.CENTRAL_DIFFERENCE, dt, 1001, ())["final_U"]
    out.update(c2s_rec=np.array(rec, np.int32), c2s_U=r["U"][3::4], c2s_mid=mid, c2s_sig=signature(m))
    print(f"half-space 24x24x12, 4000 steps: {time.time() - t0:.0f} s", flush=True)
    t0 = time.time()
    models, dts, _ = W.batch2d(nt=51)
    recs, hists, sigs = [], [], []
    for m, dt in zip(models, dts):
        rec = rec_dofs(m, 4, [int(m.free_dof()[int(m.load_node[0])][0])])
        recs.append(rec)
        hists.append(oracle.run(m, oracle.CENTRAL_DIFFERENCE, dt, 51, rec)["U"])
        sigs.append(signature(m))
    out.update(c4_rec=np.array(recs, np.int32), c4_U=np.array(hists), c4_sig=np.array(sigs))
    print(f"64 realizations, 50 steps: {time.time() - t0:.0f} s", flush=True)
    np.savez_compressed(HERE / "config_refs.npz", **out)


if __name__ == "__main__":
    main()
