"""Benchmark: explicit CentralDifference Hexa8 throughput in DOF-steps/s on one MI355X.

Workload (default): 3D elastic half-space, lin3DHexa8 100 x 100 x 50 elements of 2 m
(520,251 nodes, 1,560,753 free DOFs), six layers Vs 150-800 m/s with a seeded
lognormal perturbation, Lysmer dashpots, Ricker point source, lumped mass, FP64.
A "step" is one CentralDifference time step over the whole mesh with all state
resident in HBM. value = nfree * steps / wall seconds of the timed region.

The roofline object prices the explicit step against HBM bandwidth with the
algorithmic bytes per step 32 B per free DOF + (4 nen + 16) B per element + 8 dim B
per node (read U_k-1, U_k-2, keff; write U_k; connectivity, lambda, mu; coordinates);
the average duration of one step is measured with HIP events on the model's stream.

The cpu_baseline is the FP64 oracle (a line-by-line restatement of SVL's per-step
algorithm, single-threaded like SVL's serial build) timed on a few steps of the same
workload.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


class HipEvents:
    """hipEvent timing on an explicit stream (libamdhip64 through ctypes)."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.L.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.L.hipEventSynchronize.argtypes = [C.c_void_p]
        self.L.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.L.hipEventDestroy.argtypes = [C.c_void_p]

    def event(self):
        e = C.c_void_p()
        assert self.L.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.L.hipEventRecord(e, C.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.L.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.L.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return float(ms.value)


def algorithmic_bytes_per_step(m) -> int:
    return 32 * m.nfree + (4 * m.nen + 16) * m.nelem + 8 * m.ndim * m.nnodes


def cpu_baseline(m, dt, steps):
    import oracle
    r = oracle.run(m, oracle.CENTRAL_DIFFERENCE, dt, steps + 1, ())
    secs = r["seconds"]
    return {"value": m.nfree * steps / secs, "unit": "DOF-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle (SVL restatement, g++ -O3, 1 thread): {steps} CentralDifference steps of the "
                      f"same {m.name} model, time loop only ({secs:.2f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--ne", type=str, default="100,100,50", help="hexa elements nx,ny,nz")
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile-nodes", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--variant", type=int, default=0, help="1 fused node tiles (default), 2 element-colored")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(the multi-GPU sharding is described in the design notes, not built)"}))
        sys.exit(2)

    from svl_amd import workloads as W
    from svl_amd.integrators import CentralDifference, DeviceModel

    nx, ny, nz = (int(v) for v in args.ne.split(","))
    t0 = time.time()
    m, dt, nt = W.halfspace3d(nx, ny, nz, h=2.0, nt=max(4001, args.steps + args.warmup + 2))
    print(f"# model {m.name}: nodes {m.nnodes} elems {m.nelem} nfree {m.nfree} dashpots {m.dash_node.size} "
          f"dt {dt:.4e} (built in {time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    dm = DeviceModel(m, tile_nodes=args.tile_nodes, block=args.block, variant=args.variant)
    info = dm.info()
    cd = CentralDifference(dm, dt)
    cd.Initialize()
    ev = HipEvents()
    stream = dm.stream()
    cd.advance_async(1, args.warmup)
    dm.sync()
    e0, e1 = ev.event(), ev.event()
    t_start = time.perf_counter()
    ev.record(e0, stream)
    cd.advance_async(1 + args.warmup, args.steps)
    ev.record(e1, stream)
    dm.sync()
    wall = time.perf_counter() - t_start
    gpu_ms = ev.elapsed_ms(e0, e1)
    step_s = gpu_ms / 1e3 / args.steps
    bytes_step = algorithmic_bytes_per_step(m)
    achieved = bytes_step / step_s / 1e9
    value = m.nfree * args.steps / wall
    print(f"# wall {wall:.4f}s  event {gpu_ms / 1e3:.4f}s  step {step_s * 1e6:.2f} us  "
          f"alg bytes/step {bytes_step / 1e6:.1f} MB  achieved {achieved:.0f} GB/s  variant {info['variant']}",
          file=sys.stderr, flush=True)
    out = {
        "metric": "DOF-steps/s explicit Hexa8 (CentralDifference, lumped, FP64)",
        "value": value,
        "unit": "DOF-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"halfspace3d lin3DHexa8 {nx}x{ny}x{nz} h=2m layered+lognormal Vs, Lysmer, "
                               f"Ricker point load, CentralDifference",
                   "nfree": m.nfree, "nelem": m.nelem, "dt": dt, "kernel_variant": info["variant"],
                   "ntiles": info["ntiles"], "lds_bytes": info["lds_bytes"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "one CentralDifference step (all kernels of the step)",
                     "alg_bytes_per_step": bytes_step},
    }
    dm.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(m, dt, args.cpu_steps)
        out["cpu_baseline"]["cores"] = 1
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
