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


def copy_calibration(nbytes: int, reps: int = 20):
    """Times svlg_stream_copy of `nbytes` (device to device) on its own stream; under
    rocprofv3 --pmc the same launches give the counter reading for a known transfer."""
    from svl_amd import _capi
    L = C.CDLL("libamdhip64.so")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    L.hipStreamSynchronize.argtypes = [C.c_void_p]
    a, b, st = C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert L.hipMalloc(C.byref(a), nbytes) == 0 and L.hipMalloc(C.byref(b), nbytes) == 0
    assert L.hipMemset(a, 0, nbytes) == 0 and L.hipStreamCreate(C.byref(st)) == 0
    n = nbytes // 8
    lib = _capi.lib()
    ev = HipEvents()
    _capi.check(lib.svlg_stream_copy(b, a, n, st), "svlg_stream_copy")
    L.hipStreamSynchronize(st)
    e0, e1 = ev.event(), ev.event()
    ev.record(e0, st.value)
    for _ in range(reps):
        _capi.check(lib.svlg_stream_copy(b, a, n, st), "svlg_stream_copy")
    ev.record(e1, st.value)
    ms = ev.elapsed_ms(e0, e1)
    L.hipStreamSynchronize(st)
    L.hipFree(a)
    L.hipFree(b)
    return {"bytes_per_launch": 2 * nbytes, "launches": reps + 1, "avg_us": ms * 1e3 / reps,
            "copy_GBs": 2 * nbytes * reps / (ms / 1e3) / 1e9}


def traffic_from_profile(variant: int, tile_nodes: int, name: str = "halfspace3d_fused_step.md"):
    """HBM bytes per launch of the fused step kernel from a committed rocprofv3 PMC summary
    under profiles/, if it covers this configuration."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if line.startswith("traffic_"):
            k, _, v = line.partition(":")
            vals[k.strip()] = v.strip()
    # Several tile sizes: one "traffic_bytes_t<tile nodes>: <bytes>" line each.
    if f"traffic_bytes_t{tile_nodes}" in vals and int(vals.get("traffic_variant", variant)) == variant:
        return float(vals[f"traffic_bytes_t{tile_nodes}"])
    try:
        if int(vals.get("traffic_variant", variant)) == variant and int(vals["traffic_tile_nodes"]) == tile_nodes:
            return float(vals["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None
    return None


def algorithmic_bytes_per_step(m) -> int:
    return 32 * m.nfree + (4 * m.nen + 16) * m.nelem + 8 * m.ndim * m.nnodes


def cpu_baseline(m, dt, steps):
    import oracle
    r = oracle.run(m, oracle.CENTRAL_DIFFERENCE, dt, steps + 1, ())
    secs = r["seconds"]
    return {"value": m.nfree * steps / secs, "unit": "DOF-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle (SVL restatement, g++ -O3, 1 thread): {steps} CentralDifference steps of the "
                      f"same {m.name} model, time loop only ({secs:.2f} s)"}


def cpu_baseline_parallel(m, dt, steps, procs):
    """`procs` single-threaded oracle processes run the same sample at once on the host's cores:
    the aggregate DOF-steps/s is what the host delivers when every core runs one SVL process
    (an upper bound of an SVL run on that many MPI ranks, which would also exchange halos)."""
    import multiprocessing as mp
    ctx = mp.get_context("fork")
    with ctx.Pool(procs) as pool:
        t = time.perf_counter()
        secs = pool.starmap(_oracle_sample, [(m, dt, steps)] * procs)
        wall = time.perf_counter() - t
    return {"value": procs * m.nfree * steps / max(secs), "unit": "DOF-steps/s", "cores": procs, "kind": "port",
            "sample": f"{procs} concurrent single-threaded oracle processes, {steps} CentralDifference steps "
                      f"of the same {m.name} model each, time loops only ({min(secs):.2f}-{max(secs):.2f} s; "
                      f"wall {wall:.2f} s with the model set-up)"}


def _oracle_sample(m, dt, steps):
    import oracle
    return oracle.run(m, oracle.CENTRAL_DIFFERENCE, dt, steps + 1, ())["seconds"]


def host_cores():
    """Cores the host offers (nproc) and the CPU model (lscpu / /proc/cpuinfo)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return n, model


def layer2d_bench(args):
    """The 2D plane-strain layer (lin2DQuad4 200 x 100, Lysmer, Ricker point load,
    CentralDifference lumped FP64): DOF-steps/s and the roofline of its fused step, eager
    launches and launches replayed from a captured hipGraph."""
    from svl_amd import batch as B
    from svl_amd import workloads as W
    from svl_amd.integrators import CentralDifference, DeviceModel
    steps = args.layer_steps
    m, dt, nt = W.layer2d(200, 100, nt=steps + args.warmup + 2)
    # 20,301 nodes: the default 448-node tiles give 51 workgroups for 256 CUs, so the tile
    # size is picked by the autotuner (timed fused steps per candidate).
    tile, tune = B.autotune_tile_nodes(m, [0, m.nnodes], [dt], candidates=(24, 32, 48, 64, 96, 128, 192, 448),
                                       steps=50)
    dm = DeviceModel(m, tile_nodes=tile)
    info = dm.info()
    ev = HipEvents()
    stream = dm.stream()
    res = {}
    for mode in ("eager", "graph"):
        if mode == "eager":
            cd = CentralDifference(dm, dt)
            cd.Initialize()
            run = lambda k0, n: cd.advance_async(k0, n)  # noqa: E731
        else:
            fp = [0, m.nfree]
            cd = B.BatchedCentralDifference(dm, [dt], [0, m.nnodes], fp)
            cd.Initialize()
            run = lambda k0, n: cd.advance(k0, n, graph_steps=30)  # noqa: E731
        run(1, args.warmup)
        dm.sync()
        e0, e1 = ev.event(), ev.event()
        t = time.perf_counter()
        ev.record(e0, stream)
        run(1 + args.warmup, steps)
        ev.record(e1, stream)
        dm.sync()
        wall = time.perf_counter() - t
        step_s = ev.elapsed_ms(e0, e1) / 1e3 / steps
        res[mode] = {"us_per_step": step_s * 1e6, "dof_steps_per_s": m.nfree * steps / wall}
    bytes_step = algorithmic_bytes_per_step(m)
    best = min(res.values(), key=lambda r: r["us_per_step"])
    achieved = bytes_step / (best["us_per_step"] * 1e-6) / 1e9
    out = {"workload": "layer2d lin2DQuad4 200x100 h=1m, rho 2000, Vs 200, nu 0.25, Lysmer sides and base, Ricker "
                       "5 Hz surface point load, CentralDifference dt=0.5 h/Vp, FP64",
           "nfree": m.nfree, "steps": steps, "ntiles": info["ntiles"], "tile_nodes": tile,
           "autotune_us_per_step": tune, **res,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "alg_bytes_per_step": bytes_step,
                        "alg_bytes_per_dof_step": bytes_step / m.nfree}}
    print(f"# layer2d: {m.nfree} DOFs, tile {tile} ({info['ntiles']} tiles), eager {res['eager']['us_per_step']:.2f} "
          f"us/step, graph {res['graph']['us_per_step']:.2f} us/step, {achieved:.0f} GB/s", file=sys.stderr, flush=True)
    dm.close()
    return out


def precision_bench(args, m, dt, us_fp64):
    """Element-step precision variants on the full half-space: step time and the relative L2
    difference of a 400-step displacement history (every 997th free DOF) from the FP64 path."""
    import numpy as np
    from svl_amd.integrators import CentralDifference, DeviceModel
    rec = list(range(5, m.nfree, 997))
    out = {}
    hist = {}
    for name, prec in (("fp64", 0), ("fp32", 2), ("bf16_mfma", 1)):
        dm = DeviceModel(m, elem_precision=prec)
        cd = CentralDifference(dm, dt)
        cd.Initialize()
        hist[name] = cd.advance(1, 400, rec)
        ev = HipEvents()
        stream = dm.stream()
        e0, e1 = ev.event(), ev.event()
        ev.record(e0, stream)
        cd.advance_async(401, args.steps // 4)
        ev.record(e1, stream)
        dm.sync()
        out[name] = {"us_per_step": ev.elapsed_ms(e0, e1) * 1e3 / (args.steps // 4)}
        dm.close()
    for name in out:
        h = hist[name]
        out[name]["rel_l2_vs_fp64_400_steps"] = float(np.linalg.norm(h - hist["fp64"]) / np.linalg.norm(hist["fp64"]))
    print(f"# precision: " + ", ".join(f"{k} {v['us_per_step']:.2f} us/step (rel {v['rel_l2_vs_fp64_400_steps']:.1e})"
                                      for k, v in out.items()), file=sys.stderr, flush=True)
    return out


def implicit_bench(args):
    """Implicit ExtendedNewmarkBeta step time on the 64^3 block: Jacobi-PCG against flexible CG
    with the trained graph-convolution preconditioner, both to the same relative residual."""
    import numpy as np
    from svl_amd import workloads as W
    from svl_amd.integrators import DeviceModel, ExtendedNewmarkBeta, LearnedPreconditioner
    n = args.implicit_ne
    m, dt, nt = W.block3d_implicit(n, n, n, nt=max(501, 2 * args.implicit_steps + 10))
    # 330-node tiles give 1021 tiles on the 64^3 block: the operator (2 workgroups per CU) and the
    # network layers (4 per CU) then run in whole rounds of resident workgroups.
    dm = DeviceModel(m, tile_nodes=args.implicit_tile_nodes)
    nm = ExtendedNewmarkBeta(dm, dt, rtol=args.implicit_rtol, maxit=5000)
    out = {"workload": f"block3d lin3DHexa8 {n}x{n}x{n} h=1m, Vs~U[200,1000] (seed 20260302), Rayleigh 5% at "
                       f"1 and 10 Hz, fixed base, Ricker point load, ExtendedNewmarkBeta dt={dt}",
           "nfree": m.nfree, "rtol": args.implicit_rtol, "steps": args.implicit_steps,
           "tile_nodes": args.implicit_tile_nodes, "ntiles": dm.info()["ntiles"]}

    def timed(label):
        dm.sync()
        t = time.perf_counter()
        nm.advance(1, args.implicit_steps)
        el = (time.perf_counter() - t) / args.implicit_steps
        its = nm.iters.astype(float)
        print(f"# implicit {label}: {el * 1e3:.3f} ms/step, CG iterations mean {its.mean():.2f}",
              file=sys.stderr, flush=True)
        return {"s_per_step": el, "iters_mean": float(its.mean()), "iters_max": int(its.max()),
                "relres_max": float(nm.relres.max())}

    nm.Initialize()
    nm.advance(1, 2)
    nm.Initialize()
    out["jacobi_pcg"] = timed("Jacobi-PCG")
    nm.Initialize()
    t0 = time.perf_counter()
    pc = LearnedPreconditioner(nm, channels=args.gnn_channels, layers=args.gnn_layers, mfma=-1)
    loss = pc.train(nsteps=24, every=3, sweeps=2, epochs=args.gnn_epochs, lr=2e-3)
    train_s = time.perf_counter() - t0
    pc.use(True)
    info = pc.info()
    lf = timed("learned FCG")
    lf.update({"channels": info["channels"], "layers": info["layers"], "mfma_bf16": bool(info["mfma"]),
               "train_s": train_s, "loss_first": float(loss[0]) if len(loss) else None,
               "loss_last": float(loss[-1]) if len(loss) else None,
               "trained_params_accepted": bool(pc.accepted),
               "validation_iters_init_vs_trained": [info["val_iters_before"], info["val_iters_after"]],
               "spectrum_ehat_lanczos": [info["lmin"], info["lmax"]]})
    out["learned_fcg"] = lf
    out["speedup_vs_jacobi"] = out["jacobi_pcg"]["s_per_step"] / lf["s_per_step"]
    dm.close()
    return out


def nonlinear_bench(args):
    """Newton-Raphson + NewmarkBeta step time on the 80x80x40 J2 soil-structure block: every
    Newton solve by Jacobi-PCG against flexible CG with the graph-convolution preconditioner
    (Chebyshev-initialised, BF16 MFMA layers) of the initial effective stiffness, both to the
    same relative residual. Both runs start from rest and time the same window of steps."""
    import numpy as np
    from svl_amd import workloads as W
    from svl_amd.integrators import DeviceModel, LearnedPreconditioner, NewmarkBeta
    skip, steps = args.nonlinear_skip, args.nonlinear_steps
    m, dt, nt = W.soil_structure_j2(nt=max(1001, skip + steps + 2))
    out = {"workload": "soil_structure_j2 lin3DHexa8 80x80x40 h=2m, 4 Plastic3DJ2 layers (c~U[50,150] kPa, seed "
                       "20260305, H=0.1G, beta=1) + elastic 10x10x8 block (30 GPa), Rayleigh 5% at 1 and 10 Hz, "
                       f"fixed base, horizontal Ricker 3 Hz on the block; NewmarkBeta + Newton (test 5, 1e-6, "
                       f"nmax 20), dt={dt}", "nfree": m.nfree, "rtol": args.implicit_rtol,
           "steps_timed": [skip + 1, skip + steps]}

    def run(learned):
        dm = DeviceModel(m)
        nl = NewmarkBeta(dm, dt, cnvgtol=1e-6, nmax=20, cnvgtest=5, rtol=args.implicit_rtol, maxit=5000)
        nl.Initialize()
        info = None
        if learned:
            pc = LearnedPreconditioner(nl, channels=1, layers=args.gnn_layers, mfma=-1)
            pc.use(True)
            info = pc.info()
        if skip:
            nl.advance(1, skip)
        dm.sync()
        t = time.perf_counter()
        nl.advance(skip + 1, steps)
        el = (time.perf_counter() - t) / steps
        r = {"s_per_step": el, "newton_iters_mean": float(nl.iters.mean()),
             "cg_iters_per_solve": float(nl.cg_iters.sum() / max(nl.iters.sum(), 1))}
        if info:
            r.update({"layers": info["layers"], "mfma_bf16": bool(info["mfma"])})
        print(f"# nonlinear {'learned FCG' if learned else 'Jacobi-PCG'}: {el * 1e3:.3f} ms/step, Newton "
              f"{r['newton_iters_mean']:.2f}, CG per solve {r['cg_iters_per_solve']:.1f}", file=sys.stderr, flush=True)
        dm.close()
        return r

    out["jacobi_pcg"] = run(False)
    out["learned_fcg"] = run(True)
    out["speedup_vs_jacobi"] = out["jacobi_pcg"]["s_per_step"] / out["learned_fcg"]["s_per_step"]
    return out


def batched_bench(args):
    """64 independent 2D realizations (uneven sizes, own Vs fields, dt and plane-wave input)
    advanced together: tile size from the autotuner, one fused launch per step for all
    realizations, replayed from a captured hipGraph. value = total free DOFs x steps / s."""
    import numpy as np
    from svl_amd import batch as B
    from svl_amd import workloads as W
    from svl_amd.integrators import DeviceModel
    t0 = time.perf_counter()
    nt = args.batch_steps + args.batch_warmup + 2
    models, dts, nt = W.batch2d(args.batch_n, nt=nt)
    u, nptr, fptr = B.union(models)
    build_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    if args.batch_tile_nodes > 0:
        best, times = args.batch_tile_nodes, {}
    else:
        best, times = B.autotune_tile_nodes(u, nptr, dts, candidates=(96, 128, 160, 192, 256, 320, 448, 640),
                                            steps=20)
    tune_s = time.perf_counter() - t0
    dm = DeviceModel(u, tile_nodes=best)
    info = dm.info()
    cd = B.BatchedCentralDifference(dm, dts, nptr, fptr)
    cd.Initialize()
    g = args.batch_graph_steps
    cd.advance(1, args.batch_warmup, graph_steps=g)
    ev = HipEvents()
    stream = dm.stream()
    e0, e1 = ev.event(), ev.event()
    t = time.perf_counter()
    ev.record(e0, stream)
    capi_steps = args.batch_steps
    from svl_amd import _capi as capi
    capi.check(capi.lib().svlg_cd_graph_steps(dm.h, 1 + args.batch_warmup, capi_steps, g), "svlg_cd_graph_steps")
    ev.record(e1, stream)
    dm.sync()
    wall = time.perf_counter() - t
    step_s = ev.elapsed_ms(e0, e1) / 1e3 / capi_steps
    nnodes = u.nnodes
    alg = 32 * u.nfree + (4 * 4 + 16) * u.nelem + 8 * 2 * nnodes
    out = {"workload": f"{len(models)} lin2DQuad4 realizations nx~U{{50..400}} ny~U{{50..200}} h=1m, lognormal Vs "
                       f"(mean 300, sigma 0.2), own dt, SV plane-wave base forces + Lysmer, CentralDifference "
                       f"(seeds 20260304+i)", "realizations": len(models), "nfree_total": u.nfree,
           "nelem_total": u.nelem, "dt_range": [min(dts), max(dts)], "steps": capi_steps,
           "dof_steps_per_s": u.nfree * capi_steps / wall, "us_per_step": step_s * 1e6,
           "graph_steps": g, "tile_nodes": best, "autotune_us_per_step": times, "kernel_variant": info["variant"],
           "alg_bytes_per_step": alg, "achieved_GBs": alg / step_s / 1e9,
           "traffic": traffic_from_profile(info["variant"], best, "batch2d_fused_step.md"),
           "roofline_frac": alg / step_s / 1e9 / HBM_PEAK_GBS, "build_s": build_s, "autotune_s": tune_s}
    print(f"# batched: {len(models)} realizations, {u.nfree} DOF, {step_s * 1e6:.1f} us/step, "
          f"{out['dof_steps_per_s']:.3e} DOF-steps/s, {out['achieved_GBs']:.0f} GB/s, tile {best}",
          file=sys.stderr, flush=True)
    dm.close()
    return out


def drm_bench(args, us_per_step_plain):
    """The headline half-space loaded through a domain reduction layer (synthetic SV
    plane-wave signals) instead of the point source. The layer's nodal forces of every step
    are evaluated when the model is created (they depend only on the signals); the fused step
    adds them at the layer's nodes. Reported beside the plain step time of the same mesh."""
    from svl_amd import workloads as W
    from svl_amd.integrators import CentralDifference, DeviceModel
    nx, ny, nz = (int(v) for v in args.ne.split(","))
    t0 = time.perf_counter()
    m, dt, nt = W.halfspace3d_drm(nx, ny, nz, h=2.0, nt=args.drm_steps + args.drm_warmup + 2)
    build_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    dm = DeviceModel(m, tile_nodes=args.tile_nodes, block=args.block, variant=args.variant)
    create_s = time.perf_counter() - t0
    cd = CentralDifference(dm, dt)
    cd.Initialize()
    ev = HipEvents()
    stream = dm.stream()
    cd.advance_async(1, args.drm_warmup)
    dm.sync()
    e0, e1 = ev.event(), ev.event()
    t = time.perf_counter()
    ev.record(e0, stream)
    cd.advance_async(1 + args.drm_warmup, args.drm_steps)
    ev.record(e1, stream)
    dm.sync()
    wall = time.perf_counter() - t
    step_s = ev.elapsed_ms(e0, e1) / 1e3 / args.drm_steps
    # Per step the fused kernel reads the slot (4 B) and the force (24 B) of every DRM node.
    drm_bytes = 28 * m.drm_node.size
    out = {"workload": f"halfspace3d lin3DHexa8 {nx}x{ny}x{nz} h=2m, Lysmer, DRM box 0.8 L (open at the surface), "
                       f"synthetic SV Ricker signals, CentralDifference",
           "ndrm_elem": int(m.drm_elem.size), "ndrm_node": int(m.drm_node.size), "nfree": m.nfree,
           "steps": args.drm_steps, "us_per_step": step_s * 1e6, "us_per_step_point_source": us_per_step_plain,
           "drm_overhead_us": step_s * 1e6 - us_per_step_plain, "dof_steps_per_s": m.nfree * args.drm_steps / wall,
           "drm_alg_bytes_per_step": drm_bytes, "signal_rows": nt, "build_s": build_s,
           "model_create_s": create_s}
    print(f"# drm: {m.drm_elem.size} DRM elements, {step_s * 1e6:.2f} us/step "
          f"(point source {us_per_step_plain:.2f})", file=sys.stderr, flush=True)
    dm.close()
    return out


def pml_bench(args):
    """2D half-space wrapped by a 2D PML (PML2DQuad4 + lin2DQuad4, EQUAL constraints) under a
    surface Ricker load, ExtendedNewmarkBeta with a scaled MINRES solve per step. Per MINRES
    iteration the algorithmic traffic is 12 B per nonzero (value + column) and 116 B per
    free DOF (row pointer, the gathered vector, and the three-term vector updates)."""
    import numpy as np
    from svl_amd.pml import PmlNewmark, halfspace_box
    nx, ny, npml = (int(v) for v in args.pml_ne.split(","))
    dt = 0.005
    nt = args.pml_steps + args.pml_warmup + 2
    tt = np.arange(nt) * dt
    a = (np.pi * 10.0 * (tt - 0.15)) ** 2
    t0 = time.perf_counter()
    m = halfspace_box(nx, ny, npml, load_series=(1 - 2 * a) * np.exp(-a))
    build_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    pm = PmlNewmark(m, dt)
    create_s = time.perf_counter() - t0
    info = pm.info()
    pm.advance(1, args.pml_warmup)
    t0 = time.perf_counter()
    _, _, _, its = pm.advance(1 + args.pml_warmup, args.pml_steps)
    wall = time.perf_counter() - t0
    us_step = wall / args.pml_steps * 1e6
    # The same window again with eagerly launched MINRES iterations (no graph replay).
    pm.set_graph(False)
    t0 = time.perf_counter()
    pm.advance(1 + args.pml_warmup + args.pml_steps, args.pml_steps // 2)
    us_step_eager = (time.perf_counter() - t0) / (args.pml_steps // 2) * 1e6
    iters = float(its.mean())
    it_bytes = 12 * info["nnz"] + 116 * info["nfree"]
    out = {"workload": f"2D half-space lin2DQuad4 {nx}x{ny} h=1m wrapped by PML2DQuad4 {npml} elements thick "
                       f"(left, right, bottom), EQUAL ties, Ricker surface load, ExtendedNewmarkBeta dt={dt}",
           "nelem": int(len(m.conn)), **info, "steps": args.pml_steps, "us_per_step": us_step,
           "minres_iters_mean": iters, "us_per_minres_iter": us_step / max(iters, 1.0),
           "alg_bytes_per_minres_iter": it_bytes, "us_per_step_eager_launches": us_step_eager,
           "achieved_GBs_per_iter": it_bytes / (us_step / max(iters, 1.0) * 1e-6) / 1e9,
           "build_s": build_s, "model_create_s": create_s}
    print(f"# pml: {info['nfree']} free DOFs, {us_step:.0f} us/step, {iters:.1f} MINRES iterations",
          file=sys.stderr, flush=True)
    del pm
    return out


def pml3d_bench(args):
    """3D half-space wrapped by PML3DHexa8 on five faces (edge and corner elements carry the
    history matrix G and the integrated displacement Ubar), ExtendedNewmarkBeta with the scaled
    MINRES solve; eagerly launched MINRES iterations. Per MINRES iteration the algorithmic traffic
    is 12 B per nonzero and 116 B per free DOF, as in 2D."""
    import numpy as np
    from svl_amd.pml import PmlNewmark, halfspace_box3d
    nx, ny, nz, npml = (int(v) for v in args.pml3d_ne.split(","))
    dt = 0.005
    nt = args.pml_steps + args.pml_warmup + 2
    tt = np.arange(nt) * dt
    a = (np.pi * 10.0 * (tt - 0.15)) ** 2
    t0 = time.perf_counter()
    m = halfspace_box3d(nx, ny, nz, npml, load_series=(1 - 2 * a) * np.exp(-a))
    build_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    pm = PmlNewmark(m, dt)
    pm.set_graph(False)
    create_s = time.perf_counter() - t0
    info = pm.info()
    pm.advance(1, args.pml_warmup)
    t0 = time.perf_counter()
    _, _, _, its = pm.advance(1 + args.pml_warmup, args.pml_steps)
    wall = time.perf_counter() - t0
    us_step = wall / args.pml_steps * 1e6
    iters = float(its.mean())
    it_bytes = 12 * info["nnz"] + 116 * info["nfree"]
    out = {"workload": f"3D half-space lin3DHexa8 {nx}x{ny}x{nz} h=1m wrapped by PML3DHexa8 {npml} elements thick "
                       f"(four sides and bottom), EQUAL ties, Ricker surface load, ExtendedNewmarkBeta dt={dt}, "
                       f"eager MINRES launches",
           "nelem": int(len(m.conn)), **info, "steps": args.pml_steps, "us_per_step": us_step,
           "minres_iters_mean": iters, "us_per_minres_iter": us_step / max(iters, 1.0),
           "alg_bytes_per_minres_iter": it_bytes,
           "achieved_GBs_per_iter": it_bytes / (us_step / max(iters, 1.0) * 1e-6) / 1e9,
           "build_s": build_s, "model_create_s": create_s}
    print(f"# pml3d: {info['nfree']} free DOFs, {us_step:.0f} us/step, {iters:.1f} MINRES iterations",
          file=sys.stderr, flush=True)
    del pm
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--ne", type=str, default="100,100,50", help="hexa elements nx,ny,nz")
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile-nodes", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--variant", type=int, default=0,
                    help="0 auto, 1 fused + colour phases, 2 element-colored, 3 fused + corner slots")
    ap.add_argument("--implicit-steps", type=int, default=30)
    ap.add_argument("--implicit-ne", type=int, default=64, help="elements per side of the implicit block")
    ap.add_argument("--implicit-rtol", type=float, default=1e-8)
    ap.add_argument("--implicit-tile-nodes", type=int, default=330)
    ap.add_argument("--gnn-channels", type=int, default=1)
    ap.add_argument("--gnn-layers", type=int, default=8)
    ap.add_argument("--gnn-epochs", type=int, default=30)
    ap.add_argument("--no-implicit", action="store_true")
    ap.add_argument("--nonlinear-steps", type=int, default=20)
    ap.add_argument("--nonlinear-skip", type=int, default=150,
                    help="untimed steps before the timed window of the J2 soil-structure run")
    ap.add_argument("--no-nonlinear", action="store_true")
    ap.add_argument("--batch-n", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=600)
    ap.add_argument("--batch-warmup", type=int, default=30)
    ap.add_argument("--batch-graph-steps", type=int, default=30)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-tile-nodes", type=int, default=0, help="fixed node-tile size (0: autotune)")
    ap.add_argument("--drm-steps", type=int, default=300)
    ap.add_argument("--drm-warmup", type=int, default=30)
    ap.add_argument("--no-drm", action="store_true")
    ap.add_argument("--pml-ne", type=str, default="300,150,15", help="PML leg: nx,ny,PML elements")
    ap.add_argument("--pml-steps", type=int, default=40)
    ap.add_argument("--pml-warmup", type=int, default=5)
    ap.add_argument("--no-pml", action="store_true")
    ap.add_argument("--pml3d-ne", type=str, default="24,24,12,4", help="3D PML leg: nx,ny,nz,PML elements")
    ap.add_argument("--no-pml3d", action="store_true")
    ap.add_argument("--layer-steps", type=int, default=2000)
    ap.add_argument("--no-layer", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--cpu-procs", type=int, default=0, help="concurrent oracle processes (0: one per host core)")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MB",
                    help="only run the device copy of MB megabytes (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calib:
        print(json.dumps({"copy_calibration": copy_calibration(args.copy_calib * 1000000)}), flush=True)
        return
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
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_from_profile(info["variant"], args.tile_nodes or 160),
                     "kernel": "k_cd_corner / k_cd_fused: one launch = one CentralDifference step",
                     "avg_launch_us": step_s * 1e6,
                     "alg_bytes_per_step": bytes_step},
    }
    dm.close()
    if not args.no_precision:
        out["precision"] = precision_bench(args, m, dt, step_s * 1e6)
    if not args.no_layer:
        out["layer2d"] = layer2d_bench(args)
    if not args.no_implicit:
        out["implicit"] = implicit_bench(args)
    if not args.no_nonlinear:
        out["nonlinear"] = nonlinear_bench(args)
    if not args.no_batch:
        out["batched"] = batched_bench(args)
    if not args.no_drm:
        out["drm"] = drm_bench(args, step_s * 1e6)
    if not args.no_pml:
        out["pml2d"] = pml_bench(args)
    if not args.no_pml3d:
        out["pml3d"] = pml3d_bench(args)
    if not args.no_cpu:
        ncores, cpu_model = host_cores()
        out["cpu_baseline"] = cpu_baseline(m, dt, args.cpu_steps)
        out["cpu_baseline"]["cores"] = 1
        out["cpu_baseline"]["host"] = f"{ncores} cores, {cpu_model}"
        out["cpu_baseline_all_cores"] = cpu_baseline_parallel(m, dt, max(1, args.cpu_steps // 2),
                                                             args.cpu_procs or ncores)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
