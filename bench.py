#!/usr/bin/env python3
"""Benchmark of the gfx950 Shakhov CDUGKS step on the Kn = 1 cavity.

Workload (one GPU): 200x200 quads jittered by 10% (40,000 cells), 101x101
Newton-Cotes velocities on [-5, 5]^2 (10,201), Shakhov model. A step = one
full time step (all kernels) of the whole mesh; value = DV-cell updates/s =
NC * NV * steps / seconds, inputs resident in HBM, timed with a device
synchronisation on both sides.

The step runs in three storages of the distributions g, h:
  fp32   FP32 words (deviation from the rest Maxwellian, stochastically
         rounded), FP64 transport arithmetic, FP32 equilibria
  fp32e  an FP32 word plus a byte extending its significand to 31 bits (5 B),
         the FP32 kernels
  fp32x  the same format, the FP64 kernels
  fp64   meso's own precision
The line's value is the fastest storage whose every macro field is within
1e-6 relative L2 of the FP64 run at meso's convergence, in this run's
reduced-resolution leg (at_convergence) and in the committed full-resolution
measurement (profiles/convergence_<workload>_full.json,
at_convergence_full_resolution); with no convergence leg, FP32 (--headline
picks one).
Reported per storage: ms/step, DV-cell updates/s, the roofline of the
dominant kernel (algorithmic bytes: read + write of every stored g, h value of
the DV-cells the launch owns, over its average launch time from HIP events),
per-kernel times, and the relative L2 of every macro against FP64 after the
timed run. at_convergence: the same workload at --converge-scale run to meso's
stopping rule in all three storages, with the errors there. cpu_baseline: the
FP64 oracle with OpenMP on this host (one step of the same workload) and its
measured speed ratio to meso itself.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--scale s]
                       [--storage all|fp32x|fp32e|fp32|fp64] [--headline auto|fp32x|fp32e|fp32|fp64]
                       [--converge-scale s] [--copy-calibration MIB]
                       [--autotune] [--graphs] [--batch N]
       python bench.py --converge [--workload W] [--scale s] [--max-steps N] [--time-limit S] [--run-on]

With --autotune each context first runs meso_hip_autotune (trial steps of the
initial state, not committed) to choose its patch and velocity-chunk tiles.
The JSON line records the tiles in use (config.tiles).
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "meso-lts_amd"))

from meso_hip import capi, configs  # noqa: E402

METRIC = "DV-cell updates/s, HBM GB/s % of peak; FP32-store vs FP64 at <=1e-6 rel err"
HBM_PEAK_GBS = 8000.0
# Speed of the restatement (the oracle, OpenMP threads) over meso itself (MPI
# ranks over the velocity set, src/solver/dvs/dvs_mpi.cpp:5-28), compiled from
# its sources with -O2 (oracle/Makefile.ref), timed on identical inputs in the
# 8-core build container: the Kn = 1 cavity at half resolution (100^2 quads x
# 101^2 velocities), seconds per step from 1- and 3-step runs.
CALIBRATION = {"oracle_over_reference_1core": 1.91, "oracle_over_reference_8cores": 1.06,
               "reference_8ranks_dv_cell_per_s": 7.73e6, "oracle_8threads_dv_cell_per_s": 8.21e6,
               "where": "8-core Intel Xeon build container, not the GPU host",
               "workload": "cavity_kn1_quad at scale 0.5 (10,000 cells x 10,201 velocities)"}
WORKLOAD = "cavity_kn1_quad"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a - b))


def time_steps(dev, steps, warmup):
    dev.step(warmup)
    dev.sync()
    t0 = time.perf_counter()
    rep = dev.step(steps)
    dev.sync()
    t1 = time.perf_counter()
    if rep.bad_cells:
        raise RuntimeError(f"numerical failure in the timed steps (cell {rep.first_bad_cell})")
    return t1 - t0


def kernel_profile(dev, steps):
    dev.profile(True)
    dev.step(steps)
    kt = dev.kernel_times()
    dev.profile(False)
    return {k: {"avg_ms": v[0] / max(v[1], 1), "launches": v[1]} for k, v in kt.items()}


def cpu_baseline(case_path, steps):
    sys.path.insert(0, ROOT)
    from oracle.oracle import Oracle, get_threads

    cs = capi.Case(case_path)
    orc = Oracle(cs)
    orc.init_equilibrium()
    t0 = time.perf_counter()
    orc.step(steps)
    dt = time.perf_counter() - t0
    ups = cs.n_cells * cs.n_vel * steps / dt
    return {"value": ups, "unit": "DV-cell updates/s", "cores": get_threads(), "kind": "port",
            "label": "restatement",
            "sample": f"{steps} full step(s) of the same workload ({cs.n_cells} cells x {cs.n_vel} velocities), "
                      f"FP64 oracle with OpenMP, {dt:.1f} s",
            "calibration": CALIBRATION}


def copy_calibration(mib):
    """A plain float32 device-to-device copy of `mib` MiB, three launches: the
    HBM byte counters of one launch should read `mib` MiB fetched and written."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"[bench] copy calibration: 3 x {mib} MiB copies done")


def batch_main(args):
    """The batched cavity sweep: N independent Shakhov cavities (quads 32^2 to
    128^2, Kn log-uniform in [0.01, 10], lid speed U[0.05, 0.3], Newton-Cotes
    16^2 to 64^2 velocities, seeds 1000 + i), FP32 storage; value = sum of
    NC * NV * steps / seconds of the fused batch (meso_hip_batch_step: one
    launch per stage over all instances, steps replayed as captured pairs
    with --graphs). Also timed on the same devices: round-robin issue on the
    instances' own streams (meso_hip_step_batch) and one instance after the
    other (meso_hip_step)."""
    work = os.path.join(tempfile.gettempdir(), "meso_hip_bench")
    insts = configs.batch_instances(args.batch)
    cases = [capi.Case(configs.build_instance(i, work, args.scale)) for i in insts]
    devs = []
    chunks = []
    for cs in cases:
        d = capi.Device(cs, capi.STORAGE_FP32)
        d.init_equilibrium()
        chunks.append(d.autotune(3) if args.autotune else 0)
        devs.append(d)
    dvc = sum(cs.n_cells * cs.n_vel for cs in cases)
    log(f"[bench] batch of {len(cases)}: {dvc:.4e} DV-cells in total "
        f"(cells {min(c.n_cells for c in cases)}..{max(c.n_cells for c in cases)}, "
        f"velocities {min(c.n_vel for c in cases)}..{max(c.n_vel for c in cases)})")

    def timed(run):
        run(args.warmup)
        for d in devs:
            d.sync()
        t0 = time.perf_counter()
        reps = run(args.steps)
        for d in devs:
            d.sync()
        secs = time.perf_counter() - t0
        if reps is not None and any(r.bad_cells for r in reps):
            raise RuntimeError("numerical failure in the batch")
        return secs

    batch = capi.Batch(devs)
    batch.set_graphs(args.graphs)
    batch.set_streams(args.batch_streams)
    fused_info = batch.info()
    tl = [d.tiles() for d in devs]
    log("[bench] tiles after the batch plan: vchunk " + " ".join(str(t["vchunk"]) for t in tl))
    log("[bench] LDS per block: moments " + " ".join(str(t["moment_lds"]) for t in tl) +
        " | update " + " ".join(str(t["update_lds"]) for t in tl))
    secs = timed(batch.step)
    log(f"[bench] fused batch ({fused_info[0]} instances in {fused_info[1]} group(s), {fused_info[2]} launches "
        f"per step): {secs / args.steps * 1e3:.3f} ms per step of all instances, {dvc * args.steps / secs:.4e} "
        f"DV-cell updates/s")
    batch.close()
    for d in devs:
        d.set_graphs(args.graphs)
    rr = timed(lambda n: capi.step_batch(devs, n))
    log(f"[bench] round-robin (meso_hip_step_batch): {rr / args.steps * 1e3:.3f} ms per step")

    def one_by_one(n):
        for d in devs:
            d.step(n)
    seq = timed(one_by_one)
    log(f"[bench] one after the other: {seq / args.steps * 1e3:.3f} ms per step")
    ups = dvc * args.steps / secs
    line = {
        "metric": METRIC, "value": ups, "unit": "DV-cell updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": secs / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"batch_{len(cases)}", "instances": len(cases), "dv_cells": dvc,
                   "autotune": bool(args.autotune), "graphs": bool(args.graphs),
                   "storage": "fp32 g,h; fp64 moments/macros",
                   "fused": {"instances": fused_info[0], "groups": fused_info[1],
                             "launches_per_step": fused_info[2], "streams": args.batch_streams}},
        "round_robin_ms_per_step": rr / args.steps * 1e3,
        "sequential_ms_per_step": seq / args.steps * 1e3,
        "fused_over_round_robin": rr / secs,
        "fused_over_sequential": seq / secs,
        "vchunks": chunks if args.autotune else None,
    }
    for d in devs:
        d.close()
    print(json.dumps(line), flush=True)


def macro_errors(a, b, dim, shakhov):
    """Relative L2 over cells (unweighted) of every macro component of a against b."""
    e = {"rho": rel_l2(a["rho"], b["rho"])}
    for i, c in enumerate("UVW"[:dim]):
        e[c] = rel_l2(a["u"][:, i], b["u"][:, i])
    if shakhov:
        e["T"] = rel_l2(a["T"], b["T"])
        for i, c in enumerate("xyz"[:dim]):
            e["q" + c] = rel_l2(a["q"][:, i], b["q"][:, i])
    return e


def converge_one(args):
    """One storage of the workload stepped alone (at --scale): to meso's
    stopping rule (every 100 steps, ten checks in a row with every residual
    <= the case's residual-limit) or for --converge-steps steps; the final
    macros and the residual history go to --out."""
    import numpy as np
    work = os.path.join(tempfile.gettempdir(), "meso_hip_bench")
    cs = capi.Case(configs.build(args.workload, work, args.scale))
    dim = cs.mesh.dim
    limit = cs.setting_double("residual-limit", 1e-6)
    idx = [0, 4, 1, 2, 5, 6] if dim == 2 else [0, 4, 1, 2, 3, 5, 6, 7]
    st = {"fp64": capi.STORAGE_FP64, "fp32": capi.STORAGE_FP32, "fp32e": capi.STORAGE_FP32E,
          "fp32x": capi.STORAGE_FP32X}[args.converge_one]
    dev = capi.Device(cs, st)
    dev.init_equilibrium()
    dev.residual()
    t0, step, run, hist = time.time(), 0, 0, []
    while step < args.max_steps and time.time() - t0 < args.time_limit:
        rep = dev.step(100)
        if rep.bad_cells:
            raise RuntimeError(f"numerical failure at step {step + rep.steps_done}")
        step += 100
        r = np.asarray(dev.residual())[idx]
        hist.append(float(r.max()))
        run = run + 1 if bool(np.all(r <= limit)) else 0
        if step % 1000 == 0:
            log(f"[converge-one] {args.converge_one} step {step} residual {r.max():.3e} ({run}) t={time.time() - t0:.0f}s")
        if (args.converge_steps == 0 and run >= 10) or (args.converge_steps and step >= args.converge_steps):
            break
    m = dev.macros()
    if args.out:
        np.savez(args.out, step=step, rho=m["rho"], u=m["u"], T=m["T"], q=m["q"], residual=np.array(hist), dim=dim,
                 shakhov=int(cs.model == capi.MODEL_CDUGKS_SHAKHOV))
    print(json.dumps({"storage": args.converge_one, "workload": args.workload, "mesh_scale": args.scale,
                      "steps": step, "converged": run >= 10, "seconds": time.time() - t0,
                      "last_residual": hist[-1] if hist else None}), flush=True)


def converge_compare(paths):
    """Relative L2 over cells of every macro of each npz against the first."""
    import numpy as np
    ref = np.load(paths[0])
    out = {"reference": os.path.basename(paths[0]), "steps": int(ref["step"]), "errors": {}}
    for p in paths[1:]:
        b = np.load(p)
        if int(b["step"]) != int(ref["step"]):
            raise SystemExit(f"{p}: {int(b['step'])} steps against {int(ref['step'])}")
        out["errors"][os.path.basename(p)] = macro_errors(
            {"rho": b["rho"], "u": b["u"], "T": b["T"], "q": b["q"]},
            {"rho": ref["rho"], "u": ref["u"], "T": ref["T"], "q": ref["q"]}, int(ref["dim"]), bool(ref["shakhov"]))
    print(json.dumps(out), flush=True)


def converge_main(args):
    line = converge_run(args.workload, args.scale, args.max_steps, args.time_limit, args.residual_limit,
                        args.curve_every, args.run_on)
    print(json.dumps(line), flush=True)


def converge_run(workload, scale, max_steps, time_limit, residual_limit=0.0, curve_every=1000, run_on=False):
    """The workload run to meso's stopping rule in FP64, FP32 and FP32X storage
    side by side: every 100 steps meso's residuals (meso_hip_residual), converged
    after 10 consecutive checks with every residual <= the case's
    residual-limit (cdugks_shakhov.cpp:730-756). Reports the step each storage
    converged at and every macro's FP32 / FP32X error against the FP64 run at
    the FP64 convergence step (and at the end of the run). The run ends when
    FP64 has converged (with run_on: when every storage has), or at
    max_steps / time_limit."""
    work = os.path.join(tempfile.gettempdir(), "meso_hip_bench")
    cs = capi.Case(configs.build(workload, work, scale))
    dim = cs.mesh.dim
    shakhov = cs.model == capi.MODEL_CDUGKS_SHAKHOV
    limit = cs.setting_double("residual-limit", 1e-6)
    if residual_limit > 0:
        limit = residual_limit
    idx = ([0, 4, 1, 2, 5, 6] if dim == 2 else [0, 4, 1, 2, 3, 5, 6, 7]) if shakhov else list(range(1 + dim))
    kinds = [("fp64", capi.STORAGE_FP64), ("fp32", capi.STORAGE_FP32), ("fp32e", capi.STORAGE_FP32E),
             ("fp32x", capi.STORAGE_FP32X)]
    devs = {}
    for tag, st in kinds:
        try:
            d = capi.Device(cs, st)
        except capi.MesoError as e:
            log(f"[converge] {tag}: {e}")
            continue
        d.init_equilibrium()
        d.residual()
        devs[tag] = d
    log(f"[converge] {workload} scale {scale}: NC={cs.n_cells} NV={cs.n_vel} dt={cs.params.dt:.4e} "
        f"residual-limit {limit:g}")
    state = {t: 0 for t in devs}
    conv_at = {t: None for t in devs}
    history = {t: [] for t in devs}
    at_fp64 = None
    curve = []
    t0 = time.time()
    step = 0
    while step < max_steps and time.time() - t0 < time_limit:
        live = [t for t in devs if conv_at[t] is None]
        if not live:
            break
        for t in live:
            rep = devs[t].step(100)
            if rep.bad_cells:
                raise RuntimeError(f"{t}: numerical failure at step {step + rep.steps_done}")
        step += 100
        for t in live:
            r = np.asarray(devs[t].residual())[idx]
            rmax = float(np.max(r))
            history[t].append(rmax)
            state[t] = state[t] + 1 if bool(np.all(r <= limit)) else 0
            if state[t] >= 10:
                conv_at[t] = step
        errs = None
        if step % curve_every == 0 or conv_at["fp64"] == step:
            ref = devs["fp64"].macros()
            errs = {t: macro_errors(devs[t].macros(), ref, dim, shakhov) for t in devs if t != "fp64"}
            curve.append({"step": step, **{t: max(e.values()) for t, e in errs.items()}})
        if conv_at["fp64"] == step:
            at_fp64 = {"step": step, "errors": errs}
        log(f"[converge] step {step} " + " ".join(f"{t} res {history[t][-1]:.2e}({state[t]})" for t in live) +
            ("" if errs is None else " | err " + " ".join(f"{t} {max(e.values()):.2e}" for t, e in errs.items())) +
            f" t={time.time() - t0:.0f}s")
        if conv_at["fp64"] is not None and not run_on:
            break
    ref = devs["fp64"].macros()
    final = {t: macro_errors(devs[t].macros(), ref, dim, shakhov) for t in devs if t != "fp64"}
    line = {
        "metric": "FP32-store vs FP64 rel L2 at convergence", "workload": workload, "mesh_scale": scale,
        "cells": cs.n_cells, "velocities": cs.n_vel, "dt": cs.params.dt, "residual_limit": limit,
        "rule": "10 consecutive checks every 100 steps, every residual <= limit (cdugks_shakhov.cpp:730-756)",
        "converged_at": conv_at, "steps_run": step, "seconds": time.time() - t0,
        "at_fp64_convergence": at_fp64, "at_end": {"step": step, "errors": final},
        "max_error_curve": curve,
        "residual_history_every_100": {t: h[-50:] for t, h in history.items()},
    }
    for d in devs.values():
        d.close()
    return line


def renumbered_case(case_path, how):
    """A copy of the case directory whose mesh file has its cells renumbered."""
    import shutil
    src = os.path.dirname(case_path)
    dst = f"{src}_{how}"
    mesh = capi.Case(case_path).setting_str("mesh-file")
    if not os.path.exists(os.path.join(dst, mesh)):
        shutil.rmtree(dst, ignore_errors=True)
        shutil.copytree(src, dst)
        method = capi.RENUMBER_RCB if how == "rcb" else capi.RENUMBER_RANDOM
        capi.renumber_mesh(os.path.join(src, mesh), os.path.join(dst, mesh), method, 12345)
    log(f"[bench] mesh renumbered ({how}): {os.path.join(dst, mesh)}")
    return os.path.join(dst, os.path.basename(case_path))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--scale", type=float, default=1.0, help="mesh resolution scale (<1: quick runs)")
    ap.add_argument("--workload", default=WORKLOAD)
    ap.add_argument("--storage", choices=("all", "fp32x", "fp32e", "fp32", "fp64"), default="all",
                    help="storage kinds to run (profiling runs use one; the JSON line needs all)")
    ap.add_argument("--headline", choices=("auto", "fp32x", "fp32e", "fp32", "fp64"), default="auto",
                    help="storage whose rate is the line's value (auto: the fastest one within 1e-6 of FP64 "
                         "on every macro field at convergence)")
    ap.add_argument("--converge-scale", type=float, default=0.25,
                    help="also run the workload at this mesh scale to meso's convergence rule in all storages "
                         "and report the errors there (0: skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (HBM counter calibration)")
    ap.add_argument("--batch", type=int, default=0, metavar="N",
                    help="instead: the batched cavity sweep of N seeded instances (FP32 storage)")
    ap.add_argument("--batch-streams", type=int, default=4, metavar="S",
                    help="--batch: sub-batches of the fused batch, one stream each (meso_hip_batch_set_streams)")
    ap.add_argument("--vchunk", type=int, default=0,
                    help="velocities per block of the patch kernels (meso_hip_set_vchunk; 0: the default)")
    ap.add_argument("--autotune", action="store_true",
                    help="run meso_hip_autotune before timing (off by default: the default tiles are "
                         "deterministic, so that profiles of the same command see the same kernels)")
    ap.add_argument("--graphs", action="store_true", help="replay captured step pairs (meso_hip_set_graphs)")
    ap.add_argument("--renumber", choices=("rcb", "random"), default=None,
                    help="run the workload on its mesh file renumbered (meso_mesh_renumber: locality order, or a "
                         "seeded random order to measure what a poor order costs)")
    ap.add_argument("--converge", action="store_true",
                    help="instead: run FP64, FP32 and FP32X storage to meso's convergence rule and report the "
                         "FP32 / FP32X errors against FP64 there")
    ap.add_argument("--max-steps", type=int, default=200000, help="--converge: step cap")
    ap.add_argument("--time-limit", type=float, default=900.0, help="--converge: wall-time cap in seconds")
    ap.add_argument("--residual-limit", type=float, default=0.0, help="--converge: override the case's limit")
    ap.add_argument("--curve-every", type=int, default=1000, help="--converge: steps between error samples")
    ap.add_argument("--run-on", action="store_true",
                    help="--converge: keep stepping FP32 / FP32X after FP64 converged until they do")
    ap.add_argument("--converge-one", choices=("fp64", "fp32", "fp32e", "fp32x"), default=None,
                    help="instead: run one storage alone (the full-resolution runs do not fit one call together): "
                         "to meso's convergence rule, or for --converge-steps steps; its macros go to --out (npz)")
    ap.add_argument("--converge-steps", type=int, default=0, help="--converge-one: steps (0: to convergence)")
    ap.add_argument("--out", default="", help="--converge-one: npz file of the final macros")
    ap.add_argument("--converge-compare", nargs="+", default=None, metavar="NPZ",
                    help="instead: relative L2 of every macro of the npz files (after the first) against the "
                         "first (the FP64 run), as one JSON line")
    args = ap.parse_args()
    if args.converge:
        converge_main(args)
        return
    if args.converge_one:
        converge_one(args)
        return
    if args.converge_compare:
        converge_compare(args.converge_compare)
        return
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.batch:
        batch_main(args)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the step is benchmarked on one MI355X (--gpus {args.gpus} "
              "is not supported; the multi-GPU velocity split is a design, not a measured path).", file=sys.stderr)
        sys.exit(2)

    work = os.path.join(tempfile.gettempdir(), "meso_hip_bench")
    case_path = configs.build(args.workload, work, args.scale)
    if args.renumber:
        case_path = renumbered_case(case_path, args.renumber)
    cs = capi.Case(case_path)
    nc, nv = cs.n_cells, cs.n_vel
    shakhov = cs.model == capi.MODEL_CDUGKS_SHAKHOV
    nd = 2 if shakhov else 1
    log(f"[bench] {args.workload}: NC={nc} NF={cs.n_faces} NV={nv} model={cs.model}")

    res = {}
    macros = {}
    kinds = [(capi.STORAGE_FP32X, "fp32x"), (capi.STORAGE_FP32E, "fp32e"), (capi.STORAGE_FP32, "fp32"),
             (capi.STORAGE_FP64, "fp64")]
    if args.storage != "all":
        kinds = [k for k in kinds if k[1] == args.storage]
    tiles = {}
    for storage, tag in kinds:
        dev = capi.Device(cs, storage)
        dev.init_equilibrium()
        if args.vchunk:
            dev.set_vchunk(args.vchunk)
        if args.autotune:
            dev.autotune(3)
        tiles[tag] = dev.tiles()
        log(f"[bench] {tag} tiles: {tiles[tag]}")
        dev.set_graphs(args.graphs)
        secs = time_steps(dev, args.steps, args.warmup)
        kt = kernel_profile(dev, min(args.steps, 5))
        macros[tag] = dev.macros()
        res[tag] = dict(seconds=secs, ups=nc * nv * args.steps / secs, kernels=kt, bytes=dev.device_bytes())
        log(f"[bench] {tag}: {secs / args.steps * 1e3:.3f} ms/step, {res[tag]['ups']:.4e} DV-cell updates/s")
        for k, v in sorted(kt.items(), key=lambda x: -x[1]["avg_ms"]):
            log(f"          {k:14s} {v['avg_ms']:.4f} ms x{v['launches']}")
        dev.close()

    if args.storage != "all":
        return
    dim = cs.mesh.dim
    # bytes per stored distribution value: FP32X keeps an FP32 word and a byte
    sbytes = {"fp32": 4, "fp32x": 5, "fp32e": 5, "fp64": 8}

    def roofline(tag):
        """Dominant kernel of a storage: algorithmic bytes (read + write of
        every distribution value of the DV-cells the launch owns) over its
        average launch time."""
        kk = res[tag]["kernels"]
        dom = max(kk, key=lambda k: kk[k]["avg_ms"])
        dom_ms = kk[dom]["avg_ms"]
        # DV-cells one launch of the dominant kernel processes, as the library
        # reports them: the update launch owns update_owned cells (every cell
        # without a mirror boundary); the mirror-boundary cells have their own launch
        owned = tiles[tag]["update_owned"]
        units = {"cell_update": owned * nv, "cell_update_bnd": (nc - owned) * nv}.get(dom, nc * nv)
        b_alg = 2 * nd * sbytes[tag]
        achieved = b_alg * units / (dom_ms * 1e-3) / 1e9
        traffic = None
        tpath = os.path.join(ROOT, "profiles", f"traffic_{args.workload}_{tag}.json")
        if os.path.exists(tpath):
            with open(tpath) as fp:
                traffic = json.load(fp).get(dom)
        return {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "alg_bytes_per_dv_cell": b_alg,
                "alg_bytes_per_launch": b_alg * units, "avg_launch_ms": dom_ms}

    per = {}
    for tag in res:
        ms = res[tag]["seconds"] / args.steps * 1e3
        b_alg = 2 * nd * sbytes[tag]
        gbs = b_alg * nc * nv / (ms * 1e-3) / 1e9
        per[tag] = {"ms_per_step": ms, "value": res[tag]["ups"], "roofline": roofline(tag),
                    "step_hbm": {"alg_bytes_per_step": b_alg * nc * nv, "achieved_GBs": gbs,
                                 "frac": gbs / HBM_PEAK_GBS},
                    "kernels_ms": {k: v["avg_ms"] for k, v in res[tag]["kernels"].items()}}
        if tag != "fp64":
            per[tag]["vs_fp64_rel_l2_after_run"] = macro_errors(macros[tag], macros["fp64"], dim, shakhov)
            per[tag]["over_fp64"] = res[tag]["ups"] / res["fp64"]["ups"]

    conv = None
    if args.converge_scale > 0:
        conv = converge_run(args.workload, args.converge_scale, 20000, 240.0)

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(case_path, args.cpu_steps)
            log(f"[bench] cpu baseline: {cpu['value']:.4e} DV-cell updates/s on {cpu['cores']} threads")
        except Exception as e:  # the baseline is reported, never required
            log(f"[bench] cpu baseline failed: {e}")

    # the full-resolution run of the workload to meso's convergence (about 25
    # minutes of GPU time for the four storages, so a committed measurement,
    # not part of this run)
    full = None
    fpath = os.path.join(ROOT, "profiles", f"convergence_{args.workload}_full.json")
    if args.scale == 1.0 and os.path.exists(fpath):
        with open(fpath) as fp:
            full = json.load(fp)
    head = args.headline
    if head == "auto":
        # the fastest storage that holds the precision target at convergence:
        # in this run's reduced-resolution leg and in the committed
        # full-resolution measurement
        head = "fp32"
        if conv and conv.get("at_fp64_convergence"):
            errs = conv["at_fp64_convergence"]["errors"]
            ok = [t for t in ("fp32", "fp32e", "fp32x") if t in errs and t in res and max(errs[t].values()) <= 1e-6]
            if full:
                ok = [t for t in ok if t not in full["errors"] or max(full["errors"][t].values()) <= 1e-6]
            head = max(ok + ["fp64"], key=lambda t: res[t]["ups"])
    storage_txt = {"fp32x": "fp32 word + 8-bit significand extension per g,h value (5 B); fp64 arithmetic",
                   "fp32e": "fp32 word + 8-bit significand extension per g,h value (5 B); the fp32 kernels (fp32 "
                            "face values and fluxes, stochastically rounded; fp64 transport arithmetic and macros)",
                   "fp32": "fp32 g,h (deviation from the rest Maxwellian, stochastically rounded); fp64 transport "
                           "arithmetic and moments/macros",
                   "fp64": "fp64"}
    line = {
        "metric": METRIC,
        "value": res[head]["ups"],
        "unit": "DV-cell updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": per[head]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "cells": nc, "faces": cs.n_faces, "velocities": nv,
                   "distributions": nd, "storage": storage_txt[head], "headline_storage": head,
                   "mesh_scale": args.scale, "graphs": bool(args.graphs), "autotune": bool(args.autotune),
                   "tiles": tiles.get(head)},
        "roofline": per[head]["roofline"],
        "step_hbm": per[head]["step_hbm"],
        "storages": per,
        "at_convergence": conv,
        "at_convergence_full_resolution": full,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
