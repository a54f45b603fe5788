#!/usr/bin/env python3
"""Throughput of the matrix-free tetrahedral vmult on one MI355X.

A step is one vmult over the whole mesh with the input already resident in
HBM.  Default workload: CG Poisson P2, FP64, unit cube 128^3 hexes x 6
perturbed Kuhn tets (12,582,912 cells, 16,974,593 DoFs) -- the HBM-resident
size of the reference's CG cells-batched operator
(CG/cells_batched/application.cc:1236-1416; metric definition GDoF/s =
n_dofs * steps / time, application.cc:1871-1889).

Prints one JSON line with the metric, a roofline object (algorithmic bytes of
the reference data model / event-timed vmult duration vs 8 TB/s) and a CPU
baseline (the oracle's OpenMP vmult on the host, bounded sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "simplex-benchmarks_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

METRIC = "vmult GDoF/s + %HBM peak; time to 1e-10 solve: learned vs plain hybrid MG"
HBM_PEAK_GBS = 8000.0
KERNEL = {"cg": "cg_wave_kernel + cg_fixup_kernel (one launch each per vmult)",
          "dg": "dg_cell_kernel (one launch per vmult)", "dg_vector": "dg_cell_kernel",
          "dg_helmholtz": "dg_cell_kernel"}
WORKLOAD = {"cg": "CG Poisson", "dg": "DG SIPG Poisson", "dg_vector": "DG SIPG vector Laplace",
            "dg_helmholtz": "DG Helmholtz-like vector (mass 1 + Laplace 1)"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=128, help="hexes per cube edge")
    ap.add_argument("--degree", type=int, default=2)
    ap.add_argument("--kind", default="cg")
    ap.add_argument("--precision", default="f64")
    ap.add_argument("--no-reorder", action="store_true")
    ap.add_argument("--patch-cells", type=int, default=0)
    ap.add_argument("--accumulate", type=int, default=0, help="CG LDS accumulation: 0 auto, 1 atomics, 2 read-add-write")
    ap.add_argument("--cpu-baseline", type=int, default=1, help="0 to skip the CPU leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--seed", type=int, default=7)
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(op, x, seconds):
    """Oracle (CPU restatement of the reference vmult), OpenMP over cell
    colours on the host cores; timed over whole vmults until `seconds`."""
    import oracle_py

    pr = oracle_py.OracleProblem.from_operator(op)
    colors = pr.colors()
    dst = np.empty(pr.n_dofs)
    pr.vmult_omp(x, colors, dst)  # warm-up
    calls, t0 = 0, time.perf_counter()
    while True:
        pr.vmult_omp(x, colors, dst)
        calls += 1
        el = time.perf_counter() - t0
        if el >= seconds or calls >= 50:
            break
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": pr.n_dofs * calls / el / 1e9, "unit": "GDoF/s", "cores": cores,
            "kind": "port",
            "sample": f"{calls} full vmults of the same mesh and vector on {cores} host threads "
                      f"({el:.1f} s); oracle uses a collapsed Gauss rule with "
                      f"{(op.info['degree'] + 2) ** 3} points per cell"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; "
                                   "multi-GPU sharding is a design note only"}))
        sys.exit(2)
    import simplexmf as smf

    if smf.device_count() < 1:
        raise SystemExit("no HIP device visible")
    t0 = time.time()
    mesh = smf.Mesh.kuhn_cube(args.n, 0.2, args.seed)
    op = smf.Operator(mesh, args.degree, kind=args.kind, precision=args.precision,
                      reorder=not args.no_reorder, patch_cells=args.patch_cells,
                      accumulate=args.accumulate)
    info = op.info
    log(f"setup {time.time() - t0:.1f}s: {info}")
    x = np.random.default_rng(args.seed).uniform(-1, 1, op.m())
    src = smf.DeviceVector.from_numpy(x.astype(op.dtype))
    dst = op.initialize_dof_vector()
    stream = smf.Stream()
    for _ in range(args.warmup):
        op.vmult(dst, src, stream)
    stream.synchronize()
    smf.synchronize()
    ev0, ev1 = smf.Event(), smf.Event()
    ev0.record(stream)
    for _ in range(args.steps):
        op.vmult(dst, src, stream)
    ev1.record(stream)
    stream.synchronize()
    smf.synchronize()
    ms = ev0.elapsed_ms(ev1)
    t_step = ms / 1e3 / args.steps
    gdofs = op.m() / t_step / 1e9
    bytes_alg = info["algorithmic_bytes"]
    achieved = bytes_alg / t_step / 1e9
    traffic = None
    prof = os.path.join(ROOT, "profiles", f"traffic_{args.kind}_p{args.degree}_{args.precision}_n{args.n}.json")
    if os.path.exists(prof):
        with open(prof) as f:
            traffic = json.load(f).get("bytes_per_vmult")
    base = None
    if args.cpu_baseline and args.kind == "cg":
        log("cpu baseline ...")
        base = cpu_baseline(op, x, args.cpu_seconds)
    prec = {"f64": "f64", "f32": "f32"}[args.precision]
    ncell = info["n_cells"]
    out = {
        "metric": METRIC,
        "value": round(gdofs, 4),
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": prec,
        "data": "synthetic: seeded perturbed Kuhn mesh, x ~ U[-1,1]",
        "config": {
            "workload": f"{WORKLOAD[args.kind]} P{args.degree} vmult, {args.n}^3 x 6 perturbed "
                        f"Kuhn tets ({ncell} cells, {op.m()} DoFs), {prec.upper()}",
            "kind": args.kind, "degree": args.degree, "n": args.n, "cells": ncell,
            "dofs": op.m(), "reorder": not args.no_reorder, "parallelism": "single GPU",
            "patches": info["n_patches"], "patch_colors": info["n_patch_colors"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes": bytes_alg,
            "kernel": KERNEL[args.kind],
        },
        "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
