#!/usr/bin/env python3
"""Throughput of the matrix-free tetrahedral vmult on one MI355X.

A step is one vmult over the whole mesh with the input already resident in
HBM.  Default workload: DG SIPG Poisson P2, FP64, unit cube 64^3 hexes x 6
perturbed Kuhn tets (1,572,864 cells, 15,728,640 DoFs), the reference's DG
cells-batched operator (DG/cells_batched/application.cc:1318-1332; metric
GDoF/s = n_dofs * steps / time, CG/cells_batched/application.cc:1871-1889).
The other degrees and precisions of that config, the Helmholtz-like vector
operator (32^3) and the CG P2 operator at 128^3 (the HBM-resident size of
CG/cells_batched/application.cc:1236-1416) are reported under "variants";
the CG operator against its assembled CSR matrix (the matrix-based timing
of CG/cells_batched/application.cc:1944-1988) under "matrix_based"; CG and
DG P2 on a mesh in a scrambled input order with the hierarchical reordering
on and off (CG/cells_batched/application.cc:1699-1756) under "external_mesh".

Prints one JSON line with the metric, a roofline object (algorithmic bytes of
the reference data model / event-timed vmult duration vs 8 TB/s) and a CPU
baseline (oracle/cpu_baseline.cpp, the reference's optimised DG vmult
restated for vectorised multi-core CPUs, on the same mesh and vector).
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
KERNEL = {"cg": "cg_wave_kernel + cg_fixup_kernel (one launch each per vmult)", "cg_vector": "cg_wave_kernel (NC=3) + cg_fixup_kernel",
          "dg": "dg_quad_kernel (four lanes per cell, block values staged in LDS; one launch per vmult)",
          "dg_vector": "dg_quad_kernel (NC=3)", "dg_helmholtz": "dg_quad_kernel (NC=3, mass term)"}
WORKLOAD = {"cg": "CG Poisson", "cg_vector": "CG vector Laplace (3 components)", "dg": "DG SIPG Poisson", "dg_vector": "DG SIPG vector Laplace",
            "dg_helmholtz": "DG Helmholtz-like vector (mass 1 + Laplace 1)"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=64, help="hexes per cube edge")
    ap.add_argument("--degree", type=int, default=2)
    ap.add_argument("--kind", default="dg")
    ap.add_argument("--variants", type=int, default=1, help="0 to skip the other workloads' lines")
    ap.add_argument("--external-n", type=int, default=64,
                    help="CG/DG P2 on an n^3 cube in a scrambled input order, reordering on and off (0 skips)")
    ap.add_argument("--matrix-n", type=int, default=64,
                    help="matrix-free vs assembled CSR comparison, P2 on an n^3 cube, P1 on (2n)^3, "
                         "P3 on (n/2)^3 (0 skips)")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="MiB copied once by a full-chip copy kernel before timing (scale for rocprof counters)")
    ap.add_argument("--precision", default="f64")
    ap.add_argument("--no-reorder", action="store_true")
    ap.add_argument("--manifold", default="none", choices=["none", "spherical", "straight"],
                    help="CG: curved cells on this manifold (P2 geometry map)")
    ap.add_argument("--patch-cells", type=int, default=0)
    ap.add_argument("--cells-per-wave", type=int, default=0, help="CG: cells per wave of a patch workgroup")
    ap.add_argument("--accumulate", type=int, default=0, help="CG LDS accumulation: 0 auto, 1 atomics, 2 read-add-write")
    ap.add_argument("--dg-kernel", type=int, default=0,
                    help="DG kernel: 0 auto, 1 LDS tables, 2 compile-time tables, 3 four lanes per cell, "
                         "4 four lanes per cell with the block's values staged in LDS")
    ap.add_argument("--cpu-baseline", type=int, default=1, help="0 to skip the CPU leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--mode", default="vmult", choices=["vmult", "solve"])
    ap.add_argument("--solves", type=int, default=5, help="repeated solves of the solver comparison")
    ap.add_argument("--solve-compare", type=int, default=1,
                    help="0 skips the learned-vs-plain comparison (the external-mesh solve still runs)")
    ap.add_argument("--solve-n", type=int, default=48, help="mesh of the solver comparison (0: skip it)")
    ap.add_argument("--solve-kind", default="cg", choices=["cg", "dg"], help="operator of --mode solve")
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


def _time_solves(smf, fn, solves):
    fn()  # warm-up
    times = []
    for _ in range(solves):
        smf.synchronize()
        t = time.perf_counter()
        st = fn()
        smf.synchronize()
        times.append(time.perf_counter() - t)
    return st, times


def _rates(st, tmin, dofs):
    n = st["iterations"]
    rho = (st["final_residual"] / st["initial_residual"]) ** (1.0 / n)
    n10 = -10 * np.log(10) / np.log(rho)
    t10 = tmin * n10 / n
    return {"iterations": n, "rho": rho, "n10": n10, "t10_s": t10, "E10_dofs_per_s": dofs / t10}


def solve_compare(smf, n, degree, seed, solves, learned_degree=3, kind="cg"):
    """Time to 1e-10 of mixed-precision hybrid-MG CG (FP64 outer CG, FP32
    V-cycle, Chebyshev(5)-Jacobi: the reference's MultigridPreconditioner +
    SolverCG, application_ph.cc:1754-2026, metrics of application_hp.cc:
    1954-1988) against flexible CG with the learned post-smoothing step
    behind a Chebyshev(learned_degree) V-cycle, on a Gaussian-random-field
    right-hand side.  Errors against an FP64 reference solve to 1e-13 whose
    residual is checked with the FP64 oracle's operator.
    kind "dg": DG SIPG with the DG -> CG (Nitsche) -> P1 -> h hierarchy
    (application_phc.cc:3416-3552)."""
    t0 = time.time()
    mesh = smf.Mesh.kuhn_cube(n, 0.2, seed)
    op = smf.Operator(mesh, degree, kind=kind, precision="f64")
    mg = smf.Multigrid(op)
    b = smf.DeviceVector.from_numpy(smf.assemble_rhs(op, "grf", seed=seed, n_features=256, corr_length=0.1))
    x = op.initialize_dof_vector()
    stream = smf.Stream()
    smf.solve_cg(op, x, b, mg, rel_tol=1e-13, max_iter=300, stream=stream)
    xref = x.numpy()
    # the reference solution checked with the FP64 oracle's operator (OpenMP
    # on the host): its relative residual bounds how far it is from the oracle
    # solve, so the errors below are errors against the FP64 oracle
    import oracle_py

    t1 = time.time()
    pr = oracle_py.OracleProblem.from_operator(op) if kind == "cg" else oracle_py.DGOracleProblem.from_operator(op)
    bh = b.numpy()
    ref_res = float(np.linalg.norm(bh - pr.vmult_omp(xref)) / np.linalg.norm(bh))
    log(f"reference solution: oracle residual {ref_res:.2e} ({time.time() - t1:.1f}s)")
    del pr

    def errors(xv):
        e = smf.DeviceVector.from_numpy(xv.numpy() - xref)
        ae = op.initialize_dof_vector()
        op.vmult(ae, e)
        scale = np.max(np.abs(xref))
        return {"max_rel": float(np.max(np.abs(e.numpy())) / scale),
                "energy": float(np.sqrt(max(np.dot(e.numpy(), ae.numpy()), 0.0)))}

    name = "MG-CG Poisson" if kind == "cg" else "DG SIPG hybrid h/p/c-MG-CG"
    out = {"workload": f"{name} P{degree}, {n}^3 x 6 perturbed Kuhn tets, FP64 CG + FP32 V-cycle, "
                       f"GRF right-hand side (l=0.1), stop at 1e-10", "dofs": op.m(),
           "levels": mg.info()["n_dofs"], "solves": solves,
           "reference_solution": {"solve": "plain MG-CG to 1e-13", "oracle_relative_residual": ref_res,
                                  "note": "errors below are against this solution; its residual with the FP64 "
                                          "oracle's operator is given"}}
    st, times = _time_solves(smf, lambda: smf.solve_cg(op, x, b, mg, rel_tol=1e-10, stream=stream), solves)
    plain = {"time_s": min(times), "mean_s": float(np.mean(times)), **_rates(st, min(times), op.m()),
             "error": errors(x), "smoother": "Chebyshev(5)-Jacobi"}
    out["plain"] = plain
    log(f"plain {kind} MG-CG {n}^3: {plain['iterations']} it, {plain['time_s'] * 1e3:.2f} ms")
    if degree == 2:
        mgl = smf.Multigrid(op, smoother_degree=learned_degree)
        ls = smf.LearnedSmoother(op, seed=0)
        tag = "" if kind == "cg" else "_dg"
        wpath = os.path.join(ROOT, "simplex-benchmarks_amd", "weights",
                             f"learned_smoother_p2{tag}_cheb{learned_degree}.bin")
        if os.path.exists(wpath):
            with open(wpath, "rb") as f:
                ls.load(f.read())
            origin = os.path.relpath(wpath, ROOT)
        else:
            ls.train(mgl, n_rhs=256, pairs_per_rhs=12, steps_per_pair=2, loss=0)
            origin = "trained in this run on the bench mesh (256 right-hand sides)"
        st, times = _time_solves(smf, lambda: smf.solve_fcg_learned(op, x, b, mgl, ls, 1e-10, 300, stream=stream),
                                 solves)
        learned = {"time_s": min(times), "mean_s": float(np.mean(times)), **_rates(st, min(times), op.m()),
                   "error": errors(x), "fallback": st["fallback"], "learned_steps": st["learned_steps"],
                   "mean_omega": st["mean_omega"], "smoother": f"Chebyshev({learned_degree})-Jacobi + learned step",
                   "weights": origin}
        st, times = _time_solves(smf, lambda: smf.solve_cg(op, x, b, mgl, rel_tol=1e-10, stream=stream), solves)
        ablation = {"time_s": min(times), **_rates(st, min(times), op.m()),
                    "smoother": f"Chebyshev({learned_degree})-Jacobi, no learned step"}
        out["learned"] = learned
        out["plain_same_vcycle"] = ablation
        out["speedup_learned_vs_plain"] = plain["time_s"] / learned["time_s"]
        out["error_ratio_energy"] = learned["error"]["energy"] / max(plain["error"]["energy"], 1e-300)
        log(f"learned {kind} FCG {n}^3: {learned['iterations']} it, {learned['time_s'] * 1e3:.2f} ms, "
            f"speedup {out['speedup_learned_vs_plain']:.3f}")
    log(f"solve comparison {time.time() - t0:.1f}s")
    return out


def solve_mode(args):
    """One JSON line for the solver comparison alone (value: plain MG-CG
    seconds to 1e-10)."""
    import simplexmf as smf

    r = solve_compare(smf, args.solve_n, args.degree, args.seed, args.solves, kind=args.solve_kind)
    p = r["plain"]
    out = {
        "metric": METRIC, "value": round(p["time_s"], 6), "unit": "s to 1e-10 (min over solves)",
        "n_gpus": 1, "steps": args.solves, "warmup": 1, "ms_per_step": round(p["time_s"] * 1e3, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32",
        "data": "synthetic: seeded perturbed Kuhn mesh, Gaussian random field RHS (l=0.1)",
        "config": {"workload": r["workload"], "dofs": r["dofs"], "levels": r["levels"]},
        "solve": r,
    }
    print(json.dumps(out), flush=True)


def measure(smf, kind, n, degree, precision, steps, warmup, seed, reorder=True, patch_cells=0, cells_per_wave=0,
            accumulate=0, keep=False, dg_kernel=0, manifold="none"):
    """Event-timed vmults of one operator on `stream`; returns the numbers of
    one bench line (and the operator and input when keep=True)."""
    t0 = time.time()
    mesh = smf.Mesh.kuhn_cube(n, 0.2, seed)
    op = smf.Operator(mesh, degree, kind=kind, precision=precision, reorder=reorder,
                      patch_cells=patch_cells, accumulate=accumulate, dg_kernel=dg_kernel, mass_coefficient=1.0,
                      cells_per_wave=cells_per_wave, manifold=manifold,
                      laplace_coefficient=1.0)
    info = op.info
    log(f"setup {kind} p{degree} {precision} n{n}: {time.time() - t0:.1f}s")
    x = np.random.default_rng(seed).uniform(-1, 1, op.m())
    src = smf.DeviceVector.from_numpy(x.astype(op.dtype))
    dst = op.initialize_dof_vector()
    stream = smf.Stream()
    for _ in range(warmup):
        op.vmult(dst, src, stream)
    stream.synchronize()
    smf.synchronize()
    ev0, ev1 = smf.Event(), smf.Event()
    ev0.record(stream)
    for _ in range(steps):
        op.vmult(dst, src, stream)
    ev1.record(stream)
    stream.synchronize()
    smf.synchronize()
    t_step = ev0.elapsed_ms(ev1) / 1e3 / steps
    res = {"kind": kind, "degree": degree, "precision": precision, "n": n, "cells": info["n_cells"],
           "dofs": op.m(), "ms_per_step": t_step * 1e3, "gdofs": op.m() / t_step / 1e9,
           "algorithmic_bytes": info["algorithmic_bytes"],
           "achieved_gbs": info["algorithmic_bytes"] / t_step / 1e9, "flops": info["flops"],
           "info": info}
    res["frac"] = res["achieved_gbs"] / HBM_PEAK_GBS
    if keep:
        res["op"], res["x"] = op, x
    return res


def _event_time(smf, fn, steps, warmup):
    """seconds per call of fn(stream), HIP events on the stream the kernels run on"""
    stream = smf.Stream()
    for _ in range(warmup):
        fn(stream)
    stream.synchronize()
    ev0, ev1 = smf.Event(), smf.Event()
    ev0.record(stream)
    for _ in range(steps):
        fn(stream)
    ev1.record(stream)
    stream.synchronize()
    return ev0.elapsed_ms(ev1) / 1e3 / steps


def matrix_compare(smf, n, degree, precision, steps, seed):
    """Matrix-free vmult against the assembled CSR matrix of the same CG
    operator (the comparison of CG/cells_batched/application.cc:1944-1988),
    both event-timed on the GPU at the same size."""
    t0 = time.time()
    op = smf.Operator(smf.Mesh.kuhn_cube(n, 0.2, seed), degree, kind="cg", precision=precision)
    M = smf.SystemMatrix(op)
    log(f"setup matrix comparison p{degree} {precision} n{n}: {time.time() - t0:.1f}s")
    x = np.random.default_rng(seed).uniform(-1, 1, op.m()).astype(op.dtype)
    src = smf.DeviceVector.from_numpy(x)
    dst = op.initialize_dof_vector()
    t_mf = _event_time(smf, lambda st: op.vmult(dst, src, st), steps, 3)
    t_mb = _event_time(smf, lambda st: M.vmult(dst, src, st), steps, 3)
    mi = M.info
    return {"kind": "cg", "degree": degree, "precision": precision, "n": n, "dofs": op.m(),
            "nnz": mi["nnz"], "nnz_per_row": round(mi["nnz"] / op.m(), 3), "row_blocks": mi["n_row_blocks"],
            "matrix_free": {"ms_per_step": round(t_mf * 1e3, 5), "gdofs": round(op.m() / t_mf / 1e9, 4),
                            "algorithmic_bytes": op.info["algorithmic_bytes"],
                            "frac": round(op.info["algorithmic_bytes"] / t_mf / 1e9 / HBM_PEAK_GBS, 4)},
            "csr_spmv": {"ms_per_step": round(t_mb * 1e3, 5), "gdofs": round(op.m() / t_mb / 1e9, 4),
                         "algorithmic_bytes": mi["algorithmic_bytes"],
                         "frac": round(mi["algorithmic_bytes"] / t_mb / 1e9 / HBM_PEAK_GBS, 4),
                         "kernel": "spmv_stream_kernel (csrc/spmv_kernels.hip)"},
            "speedup_matrix_free": round(t_mb / t_mf, 3),
            "note": "matrix-based = assembled CSR (int32 columns, values in the operator precision); "
                    "bytes: nnz (4 + w) + rows (8 + 2 w); the matrix-free operator's data at these sizes is "
                    "partly resident in the 256 MB Infinity Cache when its model bytes are below that size, "
                    "the CSR matrices are not"}


def external_compare(smf, n, steps, seed, solve_n=0):
    """Operators on a mesh in an unstructured input order (the Kuhn cube with
    vertices and cells randomly permuted, as an external mesh would arrive),
    with the hierarchical reordering on and off: the reference's reorder_grid
    study (CG/cells_batched/application.cc:1699-1756)."""
    mesh = smf.Mesh.kuhn_cube(n, 0.2, seed).scrambled(seed=seed + 1)
    out = []
    for kind in ("cg", "dg"):
        for reorder in (True, False):
            t0 = time.time()
            op = smf.Operator(mesh, 2, kind=kind, reorder=reorder)
            log(f"setup external {kind} reorder={reorder} n{n}: {time.time() - t0:.1f}s")
            x = np.random.default_rng(seed).uniform(-1, 1, op.m())
            src = smf.DeviceVector.from_numpy(x)
            dst = op.initialize_dof_vector()
            t = _event_time(smf, lambda st: op.vmult(dst, src, st), steps, 3)
            out.append({"kind": kind, "degree": 2, "precision": "f64", "n": n, "reorder": reorder,
                        "dofs": op.m(), "ms_per_step": round(t * 1e3, 5), "gdofs": round(op.m() / t / 1e9, 4),
                        "frac": round(op.info["algorithmic_bytes"] / t / 1e9 / HBM_PEAK_GBS, 4)})
            del op, src, dst
    res = {"mesh": f"{n}^3 x 6 perturbed Kuhn tets with vertices and cells in a seeded random order, "
                   "local vertices rotated", "runs": out}
    if solve_n > 0:
        # MG-CG to 1e-10 without a lattice: p-levels P2 -> P1 on the mesh itself,
        # then the smoothed-aggregation AMG coarse solver on the P1 level
        # (the reference's MGCoarseAMG, application_hp.cc:1684-1731)
        emesh = smf.Mesh.kuhn_cube(solve_n, 0.2, seed).scrambled(seed=seed + 1)
        op = smf.Operator(emesh, 2)
        t0 = time.time()
        mg = smf.Multigrid(op)
        setup = time.time() - t0
        b = smf.DeviceVector.from_numpy(smf.assemble_rhs(op, "grf", seed=seed, n_features=256, corr_length=0.1))
        x = op.initialize_dof_vector()
        stream = smf.Stream()
        st, times = _time_solves(smf, lambda: smf.solve_cg(op, x, b, mg, rel_tol=1e-10, stream=stream), 5)
        inf = mg.info()
        res["solve"] = {"workload": f"MG-CG Poisson P2 on the scrambled {solve_n}^3 mesh, FP64 CG + FP32 V-cycle, "
                                    "AMG coarse solve (one FP64 V-cycle), GRF right-hand side, stop at 1e-10",
                        "dofs": op.m(), "levels": inf["n_dofs"], "amg_rows": inf["amg_rows"],
                        "setup_s": round(setup, 3), "time_s": min(times), **_rates(st, min(times), op.m())}
        log(f"external MG-CG {solve_n}^3: {st['iterations']} it, {min(times) * 1e3:.2f} ms, AMG {inf['amg_rows']}")
    return res


def cpu_baseline_dg(op, x, y_gpu, seconds):
    """The reference's optimised DG SIPG vmult restated for CPUs
    (oracle/cpu_baseline.cpp: 8-lane vectors of cells or faces, four vectors
    per shape matrix, OpenMP over cell batches and coloured face batches) on
    the same mesh and input vector as the GPU line, timed with the
    reference's protocol (repetitions of back-to-back products,
    CG/cells_batched/application.cc:1870-1889), sized to about `seconds`."""
    import oracle_py

    t0 = time.perf_counter()
    cpu = oracle_py.DGBaseline(op.mesh.vertices, op.cell_vertices(), op.info["degree"])
    setup = time.perf_counter() - t0
    inf = cpu.info()
    y = cpu.vmult(x)  # warm-up, and the cross-check against the GPU product
    err = float(np.max(np.abs(y - y_gpu)) / np.max(np.abs(y)))
    one = float(cpu.time(x, reps=1, n=1)[0])
    reps = 5
    per_rep = max(1, min(100, int(seconds / reps / max(one, 1e-9))))
    sec = cpu.time(x, reps=reps, n=per_rep)
    best, med = float(sec.min()), float(np.median(sec))
    return {"value": op.m() * per_rep / best / 1e9, "unit": "GDoF/s", "cores": inf["threads"], "kind": "port",
            "median": op.m() * per_rep / med / 1e9,
            "sample": f"{reps} repetitions x {per_rep} back-to-back vmults (best shown, median also given) of "
                      f"the same DG SIPG P{op.info['degree']} FP64 operator, mesh and input vector as the GPU "
                      f"line ({op.m()} DoFs), {inf['threads']} host threads, {inf['vector_bits']}-bit "
                      f"vectors of 8 cells/faces, {inf['n_q']} cell and {inf['n_qf']} face quadrature points; "
                      f"plan set-up {setup:.1f} s",
            "implementation": "oracle/cpu_baseline.cpp (vectorised restatement of DG/cells_batched/"
                              "application.cc:1493-2845)",
            "rel_diff_vs_gpu": err}


def main():
    args = parse()
    if args.mode == "solve":
        if args.gpus != 1:
            print(json.dumps({"error": "this benchmark runs on exactly one GPU"}))
            sys.exit(2)
        return solve_mode(args)
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; "
                                   "multi-GPU sharding is a design note only"}))
        sys.exit(2)
    import simplexmf as smf

    if smf.device_count() < 1:
        raise SystemExit("no HIP device visible")
    if args.calibrate_copy > 0:
        nb = args.calibrate_copy << 20
        a, b = smf.DeviceVector(nb // 8), smf.DeviceVector(nb // 8)
        a.zero()
        # full-chip copy kernel (convert_kernel<double, double>): nb bytes read, nb written
        smf.check(smf.lib().mf_copy_kernel(b.ptr, a.ptr, nb // 8, None))
        smf.synchronize()
        del a, b
    head = measure(smf, args.kind, args.n, args.degree, args.precision, args.steps, args.warmup,
                   args.seed, reorder=not args.no_reorder, patch_cells=args.patch_cells,
                   cells_per_wave=args.cells_per_wave,
                   accumulate=args.accumulate,
                   keep=args.cpu_baseline and args.kind in ("cg", "dg") and args.manifold == "none",
                   dg_kernel=args.dg_kernel, manifold=args.manifold)
    info = head["info"]
    traffic = None
    curved = "" if args.manifold == "none" else f"_{args.manifold}"
    prof = os.path.join(ROOT, "profiles",
                        f"traffic_{args.kind}_p{args.degree}_{args.precision}_n{args.n}{curved}.json")
    if os.path.exists(prof):
        with open(prof) as f:
            traffic = json.load(f).get("bytes_per_vmult")
    variants = []
    if args.variants:
        # (kind, n, degree, precision, hierarchical reordering)
        todo = [("dg", 64, 1, "f64", True), ("dg", 64, 3, "f64", True), ("dg", 64, 1, "f32", True),
                ("dg", 64, 2, "f32", True), ("dg", 64, 3, "f32", True), ("dg", 64, 2, "f64", False),
                ("dg_helmholtz", 32, 2, "f64", True), ("dg_helmholtz", 32, 2, "f32", True),
                ("cg", 128, 2, "f64", True), ("cg", 128, 2, "f32", True), ("cg_vector", 64, 2, "f64", True),
                ("cg", 128, 2, "f64", "spherical"), ("dg", 64, 2, "f64", "spherical"),
                ("cg", 16, 2, "f64", True), ("cg", 48, 2, "f64", True)]
        for kind, n, deg, prec, reorder in todo:
            # a string in the last slot: curved cells on that manifold (reordering on)
            manifold = reorder if isinstance(reorder, str) else "none"
            reorder = True if isinstance(reorder, str) else reorder
            if (kind, n, deg, prec, reorder, manifold) == (args.kind, args.n, args.degree, args.precision,
                                                           not args.no_reorder, "none"):
                continue
            r = measure(smf, kind, n, deg, prec, max(10, args.steps // 2), 3, args.seed, reorder=reorder,
                        manifold=manifold)
            r["reorder"] = reorder
            r["manifold"] = manifold
            # operator data below the 256 MB Infinity Cache: not an HBM figure
            r["cache_resident"] = r["algorithmic_bytes"] < 256e6
            if manifold != "none":
                # the curved kernel's own data model: src + dst, packed local indices (4 B per
                # local DoF) and the 27 FP64 values of the P2 geometry map per cell (216 B);
                # it forms J, adj J and det J per point in registers (FP64-VALU-bound)
                w = 8 if prec == "f64" else 4
                if kind == "cg":
                    own = 2 * w * r["dofs"] + 4 * r["info"]["dofs_per_cell"] * r["cells"] + 27 * w * r["cells"]
                    what = "4 B per local DoF + 27 w of P2 geometry per cell"
                else:
                    own = 2 * w * r["dofs"] + (28 * w + 20) * r["cells"]
                    what = "27 w of P2 geometry + tau + 4 neighbour ids + face numbers per cell"
                r["algorithmic_bytes"] = own
                r["achieved_gbs"] = own / (r["ms_per_step"] * 1e-3) / 1e9
                r["frac"] = r["achieved_gbs"] / HBM_PEAK_GBS
                r["note"] = (f"bytes: the curved kernel's own data model (2 w per DoF + {what}); it forms J, "
                             "cof J and det J per quadrature point in registers (FP64-VALU-bound); the "
                             "reference stores J^-T and JxW per quadrature point instead")
            variants.append({k: (round(v, 5) if isinstance(v, float) else v) for k, v in r.items()
                             if k in ("kind", "degree", "precision", "n", "reorder", "manifold", "dofs", "ms_per_step",
                                      "gdofs", "achieved_gbs", "frac", "note", "cache_resident",
                                      "algorithmic_bytes")})
    base = None
    if args.cpu_baseline and args.manifold == "none":  # the oracle baseline covers affine cells
        log("cpu baseline ...")
        if args.kind == "cg":
            base = cpu_baseline(head["op"], head["x"], args.cpu_seconds)
        elif args.kind == "dg" and args.precision == "f64":
            op, x = head["op"], head["x"]
            y = op.initialize_dof_vector()
            op.vmult(y, smf.DeviceVector.from_numpy(x))
            base = cpu_baseline_dg(op, x, y.numpy(), args.cpu_seconds)
    prec = args.precision
    out = {
        "metric": METRIC,
        "value": round(head["gdofs"], 4),
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": prec,
        "data": "synthetic: seeded perturbed Kuhn mesh, x ~ U[-1,1]",
        "config": {
            "workload": f"{WORKLOAD[args.kind]} P{args.degree} vmult, {args.n}^3 x 6 perturbed "
                        f"Kuhn tets ({head['cells']} cells, {head['dofs']} DoFs), {prec.upper()}"
                        + ("" if args.manifold == "none" else f", curved cells ({args.manifold} manifold)"),
            "kind": args.kind, "degree": args.degree, "n": args.n, "cells": head["cells"],
            "dofs": head["dofs"], "reorder": not args.no_reorder, "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(head["achieved_gbs"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(head["frac"], 4),
            "traffic": traffic,
            "algorithmic_bytes": head["algorithmic_bytes"],
            "kernel": KERNEL[args.kind],
        },
        "cpu_baseline": base,
        "variants": variants,
    }
    if args.matrix_n > 0:
        # P1 at 2n, P2 at n, P3 at n/2: about the same number of DoFs
        out["matrix_based"] = [matrix_compare(smf, m, deg, "f64", max(10, args.steps // 2), args.seed)
                               for deg, m in ((1, 2 * args.matrix_n), (2, args.matrix_n),
                                              (3, max(1, args.matrix_n // 2)))]
    if args.external_n > 0:
        out["external_mesh"] = external_compare(smf, args.external_n, max(10, args.steps // 2), args.seed,
                                                solve_n=args.solve_n)
    if args.solve_n > 0 and args.solve_compare:
        out["solve"] = solve_compare(smf, args.solve_n, 2, args.seed, args.solves)
        out["solve_dg"] = solve_compare(smf, args.solve_n, 2, args.seed, args.solves, kind="dg")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
