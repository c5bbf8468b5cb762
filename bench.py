#!/usr/bin/env python3
"""Benchmark of the FEEC assembly-and-solve hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): 3D mixed
Hodge-Laplace source problem, k = 1, full complex, on the Kuhn subdivision of
a 64^3 cube grid (1,572,864 tets), interior vertices jittered by
U(+-0.05 h) (SplitMix64 seed 2), Euclidean squared edge lengths from the
coordinates, source cochain N(0,1) on the edges (seed 2).

One step = re-assemble every KKT element block and the CSR values from the
resident geometry (pattern reused), then solve the KKT system to a true
relative residual |b - A x|/|b| <= 1e-10. The headline solve is mixed
precision: FP64 MINRES and KKT SpMV with the Chebyshev block preconditioner
run in FP32 arithmetic on its FP16 SELL-64 copy (FEEC_PREC_FP64_PC32;
`--mixed-mode refine` selects FP32 MINRES inside FP64 iterative refinement,
FEEC_PREC_MIXED, which is timed beside it either way). `value` is completed
steps per second (solves/s) with all inputs resident in HBM. The plain FP64
GPU path (FP64 matrix, vectors, MINRES and polynomial, same 1e-10
certificate) is timed beside it.

  python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import math
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "elmats/s; solves/s @1e-10 resid, mixed vs FP64 GPU; p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_problem(fa, n, precision, ctx):
    g = fa.CartesianGrid(3, n)
    cells, coords = g.triangulate()
    coords = g.jitter(coords, 0.05, 2)
    topo = fa.Complex(3, cells, nverts=len(coords), ctx=ctx)
    geo = fa.MeshLengthsSq.from_coords(topo, coords)
    wc = fa.WhitneyComplex(geo)
    return topo, geo, wc


def cpu_baseline(n_sample, gpu_iters, c2_cells, c2_nnz, cheb_degree):
    """The oracle (C restatement of the reference algorithm) timed on this
    host on a bounded sample: the element matrices and CSR assembly of the KKT
    blocks of an n_sample^3 Kuhn mesh, and FP64 MINRES iterations with the
    same Chebyshev block preconditioner as the GPU, on its KKT system; scaled
    to the 64^3 workload per cell / per nonzero, with the iteration count of
    the GPU FP64 path (same algorithm and preconditioner)."""
    from oracle import oracle as o
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cells, coords = o.kuhn(3, n_sample)
    rng = np.random.default_rng(2)
    inter = np.all((coords > 1e-12) & (coords < 1 - 1e-12), axis=1)
    coords[inter] += 0.05 / n_sample * (2 * rng.random((inter.sum(), 3)) - 1)
    cx = o.Complex(3, cells)
    L = cx.lengths_from_coords(coords)
    nc = cx.nsimplices(3)
    t0 = time.perf_counter()
    for kind, k in ((o.MASS, 0), (o.DIF_TEST, 1), (o.DIF_BOTH, 2), (o.MASS, 1), (o.DIF_BOTH, 1)):
        o.element_matrices(cx, L, kind, k, cores)
    t_el = time.perf_counter() - t0
    c = o.WhitneyComplex(cx, L)
    t0 = time.perf_counter()
    f = np.random.default_rng(2).standard_normal(cx.nsimplices(1))
    kkt, rhs, ns, nu = o.assemble_mixed_kkt(c, f, 1, c.harmonics(1))
    t_asm = time.perf_counter() - t0
    prec = o.chebyshev_jacobi(c, 1, 0, cheb_degree)
    iters = 60
    t0 = time.perf_counter()
    o.minres_cheb(kkt, rhs, prec, 1e-30, iters)
    t_it = (time.perf_counter() - t0) / iters
    scale_cells = c2_cells / nc
    t_step = (t_el + t_asm) * scale_cells + t_it * (c2_nnz / kkt.nnz) * gpu_iters
    return {
        "value": 1.0 / t_step, "unit": "solves/s", "cores": cores, "kind": "port",
        "sample": (f"oracle on a {n_sample}^3 Kuhn mesh ({nc} tets, KKT nnz {kkt.nnz}): element matrices "
                   f"{t_el:.2f}s + assembly {t_asm:.2f}s scaled x{scale_cells:.1f} by cells, and {iters} "
                   f"FP64 MINRES iterations with the degree-{cheb_degree} Chebyshev block preconditioner "
                   f"({t_it * 1e3:.2f} ms each) scaled by nnz to the 64^3 KKT and multiplied by the GPU "
                   f"FP64 path's {gpu_iters} iterations"),
        "seconds_per_step_estimate": t_step,
        "elmat_cells_per_s": nc / t_el,
    }


def kuhn_n_for(d, cells):
    """Cells per axis of a d-dim Kuhn box with about `cells` simplices."""
    return max(1, int(round((cells / math.factorial(d)) ** (1.0 / d))))


def run_c3(args):
    """4D Lorentzian Hodge-Dirac element matrices on the Kuhn 16^4 box
    (1,572,864 4-simplices): time axis 0 unscaled, axes 1..3 jittered by
    U(+-0.02 h) (seed 3), squared lengths in the Minkowski signature
    (-,+,+,+). Throughput of the all-grade element kernel, and the error of
    every block on a sample of cells against the FP64 oracle."""
    import formoniq_amd as fa
    from oracle import oracle as o
    n = args.n if args.n != 64 else 16
    g = fa.CartesianGrid(4, n)
    cells, coords = g.triangulate()
    coords = g.jitter(coords, 0.02, 3, axes=[1, 2, 3])
    topo = fa.Complex(4, cells, nverts=len(coords))
    geo = fa.MeshLengthsSq.from_coords(topo, coords, minkowski=True)
    nc = topo.nsimplices(4)
    census, _ = geo.causal_census()
    sec, per = geo.dirac_element_time(reps=args.steps)
    blocks = [(fa.MASS, k) for k in range(5)] + [(fa.DIF_TRIAL, k) for k in range(1, 5)]
    rng = np.random.default_rng(3)
    sample = np.sort(rng.choice(nc, size=min(nc, 2048), replace=False))
    L = geo.lengths()
    edges = topo.cell_faces(1)
    worst = 0.0
    for kind, k in blocks:
        el = fa.WhitneyComplex(geo).element(kind, k)
        for c in sample:
            want = o.element_matrix(4, kind, k, L[edges[c]])
            got = el[c]
            scale = max(np.abs(want).max(), 1e-300)
            worst = max(worst, float(np.abs(got - want).max() / scale))
        del el
    bytes_per_cell = 8 * 10 + 4 * 10 + 8 * per
    line = {
        "config": "C3", "workload": f"4D Lorentzian Hodge-Dirac element matrices, Kuhn {n}^4 ({nc} cells), "
                                    "axes 1..3 jittered 0.02h seed 3, Minkowski (-,+,+,+)",
        "edges_timelike_null_spacelike": census.tolist(),
        "seconds_per_launch": sec, "values_per_cell": per,
        "element_matrices_per_s": 9 * nc / sec, "cells_per_s": nc / sec,
        "hbm_gbs": bytes_per_cell * nc / sec / 1e9,
        "max_rel_err_vs_oracle": worst, "sampled_cells": int(sample.size), "dtype": "f64",
    }
    print(json.dumps(line), flush=True)


def run_leapfrog(args):
    """Explicit Hodge-Dirac leapfrog (solve_dirac_leapfrog's loop) on the full
    complex of a jittered 3D Kuhn mesh (default 32^3, seed 6), dt = 0.1 h_min,
    all grades: GPU steps/s with the CG mass solves, the staggered invariant's
    drift, and the oracle (SuperLU-factored color masses) timed on a 12^3
    sample of the same workload. Beside it, a few implicit Gauss-Legendre
    steps (solve_dirac's integrator) on the same mesh and dt."""
    import formoniq_amd as fa
    from oracle import oracle as o
    n = args.n if args.n != 64 else 32
    nsteps = max(args.steps, 1) * 10

    def mesh(nn):
        g = fa.CartesianGrid(3, nn)
        cells, coords = g.triangulate()
        coords = g.jitter(coords, 0.05, 6)
        return cells, coords

    cells, coords = mesh(n)
    topo = fa.Complex(3, cells, nverts=len(coords))
    geo = fa.MeshLengthsSq.from_coords(topo, coords)
    wc = fa.WhitneyComplex(geo)
    ndofs = int(sum(topo.counts))
    y0 = np.sin(0.7 * np.arange(ndofs) + 0.3) + 0.1
    dt = 0.1 * float(np.sqrt(geo.lengths().min()))
    fa.dirac_leapfrog(wc, dt, 2, y0)  # warm-up
    t0 = time.perf_counter()
    y, e, rep = fa.dirac_leapfrog(wc, dt, nsteps, y0)
    wall = time.perf_counter() - t0
    drift = float(np.abs(e - e[0]).max() / e[0])
    # implicit Gauss-Legendre (2 stages) on the same mesh and dt, a few steps
    irk_steps = 3
    fa.dirac_irk(wc, dt, 1, y0)  # warm-up
    yi, ei, repi = fa.dirac_irk(wc, dt, irk_steps, y0)
    irk_drift = float(np.abs(ei - ei[0]).max() / ei[0])
    # the oracle on a 12^3 sample, same dt rule
    cs, xs = mesh(12)
    ocx = o.Complex(3, cs)
    oL = ocx.lengths_from_coords(xs)
    oc = o.WhitneyComplex(ocx, oL)
    ys = np.sin(0.7 * np.arange(int(sum(ocx.nsimplices(k) for k in range(4)))) + 0.3) + 0.1
    offs, mass, op = o.hodge_dirac(oc, -1.0)
    lf = o.Leapfrog(mass, op, o.grade_parity_coloring(offs), 0.1 * float(np.sqrt(oL.min())))
    t1 = time.perf_counter()
    for _ in range(20):
        ys = lf.step(ys)
    cpu = time.perf_counter() - t1
    line = {
        "config": "leapfrog", "workload": f"Hodge-Dirac leapfrog, 3D Kuhn {n}^3 ({topo.nsimplices(3)} tets), "
                                          f"jitter 0.05h seed 6, full complex, {ndofs} DOFs, dt = 0.1 h_min",
        "steps": nsteps, "steps_per_s": nsteps / rep.seconds, "device_seconds": rep.seconds,
        "wall_seconds": wall, "cg_iters_per_solve": rep.iters / (2 * nsteps + 1),
        "energy_rel_drift": drift, "dtype": "f64",
        "gauss_legendre_2": {"steps": irk_steps, "steps_per_s": irk_steps / repi.seconds,
                             "gmres_iters_per_step": repi.iters / irk_steps,
                             "energy_rel_drift": irk_drift},
        "cpu_baseline": {"steps_per_s": 20 / cpu, "kind": "port", "cores": 1,
                         "sample": "oracle leapfrog (SuperLU-factored color masses), 20 steps on a 12^3 "
                                   "Kuhn mesh, stepping only", "sample_dofs": int(ys.size)},
    }
    print(json.dumps(line), flush=True)


def config_requests(fa, count, seed, dims, cells_lo, cells_hi):
    """The request mix of the batched and serving configurations: d uniform in
    dims, k uniform in 0..d-1, cells log-uniform in [cells_lo, cells_hi], jitter
    0.1h / 0.05h / 0.02h by dimension, mesh and source seeds from `seed`."""
    rng = np.random.default_rng(seed)
    amp = {2: 0.1, 3: 0.05, 4: 0.02}
    out = []
    for i in range(count):
        d = int(rng.choice(dims))
        k = int(rng.integers(0, d))
        cells = math.exp(rng.uniform(math.log(cells_lo), math.log(cells_hi)))
        r = fa.Request()
        r.dim, r.grade, r.n, r.bc = d, k, kuhn_n_for(d, cells), fa.FULL
        r.jitter = amp[d]
        r.mesh_seed, r.source_seed = seed * 100000 + i, seed * 100000 + 50000 + i
        r.precision = fa.FP64_PC32
        out.append(r)
    return out


def _mode_names(fa):
    # "mixed": FP64 MINRES with the FP32 polynomial on FP16 storage
    return {fa.FP64_PC32: "mixed", fa.MIXED: "fp32_refine", fa.FP64: "fp64"}


def run_c4(args):
    """512 independent jittered Kuhn meshes (seed 4): d in {2,3,4}, k in
    0..d-1, cells log-uniform in 1e3..1e5; all element matrices in one
    persistent launch, then mixed-precision solves to 1e-10 in lockstep.
    The same batch with FP32 MINRES inside FP64 refinement and in FP64
    beside it."""
    import formoniq_amd as fa
    reqs = config_requests(fa, args.batch, 4, [2, 3, 4], 1e3, 1e5)
    opts = fa.solve_opts(use_graph=0 if args.no_graph else 1)
    names = _mode_names(fa)
    out = {}
    for prec in (fa.FP64_PC32, fa.MIXED, fa.FP64):
        for r in reqs:
            r.precision = prec
        fa.batch_run(reqs[:8], opts)  # warm up tables and the autotuner
        t0 = time.perf_counter()
        res = fa.batch_run(reqs, opts)
        dt = time.perf_counter() - t0
        log(f"c4 {names[prec]}: {dt:.2f}s")
        ok = all(x.status == 0 and x.residual <= 1e-10 for x in res)
        out[names[prec]] = {
            "seconds": dt, "solves_per_s": len(res) / dt, "all_converged_1e-10": ok,
            "setup_and_elements_seconds": float(res[0].admit_time),
            "solve_phase_seconds": float(res[0].solve_seconds),
            "results_done_seconds": float(max(x.finish_time for x in res)),
            "cells": int(sum(x.ncells for x in res)), "dofs": int(sum(x.ndofs for x in res)),
            "iters": int(sum(x.iters for x in res)),
        }
    line = {"config": "C4", "workload": f"{len(reqs)} jittered Kuhn meshes, d in 2..4, k in 0..d-1, "
                                        "cells logU[1e3,1e5], seed 4", **out,
            "mixed_over_fp64": out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]}
    print(json.dumps(line), flush=True)


def run_c5(args):
    """Serving stream (seed 5): Poisson arrivals at a fixed rate of
    Hodge-Laplace solve requests, d in {2,3}, k in 0..d-1, cells log-uniform
    in 1e3..2e6; admitted by HBM budget, solved mixed precision (FP64 MINRES,
    FP32 polynomial on FP16 storage) to 1e-10. Throughput and p50/p99
    latency, and the same stream on the FP64 path."""
    import formoniq_amd as fa
    reqs = config_requests(fa, args.requests, 5, [2, 3], 1e3, 2e6)
    gaps = np.random.default_rng(5).exponential(1.0 / args.rate, size=len(reqs))
    names = _mode_names(fa)
    out = {}
    for prec in (fa.FP64_PC32, fa.FP64):
        for r in reqs:
            r.precision = prec
        opts = fa.ServerOpts()
        opts.hbm_budget_bytes = 200 << 30
        opts.max_inflight = args.slots
        opts.solve = fa.solve_opts(precision=prec, graph_iters=args.graph_iters)
        srv = fa.Server(opts)
        if not args.no_warmup:
            # a server's start-up: one small request of every (d, k) the stream holds
            warm = [r for r in config_requests(fa, 64, 55, [2, 3], 1e3, 2e3)]
            seen = {}
            for r in warm:
                seen.setdefault((r.dim, r.grade), r)
            for r in seen.values():
                r.precision = prec
                srv.submit(r)
            srv.drain()
        t0 = time.perf_counter()
        tickets = []
        for i, (r, gap) in enumerate(zip(reqs, gaps)):
            time.sleep(gap)
            tickets.append(srv.submit(r))
            if i % 8 == 7:
                log(f"c5 {names[prec]}: {i + 1} submitted at "
                    f"{time.perf_counter() - t0:.1f}s")
        srv.drain()
        dt = time.perf_counter() - t0
        res = [srv.poll(t) for t in tickets]
        srv.close()
        lat = np.array([x.latency for x in res])
        ok = all(x.status == 0 and x.residual <= 1e-10 for x in res)
        slow = sorted(zip(lat, reqs, res), key=lambda t: -t[0])[:3]
        out[names[prec]] = {
            "seconds": dt, "solves_per_s": len(res) / dt, "all_converged_1e-10": ok,
            "latency_p50_ms": float(np.percentile(lat, 50) * 1e3),
            "latency_p99_ms": float(np.percentile(lat, 99) * 1e3),
            "cells": int(sum(x.ncells for x in res)),
            "build_ms_max": float(max(x.build_seconds for x in res) * 1e3),
            "slowest": [{"dim": q.dim, "grade": q.grade, "n": q.n, "cells": int(x.ncells),
                         "dofs": int(x.ndofs), "iters": int(x.iters), "latency_ms": float(l * 1e3),
                         "service_ms": float((x.finish_time - x.admit_time) * 1e3),
                         "build_ms": float(x.build_seconds * 1e3), "solve_ms": float(x.solve_seconds * 1e3)}
                        for l, q, x in slow],
        }
    line = {"config": "C5", "workload": f"{len(reqs)} requests, Poisson rate {args.rate}/s, d in 2..3, "
                                        f"k in 0..d-1, cells logU[1e3,2e6], seed 5, {args.slots} slots", **out}
    print(json.dumps(line), flush=True)


def run_kernels(args):
    """Short run for hardware-counter passes: the 3D 64^3 k=1 mixed problem's element
    kernel, KKT scatter and fused MINRES SpMV (FP32, FP64) launched a few
    times each, then a streaming kernel of known size (torch multiply of
    256 Mi floats: 1 GiB read, 1 GiB written) to calibrate byte counters."""
    import torch
    import formoniq_amd as fa
    ctx = fa.Context(0)
    topo, geo, wc = build_problem(fa, args.n, fa.MIXED, ctx)
    prob = fa.Problem(wc, 1, fa.solve_opts(precision=fa.MIXED))
    prob.assemble()
    times = prob.kernel_times(reps=args.steps)
    a = torch.ones(256 << 20, device="cuda:0", dtype=torch.float32)
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 2.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"config": "kernels", "n_kkt": prob.sizes[0] + prob.sizes[1], "nnz": prob.nnz,
                      "kernel_seconds": times, "calibration_bytes_read": a.numel() * 4,
                      "calibration_bytes_written": b.numel() * 4}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64, help="cells per axis (the workload is 64)")
    ap.add_argument("--fp64-steps", type=int, default=1,
                    help="timed steps of the FP64 path and of the other mixed-precision mode")
    ap.add_argument("--mixed-mode", default="pc32", choices=["pc32", "refine"],
                    help="headline solve: FP64 MINRES with the FP32 polynomial on FP16 storage (pc32), "
                         "or FP32 MINRES inside FP64 refinement (refine)")
    ap.add_argument("--fp64-cheb", type=int, default=24, help="Chebyshev degree of the FP64 path")
    ap.add_argument("--cpu-n", type=int, default=32)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-jacobi", action="store_true", help="skip the FP64 Jacobi-block solve")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "c2_kernel_traffic.json"))
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the Krylov kernels one by one instead of replaying graph segments "
                         "(same kernels; for tracing tools)")
    ap.add_argument("--graph-iters", type=int, default=32,
                    help="MINRES iterations per replayed graph segment")
    ap.add_argument("--no-warmup", action="store_true",
                    help="c5: start the timed stream on a cold server (no start-up requests)")
    ap.add_argument("--config", default="c2", choices=["c2", "c3", "c4", "c5", "kernels", "leapfrog"],
                    help="c2 (default): the headline line; c3/c4/c5: the 4D Dirac element, batched "
                         "and serving workloads, each printed as its own JSON line")
    ap.add_argument("--batch", type=int, default=512, help="c4: number of meshes")
    ap.add_argument("--requests", type=int, default=48, help="c5: number of requests")
    ap.add_argument("--rate", type=float, default=0.5, help="c5: Poisson arrival rate per second")
    ap.add_argument("--slots", type=int, default=4, help="c5: requests resident at once")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (the multi-GPU "
              "sharding of assembly, MINRES and the batches is designed but not built)", file=sys.stderr)
        sys.exit(2)
    if args.config != "c2":
        {"c3": run_c3, "c4": run_c4, "c5": run_c5, "kernels": run_kernels,
         "leapfrog": run_leapfrog}[args.config](args)
        return

    import formoniq_amd as fa
    ctx = fa.Context(0)
    t0 = time.perf_counter()
    topo, geo, wc = build_problem(fa, args.n, fa.MIXED, ctx)
    t_mesh = time.perf_counter() - t0
    ncells = topo.nsimplices(3)
    log(f"mesh {args.n}^3: {ncells} tets, counts {topo.counts.tolist()}, numbering {t_mesh:.2f}s")
    f = fa.normal_vector(topo.nsimplices(1), 2)

    t0 = time.perf_counter()
    use_graph = 0 if args.no_graph else 1
    head_prec = fa.FP64_PC32 if args.mixed_mode == "pc32" else fa.MIXED
    other_prec = fa.MIXED if head_prec == fa.FP64_PC32 else fa.FP64_PC32
    mode_name = {fa.FP64_PC32: "fp64_pc32", fa.MIXED: "fp32_refine"}
    mixed = fa.Problem(wc, 1, fa.solve_opts(precision=head_prec, use_graph=use_graph))
    log(f"mixed problem ({mode_name[head_prec]}): sizes {mixed.sizes}, nnz {mixed.nnz}, "
        f"setup {time.perf_counter() - t0:.2f}s")
    n_kkt = mixed.sizes[0] + mixed.sizes[1]

    # source resident on the device before timing
    mixed.upload_source(f)

    for i in range(args.warmup):
        mixed.assemble()
        r = mixed.solve()
        log(f"warmup {i}: iters {r.iters} refinements {r.refinements} residual {r.residual:.2e} "
            f"solve {r.seconds:.3f}s")
    ctx.synchronize()
    lat, reps = [], []
    t_start = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        mixed.assemble()
        r = mixed.solve()
        ctx.synchronize()
        lat.append(time.perf_counter() - ts)
        reps.append(r)
        log(f"step {i}: {lat[-1]:.3f}s iters {r.iters} refinements {r.refinements} residual {r.residual:.2e}")
    ctx.synchronize()
    elapsed = time.perf_counter() - t_start
    assert all(r.converged and r.residual <= 1e-10 for r in reps)
    ms_per_step = elapsed / args.steps * 1e3

    times = mixed.kernel_times(reps=10)
    log(f"kernel times: {times}")

    # the other mixed-precision mode on the same workload
    alt = fa.Problem(wc, 1, fa.solve_opts(precision=other_prec, use_graph=use_graph))
    alt.upload_source(f)
    alt.assemble()
    r_alt = alt.solve()
    t_alt = []
    for i in range(args.fp64_steps):
        ts = time.perf_counter()
        alt.assemble()
        r_alt = alt.solve()
        ctx.synchronize()
        t_alt.append(time.perf_counter() - ts)
        log(f"{mode_name[other_prec]} step {i}: {t_alt[-1]:.3f}s iters {r_alt.iters} "
            f"refinements {r_alt.refinements} residual {r_alt.residual:.2e}")
    assert r_alt.converged and r_alt.residual <= 1e-10
    alt_rate = 1.0 / statistics.mean(t_alt)
    del alt

    # plain FP64 GPU path
    # each path at the Chebyshev degree that is fastest for it on this
    # workload (FP64: 24; the mixed modes run the default, 64)
    fp64 = fa.Problem(wc, 1, fa.solve_opts(precision=fa.FP64, cheb_degree=args.fp64_cheb,
                                           use_graph=use_graph))
    fp64.upload_source(f)
    fp64.assemble()
    r64 = fp64.solve()
    log(f"fp64 warmup: iters {r64.iters} residual {r64.residual:.2e} {r64.seconds:.3f}s")
    t64 = []
    for i in range(args.fp64_steps):
        ts = time.perf_counter()
        fp64.assemble()
        r64 = fp64.solve()
        ctx.synchronize()
        t64.append(time.perf_counter() - ts)
        log(f"fp64 step {i}: {t64[-1]:.3f}s iters {r64.iters} residual {r64.residual:.2e}")
    assert r64.converged and r64.residual <= 1e-10
    fp64_rate = 1.0 / statistics.mean(t64)
    fp64_cheb = int(fp64.opts.cheb_degree)
    del fp64

    # the reference's block structure with Jacobi blocks (no polynomial): the
    # plainest FP64 GPU path
    jac_rate, r_jac = None, None
    if not args.no_jacobi:
        jac = fa.Problem(wc, 1, fa.solve_opts(precision=fa.FP64, cheb_degree=0, use_graph=use_graph))
        jac.upload_source(f)
        ts = time.perf_counter()
        jac.assemble()
        r_jac = jac.solve()
        ctx.synchronize()
        t_jac = time.perf_counter() - ts
        log(f"fp64 jacobi: {t_jac:.3f}s iters {r_jac.iters} residual {r_jac.residual:.2e}")
        assert r_jac.converged and r_jac.residual <= 1e-10
        jac_rate = 1.0 / t_jac
        del jac
    value = args.steps / elapsed

    # roofline of the dominant kernel of the mixed solve: the FP32 Chebyshev
    # step (one preconditioner SpMV fused with the polynomial recurrence);
    # the fused MINRES SpMVs are reported beside it
    nnz = mixed.nnz
    pc = mixed.preconditioner()
    p_nnz, cheb_deg, p_col_bytes = pc["nnz"], pc["degree"], pc["column_bytes"]
    spmv_bytes = 8 * nnz + 8 * (n_kkt + 1) + 12 * n_kkt  # FP32 values, int32 columns
    spmv64_bytes = 12 * nnz + 8 * (n_kkt + 1) + 24 * n_kkt
    if pc["format"] in ("sell64_f16", "sell64_f16_c32"):
        # three-term step on the SELL-64 copy: 4 B per off-diagonal nonzero
        # (FP16 value + int16 offset; 6 B with int32 columns; padding is not
        # counted), 8 B per 64-row slice offset, x_{k+1} read (4n), x_k read
        # and overwritten (8n), f read (4n)
        p_val_bytes = 2
        p_col_bytes = 2 if pc["format"] == "sell64_f16" else 4
        cheb_bytes = ((2 + p_col_bytes) * pc["offdiag_nnz"] + 8 * ((n_kkt + 63) // 64 + 1)
                      + 16 * n_kkt)
    else:
        # two-term CSR step: P values (2 B as scaled FP16, else 4 B) + columns
        # (2 B as band offsets, else 4 B), row pointers, d (4n), res and z read
        # and written (16n), the next direction written (4n); dinv (4n) only
        # when unscaled
        p_val_bytes = 2 if pc["format"] == "csr_f16" else 4
        cheb_bytes = ((p_val_bytes + p_col_bytes) * p_nnz + 8 * (n_kkt + 1) + 24 * n_kkt
                      + (4 * n_kkt if p_val_bytes == 4 else 0))
    achieved = cheb_bytes / times["cheb32"] / 1e9
    traffic = None
    if os.path.exists(args.traffic):
        with open(args.traffic) as fh:
            tr = json.load(fh)
        # only a measurement of the same kernel on the same workload applies
        if tr.get("n_kkt") == n_kkt and tr.get("cheb_step_f32_format") == pc["format"]:
            traffic = tr.get("cheb_step_f32_hbm_bytes_per_launch")
    # per tet: 6 edge lengths + 4 vertex indices in; M_0, dif_test(1), dif_both(2),
    # two diagonals, M_1 and dif_both(1) out (138 doubles)
    elmat_bytes = ncells * (48 + 8 * (16 + 24 + 36 + 6 + 4 + 36 + 16))

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f64 assembly, f64 minres, f32 preconditioner polynomial on f16 storage"
                  if head_prec == fa.FP64_PC32 else "f64 assembly, f32 minres + f64 refinement"),
        "data": "synthetic",
        "config": {
            "workload": f"3D mixed Hodge-Laplace k=1, Kuhn {args.n}^3 ({ncells} tets), jitter 0.05h seed 2, "
                        "N(0,1) edge source seed 2, full complex, true residual <= 1e-10",
            "kkt_n": n_kkt, "kkt_nnz": nnz,
        },
        "mixed_mode": mode_name[head_prec],
        f"solves_per_s_{mode_name[other_prec]}": alt_rate,
        f"iters_{mode_name[other_prec]}": int(r_alt.iters),
        f"refinements_{mode_name[other_prec]}": int(r_alt.refinements),
        "solves_per_s_fp64": fp64_rate,
        "mixed_over_fp64": value / fp64_rate,
        "solves_per_s_fp64_jacobi": jac_rate,
        "mixed_over_fp64_jacobi": value / jac_rate if jac_rate else None,
        "iters_fp64_jacobi": int(r_jac.iters) if r_jac is not None else None,
        "latency_p50_ms": float(np.percentile(np.array(lat) * 1e3, 50)),
        "latency_p99_ms": float(np.percentile(np.array(lat) * 1e3, 99)),
        "iters_mixed": int(reps[-1].iters), "refinements_mixed": int(reps[-1].refinements),
        "iters_fp64": int(r64.iters),
        "residual_mixed": float(reps[-1].residual), "residual_fp64": float(r64.residual),
        "elmat_blocks_per_s": 7 * ncells / times["elmat"],
        "elmat_cells_per_s": ncells / times["elmat"],
        "elmat_hbm_gbs": elmat_bytes / times["elmat"] / 1e9,
        "kernel_seconds": times,
        "roofline": {
            "kernel": ("k_cheb3_step (Chebyshev step on the SELL-64 FP16 preconditioner: SpMV + three-term "
                       "recurrence, FP32 arithmetic)" if pc["format"].startswith("sell64") else
                       "k_cheb_step<float> (Chebyshev step: preconditioner SpMV + recurrence, FP32 arithmetic)"),
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "algorithmic_bytes_per_launch": cheb_bytes,
            "launch_seconds": times["cheb32"], "precond_nnz": p_nnz, "cheb_degree": cheb_deg,
            "precond_column_bytes": p_col_bytes, "precond_value_bytes": p_val_bytes,
            "precond_format": pc["format"], "precond_offdiag_nnz": pc["offdiag_nnz"],
            "precond_stored_entries": pc["stored"],
        },
        "mr_spmv_f32": {"launch_seconds": times["spmv32"], "algorithmic_bytes_per_launch": spmv_bytes,
                        "achieved_gbs": spmv_bytes / times["spmv32"] / 1e9},
        "mr_spmv_f64": {"launch_seconds": times["spmv64"], "algorithmic_bytes_per_launch": spmv64_bytes,
                        "achieved_gbs": spmv64_bytes / times["spmv64"] / 1e9},
    }
    if not args.no_cpu_baseline:
        try:
            line["cpu_baseline"] = cpu_baseline(args.cpu_n, int(r64.iters), ncells, nnz,
                                                fp64_cheb)
        except Exception as e:  # reported, never fatal for the GPU line
            line["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
