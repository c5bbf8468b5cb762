#!/usr/bin/env python3
"""Benchmark of the FEEC assembly-and-solve hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): 3D mixed
Hodge-Laplace source problem, k = 1, full complex, on the Kuhn subdivision of
a 64^3 cube grid (1,572,864 tets), interior vertices jittered by
U(+-0.05 h) (SplitMix64 seed 2), Euclidean squared edge lengths from the
coordinates, source cochain N(0,1) on the edges (seed 2).

One step = re-assemble every KKT element block and the CSR values from the
resident geometry (pattern reused), refresh the block preconditioner from the
new values, then solve the KKT system to a true relative residual
|b - A x|/|b| <= 1e-10. The headline solve is mixed precision: FP64 MINRES and
KKT SpMV with the multigrid block preconditioner (Hiptmair-Xu blocks over a
geometric V-cycle, each block accelerated by a Chebyshev polynomial) applied
in FP32 (FEEC_PREC_FP64_PC32; `--mixed-mode refine` selects FP32 MINRES
inside FP64 iterative refinement, FEEC_PREC_MIXED, which is timed beside it
either way). `value` is completed steps per second (solves/s) with all inputs
resident in HBM. Beside it: the plain FP64 GPU path (FP64 matrix, vectors,
MINRES and preconditioner, same certificate), the Chebyshev-Jacobi
polynomial preconditioner on the same problem, the 4D Lorentzian element
kernel (C3) and a short serving stream (C5) as sub-objects, and the oracle
timed on the host on the same 64^3 workload.

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


def cpu_model():
    """The host CPU's model name and logical CPU count from /proc/cpuinfo."""
    name, count = "unknown", 0
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name") and name == "unknown":
                    name = line.split(":", 1)[1].strip()
                if line.startswith("processor"):
                    count += 1
    except OSError:
        pass
    return name, count


def cpu_baseline(fa, n, cheb_m, cores):
    """The oracle timed on this host on the C2 workload itself: the same 64^3
    jittered mesh and N(0,1) source (built by the library's host functions,
    as the GPU run builds them), the KKT system assembled by the oracle's C
    element kernels (OpenMP, `cores` threads) and its serial COO->CSR merge,
    the oracle's restatement of the GPU's preconditioner (HX blocks with the
    Chebyshev acceleration of the same degree, lambda_max by the same Lanczos
    steps) and its Python restatement of MINRES (scipy SpMV) run until the
    true relative residual is <= 1e-10, as the GPU's certificate. Measured, not extrapolated, except the 1-core figure's element
    kernels: those are timed on 1 thread over the first 1/8 of the cells and
    scaled by 8 (everything after the element kernels is single-threaded
    in both figures)."""
    from oracle import oracle as o
    g = fa.CartesianGrid(3, n)
    cells, coords = g.triangulate()
    coords = g.jitter(coords, 0.05, 2)
    cx = o.Complex(3, cells)
    L = cx.lengths_from_coords(coords)
    f = fa.normal_vector(cx.nsimplices(1), 2)
    blocks = ((o.MASS, 0), (o.DIF_TEST, 1), (o.DIF_BOTH, 2), (o.MASS, 1), (o.DIF_BOTH, 1))
    t0 = time.perf_counter()
    for kind, k in blocks:
        o.element_matrices(cx, L, kind, k, cores)
    t_el = time.perf_counter() - t0
    part = cells[: len(cells) // 8]
    pcx = o.Complex(3, part)
    pL = pcx.lengths_from_coords(coords)
    t0 = time.perf_counter()
    for kind, k in blocks:
        o.element_matrices(pcx, pL, kind, k, 1)
    t_el1 = (time.perf_counter() - t0) * len(cells) / len(part)
    del pcx, pL
    t0 = time.perf_counter()
    c = o.WhitneyComplex(cx, L, memo=True)
    kkt, rhs, ns, nu = o.assemble_mixed_kkt(c, f, 1, c.harmonics(1))
    A = kkt.to_scipy()
    c.hdif_gram(0)  # the nodal block the preconditioner needs (dif_both(1))
    t_asm = time.perf_counter() - t0
    t0 = time.perf_counter()
    hx = o.hx_blocks(c, coords, 1, n)
    lam = o.hx_block_lambdas(c, coords, 1, n, blocks=hx)
    prec = o.hx_block_preconditioner(c, coords, 1, n, 0, cheb_m=cheb_m, lam=lam, blocks=hx)
    t_setup = time.perf_counter() - t0
    # MINRES stops on the preconditioner-norm residual; a quarter of the
    # target there, then correction runs until the true residual meets it
    t0 = time.perf_counter()
    x, rep = o.minres_py(lambda v: A @ v, prec, rhs, 0.25e-10, 1000)
    iters = rep.iters
    bnorm = np.linalg.norm(rhs)
    res = rhs - A @ x
    while np.linalg.norm(res) > 1e-10 * bnorm and iters < 1000:
        dx, rep = o.minres_py(lambda v: A @ v, prec, res, 0.25e-10 * bnorm / np.linalg.norm(res), 1000)
        x, iters = x + dx, iters + rep.iters
        res = rhs - A @ x
    t_solve = time.perf_counter() - t0
    resid = float(np.linalg.norm(res) / bnorm)
    t_all = t_asm + t_setup + t_solve
    t_one = t_all - t_el + t_el1
    model, ncpu = cpu_model()
    return {
        "value": 1.0 / t_all, "unit": "solves/s", "cores": cores, "kind": "port",
        "sample": (f"the C2 workload itself (64^3, {len(cells)} tets, KKT nnz {A.nnz}) through the oracle: "
                   f"KKT assembly {t_asm:.1f}s (element kernels {t_el:.1f}s of it on {cores} threads), "
                   f"HX + Chebyshev (m={cheb_m}) set-up {t_setup:.1f}s, MINRES {iters} iterations "
                   f"{t_solve:.1f}s to true residual {resid:.1e}"),
        "one_core": {"value": 1.0 / t_one, "seconds_per_solve": t_one,
                     "element_kernels_seconds": t_el1,
                     "note": "element kernels timed on 1 thread over 1/8 of the cells, x8"},
        "seconds_per_solve": t_all, "cpu_model": model, "host_logical_cpus": ncpu,
        "iters": int(iters), "true_residual": resid,
    }


def kuhn_n_for(d, cells):
    """Cells per axis of a d-dim Kuhn box with about `cells` simplices."""
    return max(1, int(round((cells / math.factorial(d)) ** (1.0 / d))))


def c3_measure(n=16, reps=3, sample=2048):
    """4D Lorentzian Hodge-Dirac element matrices on the Kuhn n^4 box (16^4:
    1,572,864 4-simplices): time axis 0 unscaled, axes 1..3 jittered by
    U(+-0.02 h) (seed 3), squared lengths in the Minkowski signature
    (-,+,+,+). Throughput of the all-grade element kernel (nine blocks per
    cell: the mass of grades 0..4 and the exterior derivative of grades
    1..4), and the error of every block on a sample of cells against the FP64
    oracle."""
    import formoniq_amd as fa
    from oracle import oracle as o
    g = fa.CartesianGrid(4, n)
    cells, coords = g.triangulate()
    coords = g.jitter(coords, 0.02, 3, axes=[1, 2, 3])
    topo = fa.Complex(4, cells, nverts=len(coords))
    geo = fa.MeshLengthsSq.from_coords(topo, coords, minkowski=True)
    nc = topo.nsimplices(4)
    census, _ = geo.causal_census()
    sec, per = geo.dirac_element_time(reps=reps)
    blocks = [(fa.MASS, k) for k in range(5)] + [(fa.DIF_TRIAL, k) for k in range(1, 5)]
    rng = np.random.default_rng(3)
    pick = np.sort(rng.choice(nc, size=min(nc, sample), replace=False))
    L = geo.lengths()
    edges = topo.cell_faces(1)
    worst = 0.0
    for kind, k in blocks:
        el = fa.WhitneyComplex(geo).element(kind, k)
        for c in pick:
            want = o.element_matrix(4, kind, k, L[edges[c]])
            scale = max(np.abs(want).max(), 1e-300)
            worst = max(worst, float(np.abs(el[c] - want).max() / scale))
        del el
    bytes_per_cell = 8 * 10 + 4 * 10 + 8 * per
    # the kernel writes 30x what it reads: read its rate against what plain
    # streaming writes sustain on this device, beside the 8 TB/s peak
    stream = topo.ctx.stream_rates(1 << 31)
    return {
        "workload": f"4D Lorentzian Hodge-Dirac element matrices, Kuhn {n}^4 ({nc} cells), axes 1..3 "
                    "jittered 0.02h seed 3, Minkowski (-,+,+,+)",
        "edges_timelike_null_spacelike": census.tolist(),
        "seconds_per_launch": sec, "values_per_cell": per,
        "element_matrices_per_s": len(blocks) * nc / sec, "cells_per_s": nc / sec,
        "hbm_gbs": bytes_per_cell * nc / sec / 1e9,
        "frac_of_peak": bytes_per_cell * nc / sec / 1e9 / HBM_PEAK_GBS,
        "stream_rates_gbs": stream,
        "frac_of_write_stream": bytes_per_cell * nc / sec / 1e9 / stream["write_gbs"],
        "max_rel_err_vs_oracle": worst, "sampled_cells": int(pick.size), "dtype": "f64",
    }


def analytic_measure(fa, topo, geo, reps=5):
    """The analytic-data path on the headline mesh (3D Kuhn 64^3, 1,572,864
    tets): a smooth ambient 1-form sampled at the degree-3 cell nodes (5 per
    tet) and resident on the device; per call, its SourceForm load (the
    element kernel plus the per-edge sum), its L2 projection (load, mass
    assembly, Jacobi CG to 1e-12; and with the mass assembled once and
    reused) and the fe_l2_error of the projection; the
    de Rham map of the same form on the edges (degree-3 rule, 3 nodes per
    edge). Wall time per call with the device synchronized, median of reps."""
    nc = topo.nsimplices(3)
    ne = topo.nsimplices(1)

    def field(p):
        return np.stack([np.sin(np.pi * p[..., 1]), np.cos(np.pi * p[..., 2]),
                         np.sin(np.pi * p[..., 0]) * p[..., 1]], -1)
    vals = fa.DeviceArray.from_numpy(field(fa.cell_quadrature_points(geo, 3)))
    evals = fa.DeviceArray.from_numpy(field(fa.simplex_quadrature_points(geo, 1, 3)))

    def timed(fn):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            out = fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)), out
    t_load, _ = timed(lambda: fa.source_load(geo, 1, vals, 3, "ambient"))
    t_proj, (x, rep) = timed(lambda: fa.l2_projection(geo, 1, vals, 3, "ambient"))
    mass = fa.WhitneyComplex(geo).mass(1)
    t_proj_m, _ = timed(lambda: fa.l2_projection(geo, 1, vals, 3, "ambient", mass=mass))
    t_err, err = timed(lambda: fa.fe_l2_error(x, vals, geo, 1, "ambient"))
    t_dr, _ = timed(lambda: fa.derham_map(geo, 1, evals, 3))
    # per tet: 5 nodes x 3 components of the field, 6 edge ids, 4 vertex ids,
    # 6 element loads out; once per mesh: the squared lengths and coordinates
    load_bytes = nc * (5 * 3 * 8 + 6 * 4 + 4 * 4 + 6 * 8) + ne * 8 + topo.nsimplices(0) * 3 * 8
    return {
        "workload": f"ambient 1-form (sin pi y, cos pi z, y sin pi x) on Kuhn 64^3 ({nc} tets), "
                    "degree-3 cell rule (5 nodes), field resident on the device",
        "source_load_ms": t_load * 1e3, "source_load_cells_per_s": nc / t_load,
        "source_load_algorithmic_gbs": load_bytes / t_load / 1e9,
        "l2_projection_ms": t_proj * 1e3, "l2_projection_cg_iters": int(rep.iters),
        "l2_projection_reused_mass_ms": t_proj_m * 1e3,
        "fe_l2_error_ms": t_err * 1e3, "projection_l2_error": err,
        "derham_map_edges_ms": t_dr * 1e3, "derham_map_edges_per_s": ne / t_dr,
        "dtype": "f64",
    }


def run_analytic(args):
    import formoniq_amd as fa
    ctx = fa.Context(0)
    topo, geo, wc = build_problem(fa, args.n, fa.FP64_PC32, ctx)
    print(json.dumps({"config": "analytic", **analytic_measure(fa, topo, geo, args.steps)}), flush=True)


def run_c3(args):
    import formoniq_amd as fa
    n = args.n if args.n != 64 else 16
    print(json.dumps({"config": "C3", **c3_measure(n, args.steps), "tuning": fa.tune_table()}),
          flush=True)


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


def c5_measure(count, rate, slots, precisions, graph_iters=32, warmup=True, seed=5):
    """Serving stream (seed 5): Poisson arrivals at a fixed rate of
    Hodge-Laplace solve requests, d in {2,3}, k in 0..d-1, cells log-uniform
    in 1e3..2e6; admitted by HBM budget, each solved to 1e-10 with the
    library's default preconditioner choice (the multigrid blocks for the
    large Kuhn-grid requests). Throughput and p50/p99 latency per precision."""
    import formoniq_amd as fa
    reqs = config_requests(fa, count, seed, [2, 3], 1e3, 2e6)
    gaps = np.random.default_rng(seed).exponential(1.0 / rate, size=len(reqs))
    names = _mode_names(fa)
    out = {}
    for prec in precisions:
        for r in reqs:
            r.precision = prec
        opts = fa.ServerOpts()
        opts.hbm_budget_bytes = 200 << 30
        opts.max_inflight = slots
        opts.solve = fa.solve_opts(precision=prec, graph_iters=graph_iters)
        srv = fa.Server(opts)
        if warmup:
            # a server's start-up: one small request of every (d, k) the stream holds
            seen = {}
            for r in config_requests(fa, 64, 55, [2, 3], 1e3, 2e3):
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
                log(f"c5 {names[prec]}: {i + 1} submitted at {time.perf_counter() - t0:.1f}s")
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
            "cells": int(sum(x.ncells for x in res)), "largest_cells": int(max(x.ncells for x in res)),
            "build_ms_max": float(max(x.build_seconds for x in res) * 1e3),
            "slowest": [{"dim": q.dim, "grade": q.grade, "n": q.n, "cells": int(x.ncells),
                         "dofs": int(x.ndofs), "iters": int(x.iters), "latency_ms": float(l * 1e3),
                         "service_ms": float((x.finish_time - x.admit_time) * 1e3),
                         "build_ms": float(x.build_seconds * 1e3), "solve_ms": float(x.solve_seconds * 1e3)}
                        for l, q, x in slow],
        }
    return {"workload": f"{len(reqs)} requests, Poisson rate {rate}/s, d in 2..3, k in 0..d-1, "
                        f"cells logU[1e3,2e6], seed {seed}, {slots} slots", **out}


def run_c5(args):
    import formoniq_amd as fa
    line = c5_measure(args.requests, args.rate, args.slots, (fa.FP64_PC32, fa.FP64), args.graph_iters,
                      not args.no_warmup)
    print(json.dumps({"config": "C5", **line}), flush=True)


def run_kernels(args):
    """Short run for hardware-counter passes: the headline 64^3 k=1 problem's
    element kernel, KKT scatter, fused MINRES SpMV (FP32, FP64) and the
    multigrid preconditioner's kernels (finest-level sweep, product by the
    preconditioner matrix, a B application, a block preconditioner
    application), each launched one by one a few times, then a streaming
    kernel of known size (torch multiply of 256 Mi floats: 1 GiB read, 1 GiB
    written) to calibrate byte counters."""
    import torch
    import formoniq_amd as fa
    ctx = fa.Context(0)
    topo, geo, wc = build_problem(fa, args.n, fa.FP64_PC32, ctx)
    prob = fa.Problem(wc, 1, fa.solve_opts(precision=fa.FP64_PC32))
    prob.assemble()
    times = prob.kernel_times(reps=args.steps)
    mg = prob.mg_kernel_times(reps=args.steps, use_graph=False)
    a = torch.ones(256 << 20, device="cuda:0", dtype=torch.float32)
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 2.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"config": "kernels", "n_kkt": prob.sizes[0] + prob.sizes[1], "nnz": prob.nnz,
                      "kernel_seconds": times, "multigrid_seconds": mg,
                      "calibration_bytes_read": a.numel() * 4,
                      "calibration_bytes_written": b.numel() * 4}), flush=True)


def run_step(args):
    """One headline step (64^3 k=1, FP64_PC32, the default preconditioner) with
    graphs off, for kernel tracing: a warm-up step, then one assemble() and one
    solve(), each between 0.3 s pauses so that a trace splits at them."""
    import formoniq_amd as fa
    ctx = fa.Context(0)
    topo, geo, wc = build_problem(fa, args.n, fa.FP64_PC32, ctx)
    f = fa.normal_vector(topo.nsimplices(1), 2)
    prob = fa.Problem(wc, 1, fa.solve_opts(precision=fa.FP64_PC32, use_graph=0))
    prob.upload_source(f)
    prob.assemble()
    prob.solve()
    ctx.synchronize()
    time.sleep(0.3)
    t0 = time.perf_counter()
    prob.assemble()
    ctx.synchronize()
    t1 = time.perf_counter()
    time.sleep(0.3)
    t2 = time.perf_counter()
    r = prob.solve()
    ctx.synchronize()
    t3 = time.perf_counter()
    print(json.dumps({"config": "step", "assemble_ms": (t1 - t0) * 1e3, "solve_ms": (t3 - t2) * 1e3,
                      "iters": int(r.iters), "residual": float(r.residual)}), flush=True)


def timed_steps(ctx, prob, steps, label):
    """Time `steps` steps (assemble + solve) of a resident problem."""
    lat, reps = [], []
    for i in range(steps):
        ts = time.perf_counter()
        prob.assemble()
        r = prob.solve()
        ctx.synchronize()
        lat.append(time.perf_counter() - ts)
        reps.append(r)
        log(f"{label} step {i}: {lat[-1] * 1e3:.1f} ms iters {r.iters} refinements {r.refinements} "
            f"residual {r.residual:.2e} (solve {r.seconds * 1e3:.1f} ms)")
    assert all(r.converged and r.residual <= 1e-10 for r in reps), label
    return lat, reps


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64, help="cells per axis (the workload is 64)")
    ap.add_argument("--fp64-steps", type=int, default=2,
                    help="timed steps of the FP64 path and of the other mixed-precision mode")
    ap.add_argument("--mixed-mode", default="pc32", choices=["pc32", "refine"],
                    help="headline solve: FP64 MINRES with the FP32 preconditioner (pc32), or FP32 MINRES "
                         "inside FP64 refinement (refine)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-c3", action="store_true", help="leave out the C3 sub-object")
    ap.add_argument("--no-c5", action="store_true", help="leave out the C5 sub-object")
    ap.add_argument("--no-analytic", action="store_true", help="leave out the analytic-data sub-object")
    ap.add_argument("--c5-requests", type=int, default=16, help="requests of the C5 sub-object's stream")
    ap.add_argument("--c5-rate", type=float, default=2.0, help="arrival rate of the C5 sub-object's stream")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "c2_kernel_traffic.json"))
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the Krylov kernels one by one instead of replaying graph segments "
                         "(same kernels; for tracing tools)")
    ap.add_argument("--graph-iters", type=int, default=32,
                    help="MINRES iterations per replayed graph segment")
    ap.add_argument("--no-warmup", action="store_true",
                    help="c5: start the timed stream on a cold server (no start-up requests)")
    ap.add_argument("--config", default="c2",
                    choices=["c2", "c3", "c4", "c5", "kernels", "step", "leapfrog", "analytic"],
                    help="c2 (default): the headline line, with C3 and C5 sub-objects; c3/c4/c5: those "
                         "workloads in full, each printed as its own JSON line")
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
        {"c3": run_c3, "c4": run_c4, "c5": run_c5, "kernels": run_kernels, "step": run_step,
         "leapfrog": run_leapfrog, "analytic": run_analytic}[args.config](args)
        return

    import formoniq_amd as fa
    ctx = fa.Context(0)
    t0 = time.perf_counter()
    topo, geo, wc = build_problem(fa, args.n, fa.FP64_PC32, ctx)
    ncells = topo.nsimplices(3)
    log(f"mesh {args.n}^3: {ncells} tets, counts {topo.counts.tolist()}, numbering "
        f"{time.perf_counter() - t0:.2f}s")
    f = fa.normal_vector(topo.nsimplices(1), 2)
    use_graph = 0 if args.no_graph else 1
    head_prec = fa.FP64_PC32 if args.mixed_mode == "pc32" else fa.MIXED
    other_prec = fa.MIXED if head_prec == fa.FP64_PC32 else fa.FP64_PC32
    mode_name = {fa.FP64_PC32: "fp64_pc32", fa.MIXED: "fp32_refine"}

    # the headline problem: the library's default preconditioner choice
    t0 = time.perf_counter()
    head = fa.Problem(wc, 1, fa.solve_opts(precision=head_prec, use_graph=use_graph,
                                           graph_iters=args.graph_iters))
    head.upload_source(f)  # source resident on the device before timing
    head.assemble()
    pc = head.preconditioner()
    log(f"headline problem ({mode_name[head_prec]}, {pc['format']}): sizes {head.sizes}, nnz {head.nnz}, "
        f"set-up {time.perf_counter() - t0:.2f}s")
    n_kkt = head.sizes[0] + head.sizes[1]
    kkt_nnz = int(head.nnz)
    for i in range(args.warmup):
        head.assemble()
        r = head.solve()
        log(f"warmup {i}: iters {r.iters} refinements {r.refinements} residual {r.residual:.2e} "
            f"solve {r.seconds:.3f}s")
    ctx.synchronize()
    t_start = time.perf_counter()
    lat, reps = timed_steps(ctx, head, args.steps, "headline")
    ctx.synchronize()
    elapsed = time.perf_counter() - t_start
    value = args.steps / elapsed
    ms_per_step = elapsed / args.steps * 1e3
    t_asm = []
    for _ in range(3):  # the assembly part of a step, on its own
        ts = time.perf_counter()
        head.assemble()
        ctx.synchronize()
        t_asm.append(time.perf_counter() - ts)
    times = head.kernel_times(reps=10)
    mg = head.mg_kernel_times(reps=20) if pc["format"].startswith("multigrid") else None
    log(f"kernel times: {times}; multigrid: {mg}")
    cheb_m, lam = pc["degree"], pc["lambda_max"]
    del head

    # the other mixed-precision mode and the plain FP64 path, same options otherwise
    rates = {}
    for prec, name in ((other_prec, mode_name[other_prec]), (fa.FP64, "fp64")):
        prob = fa.Problem(wc, 1, fa.solve_opts(precision=prec, use_graph=use_graph,
                                               graph_iters=args.graph_iters))
        prob.upload_source(f)
        prob.assemble()
        prob.solve()
        lt, rp = timed_steps(ctx, prob, args.fp64_steps, name)
        rates[name] = {"solves_per_s": 1.0 / statistics.mean(lt), "iters": int(rp[-1].iters),
                       "refinements": int(rp[-1].refinements), "residual": float(rp[-1].residual)}
        del prob
    # the polynomial (Chebyshev-Jacobi) block preconditioner on the headline mode, for comparison
    prob = fa.Problem(wc, 1, fa.solve_opts(precision=head_prec, use_graph=use_graph,
                                           preconditioner=fa.PC_CHEBYSHEV))
    prob.upload_source(f)
    prob.assemble()
    prob.solve()
    lt, rp = timed_steps(ctx, prob, 1, "chebyshev-jacobi")
    poly = {"solves_per_s": 1.0 / statistics.mean(lt), "iters": int(rp[-1].iters),
            "degree": int(prob.preconditioner()["degree"])}
    del prob

    # roofline of the dominant kernel: the product by the preconditioner
    # matrix P in FP32 (the single kernel with the largest share of a solve's
    # kernel time: one launch per Chebyshev step, 5 per preconditioner
    # application); beside it the finest-level Jacobi sweep of the V-cycle
    roof, sweep_roof = None, None
    tr = {}
    if os.path.exists(args.traffic):
        with open(args.traffic) as fh:
            tr = json.load(fh)
        if tr.get("n_kkt") != n_kkt:
            tr = {}
    if mg is not None:
        # int16 column offsets (or int32) and FP32 values per entry, 8-byte row
        # pointers, x read and y written once per row
        p_rows, p_nnz, cb = n_kkt, int(pc["nnz"]), int(pc["column_bytes"])
        spmv_bytes = p_nnz * (cb + 4) + 8 * (p_rows + 1) + 8 * p_rows
        achieved = spmv_bytes / mg["pc_spmv32"] / 1e9
        roof = {
            "kernel": "k_pc_spmv32<short,4,2> (y = P x, the block-diagonal preconditioner matrix in FP32, "
                      "4 lanes per row)",
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": tr.get("pc_spmv32_hbm_bytes_per_launch"),
            "algorithmic_bytes_per_launch": spmv_bytes, "launch_seconds": mg["pc_spmv32"],
            "rows": p_rows, "nnz": p_nnz, "column_bytes": cb,
        }
        nb0 = (args.n + 1) ** 3
        nv = 5  # slots of one 3D k = 1 cycle: sigma's V, u's gradient V and its 3 vector copies
        # 8 stencil diagonals (the diagonal and the 7 +delta ones; the -delta entries
        # are the same numbers), dinv, r and x read once, the new iterate written
        sweep_bytes = nb0 * (8 * 4 + 4 + 3 * 4 * nv)
        achieved = sweep_bytes / mg["sweep_fine"] / 1e9
        sweep_roof = {
            "kernel": "k_st_sweep<float,3,0,true> (finest-level weighted-Jacobi sweep of the multigrid V-cycle "
                      "on the Kuhn-stencil diagonals, 5 interleaved slots, FP32)",
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": tr.get("sweep_fine_hbm_bytes_per_launch"),
            "algorithmic_bytes_per_launch": sweep_bytes, "launch_seconds": mg["sweep_fine"],
            "finest_level_vertices": nb0, "slots": nv,
        }
    # per tet, the five element blocks of the KKT system: M_0 (4x4), dif_test(1)
    # (6x4), dif_both(2) (6x6), M_1 (6x6) and dif_both(1) (4x4); in: 6 squared
    # lengths and 4 vertex indices; out: the 116 values plus the two diagonal
    # blocks' 10 diagonal entries the preconditioner matrix takes
    elmat_bytes = ncells * (48 + 16 + 8 * (16 + 24 + 36 + 36 + 16 + 10))
    spmv64_bytes = 12 * kkt_nnz + 8 * (n_kkt + 1) + 24 * n_kkt  # FP64 values, int32 columns, x, y
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
        "dtype": ("f64 assembly, f64 minres and kkt spmv, f32 multigrid/hx preconditioner"
                  if head_prec == fa.FP64_PC32 else "f64 assembly, f32 minres + f64 refinement"),
        "data": "synthetic",
        "config": {
            "workload": f"3D mixed Hodge-Laplace k=1, Kuhn {args.n}^3 ({ncells} tets), jitter 0.05h seed 2, "
                       "N(0,1) edge source seed 2, full complex, true residual <= 1e-10",
            "kkt_n": n_kkt, "kkt_nnz": kkt_nnz,
            "preconditioner": pc["format"], "preconditioner_applications_per_block": int(cheb_m),
            "lambda_max": [float(x) for x in lam],
        },
        "mixed_mode": mode_name[head_prec],
        "iters": int(reps[-1].iters), "refinements": int(reps[-1].refinements),
        "residual": float(reps[-1].residual),
        "step_ms_p50": float(np.percentile(np.array(lat) * 1e3, 50)),
        "step_ms_max": float(np.max(lat) * 1e3),
        "assemble_ms": float(np.median(t_asm) * 1e3),
        f"solves_per_s_{mode_name[other_prec]}": rates[mode_name[other_prec]]["solves_per_s"],
        f"iters_{mode_name[other_prec]}": rates[mode_name[other_prec]]["iters"],
        "solves_per_s_fp64": rates["fp64"]["solves_per_s"],
        "iters_fp64": rates["fp64"]["iters"],
        "mixed_over_fp64": value / rates["fp64"]["solves_per_s"],
        "chebyshev_jacobi": poly,
        "elmat_blocks_per_s": 5 * ncells / times["elmat"],
        "elmat_cells_per_s": ncells / times["elmat"],
        "elmat_hbm_gbs": elmat_bytes / times["elmat"] / 1e9,
        "kernel_seconds": times,
        "multigrid_kernel_seconds": mg,
        "mr_spmv_f64": {"launch_seconds": times["spmv64"], "algorithmic_bytes_per_launch": spmv64_bytes,
                        "achieved_gbs": spmv64_bytes / times["spmv64"] / 1e9},
        "roofline": roof,
        "roofline_sweep_fine": sweep_roof,
    }
    if not args.no_analytic:
        try:
            line["analytic"] = analytic_measure(fa, topo, geo)
        except Exception as e:
            line["analytic"] = {"error": repr(e)}
    if not args.no_c3:
        try:
            line["c3"] = c3_measure(16, 3, 256)
        except Exception as e:  # reported, never fatal for the headline line
            line["c3"] = {"error": repr(e)}
    if not args.no_c5:
        try:
            line["c5"] = c5_measure(args.c5_requests, args.c5_rate, args.slots, (fa.FP64_PC32,),
                                    args.graph_iters, not args.no_warmup)
        except Exception as e:
            line["c5"] = {"error": repr(e)}
    # the launch choices the autotuner made during this run (element kernels fused or
    # split by grade, per size class; the batched grid): key, choice, each candidate's ms
    line["tuning"] = fa.tune_table()
    if not args.no_cpu_baseline:
        try:
            cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            line["cpu_baseline"] = cpu_baseline(fa, args.n, int(cheb_m), cores)
        except Exception as e:
            line["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
