#!/usr/bin/env python3
"""Benchmark of the FEEC assembly-and-solve hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): 3D mixed
Hodge-Laplace source problem, k = 1, full complex, on the Kuhn subdivision of
a 64^3 cube grid (1,572,864 tets), interior vertices jittered by
U(+-0.05 h) (SplitMix64 seed 2), Euclidean squared edge lengths from the
coordinates, source cochain N(0,1) on the edges (seed 2).

One step = re-assemble every KKT element block and the CSR values from the
resident geometry (pattern reused), then solve the KKT system to a true
relative residual |b - A x|/|b| <= 1e-10 with FP32 MINRES inside FP64
iterative refinement. `value` is completed steps per second (solves/s) with
all inputs resident in HBM. The plain FP64 GPU path (FP64 matrix, vectors and
MINRES, same preconditioner, same 1e-10 certificate) is timed beside it.

  python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
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


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64, help="cells per axis (the workload is 64)")
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--cpu-n", type=int, default=32)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "mr_spmv_f32_traffic.json"))
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported in this build "
              "(the path shards by design only; see the multi-GPU section of DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import formoniq_amd as fa
    ctx = fa.Context(0)
    t0 = time.perf_counter()
    topo, geo, wc = build_problem(fa, args.n, fa.MIXED, ctx)
    t_mesh = time.perf_counter() - t0
    ncells = topo.nsimplices(3)
    log(f"mesh {args.n}^3: {ncells} tets, counts {topo.counts.tolist()}, numbering {t_mesh:.2f}s")
    f = fa.normal_vector(topo.nsimplices(1), 2)

    t0 = time.perf_counter()
    mixed = fa.Problem(wc, 1, fa.solve_opts(precision=fa.MIXED))
    log(f"mixed problem: sizes {mixed.sizes}, nnz {mixed.nnz}, setup {time.perf_counter() - t0:.2f}s")
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

    # plain FP64 GPU path
    fp64 = fa.Problem(wc, 1, fa.solve_opts(precision=fa.FP64))
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
    value = args.steps / elapsed

    # roofline of the dominant kernel: the fused FP32 MINRES SpMV
    nnz = mixed.nnz
    spmv_bytes = 8 * nnz + 8 * (n_kkt + 1) + 12 * n_kkt
    achieved = spmv_bytes / times["spmv32"] / 1e9
    traffic = None
    if os.path.exists(args.traffic):
        with open(args.traffic) as fh:
            tr = json.load(fh)
        if tr.get("n_kkt") == n_kkt:
            traffic = tr.get("hbm_bytes_per_launch")
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
        "dtype": "f64 assembly, f32 krylov + f64 refinement",
        "data": "synthetic",
        "config": {
            "workload": f"3D mixed Hodge-Laplace k=1, Kuhn {args.n}^3 ({ncells} tets), jitter 0.05h seed 2, "
                        "N(0,1) edge source seed 2, full complex, true residual <= 1e-10",
            "kkt_n": n_kkt, "kkt_nnz": nnz,
        },
        "solves_per_s_fp64": fp64_rate,
        "mixed_over_fp64": value / fp64_rate,
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
            "kernel": "k_mr_spmv<float> (fused FP32 MINRES SpMV)",
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "algorithmic_bytes_per_launch": spmv_bytes,
        },
    }
    if not args.no_cpu_baseline:
        try:
            line["cpu_baseline"] = cpu_baseline(args.cpu_n, int(r64.iters), ncells, nnz,
                                                int(fp64.opts.cheb_degree))
        except Exception as e:  # reported, never fatal for the GPU line
            line["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
