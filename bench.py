"""Benchmark: solved DOFs/s of the TSEM Poisson solve on one MI355X.

Workload (one "step" = one solve from x = 0 of the resident system):
  unit-cube Poisson, 32^3 hexes x 6 Kuhn tets = 196,608 tets, seeded +-10%
  interior-vertex jitter (seed 2026), Jacobi-polynomial basis p = 6
  (7,189,057 DOFs), Dirichlet on x = 0, Neumann elsewhere, manufactured
  u = sin(pi x) sin(pi y) sin(pi z)(1 + 0.1 g) (seed 1234).
The headline value is the mixed-precision solve (FP32 matrix-free operator
and smoother inside FP64 iterative refinement) to ||b - Ax||/||b|| <= 1e-10;
the FP64 matrix-free PCG drop-in is timed on the same system beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 32] [--p 6]
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def cpu_baseline(n=8, p=6, seed_mesh=2026, seed_prob=1234, qorder=None):
    """Oracle (CPU restatement of the reference: assembled CSR + Jacobi PCG,
    one thread) on the same workload family at a bounded size."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import harness
    from tsem_amd import mesh, problems
    prob = problems.Manufactured(seed_prob)
    xyz, tet = mesh.kuhn_cube(n, jitter=0.1, seed=seed_mesh)
    S, b = harness.setup_poisson(p, xyz, tet, prob, qorder=qorder or p + 4)
    t0 = time.perf_counter()
    x, st = S.pcg(b, tol=1e-10, criterion=1, precond=1)
    dt = time.perf_counter() - t0
    return dict(value=S.ndof / dt, unit="DOFs/s", cores=1, kind="port",
                sample="n=%d (%d tets), p=%d, %d DOFs, assembled-CSR Jacobi PCG to rel-res 1e-10, %d iterations, %.2f s"
                       % (n, len(tet), p, S.ndof, st["iterations"], dt))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--p", type=int, default=6)
    ap.add_argument("--qorder", type=int, default=30)
    ap.add_argument("--precond", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-steps", type=int, default=1)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (no multi-GPU path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from tsem_amd import api, mesh, problems

    def log(*a):
        print("[bench]", *a, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    xyz, tet = mesh.kuhn_cube(args.n, jitter=0.1, seed=2026)
    ctx = api.Context(0)
    sp = api.Space(ctx, args.p, xyz, tet, qorder=args.qorder)
    log("space: %d tets, %d DOFs, setup %.1f s" % (sp.nt, sp.ndof, time.perf_counter() - t0))
    t1 = time.perf_counter()
    prob = problems.Manufactured(1234)
    b = api.poisson_setup(sp, prob)
    log("rhs + bc %.1f s" % (time.perf_counter() - t1))

    def solve(mixed):
        x = sp.zeros()
        torch.cuda.synchronize()
        t = time.perf_counter()
        x, st = sp.pcg(b, x, tol=1e-10, criterion=1, precond=args.precond, mixed=mixed, inner_tol=1e-4,
                       check_every=20, max_iter=200000)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        if not st["converged"] or st["relres"] > 1e-10:
            raise RuntimeError("solve did not reach rel-res 1e-10: %s" % st)
        return dt, st

    import ctypes
    from tsem_amd import lib as tl

    for _ in range(args.warmup):
        dt, st = solve(True)
        log("warmup mixed %.3f s %s" % (dt, st))
    times, stats = [], []
    tl.check(tl.load().tsem_timing(1))
    for _ in range(args.steps):
        dt, st = solve(True)
        times.append(dt)
        stats.append(st)
        log("mixed %.3f s %s" % (dt, st))
    kt = np.zeros(8)
    tl.check(tl.load().tsem_timing_read(tl.dptr(kt)))
    tl.check(tl.load().tsem_timing(0))
    t64 = []
    for _ in range(args.fp64_steps):
        dt, st64 = solve(False)
        t64.append(dt)
        log("fp64 %.3f s %s" % (dt, st64))
    mean = float(np.mean(times))
    # roofline of the dominant kernel of the timed region (HIP events on the
    # library stream): element kernel (MFMA, FP32) vs scatter (HBM)
    K, E, N = sp.K, sp.nt, sp.ndof
    Kd = args.p * (args.p + 1) * (args.p + 2) // 6
    elem_ms = kt[4] / max(kt[5], 1)
    scat_ms = kt[6] / max(kt[7], 1)
    flops_elem = 12.0 * Kd * K * E                        # two GEMMs through the derivative factors
    nnz_sc = float(sp.nmap)
    bytes_scat = 4.0 * E * K + 4.0 * nnz_sc + 4.0 * (N + 1) + 4.0 * N + 4.0 * N + 1.0 * N  # yl, codes, ptr, y, x, mask
    bytes_elem = 4.0 * N + 80.0 * E + 24.0 * E + 4.0 * E * K  # x (unique), descriptors, factors, yl
    elem_tot, scat_tot = kt[4], kt[6]
    kernels = {
        "element_fp32": {"avg_ms": elem_ms, "launches": int(kt[5]), "tflops": flops_elem / (elem_ms * 1e-3) / 1e12 if kt[5] else None,
                         "algorithmic_flop": flops_elem, "algorithmic_bytes": bytes_elem,
                         "gbs": bytes_elem / (elem_ms * 1e-3) / 1e9 if kt[5] else None},
        "scatter_fp32": {"avg_ms": scat_ms, "launches": int(kt[7]), "algorithmic_bytes": bytes_scat,
                         "gbs": bytes_scat / (scat_ms * 1e-3) / 1e9 if kt[7] else None},
    }
    if elem_tot >= scat_tot:
        ach = kernels["element_fp32"]["tflops"]
        roof = {"kernel": "k_apply_mfma<float> (element operator, FP32 MFMA)", "bound": "mfma", "achieved": ach,
                "peak": 157.3, "unit": "TFLOP/s", "frac": (ach / 157.3) if ach else None, "traffic": None}
    else:
        ach = kernels["scatter_fp32"]["gbs"]
        roof = {"kernel": "k_scatter<float> (deterministic scatter)", "bound": "hbm", "achieved": ach, "peak": 8000.0,
                "unit": "GB/s", "frac": (ach / 8000.0) if ach else None, "traffic": None}
    value = sp.ndof / mean
    fp64_value = sp.ndof / float(np.mean(t64)) if t64 else None
    out = {
        "metric": "solved DOFs/s to rel-res 1e-10, mixed vs FP64 PCG on 1 GPU; p50/p99 latency",
        "value": value,
        "unit": "DOFs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mean * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 operator in f64 refinement",
        "data": "synthetic (seeded jittered Kuhn mesh, manufactured solution)",
        "config": {"workload": "poisson_mixbc_kuhn%d_p%d" % (args.n, args.p), "tets": sp.nt, "dofs": sp.ndof,
                   "qorder": args.qorder, "precond": "jacobi" if args.precond == 1 else "pmg",
                   "jitter": 0.1, "mesh_seed": 2026, "problem_seed": 1234},
        "latency_ms": {"p50": float(np.percentile(times, 50)) * 1e3, "p99": float(np.percentile(times, 99)) * 1e3},
        "fp64_value": fp64_value,
        "mixed_over_fp64": (value / fp64_value) if fp64_value else None,
        "iterations_mixed": stats[-1]["iterations"],
        "outer_mixed": stats[-1]["outer"],
        "iterations_fp64": st64["iterations"] if t64 else None,
        "relres_mixed": stats[-1]["relres"],
        "roofline": roof,
        "kernels": kernels,
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(p=args.p)
            out["cpu_baseline"]["cores"] = 1
        except Exception as e:  # reported, not fatal
            log("cpu baseline failed: %r" % (e,))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
