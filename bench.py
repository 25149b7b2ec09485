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
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def host_threads():
    """Host threads this process may use: the OpenMP setting when given,
    else the CPUs of its affinity mask, at most 16 (a one-GPU box's share)."""
    env = os.environ.get("OMP_NUM_THREADS")
    n = int(env) if env and env.isdigit() else len(os.sched_getaffinity(0))
    return max(1, min(16, n))


def cpu_baseline(n=8, p=4, seed_prob=1234, repeats=5):
    """The oracle (the CPU restatement of the reference path: assembled FP64
    CSR with diagonal-first rows, solve_PCG of src/TSEMSolver.cpp:1775-1930)
    on the reference's own CPU case: the n = 8 Kuhn cube (3,072 tets), p = 4,
    mixed BC, manufactured u, to rel-res 1e-10 from x = 0.  Four variants:
    flag 1 (Jacobi, the reference default) and flag 4 (Gauss-Seidel
    p-multigrid, s = 2, one V-cycle), each on one thread (the reference is
    serial) and on all host threads (OpenMP products and vector sweeps; the
    Gauss-Seidel sweeps stay sequential).  Median solve time of `repeats`
    runs per variant (p-MG setup timed apart).  `value` is the fastest
    variant's DOFs/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import harness
    from tsem_amd import mesh, problems
    prob = problems.Manufactured(seed_prob)
    xyz, tet = mesh.kuhn_cube(n)
    t0 = time.perf_counter()
    S, b = harness.setup_poisson(p, xyz, tet, prob, qorder=p + 4)
    setup = time.perf_counter() - t0
    nth = host_threads()
    variants = {}
    for name, pc in (("jacobi", 1), ("pmg_gs", 4)):
        for th in sorted({1, nth}):
            runs = [S.pcg(b, tol=1e-10, criterion=1, precond=pc, smooth=2, threads=th)[1] for _ in range(repeats)]
            t = float(np.median([r["solve_s"] for r in runs]))
            variants["%s_%dt" % (name, th)] = {"dofs_per_s": S.ndof / t, "solve_s": t, "threads": th,
                                                "iterations": runs[0]["iterations"], "relres": runs[0]["relres"],
                                                "pmg_setup_s": runs[0]["setup_s"]}
    best = max(variants, key=lambda k: variants[k]["dofs_per_s"])
    return dict(value=variants[best]["dofs_per_s"], unit="DOFs/s", cores=variants[best]["threads"], kind="port",
                variant=best, host_threads=nth, variants=variants, assembly_setup_s=setup,
                sample="n=8 Kuhn cube (3072 tets), p=4, %d DOFs, mixed BC, PCG to rel-res 1e-10: flag 1 (Jacobi) "
                       "and flag 4 (Gauss-Seidel p-MG), 1 and %d threads, median of %d solves each"
                       % (S.ndof, nth, repeats))


def profile_traffic(kernel, n, p):
    """HBM bytes per launch of `kernel` from the committed counter summary
    profiles/hbm_traffic_kuhn<n>_p<p>.md (rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE passes, each scaled by the same counter's reading for a copy of
    known size); None when no summary covers this workload."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_kuhn%d_p%d.md" % (n, p))
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
        if len(cells) >= 5 and cells[0] == kernel:
            try:
                return float(cells[4].replace(",", ""))
            except ValueError:
                return None
    return None


def eigen_section(ctx, log, n=16, p=6, m=32, seed=7, tol=1e-8):
    """Dirichlet Laplacian on the unit cube: the m lowest eigenpairs by block
    LOBPCG (FP32 p-multigrid preconditioning, FP64 Gram and Rayleigh-Ritz),
    against the analytic values pi^2 (a^2 + b^2 + c^2)."""
    import torch
    from tsem_amd import api, mesh
    xyz, tet = mesh.kuhn_cube(n)
    sp = api.Space(ctx, p, xyz, tet, qorder=p + 3)
    sp.set_cube_marks(0b111111, 0)
    sp.zero_bc(1)
    torch.cuda.synchronize()
    t = time.perf_counter()
    lam, X, st = sp.lobpcg(m, seed=seed, tol=tol, precond=4, max_iter=500)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    ab = sorted(a * a + b * b + c * c for a in range(1, 8) for b in range(1, 8) for c in range(1, 8))[:m]
    exact = np.pi ** 2 * np.array(ab, dtype=np.float64)
    out = dict(workload="laplace_dirichlet_kuhn%d_p%d_m%d" % (n, p, m), dofs=sp.ndof, seconds=dt,
               iterations=st["iterations"], applies=st["applies"], max_resid=st["max_resid"],
               converged=st["converged"], max_rel_err_vs_analytic=float(np.max(np.abs(lam - exact) / exact)),
               lambda_over_pi2=[float(v) for v in lam / np.pi ** 2])
    log("eigen %s" % {k: v for k, v in out.items() if k != "lambda_over_pi2"})
    sp.close()
    return out


def ks_section(ctx, log, n=24, p=5, steps=20, alpha=0.3):
    """LDA helium (Z = 2, two electrons in the lowest state, 4 states per
    step) on the tanh-graded [-10, 10]^3 box: 20 SCF steps of linear density
    mixing."""
    import torch
    from tsem_amd.ks import KohnShamAtom
    t = time.perf_counter()
    atom = KohnShamAtom(ctx, n=n, p=p, Z=2.0, occ=(2.0, 0.0, 0.0, 0.0))
    setup = time.perf_counter() - t
    torch.cuda.synchronize()
    t = time.perf_counter()
    rec = atom.run(steps=steps, alpha=alpha)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    e = atom.energy()
    out = dict(workload="ks_helium_graded%d_p%d" % (n, p), dofs=atom.sp.ndof, tets=atom.sp.nt, scf_steps=len(atom.history),
               seconds=dt, setup_seconds=setup, energy=e, eigenvalues=rec["eigenvalues"],
               charge=atom.sp.integrate(atom.rho),
               reference_lda_helium={"energy": -2.8348, "eps_1s": -0.5704})
    log("kohn-sham %s" % out)
    atom.sp.close()
    return out


def serve_section(log, count=24, slots=4, seed=2026, loads=(0.8, 0.5)):
    """Served Poisson stream: count requests (n in 8..32, p in 3..8, seeded
    jitter and problem), first all at once (saturated throughput, which also
    fills the space pool), then with Poisson arrivals at each of `loads` x
    that rate (the first is the headline).  Each request's latency (arrival to
    completion) is set against its solve alone on the device with its space
    already built (one worker, a warm-up solve of the same request first)."""
    from tsem_amd import serve
    srv = serve.Server(slots=slots)
    _, sat = srv.run(serve.make_requests(count, seed=seed))
    runs = []
    for load in loads:
        rate = load * count / sat["wall_s"]
        runs.append((load, rate) + srv.run(serve.make_requests(count, seed=seed, rate=rate)))
    srv.close()
    _, rate, reqs, stream = runs[0]
    alone = serve.Server(slots=1)
    iso, iso_ph = [], []
    for r in serve.make_requests(count, seed=seed):
        alone.run([serve.Request(r.id, r.n, r.p, r.seed)])
        (done,), _ = alone.run([serve.Request(r.id, r.n, r.p, r.seed)])
        iso.append(done.done - done.start)
        iso_ph.append(done.phases)
    alone.close()
    ratio = np.array([(r.done - r.arrival) / i for r, i in zip(reqs, iso)])
    by_load = {}
    for load, rt, rq, stm in runs:
        rr = np.array([(r.done - r.arrival) / i for r, i in zip(rq, iso)])
        by_load[str(load)] = {"arrival_rate_per_s": rt, "dofs_per_s": stm["dofs_per_s"], "p50_ms": stm["p50_ms"],
                              "p99_ms": stm["p99_ms"], "latency_over_isolated_p50": float(np.percentile(rr, 50)),
                              "latency_over_isolated_p99": float(np.percentile(rr, 99))}
    # the same stream at the headline load on the C ABI server (workers,
    # lanes, pool and classes in the library; the meshes are the caller's, made
    # before the clock starts; small = below the median element work)
    from math import comb
    small = float(np.median([6 * n ** 3 * comb(p + 3, 3) for n in range(8, 33) for p in range(3, 9)]))
    nat = serve.NativeServer(slots=slots, small_cost=small)
    nat.run(serve.make_requests(count, seed=seed))
    nreqs, nstream = nat.run(serve.make_requests(count, seed=seed, rate=rate))
    nat.close()
    nratio = np.array([(r.done - r.arrival) / i for r, i in zip(nreqs, iso)])
    native = {"p50_ms": nstream["p50_ms"], "p99_ms": nstream["p99_ms"], "dofs_per_s": nstream["dofs_per_s"],
              "max_relres": nstream["max_relres"],
              "latency_over_isolated": {"p50": float(np.percentile(nratio, 50)),
                                        "p99": float(np.percentile(nratio, 99))}}
    for k in np.argsort(ratio)[::-1][:4]:
        r = reqs[k]
        log("serving worst: n %d p %d dofs %d wait %.3f s service %.3f s isolated %.3f s built %s worker %d"
            % (r.n, r.p, r.dofs, r.start - r.arrival, r.done - r.start, iso[k], r.built, r.worker))
        log("  phases in the stream %s, alone %s" % ({a: round(v, 4) for a, v in (r.phases or {}).items()},
                                                    {a: round(v, 4) for a, v in (iso_ph[k] or {}).items()}))
    out = dict(workload="served_poisson_n8-32_p3-8", requests=count, slots=slots,
               saturated={"dofs_per_s": sat["dofs_per_s"], "wall_s": sat["wall_s"],
                          "requests_per_s": count / sat["wall_s"]},
               arrival_rate_per_s=rate, stream=stream,
               latency_over_isolated={"p50": float(np.percentile(ratio, 50)), "p99": float(np.percentile(ratio, 99)),
                                      "max": float(ratio.max())},
               isolated_ms={"p50": float(np.percentile(iso, 50)) * 1e3, "p99": float(np.percentile(iso, 99)) * 1e3},
               by_load=by_load, native_c_abi=native)
    log("serving %s" % out)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--p", type=int, default=6)
    ap.add_argument("--qorder", type=int, default=30)
    ap.add_argument("--precond", type=int, default=0,
                    help="mixed-precision preconditioner: 1 Jacobi, 4 p-multigrid, 0 both (headline: the faster)")
    ap.add_argument("--mixed", type=int, default=1,
                    help="mixed-precision form: 1 FP32 PCG with reliable updates, 2 restarted refinement")
    ap.add_argument("--inner-tol", type=float, default=1e-3,
                    help="mixed 1: residual reduction between reliable updates; 2: per inner solve")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--no-eigen", action="store_true", help="skip the LOBPCG section (16^3, p = 6, 32 pairs)")
    ap.add_argument("--no-ks", action="store_true", help="skip the Kohn-Sham section (helium, 24^3, p = 5)")
    ap.add_argument("--serve", type=int, default=24, help="requests of the served-stream section (0 skips it)")
    ap.add_argument("--aux-only", action="store_true",
                    help="run only the LOBPCG and Kohn-Sham sections and print their records (for profiles); "
                         "no bench line")
    ap.add_argument("--ks-steps", type=int, default=20, help="SCF steps of the Kohn-Sham section")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="after the solves, copy a buffer of this many MiB three times (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (no multi-GPU path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from tsem_amd import api, mesh, problems

    def log(*a):
        print("[bench]", *a, file=sys.stderr, flush=True)

    if args.aux_only:
        ctx = api.Context(0)
        aux = {}
        if not args.no_eigen:
            aux["eigen"] = eigen_section(ctx, log)
        if not args.no_ks:
            aux["kohn_sham"] = ks_section(ctx, log, steps=args.ks_steps)
        print(json.dumps(aux, default=float), flush=True)
        return

    t0 = time.perf_counter()
    xyz, tet = mesh.kuhn_cube(args.n, jitter=0.1, seed=2026)
    ctx = api.Context(0)
    sp = api.Space(ctx, args.p, xyz, tet, qorder=args.qorder)
    log("space: %d tets, %d DOFs, setup %.1f s" % (sp.nt, sp.ndof, time.perf_counter() - t0))
    t1 = time.perf_counter()
    prob = problems.Manufactured(1234)
    b = api.poisson_setup(sp, prob)
    log("rhs + bc %.1f s" % (time.perf_counter() - t1))

    last = {}

    def solve(mixed, precond):
        x = sp.zeros()
        torch.cuda.synchronize()
        t = time.perf_counter()
        x, st = sp.pcg(b, x, tol=1e-10, criterion=1, precond=precond, mixed=args.mixed if mixed else 0,
                       inner_tol=args.inner_tol, check_every=20, max_iter=200000)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        if not st["converged"] or st["relres"] > 1e-10:
            raise RuntimeError("solve did not reach rel-res 1e-10: %s" % st)
        last[(mixed, precond)] = x
        return dt, st

    from tsem_amd import lib as tl

    # mixed precision with each preconditioner (Jacobi, the reference default,
    # and p-multigrid); the headline is the faster one
    pcs = (1, 4) if args.precond == 0 else (args.precond,)
    mixed_runs = {}
    for pc in pcs:
        for _ in range(args.warmup):
            dt, st = solve(True, pc)
            log("warmup mixed precond %d %.3f s %s" % (pc, dt, st))
        ts, sts = [], []
        for _ in range(args.steps):
            dt, st = solve(True, pc)
            ts.append(dt)
            sts.append(st)
            log("mixed precond %d %.3f s %s" % (pc, dt, st))
        mixed_runs[pc] = (ts, sts)
    head_pc = min(mixed_runs, key=lambda pc: float(np.mean(mixed_runs[pc][0])))
    times, stats = mixed_runs[head_pc]
    # one more solve with HIP-event timing around every full-order operator
    # kernel (the events keep its iterations out of captured graphs, so it is
    # not timed); Jacobi, so every apply is the PCG's own
    kt = np.zeros(12)
    tl.check(tl.load().tsem_timing(1))
    solve(True, 1)
    tl.check(tl.load().tsem_timing_read(tl.dptr(kt)))
    tl.check(tl.load().tsem_timing(0))
    # the same for one FP64 Jacobi solve
    kt64 = np.zeros(12)
    tl.check(tl.load().tsem_timing(1))
    solve(False, 1)
    tl.check(tl.load().tsem_timing_read(tl.dptr(kt64)))
    tl.check(tl.load().tsem_timing(0))
    # FP64 drop-in: the reference's PCG on the same system, matrix-free FP64,
    # with Jacobi (the reference default) and with p-multigrid
    fp64 = {}
    for name, pc in (("jacobi", 1), ("pmg", 4)):
        t64 = []
        for _ in range(args.fp64_steps):
            dt, st64 = solve(False, pc)
            t64.append(dt)
            log("fp64 %s %.3f s %s" % (name, dt, st64))
        if t64:
            fp64[name] = {"value": sp.ndof / float(np.mean(t64)), "ms": float(np.mean(t64)) * 1e3,
                          "iterations": st64["iterations"]}
    if args.copy_calib_mib:
        src = torch.ones(args.copy_calib_mib << 18, dtype=torch.float32, device=sp.device)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: %d MiB read + %d MiB written per copy" % (args.copy_calib_mib, args.copy_calib_mib))
        del src, dst
    mean = float(np.mean(times))
    # per-kernel HIP-event timing of the mixed solves' full-order FP32 applies
    K, E, N = sp.K, sp.nt, sp.ndof
    Kd = args.p * (args.p + 1) * (args.p + 2) // 6
    K32 = (K + 31) // 32 * 32
    nnz_sc = float(sp.nmap)
    ct = 4 if (Kd + 15) // 16 * 16 <= 64 else 2
    slab = sp.slab_state()[0]
    nch = (N + 255) // 256
    algo = {
        # k_elem_gemm_h<CT>: the gather (x through the element descriptors, into LDS), two GEMMs
        # through the derivative factors and the slab epilogue in one kernel; 12 Kd K FP32-equivalent
        # flops per element (the fp16 MFMA work is three times that: hi*hi, hi*lo and lo*hi products)
        "k_elem_gemm_h<%d>" % ct: ("mfma", 12.0 * Kd * K * E),
    }
    if slab:
        # k_scatter_slab<float>: the entity slab (E K entries), per-DOF meta and per-chunk spans, y
        # written, x and mask read (fused dot)
        algo["k_scatter_slab<float>"] = ("hbm", 4.0 * E * K + 4.0 * N + 8.0 * nch + 4.0 * N + 4.0 * N + 1.0 * N)
    else:
        # k_scatter_chunks<float>: element results, map codes + row pointers, y written, x and mask read
        # (fused dot)
        algo["k_scatter_chunks<float>"] = ("hbm", 4.0 * E * K + 4.0 * nnz_sc + 4.0 * (N + 1) + 4.0 * N + 4.0 * N
                                           + 1.0 * N)
    kernels = {}
    slots = {"k_elem_gemm_h<%d>" % ct: 8, "k_scatter_slab<float>": 10, "k_scatter_chunks<float>": 10}
    for name in algo:
        tot, cnt = kt[slots[name]], kt[slots[name] + 1]
        bound, work = algo[name]
        avg = tot / cnt if cnt else None
        rate = work / (avg * 1e-3) / (1e12 if bound == "mfma" else 1e9) if avg else None
        kernels[name] = {"avg_ms": avg, "launches": int(cnt), "total_ms": tot, "bound": bound,
                         ("algorithmic_flop" if bound == "mfma" else "algorithmic_bytes"): work,
                         ("tflops" if bound == "mfma" else "gbs"): rate}
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    kd = kernels[dom]
    if kd["bound"] == "mfma":
        # FP32-accurate arithmetic (algorithmic flops against the FP32 peak); the
        # instructions are fp16 MFMA, three products per FP32 product
        roof = {"kernel": dom, "bound": "mfma", "achieved": kd["tflops"], "peak": 157.3, "unit": "TFLOP/s",
                "fp16_mfma": {"achieved": 3 * kd["tflops"] if kd["tflops"] else None, "peak": 2500.0}}
    else:
        roof = {"kernel": dom, "bound": "hbm", "achieved": kd["gbs"], "peak": 8000.0, "unit": "GB/s"}
    roof["frac"] = roof["achieved"] / roof["peak"] if roof["achieved"] else None
    roof["traffic"] = profile_traffic(dom, args.n, args.p)
    # the whole solve against HBM, by the algorithmic bytes of one Jacobi PCG
    # iteration (s bytes per value): the operator reads p and writes A p
    # (2 N s), reads the geometric factors (6 E s) and one int32 gather slot
    # per local coefficient (4 E K), plus 12 B per face-combination term
    # (F_x); the fused CG updates move 11 N s (x += alpha p: 3, r += alpha A p
    # with z = D^-1 r and the fused r.z: 5, p = z + beta p: 3).  F_x counts
    # the terms of face modes that combine several DOFs: faces in a diagonal
    # variant (identity or reflection, every face of the Kuhn meshes) have
    # none, and the kernels read one value per such mode, so F_x = 0 here
    # (with it the 444 MB per FP32 iteration of the 32^3, p = 6 mesh).
    # Intermediates this design writes and reads back (the entity slab) are
    # not algorithmic bytes; they are reported beside it.
    def b_iter(sz):
        return 2.0 * N * sz + 6.0 * E * sz + 4.0 * E * K + 11.0 * N * sz

    it_bytes = b_iter(4.0)
    # the element results (entity slab or element-major), written and read back
    inter_bytes = 2 * 4.0 * E * K
    # the Jacobi mixed solve's iterations are exactly these bytes each
    jac_t, jac_st = mixed_runs.get(1, (None, None))
    n_it = jac_st[-1]["iterations"] if jac_st else None
    jac_mean = float(np.mean(jac_t)) if jac_t else None
    # per-kernel FP64 timing (the drop-in's full-order applies)
    k64 = {}
    for i, name in enumerate(("gather64", "element64", "scatter64")):
        tot, cnt = kt64[2 * i], kt64[2 * i + 1]
        k64[name] = {"avg_ms": tot / cnt if cnt else None, "launches": int(cnt), "total_ms": tot}
    if k64["element64"]["avg_ms"]:
        k64["element64"]["tflops"] = 12.0 * Kd * K * E / (k64["element64"]["avg_ms"] * 1e-3) / 1e12
        k64["element64"]["frac_of_fp64_peak"] = k64["element64"]["tflops"] / 78.6
    # each result's error against the FP64 solve of the same system (both at
    # rel-res 1e-10)
    xm, x64 = last.get((True, head_pc)), last.get((False, 1))
    err_vs_fp64 = float(torch.linalg.norm(xm - x64) / torch.linalg.norm(x64)) if xm is not None and x64 is not None \
        else None
    value = sp.ndof / mean
    fp64_value = max(v["value"] for v in fp64.values()) if fp64 else None
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
                   "qorder": args.qorder, "precond": "jacobi" if head_pc == 1 else "pmg",
                   "mixed_form": "reliable_updates" if args.mixed == 1 else "restarted_refinement",
                   "inner_tol": args.inner_tol, "element_results": "entity_slab" if slab else "element_major",
                   "jitter": 0.1, "mesh_seed": 2026, "problem_seed": 1234},
        "latency_ms": {"p50": float(np.percentile(times, 50)) * 1e3, "p99": float(np.percentile(times, 99)) * 1e3},
        "fp64_value": fp64_value,
        "mixed_over_fp64": (value / fp64_value) if fp64_value else None,
        "iterations_mixed": stats[-1]["iterations"],
        "outer_mixed": stats[-1]["outer"],
        "fp64": fp64,
        "mixed": {("jacobi" if pc == 1 else "pmg"): {"value": sp.ndof / float(np.mean(v[0])),
                                                      "ms": float(np.mean(v[0])) * 1e3,
                                                      "iterations": v[1][-1]["iterations"]}
                  for pc, v in mixed_runs.items()},
        "relres_mixed": stats[-1]["relres"],
        "roofline": roof,
        "solve_hbm": {"solve": "mixed precision, Jacobi PCG", "bytes_per_fp32_iteration": it_bytes,
                      "iterations": n_it, "ms": jac_mean * 1e3 if jac_mean else None,
                      "achieved_GBps": it_bytes * n_it / jac_mean / 1e9 if jac_mean else None,
                      "frac": it_bytes * n_it / jac_mean / 8e12 if jac_mean else None,
                      "intermediate_bytes_per_iteration": inter_bytes,
                      "fp64_jacobi": ({"bytes_per_iteration": b_iter(8.0), "iterations": fp64["jacobi"]["iterations"],
                                       "frac": b_iter(8.0) * fp64["jacobi"]["iterations"] / (fp64["jacobi"]["ms"] * 1e-3)
                                       / 8e12} if "jacobi" in fp64 else None)},
        "mixed_vs_fp64_rel_diff": err_vs_fp64,
        "kernels": kernels,
        "kernels_fp64": k64,
        "cpu_baseline": None,
    }
    for key, skip, fn in (("eigen", args.no_eigen, lambda: eigen_section(ctx, log)),
                          ("kohn_sham", args.no_ks, lambda: ks_section(ctx, log, steps=args.ks_steps)),
                          ("serving", args.serve <= 0, lambda: serve_section(log, count=args.serve))):
        if skip:
            continue
        try:
            out[key] = fn()
        except Exception as e:  # reported, not fatal for the headline
            out[key] = {"error": repr(e)}
            log("%s failed: %r" % (key, e))
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline()
            log("cpu baseline %s" % out["cpu_baseline"])
        except Exception as e:  # reported, not fatal
            log("cpu baseline failed: %r" % (e,))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
