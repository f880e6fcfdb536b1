"""Benchmark: assembled+solved DOFs/s of the isogeometric Poisson pipeline.

Workload (one GPU): 3D Poisson on the unit cube, degree 3, 64^3 elements
(300,763 basis functions, 274,625 free), interior control points perturbed by
seeded U(+-0.15 h), per-element kappa = exp(N(0, 0.5^2)), manufactured
f = 3 pi^2 sin sin sin, homogeneous Dirichlet.  Inputs (knots, control
points, kappa) are uploaded to HBM before the timed region.

One step = CSR pattern + element matrices + row gather (assembly) and
Jacobi-preconditioned CG to ||r||/||b|| < 1e-8 (the solve; the workload's
stated tolerance).  The mixed path
(reported as `value`) runs the CG iterations on an FP32 copy of the FP64
operator with FP64 reliable-update residuals; the plain FP64 path is
measured beside it.  The relative error of each solution against a tight
FP64 reference solve (relres 1e-13) is reported; the 1e-8 bar of the metric
is checked.

Other workloads (--workload): mg (16 glued 2D patches, MG-preconditioned
CG), mg-cube (this cube with MG-preconditioned CG and --smoother), batch
(512 independent problems), serve (a Poisson-arrival request stream).

Prints ONE JSON line.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "assembled+solved DOFs/s at rel err ≤1e-8, mixed vs plain FP64; p99 latency"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def run_step(ctx, gp, gsx, precision, tol, maxit, x, keep=False, fd=False):
    """One step: assemble, then CG to tol with the Jacobi preconditioner, or
    (fd) with the fast diagonalization built in the step."""
    A, b = gp.assemble(gsx.PATTERN_EXPR | gsx.ASM_REBUILD_PATTERN, precision)
    if fd:
        F = gsx.FastDiag(gp)
        it, rel, ok = gsx.fd_cg_solve(A, F, b, x, tol=tol, maxit=maxit, precision=precision)
        F.close()
    else:
        it, rel, ok = A.cg_solve(b, x, tol=tol, maxit=maxit, precision=precision)
    ctx.synchronize()
    if keep:
        return A, b, it, rel, ok
    A.close()
    b.close()
    return None, None, it, rel, ok


def stencil_spmv_bytes(nfree, slots, ngrid, sv, sx):
    """Algorithmic HBM bytes of one structured SpMV: every stored stencil
    coefficient once (the FP32 operator stores the upper half of the
    symmetric stencil: slots - slots // 2 per row), the row -> grid map, x
    once (grid), y once."""
    return nfree * slots * sv + nfree * 4 + ngrid * sx + nfree * sx


def cpu_threads():
    """Host threads this process may use: the CPU affinity, capped by
    OMP_NUM_THREADS when set (the GPU box gives one GPU a share of 16)."""
    n = len(os.sched_getaffinity(0))
    try:
        n = min(n, int(os.environ.get("OMP_NUM_THREADS", n)))
    except ValueError:
        pass
    return max(1, n)


def cpu_baseline(n_sample, tol):
    """The oracle built for speed (libgsx_cpuref.so: -O3, AVX2/FMA, OpenMP)
    on the same generator at n_sample^3 elements: the reference-style
    threaded assembly (element stride per thread, insertion under a lock)
    and Jacobi-CG to tol, on all host threads and on one."""
    import oracle as O
    from gismo_amd import problems
    prob = problems.poisson_noisy_cube(n=n_sample, p=3, seed=20261202)
    nb = prob.patches[0].nbasis
    nt = cpu_threads()
    runs = {}
    for threads in (nt, 1):
        ta, ts, it, nfree = O.baseline_assemble_solve(prob.patches, prob.dir_sides, prob.ifaces, tol, threads)
        runs[threads] = (ta, ts, it)
    ta, ts, it = runs[nt]
    t1a, t1s, _ = runs[1]
    return {"value": nb / (ta + ts), "unit": "DOFs/s", "cores": nt, "kind": "port",
            "one_thread_value": nb / (t1a + t1s),
            "sample": f"oracle (gcc -O3 -march=x86-64-v3 -fopenmp) assemble on {nt} threads + Jacobi-CG({tol:g}) "
                      f"of the same generator at {n_sample}^3 elements, p=3 ({nb} DOFs, {it} CG iterations): "
                      f"assembly {ta:.2f} s + solve {ts:.2f} s on {nt} threads, {t1a:.2f} s + {t1s:.2f} s on 1 "
                      f"thread, {cpu_model()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["cube", "mg", "mg-cube", "batch", "serve"], default="cube",
                    help="cube: single-patch 3D assemble + Jacobi-CG (default); mg: 16-patch 2D "
                         "assemble + multigrid-preconditioned CG; mg-cube: the cube's problem with "
                         "multigrid-preconditioned CG (--smoother); batch: 512 independent problems; "
                         "serve: Poisson-arrival request stream through the continuous-batching server")
    ap.add_argument("--smoother", choices=["scms", "jacobi", "chebyshev", "gauss_seidel"], default="scms",
                    help="mg-cube: multigrid smoother (scms: subspace corrected mass smoother)")
    ap.add_argument("--mg-kind", choices=["h", "p"], default="h",
                    help="mg: h-multigrid (knot coarsening) or p-multigrid (the coarsening of "
                         "examples/pMultiGrid_example.cpp: z transitions down to degree 1, then h)")
    ap.add_argument("--load", type=float, default=0.7, help="serve: arrival rate as a fraction of capacity")
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--p", type=int, default=3)
    ap.add_argument("--tol", type=float, default=None)
    ap.add_argument("--maxit", type=int, default=5000)
    ap.add_argument("--count", type=int, default=None,
                    help="batch: number of problems (512); serve: number of timed requests (48)")
    ap.add_argument("--cpu-sample", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-mg-min", type=int, default=20000,
                    help="serve: multigrid lane from this many free DOFs (0: no multigrid lane)")
    ap.add_argument("--serve-warm", type=int, default=None, help="serve: warm-up requests (count // 3)")
    ap.add_argument("--serve-mg-workers", type=int, default=1, help="serve: multigrid-lane threads")
    ap.add_argument("--serve-group", type=int, default=8,
                    help="serve: right-hand sides of a request sharing SpMV launches (gsx_batch_set_group)")
    ap.add_argument("--serve-mixed-only", action="store_true", help="serve: skip the FP64 run")
    ap.add_argument("--batch-grouped", action="store_true",
                    help="batch: assemble disjoint patches in a few multipatch groups (CSR blocks)")
    ap.add_argument("--batch-csr", action="store_true",
                    help="batch: members iterate on their CSR blocks instead of structured copies")
    ap.add_argument("--no-legs", action="store_true",
                    help="default workload only: skip the multigrid and batch legs")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain 1 GiB device copy (calibrates the FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the path shards by problem/patch across GPUs, which this build does not run.",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "mg":
        print(json.dumps(bench_mg(args)))
        return None
    if args.workload == "mg-cube":
        return bench_mg_cube(args)
    if args.workload == "batch":
        print(json.dumps(bench_batch(args)))
        return None
    if args.workload == "serve":
        return bench_serve(args)
    args.n = 64 if args.n is None else args.n
    args.tol = 1e-8 if args.tol is None else args.tol
    args.cpu_sample = 24 if args.cpu_sample is None else args.cpu_sample

    from gismo_amd import gsx, problems

    prob = problems.poisson_noisy_cube(n=args.n, p=args.p, seed=20261202)
    ctx = gsx.Context(0)
    gp = gsx.Problem(ctx, prob)
    ndofs = gp.nglued
    x = ctx.array(gp.nfree)

    # tight FP64 reference solution for the error check (outside the timing)
    A, b = gp.assemble(gsx.PATTERN_EXPR, gsx.FP64)
    xr = ctx.array(gp.nfree).zero()
    it_ref, rel_ref, _ = A.cg_solve(b, xr, tol=1e-13, maxit=20000)
    xref = xr.download()
    l2_ref = gp.l2_error(xr)
    A.close()
    b.close()
    xr.close()

    results = {}
    for name, prec, fd in (("fp64", gsx.FP64, False), ("mixed", gsx.MIXED, False),
                           ("fd_fp64", gsx.FP64, True), ("fd_mixed", gsx.MIXED, True)):
        for _ in range(args.warmup):
            x.zero()
            run_step(ctx, gp, gsx, prec, args.tol, args.maxit, x, fd=fd)
        # time the dominant SpMV with HIP events on the context stream
        kname = "spmv_f32" if prec == gsx.MIXED else "spmv_f64"
        ctx.enable_timing(True)
        ctx.timing_filter(kname)
        ctx.reset_timing()
        times = []
        its = []
        ctx.synchronize()
        t_all = time.perf_counter()
        for _ in range(args.steps):
            x.zero()
            ctx.synchronize()
            t0 = time.perf_counter()
            _, _, it, rel, ok = run_step(ctx, gp, gsx, prec, args.tol, args.maxit, x, fd=fd)
            times.append(time.perf_counter() - t0)
            its.append(it)
        ctx.synchronize()
        t_total = time.perf_counter() - t_all
        nl, kms = ctx.kernel_stats(kname)
        ctx.enable_timing(False)
        xs = x.download()
        err = float(np.linalg.norm(xs - xref) / np.linalg.norm(xref))
        l2 = gp.l2_error(x)
        # every CG iteration runs one SpMV; the launches a paused mixed
        # recurrence skips (about 1 us each) are counted in kms but not here
        results[name] = dict(ms=1e3 * float(np.mean(times)), ms_min=1e3 * float(np.min(times)),
                             iters=int(its[-1]), relres=float(rel), rel_err=err, l2=l2,
                             spmv_launches=nl, spmv_working_launches=int(sum(its)), spmv_ms=kms,
                             dofs_per_s=ndofs / float(np.mean(times)))

    # phase split of one mixed step (separate, untimed for `value`)
    ctx.enable_timing(True)
    ctx.timing_filter(None)
    ctx.reset_timing()
    x.zero()
    run_step(ctx, gp, gsx, gsx.MIXED, args.tol, args.maxit, x)
    keys = ("pattern_count", "pattern_scan", "pattern_fill", "sf_assemble", "sf_finish", "elem3d", "gather",
            "convert_f32",
            "stencil_build", "cg_init", "spmv_f32", "cg_vec_f32", "fd_apply_f32", "reliable_update")

    def phase_split():
        out = {}
        for k in keys:
            nl, ms = ctx.kernel_stats(k)
            if nl:
                out[k] = round(ms, 3)
        return out
    phases = phase_split()
    ctx.reset_timing()
    x.zero()
    run_step(ctx, gp, gsx, gsx.MIXED, args.tol, args.maxit, x, fd=True)
    phases_fd = phase_split()
    ctx.enable_timing(False)

    A, b = gp.assemble(gsx.PATTERN_EXPR, gsx.MIXED)
    nnz = A.nnz
    A.close()
    b.close()
    n = gp.nfree
    mixed = results["mixed"]
    avg_ms = mixed["spmv_ms"] / max(1, mixed["spmv_working_launches"])
    slots = (2 * args.p + 1) ** 3
    slots = slots - slots // 2            # FP32 operator: upper half stored
    ngrid = (args.n + 3 * args.p) ** 3   # nb = n + p per direction, padded by p each side
    traffic_alg = stencil_spmv_bytes(n, slots, ngrid, 4, 4)
    achieved = traffic_alg / (avg_ms * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": mixed["dofs_per_s"],
        "unit": "DOFs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded geometry noise and coefficients)",
        "config": {"workload": f"3D Poisson unit cube, p={args.p}, {args.n}^3 elements, "
                               f"+-0.15h control-point noise, lognormal kappa; assemble + "
                               f"Jacobi-CG to relres {args.tol:g}",
                   "dofs": ndofs, "free_dofs": n, "nnz": nnz, "parallelism": "single GPU"},
        "fp64": results["fp64"],
        "mixed": mixed,
        "speedup_mixed_vs_fp64": results["fp64"]["ms"] / mixed["ms"],
        # the same steps with CG preconditioned by the reference's
        # fastDiagonalizationOp (built in every step) instead of Jacobi
        "fd_pcg": {"fp64": results["fd_fp64"], "mixed": results["fd_mixed"],
                   "speedup_vs_jacobi_mixed": mixed["ms"] / results["fd_mixed"]["ms"],
                   "phase_ms_one_mixed_step": phases_fd},
        "reference_solution": {"relres": rel_ref, "iters": it_ref, "l2_error_vs_exact": l2_ref},
        "phase_ms_one_mixed_step": phases,
        "roofline": {"kernel": "spmv_f32 (structured (2p+1)^3 stencil, FP32 coefficients)", "bound": "hbm",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": pmc_traffic(f"sspmv_f32_cube{args.n}_p{args.p}"),
                     "algorithmic_bytes_per_launch": traffic_alg, "avg_launch_ms": avg_ms},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample, args.tol)
    gp.close()
    ctx.close()
    if not args.no_legs:
        # the multigrid and batch workloads at their own sizes, timed the same
        # way (their figures beside the headline, not in `value`)
        import copy
        legs = {}
        for key, fn in (("mg", bench_mg), ("batch", bench_batch)):
            a = copy.copy(args)
            a.n, a.tol, a.count, a.cpu_sample, a.no_cpu = None, None, None, None, True
            a.steps = max(1, min(args.steps, 2)) if key == "batch" else args.steps
            r = fn(a)
            legs[key] = {k: r[k] for k in ("metric", "value", "unit", "ms_per_step", "dtype", "config", "fp64",
                                           "mixed", "speedup_mixed_vs_fp64", "roofline") if k in r}
            log(f"leg {key}: {r['value']:.4g} DOFs/s, {r['ms_per_step']:.1f} ms per step")
        out["legs"] = legs
    print(json.dumps(out))


def copy_probe():
    """A plain device copy of known size, three launches: what the HBM
    counters read for it calibrates their bytes per KiB unit."""
    import torch
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(3):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    copy_ms = e0.elapsed_time(e1) / 3
    # read-only stream of the same buffer (a reduction)
    s = a.sum()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(3):
        s = a.sum()
    e1.record()
    torch.cuda.synchronize()
    read_ms = e0.elapsed_time(e1) / 3
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3,
                      "copy_GBps": 8 * n / (copy_ms * 1e6), "read_GBps": 4 * n / (read_ms * 1e6)}))


def pmc_traffic(name):
    """HBM bytes per launch of a kernel from its committed counter summary
    (profiles/pmc_<name>.txt, line 'traffic_bytes_per_launch N'), or None."""
    path = os.path.join(ROOT, "profiles", f"pmc_{name}.txt")
    try:
        with open(path) as f:
            for line in f:
                key, _, val = line.strip().partition(" ")
                if key == "traffic_bytes_per_launch":
                    return int(float(val))
    except OSError:
        return None
    return None


def _common(args, metric_value, ms, dtype, workload, extra):
    out = {
        "metric": METRIC,
        "value": metric_value,
        "unit": "DOFs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded geometry noise, coefficients and sources)",
        "config": {"workload": workload, "parallelism": "single GPU"},
    }
    out["config"].update(extra.pop("config", {}))
    out.update(extra)
    return out


def bench_mg(args):
    """16 degree-3 patches tiling [0,4]^2, n elements per patch direction,
    C0-glued; one step = assemble + multigrid setup (knot-coarsened
    hierarchy, Galerkin levels, dense coarse inverse; with --mg-kind p the
    p-multigrid hierarchy: levels of lower degree assembled over the fine
    geometry, lumped-mass transfers) + CG preconditioned by V(2,2) with
    damped Jacobi (omega 0.6) to relres tol."""
    from gismo_amd import gsx, problems
    n = 256 if args.n is None else args.n
    tol = 1e-8 if args.tol is None else args.tol
    levels = int(round(np.log2(n))) - 3   # coarsest 16 elements per patch direction
    prob = problems.square_grid(4, 4, n=n, p=args.p)
    pmg = args.mg_kind == "p"
    coarsening = ("h" * (levels - args.p) + "z" * (args.p - 1)) if levels > args.p else "z" * (levels - 1)

    def make_mg(A, prec=gsx.FP64):
        if pmg:
            return gsx.Multigrid.pmg(gp, A, levels, coarsening, npre=2, npost=2, damping=0.6, precision=prec)
        return gsx.Multigrid(gp, A, levels, 2, 2, 0.6, precision=prec)
    ctx = gsx.Context(0)
    gp = gsx.Problem(ctx, prob)
    ndofs = gp.nglued

    # tight FP64 reference (outside the timing)
    A, b = gp.assemble(gsx.PATTERN_FULL, gsx.FP64)
    mg = make_mg(A)
    xr = ctx.array(gp.nfree).zero()
    it_ref, rel_ref, _ = gsx.mg_cg_solve(A, mg, b, xr, tol=1e-13, maxit=500)
    xref = xr.download()
    l2_ref = gp.l2_error(xr)
    nnz = A.nnz
    mg.close(); A.close(); b.close(); xr.close()

    log(f"mg reference: {it_ref} iterations, relres {rel_ref:.2e}")
    x = ctx.array(gp.nfree)
    results = {}
    for name, prec in (("fp64", gsx.FP64), ("mixed", gsx.MIXED)):
        def step():
            A, b = gp.assemble(gsx.PATTERN_FULL | gsx.ASM_REBUILD_PATTERN | gsx.ASM_KEEP_ELEMENTS, prec)
            mg = make_mg(A, prec)
            it, rel, ok = gsx.mg_cg_solve(A, mg, b, x, tol=tol, maxit=500)
            ctx.synchronize()
            mg.close(); A.close(); b.close()
            return it, rel
        for _ in range(args.warmup):
            x.zero()
            step()
        # the outer CG's fine-level SpMV timed with HIP events on the stream
        ctx.enable_timing(True)
        ctx.timing_filter("spmv_f64")
        ctx.reset_timing()
        times = []
        for _ in range(args.steps):
            x.zero()
            ctx.synchronize()
            t0 = time.perf_counter()
            it, rel = step()
            times.append(time.perf_counter() - t0)
        nl, kms = ctx.kernel_stats("spmv_f64")
        ctx.enable_timing(False)
        log(f"mg {name}: {1e3 * np.mean(times):.1f} ms per step, {it} iterations")
        xs = x.download()
        results[name] = dict(ms=1e3 * float(np.mean(times)), ms_min=1e3 * float(np.min(times)), iters=it,
                             relres=rel, rel_err=float(np.linalg.norm(xs - xref) / np.linalg.norm(xref)),
                             l2=gp.l2_error(x), dofs_per_s=ndofs / float(np.mean(times)),
                             fine_spmv_launches=nl, fine_spmv_ms=kms)
    # phase split of one mixed step
    ctx.enable_timing(True)
    ctx.timing_filter(None)
    ctx.reset_timing()
    x.zero()
    A, b = gp.assemble(gsx.PATTERN_FULL | gsx.ASM_REBUILD_PATTERN | gsx.ASM_KEEP_ELEMENTS, gsx.MIXED)
    mg = make_mg(A, gsx.MIXED)
    gsx.mg_cg_solve(A, mg, b, x, tol=tol, maxit=500)
    ctx.synchronize()
    phases = {}
    for k in ("pattern_count", "pattern_scan", "pattern_fill", "elem2d", "gather", "mg_setup", "mg_levels",
              "mg_elem", "mg_transfer_build", "mg_galerkin", "mg_coarse_inverse", "mg_smooth", "mg_transfer",
              "mg_coarse", "cg_init", "spmv_f64", "cg_vec", "pmg_setup", "geo_elem", "pmg_mixed_elem",
              "pmg_mixed_gather", "pmg_transfer_build"):
        nl, ms = ctx.kernel_stats(k)
        if nl:
            phases[k] = round(ms, 3)
    ctx.enable_timing(False)
    mg.close(); A.close(); b.close()
    mixed = results["mixed"]
    nf = gp.nfree
    spmv_bytes = nnz * (8 + 4) + (nf + 1) * 4 + 2 * nf * 8
    avg_ms = mixed["fine_spmv_ms"] / max(1, mixed["fine_spmv_launches"])
    achieved = spmv_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    extra = {
        "roofline": {"kernel": "fine-level CSR SpMV of the outer CG (FP64)", "bound": "hbm",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "algorithmic_bytes_per_launch": spmv_bytes, "avg_launch_ms": avg_ms},
        "config": {"dofs": ndofs, "free_dofs": gp.nfree, "nnz": nnz, "levels": levels,
                   "hierarchy": f"p-multigrid, coarsening {coarsening}" if pmg else "h-multigrid",
                   "cycle": "V(2,2), damped Jacobi 0.6, exact FP64 coarse solve"},
        "fp64": results["fp64"], "mixed": mixed,
        "speedup_mixed_vs_fp64": results["fp64"]["ms"] / mixed["ms"],
        "reference_solution": {"relres": rel_ref, "iters": it_ref, "l2_error_vs_exact": l2_ref},
        "phase_ms_one_mixed_step": phases,
    }
    if not args.no_cpu:
        extra["cpu_baseline"] = cpu_baseline_mg(16 if args.cpu_sample is None else args.cpu_sample, args.p)
    out = _common(args, mixed["dofs_per_s"], mixed["ms"], "f32+f64",
                  f"2D multipatch: 16 degree-{args.p} patches on [0,4]^2, {n}x{n} elements per patch, "
                  f"assemble + {'p-' if pmg else ''}MG-preconditioned CG to relres {tol:g}", extra)
    x.close()
    gp.close()
    ctx.close()
    return out


def bench_mg_cube(args):
    """The default workload's cube (n^3 elements, degree p, noisy geometry,
    lognormal kappa, Dirichlet on every side); one step = assemble +
    multigrid setup (3 knot-coarsened levels, Galerkin operators, dense
    coarse inverse, the smoother's own setup) + CG preconditioned by V(1,1)
    with the chosen smoother to relres tol.  With scms the smoother is the
    subspace corrected mass smoother of every level, damped by its power
    iteration (the reference multigrid example's -s scms)."""
    from gismo_amd import gsx, problems
    n = 64 if args.n is None else args.n
    tol = 1e-8 if args.tol is None else args.tol
    levels = 3
    damping = 0.6 if args.smoother in ("jacobi", "chebyshev") else 1.0
    prob = problems.poisson_noisy_cube(n=n, p=args.p)
    ctx = gsx.Context(0)
    gp = gsx.Problem(ctx, prob)
    ndofs = gp.nfree
    x = ctx.array(gp.nfree)
    results = {}
    for name, prec in (("fp64", gsx.FP64), ("mixed", gsx.MIXED)):
        def step():
            A, b = gp.assemble(gsx.PATTERN_EXPR | gsx.ASM_KEEP_ELEMENTS, prec)
            mg = gsx.Multigrid(gp, A, levels, 1, 1, damping, precision=prec, smoother=args.smoother)
            it, rel, ok = gsx.mg_cg_solve(A, mg, b, x, tol=tol, maxit=500)
            ctx.synchronize()
            mg.close(); A.close(); b.close()
            return it, rel, ok
        for _ in range(args.warmup):
            x.zero()
            step()
        times = []
        for _ in range(args.steps):
            x.zero()
            ctx.synchronize()
            t0 = time.perf_counter()
            it, rel, ok = step()
            times.append(time.perf_counter() - t0)
        log(f"mg-cube {args.smoother} {name}: {1e3 * np.mean(times):.1f} ms per step, {it} iterations")
        results[name] = dict(ms=1e3 * float(np.mean(times)), ms_min=1e3 * float(np.min(times)), iters=it,
                             relres=rel, converged=bool(ok), l2=gp.l2_error(x),
                             dofs_per_s=ndofs / float(np.mean(times)))
    ctx.enable_timing(True)
    ctx.timing_filter(None)
    ctx.reset_timing()
    x.zero()
    A, b = gp.assemble(gsx.PATTERN_EXPR | gsx.ASM_KEEP_ELEMENTS, gsx.MIXED)
    mg = gsx.Multigrid(gp, A, levels, 1, 1, damping, precision=gsx.MIXED, smoother=args.smoother)
    gsx.mg_cg_solve(A, mg, b, x, tol=tol, maxit=500)
    ctx.synchronize()
    phases = {}
    for k in ("elem3d", "gather", "mg_setup", "mg_levels", "mg_elem", "mg_transfer_build", "mg_galerkin",
              "mg_coarse_inverse", "mg_smooth", "mg_transfer", "mg_coarse", "cg_init", "spmv_f64", "cg_vec"):
        nl, ms = ctx.kernel_stats(k)
        if nl:
            phases[k] = round(ms, 3)
    ctx.enable_timing(False)
    mg.close(); A.close(); b.close()
    mixed = results["mixed"]
    extra = {
        "config": {"free_dofs": gp.nfree, "levels": levels, "smoother": args.smoother,
                   "cycle": f"V(1,1), {args.smoother}, exact FP64 coarse solve"},
        "fp64": results["fp64"], "mixed": mixed,
        "phase_ms_one_mixed_step": phases,
    }
    out = _common(args, mixed["dofs_per_s"], mixed["ms"], "f32+f64",
                  f"3D Poisson unit cube, p={args.p}, {n}^3 elements, noisy geometry, lognormal kappa; "
                  f"assemble + MG({args.smoother})-preconditioned CG to relres {tol:g}", extra)
    x.close()
    gp.close()
    ctx.close()
    print(json.dumps(out))


def cpu_baseline_mg(n_sample, p):
    """Oracle assemble + MG-PCG (V(2,2), Jacobi 0.6) of the 16-patch grid at a
    reduced per-patch size."""
    import oracle as O
    from gismo_amd import problems
    t0 = time.perf_counter()
    levels = [problems.square_grid(4, 4, n=n_sample >> k, p=p) for k in range(2, -1, -1)]
    maps = [O.mapper(l.patches, l.dir_sides, l.ifaces) for l in levels]
    Af, bf = O.assemble(levels[-1].patches, maps[-1][0], maps[-1][1], 1)
    Ps = [O.transfer_multipatch(levels[i].patches, maps[i][0], maps[i][1], maps[i + 1][0], maps[i + 1][1])
          for i in range(len(levels) - 1)]
    mg = O.Multigrid(Af, Ps, npre=2, npost=2, damping=0.6)
    x, it, err = O.cg_solve(Af, bf, tol=1e-8, maxit=500, precond="mg", mg=mg)
    dt = time.perf_counter() - t0
    nb = sum(P.nbasis for P in levels[-1].patches)
    return {"value": nb / dt, "unit": "DOFs/s", "cores": 1, "kind": "port",
            "sample": f"oracle assemble + MG-PCG(1e-8) of the 16-patch grid at {n_sample}x{n_sample} "
                      f"elements per patch, 3 levels ({dt:.1f} s, {it} iterations), 1 thread of {cpu_model()}"}


def bench_batch(args):
    """Independent single-patch problems (half 2D with 8..96, half 3D with
    8..32 elements per direction, degree 2..4, +-0.1h noise, lognormal
    kappa, random per-element sources).  One step = assembly of every
    problem (each as its own patch; with --batch-grouped, disjoint patches
    grouped into a few multipatch assemblies up to an element-matrix budget)
    + one batched Jacobi-CG over all of them, members on structured operators
    with the tile shapes of the autotuner's table."""
    from gismo_amd import gsx, problems
    # relres 1e-11 keeps every member's error vs the tight solve under 1e-8
    # (at 1e-10 the worst-conditioned 3D p=4 members reach ~1.1e-8)
    tol = 1e-11 if args.tol is None else args.tol
    args.count = 512 if args.count is None else args.count
    probs = problems.batch_problems(count=args.count, seed=7)
    ctx = gsx.Context(0)
    from gismo_amd import autotune
    table = autotune.Table.load(ctx.arch())   # tile shapes per member class (defaults if not tuned)
    if args.batch_grouped:
        # disjoint patches grouped into multipatch assemblies (element path,
        # CSR), members built from the blocks
        budget = 12e9   # bytes of element matrices per assembly (the context allows 16 GiB)
        groups, cur, cur_bytes = [], [], 0.0
        for d in (2, 3):
            for pr in [q for q in probs if q.d == d]:
                P = pr.patches[0]
                bytes_ = P.nelem * (P.p[0] + 1) ** (2 * d) * 8.0
                if cur and (cur_bytes + bytes_ > budget or cur[0].d != d):
                    groups.append(cur)
                    cur, cur_bytes = [], 0.0
                cur.append(pr)
                cur_bytes += bytes_
        if cur:
            groups.append(cur)
        gps = []
        for g in groups:
            joint = problems.Problem([q.patches[0] for q in g],
                                     [(k, s) for k in range(len(g)) for s in range(1, 2 * g[0].d + 1)], [])
            gps.append((g, gsx.Problem(ctx, joint)))
    else:
        # every problem assembled as its own single patch: the assembly fills
        # its structured operator, on which the box member iterates
        gps = [([q], gsx.Problem(ctx, q)) for q in probs]
    ndofs = sum(gp.nglued for _, gp in gps)
    nfree = sum(gp.nfree for _, gp in gps)

    def run(prec, tol_, xs, keep=False):
        mats = []
        B = gsx.Batch(ctx, prec, capacity=len(probs))
        for (g, gp), xg in zip(gps, xs):
            A, b = gp.assemble(gsx.PATTERN_EXPR | gsx.ASM_REBUILD_PATTERN, prec)
            mats.append((A, b))
            off = 0
            for q in g:
                P = q.patches[0]
                m = [nb - 2 for nb in P.nbasis_dir()] + [1]
                nf = int(np.prod(m))
                kind, lanes, half = table.lookup(prec == gsx.MIXED, P.d, P.p[0], nf)
                if args.batch_csr or kind == "csr":
                    box, tile = None, (lanes if kind == "csr" else None)
                    tile = None if tile is None else (tile, 1)
                else:
                    box, tile = (P.d, P.p[0], m[0], m[1], m[2]), (lanes, half)
                B.add(A, b.view(off, nf), xg.view(off, nf), row0=off, nrows=nf, tol=tol_, maxit=args.maxit,
                      box=box, tile=tile)
                off += nf
            assert off == gp.nfree
        B.run(8)
        res = [B.result(i) for i in range(len(probs))]
        launches[0] = B.graph_launches()
        ctx.synchronize()
        B.close()
        nnz = sum(A.nnz for A, _ in mats)
        for A, b in mats:
            A.close(); b.close()
        return res, nnz

    launches = [0]
    xs = [ctx.array(gp.nfree) for _, gp in gps]
    for x in xs:
        x.zero()
    log(f"batch: {len(probs)} problems in {len(gps)} assemblies, {ndofs} DOFs")
    ref_res, nnz = run(gsx.FP64, 1e-12, xs)
    log(f"batch reference done: nnz {nnz}")
    xref = np.concatenate([x.download() for x in xs])
    results = {}
    for name, prec in (("fp64", gsx.FP64), ("mixed", gsx.MIXED)):
        for _ in range(args.warmup):
            for x in xs:
                x.zero()
            run(prec, tol, xs)
        times = []
        for _ in range(args.steps):
            for x in xs:
                x.zero()
            ctx.synchronize()
            t0 = time.perf_counter()
            res, _ = run(prec, tol, xs)
            times.append(time.perf_counter() - t0)
        log(f"batch {name}: {1e3 * np.mean(times):.1f} ms per step")
        xall = np.concatenate([x.download() for x in xs])
        # per-member relative error vs the tight FP64 solve
        errs, off = [], 0
        for g, gp in gps:
            for q in g:
                nf = int(np.prod([nb - 2 for nb in q.patches[0].nbasis_dir()]))
                a, r = xall[off:off + nf], xref[off:off + nf]
                errs.append(float(np.linalg.norm(a - r) / max(1e-300, np.linalg.norm(r))))
                off += nf
        its = [r[1] for r in res]
        results[name] = dict(ms=1e3 * float(np.mean(times)), ms_min=1e3 * float(np.min(times)),
                             graph_launches_per_step=launches[0],
                             max_rel_err=max(errs), converged=sum(r[0] == 1 for r in res),
                             iters_max=max(its), iters_mean=float(np.mean(its)),
                             dofs_per_s=ndofs / float(np.mean(times)))
    # phase split of one mixed step (launches one by one with events: no graph)
    for x in xs:
        x.zero()
    ctx.enable_timing(True)
    ctx.timing_filter(None)
    ctx.reset_timing()
    t0 = time.perf_counter()
    run(gsx.MIXED, tol, xs)
    ctx.synchronize()
    phases = {"wall_ms": round(1e3 * (time.perf_counter() - t0), 1)}
    for k in ("pattern_count", "pattern_scan", "pattern_fill", "elem2d", "elem3d", "gather", "sf_assemble",
              "sf_finish", "stencil_build", "convert_f32", "batch_box_build", "batch_step_f32"):
        nl, ms = ctx.kernel_stats(k)
        if nl:
            phases[k] = round(ms, 3)
    ctx.enable_timing(False)
    mixed = results["mixed"]
    mixed["phase_ms_one_step"] = phases
    extra = {"config": {"problems": len(probs), "dofs": ndofs, "free_dofs": nfree, "nnz": nnz,
                        "assemblies": len(gps),
                        "assembly": "multipatch groups" if args.batch_grouped else "one single-patch assembly per problem",
                        "tile_table": (f"gismo_amd/tuning/batch_tiles_{table.arch.split(':')[0]}.json"
                                       if table.entries else "defaults (not tuned)"),
                        "members": "CSR blocks" if args.batch_csr else "structured (box) members"},
             "fp64": results["fp64"], "mixed": mixed,
             "speedup_mixed_vs_fp64": results["fp64"]["ms"] / mixed["ms"],
             "reference": {"tol": 1e-12, "converged": sum(r[0] == 1 for r in ref_res)}}
    if not args.no_cpu:
        extra["cpu_baseline"] = cpu_baseline_batch(probs, tol)
    out = _common(args, mixed["dofs_per_s"], mixed["ms"], "f32+f64",
                  f"{len(probs)} independent single-patch problems (2D 8..96, 3D 8..32 elements per "
                  f"direction, p 2..4), assemble + batched Jacobi-CG to relres {tol:g}", extra)
    for x in xs:
        x.close()
    for _, gp in gps:
        gp.close()
    ctx.close()
    return out


def cpu_baseline_batch(probs, tol, budget_s=15.0):
    """Oracle assemble + Jacobi-CG of the batch's problems one after another
    (the reference's way), until budget_s seconds: DOFs/s of that sample."""
    import oracle as O
    t0 = time.perf_counter()
    dofs = 0
    k = 0
    order = sorted(range(len(probs)), key=lambda i: probs[i].patches[0].nbasis)
    for i in order[::7]:
        pr = probs[i]
        dm, nf = O.mapper(pr.patches, pr.dir_sides, pr.ifaces)
        A, b = O.assemble(pr.patches, dm, nf, 0)
        O.cg_solve(A, b, tol=tol, maxit=5000)
        dofs += pr.patches[0].nbasis
        k += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": dofs / dt, "unit": "DOFs/s", "cores": 1, "kind": "port",
            "sample": f"oracle assemble + Jacobi-CG({tol:g}) of {k} of the batch's problems "
                      f"(every 7th by size, {dofs} DOFs, {dt:.1f} s), 1 thread of {cpu_model()}"}


def bench_serve(args):
    """Stream of requests (single-patch 2D/3D problems of 1e3..2e6 DOFs,
    log-uniform, degree 2..4, 1..8 right-hand sides) with Poisson arrivals
    at `load` x the measured capacity, through the continuous-batching
    server (mixed precision; FP64 run beside it).  Capacity is measured
    first with a closed burst of warm-up requests."""
    from gismo_amd import gsx, serving
    tol = 1e-10 if args.tol is None else args.tol
    args.count = 48 if args.count is None else args.count
    nwarm = max(16, args.count // 3) if args.serve_warm is None else args.serve_warm
    reqs = serving.make_requests(nwarm + args.count, seed=20261202)
    for r in reqs:
        r.tol = tol
    warm, main_reqs = reqs[:nwarm], reqs[nwarm:]
    log(f"serve: {len(reqs)} requests generated")
    # sampled requests keep their solutions for the error check
    check = [r for r in main_reqs if r.problem.patches[0].nbasis <= 40000][:12]
    for r in check:
        r.keep = True

    def run(prec, rate, rs):
        srv = serving.Server(precision=prec, iters_per_step=8, mg_min_dofs=args.serve_mg_min,
                             group_max=args.serve_group, mg_workers=args.serve_mg_workers)
        t0 = srv.now()
        if rate is None:
            for r in rs:
                r.arrival = t0
                srv.submit(r)
        else:
            gaps = np.random.default_rng(5).exponential(1.0 / rate, len(rs))
            arr = t0 + np.cumsum(gaps)
            i = 0
            last = srv.now()
            while i < len(rs) or not srv.idle():
                now = srv.now()
                if now - last > 10.0:
                    log(f"serve: {len(srv.results)} of {len(rs)} requests done")
                    last = now
                while i < len(rs) and arr[i] <= now:
                    rs[i].arrival = float(arr[i])
                    srv.submit(rs[i])
                    i += 1
                if srv.poll() == 0 and srv.idle():
                    time.sleep(0.0005)
        srv.drain()
        log(f"serve: {len(srv.results)} requests done")
        res = srv.results
        srv.close()
        return res

    out_runs = {}
    precs = (("mixed", gsx.MIXED),) if args.serve_mixed_only else (("mixed", gsx.MIXED), ("fp64", gsx.FP64))
    for name, prec in precs:
        res = run(prec, None, warm)
        span = max(r.done for r in res) - min(r.arrival for r in res)
        work = sum(r.dofs * r.nrhs for r in res)
        cap_rps = len(res) / span
        rate = args.load * cap_rps
        log(f"serve {name}: capacity {cap_rps:.1f} requests/s ({work / span / 1e6:.1f} M DOF-solves/s)")
        res = run(prec, rate, main_reqs)
        lat = np.array([r.latency for r in res])
        span = max(r.done for r in res) - min(r.arrival for r in res)
        work = sum(r.dofs * r.nrhs for r in res)
        ok = all(all(s == 1 for s in r.status) for r in res)
        lanes = {}
        for r in res:
            ln = lanes.setdefault(r.lane, dict(requests=0, dof_solves=0, service_s=0.0, max_latency_ms=0.0))
            ln["requests"] += 1
            ln["dof_solves"] += r.dofs * r.nrhs
            ln["service_s"] = round(ln["service_s"] + (r.done - r.admitted), 3)
            ln["max_latency_ms"] = round(max(ln["max_latency_ms"], 1e3 * r.latency), 1)
        out_runs[name] = dict(requests=len(res), rate_rps=rate, capacity_rps=cap_rps, lanes=lanes,
                              p50_ms=1e3 * float(np.percentile(lat, 50)),
                              p99_ms=1e3 * float(np.percentile(lat, 99)),
                              mean_ms=1e3 * float(lat.mean()), dofs_per_s=work / span,
                              requests_per_s=len(res) / span, all_converged=ok,
                              span_s=span, results={r.rid: r for r in res if r.solutions is not None})
        log(f"serve {name}: p50 {out_runs[name]['p50_ms']:.1f} ms, p99 {out_runs[name]['p99_ms']:.1f} ms, "
            f"{work / span / 1e6:.1f} M DOF-solves/s")
    # error of the sampled mixed results vs tight FP64 solves, and the
    # sampled results of both runs against each other
    ctx = gsx.Context(0)
    errs = []
    for r in check:
        rm = out_runs["mixed"]["results"][r.rid]
        gp = gsx.Problem(ctx, r.problem)
        A, b0 = gp.assemble(gsx.PATTERN_EXPR, gsx.FP64)
        for k, seed in enumerate(r.rhs_seeds):
            b = b0 if k == 0 else ctx.to_device(serving.rhs_vector(gp.nfree, seed))
            x = ctx.array(gp.nfree).zero()
            A.cg_solve(b, x, tol=1e-13, maxit=50000)
            xr = x.download()
            errs.append(float(np.linalg.norm(rm.solutions[k] - xr) / np.linalg.norm(xr)))
            x.close()
            if k:
                b.close()
        A.close(); b0.close(); gp.close()
    ctx.close()
    for v in out_runs.values():
        v.pop("results")
    mixed = out_runs["mixed"]
    extra = {"config": {"requests": args.count, "warmup_requests": nwarm, "load": args.load,
                        "sizes": "log-uniform 1e3..2e6 DOFs, d in {2,3}, p in {2,3,4}, 1..8 rhs",
                        "tol": tol, "iterations_per_step": 8, "mg_lane_min_free_dofs": args.serve_mg_min,
                        "spmv_group_max": args.serve_group,
                        "mg_workers": args.serve_mg_workers},
             "mixed": mixed, "fp64": out_runs.get("fp64"),
             "p50_ms": mixed["p50_ms"], "p99_ms": mixed["p99_ms"],
             "sampled_max_rel_err_vs_fp64": max(errs) if errs else None,
             "sampled_solutions": len(errs)}
    out = _common(args, mixed["dofs_per_s"], None, "f32+f64",
                  f"serving: {args.count} requests, Poisson arrivals at {args.load:g} x capacity, "
                  f"continuous batching, Jacobi-CG to relres {tol:g}", extra)
    out["unit"] = "DOF-solves/s"
    print(json.dumps(out))


if __name__ == "__main__":
    main()
