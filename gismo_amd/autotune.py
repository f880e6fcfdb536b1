"""Autotuner of the batched CG's SpMV tile shapes (gsx_batch_set_tile).

The tile shape of a batch member is its SpMV's work split: for a box member
(one tensor-product patch, gsx_batch_add_box) the threads per row (1, 2, 4)
and, in mixed precision, whether the FP32 operator is the upper half of the
stencil or all of it; for a CSR member the lanes per row (8..64).  Which is
fastest depends on the degree, the dimension and the member's size (rows per
tile and the number of chunks that fill the device), so the tuner measures
it: for each class (precision, d, p, size bucket) it assembles one
batch of separately assembled single-patch problems of that class, enough
of them to fill the GPU, adds them as members once per candidate shape, and
times CG iterations of the captured batch steps (tolerance 0: every member
keeps iterating).  The table, keyed by the device architecture,
is written as JSON under gismo_amd/tuning/; `Table.lookup` serves it to the
batch and serving paths, and a class missing from it takes the default
(box member, 1 thread per row, half stencil).

Run on the GPU:  python -m gismo_amd.autotune [--out PATH] [--iters N]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_DIR = os.path.join(HERE, "tuning")

# size buckets by member rows; representative elements per direction
BUCKETS = (("s", 4096), ("m", 32768), ("l", None))
SIZES = {2: {"s": 24, "m": 96, "l": 320}, 3: {"s": 8, "m": 24, "l": 44}}
BOX_LANES = (1, 2, 4)
CSR_LANES = (8, 16, 32, 64)


def bucket(rows):
    for name, lim in BUCKETS:
        if lim is None or rows < lim:
            return name
    return BUCKETS[-1][0]


def class_key(precision_name, d, p, rows):
    return f"{precision_name}-d{d}-p{p}-{bucket(rows)}"


def candidates(mixed):
    out = []
    for lanes in BOX_LANES:
        for half in ((1, 0) if mixed else (1,)):
            out.append({"kind": "box", "lanes": lanes, "half": half})
    for lanes in CSR_LANES:
        out.append({"kind": "csr", "lanes": lanes, "half": 1})
    return out


def _name(c):
    return f"{c['kind']}-{c['lanes']}" + ("" if c["kind"] == "csr" or c["half"] else "-full")


class Table:
    """Tile shapes per class for one architecture (a dict as in the JSON)."""

    def __init__(self, entries=None, arch=None):
        self.entries = entries or {}
        self.arch = arch

    @classmethod
    def load(cls, arch, directory=DEFAULT_DIR):
        path = os.path.join(directory, f"batch_tiles_{arch.split(':')[0]}.json")
        if not os.path.exists(path):
            return cls(arch=arch)
        with open(path) as f:
            data = json.load(f)
        return cls(data.get("classes", {}), data.get("arch"))

    def lookup(self, mixed, d, p, rows):
        """(kind, lanes, half) for a member: kind 'box' or 'csr'."""
        e = self.entries.get(class_key("mixed" if mixed else "fp64", d, p, rows))
        if e is None:
            return "box", 1, 1
        b = e["best"]
        return b["kind"], b["lanes"], b["half"]


def _measure(ctx, gsx, mats, box, cand, mixed, iters, reps):
    """Microseconds per batched CG iteration of one member per (A, b, x) in
    mats, all with the candidate shape."""
    B = gsx.Batch(ctx, gsx.MIXED if mixed else gsx.FP64, capacity=len(mats))
    # tolerance 0: every member iterates through all timed steps (each
    # member keeps a residual history of maxit + 1 values)
    maxit = iters * (reps + 2)
    for A, b, x in mats:
        x.zero()
        if cand["kind"] == "box":
            B.add(A, b, x, tol=0.0, maxit=maxit, box=box, tile=(cand["lanes"], cand["half"]))
        else:
            B.add(A, b, x, tol=0.0, maxit=maxit, tile=(cand["lanes"], 1))
    B.step(iters)            # initialisation, graph capture, warm-up
    ctx.synchronize()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        B.step(iters)
        ctx.synchronize()
        times.append(time.perf_counter() - t0)
    B.close()
    return 1e6 * float(np.median(times)) / iters


def _problem(problems, d, p, n, k):
    """Representative member k of a class: a perturbed patch, lognormal kappa."""
    P = problems.box_patch(d, p, n)
    problems.perturb_interior(P, 0.1 / n, 7919 * k + 11 * n + p)
    P.kappa = problems.lognormal_kappa(P.nelem, 0.5, 104729 * k + 5 + n)
    P.felem = np.ones(P.nelem)
    P.fkind = 0
    return problems.Problem([P], problems.all_sides(d), [], name=f"tune-d{d}-p{p}-n{n}-{k}")


def tune(ctx, iters=32, reps=3, degrees=(2, 3, 4), dims=(2, 3), log=print):
    """Every member gets its own matrix (separately assembled patches with
    their own noise and kappa), as in a batch of independent problems: a
    shared matrix would be read from the Infinity Cache and favour the
    shapes that read it most often."""
    from gismo_amd import gsx, problems
    classes = {}
    for mixed in (True, False):
        pname = "mixed" if mixed else "fp64"
        prec = gsx.MIXED if mixed else gsx.FP64
        for d in dims:
            for p in degrees:
                for bname, _ in BUCKETS:
                    n = SIZES[d][bname]
                    first = _problem(problems, d, p, n, 0)
                    P = first.patches[0]
                    m = [nb - 2 for nb in P.nbasis_dir()] + [1]
                    rows = int(np.prod(m))
                    box = (d, p, m[0], m[1], m[2])
                    # enough members to give the device ~1M rows per launch
                    copies = int(max(1, min(256, round(1.0e6 / rows))))
                    gps, mats = [], []
                    for k in range(copies):
                        gp = gsx.Problem(ctx, first if k == 0 else _problem(problems, d, p, n, k))
                        A, bvec = gp.assemble(gsx.PATTERN_EXPR, prec)
                        gps.append(gp)
                        mats.append((A, bvec, ctx.array(gp.nfree)))
                    res = {}
                    for cand in candidates(mixed):
                        res[_name(cand)] = (cand, _measure(ctx, gsx, mats, box, cand, mixed, iters, reps))
                    best = min(res.values(), key=lambda t: t[1])[0]
                    key = class_key(pname, d, p, rows)
                    classes[key] = {"rows": rows, "members": copies, "best": best,
                                    "us_per_iteration": {k: round(v[1], 2) for k, v in res.items()}}
                    log(f"{key}: rows {rows} x {copies}: best {_name(best)} "
                        + " ".join(f"{k}={v[1]:.1f}" for k, v in sorted(res.items(), key=lambda t: t[1][1])))
                    for A, bvec, x in mats:
                        A.close()
                        bvec.close()
                        x.close()
                    for gp in gps:
                        gp.close()
    return classes


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default=None, help="table path (default gismo_amd/tuning/batch_tiles_<arch>.json)")
    ap.add_argument("--iters", type=int, default=32)
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args(argv)
    from gismo_amd import gsx
    ctx = gsx.Context(0)
    arch = ctx.arch()
    classes = tune(ctx, iters=args.iters, reps=args.reps,
                   log=lambda s: print(s, file=sys.stderr, flush=True))
    out = args.out or os.path.join(DEFAULT_DIR, f"batch_tiles_{arch.split(':')[0]}.json")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    with open(out, "w") as f:
        json.dump({"arch": arch, "iters": args.iters, "reps": args.reps,
                   "measure": "microseconds per batched CG iteration (SpMV + update + direction launches) "
                              "of `members` separately assembled problems of the class (own matrices), median of reps",
                   "classes": classes}, f, indent=1, sort_keys=True)
    print(out)
    ctx.close()


if __name__ == "__main__":
    main()
