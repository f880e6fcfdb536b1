"""Continuous-batching solve server on one GPU: the Python face of the
native server (gsx_server_* in the C ABI, gismo_amd/csrc/gsx_server.hip).

Requests (a single-patch Poisson/diffusion problem plus 1..8 right-hand
sides) arrive at any time and are admitted in arrival order while their
estimated device memory fits the HBM budget.  The native server runs three
threads, each with its own context (its own HIP stream):

- assembler: admits requests and assembles each one (its structured operator
  included) with its right-hand sides uploaded;
- solver (batch lane): one box member per right-hand side on the request's
  own operator, with the autotuner's tile shape, in one running gsx_batch
  whose steps advance every member together; members join and leave
  between steps;
- multigrid lane: requests of at least `mg_min_dofs` free DOFs whose patch
  coarsens dyadically to a small coarsest level: the operator is assembled,
  its hierarchy built once, and every right-hand side solved by
  MG-preconditioned CG with the hierarchy resident.

A request is complete when all its right-hand sides have finished and its
solutions are in host memory; its operator, hierarchy and workspace are then
released and its budget freed.  Every member's iterates are independent of
the other members, so a request's solutions are bitwise the same alone and
in any mix.

The reference has no serving layer: it assembles and solves one problem at
a time (src/gsAssembler/gsExprAssembler.h:754-833,
src/gsSolver/gsConjugateGradient.hpp:20-93); each member runs that
algorithm.
"""
import ctypes as C
import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import gsx, problems


@dataclass
class Request:
    rid: int
    problem: problems.Problem
    rhs_seeds: List[int]          # one per right-hand side; the first is the assembled load vector
    arrival: float = 0.0          # seconds on the server clock
    tol: float = 1e-10
    maxit: int = 20000
    keep: bool = False            # return the solutions with the result


@dataclass
class Result:
    rid: int
    dofs: int
    nfree: int
    nrhs: int
    arrival: float
    admitted: float
    assembled: float
    done: float
    iterations: List[int]
    relres: List[float]
    status: List[int]
    solutions: Optional[List[np.ndarray]] = None
    lane: str = "batch"           # "batch" (batched Jacobi-CG) or "mg" (multigrid-preconditioned CG)
    error: Optional[str] = None   # set when the request failed (status FAILED for every rhs)

    @property
    def latency(self):
        return self.done - self.arrival


def rhs_vector(nfree, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, 1.0, nfree)


# per right-hand side status of a Result: 1 converged, 2 iteration limit
# reached (the batched CG's own codes), FAILED: the request raised (for
# example a folded patch, GSX_ERR_JACOBIAN) and was dropped
FAILED = 3


class Server:
    """The native continuous-batching server (gsx_server_create ...)."""

    def __init__(self, precision=gsx.MIXED, budget_bytes=160 << 30, iters_per_step=8, capacity=4096,
                 keep_solutions=False, device=0, mg_min_dofs=20000, elem_budget=16 << 30, mg_max_coarse=4096,
                 group_max=8, mg_workers=1, group_rows=0):
        self.precision = precision
        self.keep = keep_solutions
        o = gsx.ServerOpts(device=device, precision=precision, budget_bytes=int(budget_bytes),
                           iters_per_step=iters_per_step, capacity=capacity, mg_min_free=int(mg_min_dofs),
                           mg_max_coarse=int(mg_max_coarse), elem_budget=int(elem_budget), group_max=group_max,
                           mg_workers=mg_workers, group_rows=int(group_rows))
        h = C.c_void_p()
        _check_server(None, gsx.lib().gsx_server_create(C.byref(o), C.byref(h)))
        self.h = h
        # tile shapes from the autotuner's table (box members' classes)
        from . import autotune
        ctx = gsx.Context(device)
        table = autotune.Table.load(ctx.arch())
        ctx.close()
        pname = "mixed" if precision == gsx.MIXED else "fp64"
        bounds, lo = {}, 0
        for name, lim in autotune.BUCKETS:
            bounds[name] = (lo, lim if lim is not None else 1 << 62)
            lo = lim if lim is not None else lo
        for key, e in table.entries.items():
            pr, d, p, bname = key.split("-")
            b = e["best"]
            if pr != pname or b["kind"] != "box":
                continue
            lo_, hi_ = bounds[bname]
            gsx.lib().gsx_server_set_tile(self.h, int(d[1:]), int(p[1:]), lo_, hi_, b["lanes"], b["half"])
        self.results = []
        self._inflight = {}     # native rid -> (request, host buffers)

    def now(self):
        clk = C.c_double()
        _check_server(self, gsx.lib().gsx_server_status(self.h, None, None, C.byref(clk)))
        return clk.value

    def submit(self, req):
        """Queue a request; its arrival is req.arrival (server clock)."""
        P = req.problem.patches[0]
        if len(req.problem.patches) != 1:
            raise ValueError("the server takes single-patch requests")
        keep = []
        desc = gsx.PatchDesc()
        gsx.fill_patch_desc(desc, P, keep)
        nfree = int(np.prod([nb - 2 for nb in P.nbasis_dir()]))
        nrhs = len(req.rhs_seeds)
        rhs = [None] + [np.ascontiguousarray(rhs_vector(nfree, s)) for s in req.rhs_seeds[1:]]
        xs = [np.empty(nfree) for _ in range(nrhs)]
        its = np.zeros(nrhs, dtype=np.int32)
        rels = np.zeros(nrhs)
        sts = np.zeros(nrhs, dtype=np.int32)
        rhs_p = (gsx._dp * nrhs)(*[gsx._dp() if a is None else a.ctypes.data_as(gsx._dp) for a in rhs])
        x_p = (gsx._dp * nrhs)(*[a.ctypes.data_as(gsx._dp) for a in xs])
        r = gsx.ServerRequest(patch=C.pointer(desc), nrhs=nrhs, rhs_host=rhs_p, x_host=x_p, tol=req.tol,
                              maxit=req.maxit, iters_host=its.ctypes.data_as(gsx._i32p),
                              relres_host=rels.ctypes.data_as(gsx._dp), status_host=sts.ctypes.data_as(gsx._i32p))
        rid = C.c_int64()
        _check_server(self, gsx.lib().gsx_server_submit(self.h, C.byref(r), C.byref(rid)))
        self._inflight[rid.value] = (req, xs, its, rels, sts, keep, rhs, rhs_p, x_p)

    def poll(self):
        """Collect the requests that finished; returns the number still
        outstanding."""
        buf = (gsx.ServerResult * 64)()
        n = C.c_int32()
        while True:
            _check_server(self, gsx.lib().gsx_server_poll(self.h, 64, buf, C.byref(n)))
            for k in range(n.value):
                r = buf[k]
                req, xs, its, rels, sts, *_ = self._inflight.pop(r.rid)
                failed = r.status == gsx.SERVER_FAILED
                self.results.append(Result(
                    rid=req.rid, dofs=int(r.nbasis), nfree=int(r.nfree), nrhs=int(r.nrhs), arrival=req.arrival,
                    admitted=r.t_admit, assembled=r.t_assembled if r.t_assembled > 0 else r.t_done,
                    done=r.t_done, iterations=[int(v) for v in its], relres=[float(v) for v in rels],
                    status=[FAILED] * int(r.nrhs) if failed else [int(v) for v in sts],
                    solutions=None if failed or not (self.keep or req.keep) else xs,
                    lane="mg" if r.lane == 1 else "batch", error=r.error.decode() if failed else None))
            if n.value < 64:
                break
        out = C.c_int64()
        _check_server(self, gsx.lib().gsx_server_status(self.h, C.byref(out), None, None))
        return out.value

    def idle(self):
        """No request submitted and unfinished."""
        return self.poll() == 0

    @property
    def _used(self):
        used = C.c_int64()
        _check_server(self, gsx.lib().gsx_server_status(self.h, None, C.byref(used), None))
        return used.value

    def drain(self, timeout=600.0):
        t_end = time.perf_counter() + timeout
        while self.poll() > 0:
            if time.perf_counter() > t_end:
                raise RuntimeError("server did not drain")
            gsx.lib().gsx_server_drain(self.h, 0.05)
        self.poll()

    def close(self):
        if self.h:
            gsx.lib().gsx_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _check_server(srv, st):
    if st != 0:
        msg = ""
        if srv is not None and srv.h:
            msg = (gsx.lib().gsx_server_error(srv.h) or b"").decode()
        raise gsx.GsxError(st, msg)


def dyadic_round(n):
    """n rounded to the nearest m 2^k with 8 <= m < 16 (n itself below 16):
    at most ~6% away, and such a patch coarsens dyadically k times."""
    if n < 16:
        return n
    k = int(np.floor(np.log2(n))) - 3
    return int(round(n / 2 ** k)) * 2 ** k


def make_requests(count, seed=20261202, size_range=(1e3, 2e6), nrhs_range=(1, 8), noise=0.1, sigma=0.5,
                  dyadic=True):
    """Requests of the serving workload: size log-uniform in size_range DOFs,
    d uniform in {2, 3}, p uniform in {2, 3, 4}, n = round(size^(1/d)) - p
    (at least 2; with dyadic, rounded by dyadic_round so that large requests
    admit a multigrid hierarchy), seeded geometry noise, lognormal kappa,
    random per-element sources, 1..8 right-hand sides."""
    out = []
    lo, hi = np.log(size_range[0]), np.log(size_range[1])
    for i in range(count):
        s = seed * 1000003 + i
        u = problems.uniform01(s, 60, np.arange(4))
        size = float(np.exp(lo + (hi - lo) * u[0]))
        d = 2 + int(u[1] * 2)
        p = 2 + int(u[2] * 3)
        n = max(2, int(round(size ** (1.0 / d))) - p)
        if dyadic:
            n = dyadic_round(n)
        nrhs = nrhs_range[0] + int(u[3] * (nrhs_range[1] - nrhs_range[0] + 1))
        P = problems.noisy_box_fast(d, p, n, noise, sigma, s)
        prob = problems.Problem([P], problems.all_sides(d), [], name=f"req{i}-d{d}-p{p}-n{n}")
        out.append(Request(rid=i, problem=prob, rhs_seeds=[s] + [s * 31 + k for k in range(1, nrhs)]))
    return out
