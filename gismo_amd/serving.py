"""Continuous-batching solve server on one GPU.

Requests (a single-patch Poisson/diffusion problem plus 1..8 right-hand
sides) arrive at any time and are admitted in arrival order while their
device memory fits the HBM budget.  Two lanes, each with its own context
(its own HIP stream), run beside the solver thread's batch:

- batch lane: an assembler thread assembles each request and hands the
  resident operator to the solver thread, which adds one batched-CG member
  per right-hand side to the running `gsx_batch` (box members on the
  request's structured operator, with the autotuner's tile shape) and
  advances every member together, a few iterations per step; members join
  and leave between steps;
- multigrid lane: requests of at least `mg_min_dofs` free DOFs whose patch
  coarsens dyadically to a small coarsest level go to a worker that
  assembles the operator, builds its multigrid hierarchy once, and solves
  every right-hand side by MG-preconditioned CG with the hierarchy resident.

A request is complete when all its right-hand sides have converged and its
solutions are back on the host; its operator, hierarchy and workspace are
then released and its budget freed.

Every member's iterates are independent of the other members (the batched
CG reduces per member in a fixed order; the multigrid lane solves one
request at a time), so a request's solutions are bitwise the same alone and
in any mix.

The reference has no serving layer: it assembles and solves one problem at
a time (src/gsAssembler/gsExprAssembler.h:754-833,
src/gsSolver/gsConjugateGradient.hpp:20-93); each member runs that
algorithm.
"""
import queue
import threading
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


def estimate_bytes(prob, nrhs, mixed):
    """Device bytes a request keeps resident: CSR pattern + values (+ FP32
    copy), the structured operator (FP64 slots, FP32 upper half), the load
    vector, and per right-hand side the solution, the rhs and the CG
    workspace."""
    P = prob.patches[0]
    nfree = int(np.prod([nb - 2 for nb in P.nbasis_dir()]))
    slots = (2 * P.p[0] + 1) ** P.d
    nnz = nfree * slots
    tiles = (nfree + 255) // 256 * 256
    stencil = tiles * slots * 8 + (tiles * (slots - slots // 2) * 4 if mixed else 0)
    per_rhs = nfree * (8 * 6 + (4 * 6 if mixed else 0))
    return int(nnz * (4 + 8 + (4 if mixed else 0)) + stencil + nfree * 8 * 3 + nrhs * per_rhs + (64 << 20))


def box_of(prob):
    """(d, p, m0, m1, m2): the free functions of a single patch with
    Dirichlet sides (the box of a box member)."""
    P = prob.patches[0]
    m = [nb - 2 for nb in P.nbasis_dir()] + [1]
    return (P.d, P.p[0], m[0], m[1], m[2])


def mg_plan(prob, max_coarse=4096):
    """Levels of a dyadic multigrid hierarchy for a single-patch problem, or
    0: elements per direction n = m 2^k with the coarsest (m elements per
    direction) at most max_coarse free DOFs, and element-level Galerkin
    coarsening available for (d, p) (d = 2: p <= 4; d = 3: p <= 3)."""
    P = prob.patches[0]
    d, p = P.d, P.p[0]
    if len(prob.patches) != 1 or (d == 3 and p > 3) or p > 4:
        return 0
    ne = P.nelem_dir()
    if len(set(ne)) != 1:
        return 0
    n = ne[0]
    levels = 1
    while n % 2 == 0 and n // 2 >= 2:
        if max(0, n + p - 2) ** d <= max_coarse:
            break
        n //= 2
        levels += 1
    if levels < 2 or max(0, n + p - 2) ** d > max_coarse:
        return 0
    return levels


def mg_element_bytes(prob):
    """Bytes of the fine element matrices the hierarchy is built from."""
    P = prob.patches[0]
    nb = (P.p[0] + 1) ** P.d
    return P.nelem * nb * nb * 8


def rhs_vector(nfree, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, 1.0, nfree)


# per right-hand side status of a Result: 1 converged, 2 iteration limit
# reached (the batched CG's own codes), FAILED: the request raised (for
# example a folded patch, GSX_ERR_JACOBIAN) and was dropped
FAILED = 3


class _Entry:
    def __init__(self, req, t_admit):
        self.req = req
        self.t_admit = t_admit
        self.t_asm = None
        self.gp = None
        self.A = None
        self.bs = []
        self.xs = []
        self.ids = []
        self.bytes = 0
        self.levels = 0


class Server:
    """One GPU, two contexts: one assembles, one solves."""

    def __init__(self, precision=gsx.MIXED, budget_bytes=160 << 30, iters_per_step=8, capacity=4096,
                 keep_solutions=False, device=0, mg_min_dofs=20000, elem_budget=16 << 30):
        self.precision = precision
        self.mixed = precision == gsx.MIXED
        self.budget = budget_bytes
        self.iters = iters_per_step
        self.keep = keep_solutions
        self.mg_min_dofs = mg_min_dofs
        self.elem_budget = elem_budget
        self.ctx_asm = gsx.Context(device)
        self.ctx = gsx.Context(device)
        self.ctx_mg = gsx.Context(device)
        self._mgq = queue.Queue()          # admitted multigrid-lane requests
        self.batch = gsx.Batch(self.ctx, precision, capacity=capacity)
        from . import autotune
        self.tiles = autotune.Table.load(self.ctx.arch())   # SpMV tile shape per member class
        self.capacity = capacity
        self.t0 = time.perf_counter()
        self._incoming = queue.Queue()     # submitted requests
        self._ready = queue.Queue()        # assembled entries
        self._lock = threading.Lock()
        self._used = 0                     # admitted bytes
        self._members = 0                  # batch slots in use (solver thread)
        self._outstanding = 0              # submitted and not yet finished
        self._stop = False
        self._error = None
        self.results = []
        self._active = {}                  # rid -> entry (solver thread)
        self._asm_thread = threading.Thread(target=self._assembler, daemon=True)
        self._asm_thread.start()
        self._mg_thread = threading.Thread(target=self._mg_worker, daemon=True)
        self._mg_thread.start()

    def now(self):
        return time.perf_counter() - self.t0

    def submit(self, req):
        with self._lock:
            self._outstanding += 1
        self._incoming.put(req)

    # ----- assembler thread --------------------------------------------
    def _assembler(self):
        pending = []
        try:
            while not self._stop:
                try:
                    pending.append(self._incoming.get(timeout=0.002))
                except queue.Empty:
                    pass
                while pending:
                    req = pending[0]
                    levels = self._mg_levels(req)
                    need = estimate_bytes(req.problem, len(req.rhs_seeds), self.mixed)
                    if levels:
                        need += mg_element_bytes(req.problem) + need // 2
                    with self._lock:
                        if self._used + need > self.budget and self._used > 0:
                            break
                        self._used += need
                    pending.pop(0)
                    e = _Entry(req, self.now())
                    e.bytes = need
                    if levels:
                        e.levels = levels
                        self._mgq.put(e)
                        continue
                    try:
                        e.gp = gsx.Problem(self.ctx_asm, req.problem)
                        # the assembly fills the structured operator the
                        # request's batch members iterate on
                        e.A, b0 = e.gp.assemble(gsx.PATTERN_EXPR, self.precision)
                        e.bs.append(b0)
                        for s in req.rhs_seeds[1:]:
                            e.bs.append(self.ctx_asm.to_device(rhs_vector(e.gp.nfree, s)))
                        self.ctx_asm.synchronize()
                    except gsx.GsxError as ex:
                        self._fail(e, ex, "batch")
                        continue
                    e.t_asm = self.now()
                    self._ready.put(e)
        except Exception as ex:   # not a request's own error: reported by poll()
            self._error = ex

    def _release(self, e, objs):
        for o in objs:
            if o is not None:
                try:
                    o.close()
                except Exception:
                    pass

    def _fail(self, e, ex, lane, objs=()):
        """A request that raised: free what it holds, return its budget and
        record a failed result; the server keeps serving the others."""
        self._release(e, list(objs) + list(e.xs) + list(e.bs) + [e.A, e.gp])
        e.xs, e.bs, e.A, e.gp = [], [], None, None
        req = e.req
        res = Result(rid=req.rid, dofs=0, nfree=0, nrhs=len(req.rhs_seeds), arrival=req.arrival,
                     admitted=e.t_admit, assembled=self.now(), done=self.now(), iterations=[0] * len(req.rhs_seeds),
                     relres=[float("nan")] * len(req.rhs_seeds), status=[FAILED] * len(req.rhs_seeds),
                     lane=lane, error=str(ex))
        with self._lock:
            self.results.append(res)
            self._used -= e.bytes
            self._outstanding -= 1

    def _mg_levels(self, req):
        P = req.problem.patches[0]
        nfree = int(np.prod([nb - 2 for nb in P.nbasis_dir()]))
        if nfree < self.mg_min_dofs or mg_element_bytes(req.problem) > self.elem_budget:
            return 0
        return mg_plan(req.problem)

    # ----- multigrid lane -------------------------------------------------
    def _mg_worker(self):
        """Assemble, build the hierarchy once, solve every right-hand side by
        MG-PCG (V(2,2), damped Jacobi 0.6, exact coarse solve), release."""
        try:
            while not self._stop:
                try:
                    e = self._mgq.get(timeout=0.002)
                except queue.Empty:
                    continue
                req, ctx = e.req, self.ctx_mg
                gp = A = b0 = mg = b = x = None
                try:
                    gp = gsx.Problem(ctx, req.problem)
                    A, b0 = gp.assemble(gsx.PATTERN_EXPR | gsx.ASM_KEEP_ELEMENTS, self.precision)
                    mg = gsx.Multigrid(gp, A, e.levels, 2, 2, 0.6, precision=self.precision)
                    ctx.synchronize()
                    t_asm = self.now()
                    its, rels, sts, sols = [], [], [], []
                    for k, seed in enumerate(req.rhs_seeds):
                        b = b0 if k == 0 else ctx.to_device(rhs_vector(gp.nfree, seed))
                        x = ctx.array(gp.nfree).zero()
                        it, rel, ok = gsx.mg_cg_solve(A, mg, b, x, tol=req.tol, maxit=req.maxit)
                        its.append(it)
                        rels.append(rel)
                        sts.append(1 if ok else 2)
                        sols.append(x.download())
                        x.close()
                        x = None
                        if k:
                            b.close()
                        b = None
                except gsx.GsxError as ex:
                    self._fail(e, ex, "mg", objs=[x, b if b is not b0 else None, mg, b0, A, gp])
                    continue
                t_done = self.now()
                res = Result(rid=req.rid, dofs=gp.nglued, nfree=gp.nfree, nrhs=len(req.rhs_seeds),
                             arrival=req.arrival, admitted=e.t_admit, assembled=t_asm, done=t_done,
                             iterations=its, relres=rels, status=sts,
                             solutions=sols if (self.keep or req.keep) else None, lane="mg")
                mg.close()
                A.close()
                b0.close()
                gp.close()
                with self._lock:
                    self.results.append(res)
                    self._used -= e.bytes
                    self._outstanding -= 1
        except Exception as ex:   # reported by the solver thread
            self._error = ex

    # ----- solver thread (the caller) -----------------------------------
    def poll(self):
        """Admit assembled requests into the batch, run one step, collect the
        requests that finished.  Returns the number of running members."""
        if self._error is not None:
            raise self._error
        while self._members < self.capacity:
            try:
                e = self._ready.get_nowait()
            except queue.Empty:
                break
            try:
                box = box_of(e.req.problem)
                kind, lanes, half = self.tiles.lookup(self.mixed, box[0], box[1], e.gp.nfree)
                for b in e.bs:
                    x = self.ctx.array(e.gp.nfree).zero()
                    e.xs.append(x)
                    if kind == "box":
                        e.ids.append(self.batch.add(e.A, b, x, tol=e.req.tol, maxit=e.req.maxit, box=box,
                                                    tile=(lanes, half)))
                    else:
                        e.ids.append(self.batch.add(e.A, b, x, tol=e.req.tol, maxit=e.req.maxit, tile=(lanes, 1)))
                    self._members += 1
            except gsx.GsxError as ex:
                for i in e.ids:
                    self.batch.remove(i)
                    self._members -= 1
                e.ids = []
                self._fail(e, ex, "batch")
                continue
            self._active[e.req.rid] = e
        running = self.batch.step(self.iters) if self._active else 0
        for rid in list(self._active):
            e = self._active[rid]
            res = [self.batch.result(i) for i in e.ids]
            if any(r[0] == 0 for r in res):
                continue
            sols = [x.download() for x in e.xs]   # results go back to the host
            t_done = self.now()
            for i in e.ids:
                self.batch.remove(i)
                self._members -= 1
            result = Result(
                rid=rid, dofs=e.gp.nglued, nfree=e.gp.nfree, nrhs=len(e.ids), arrival=e.req.arrival,
                admitted=e.t_admit, assembled=e.t_asm, done=t_done,
                iterations=[r[1] for r in res], relres=[r[2] for r in res], status=[r[0] for r in res],
                solutions=sols if (self.keep or e.req.keep) else None, lane="batch")
            for x in e.xs:
                x.close()
            e.A.close()
            for b in e.bs:
                b.close()
            e.gp.close()
            del self._active[rid]
            with self._lock:
                self.results.append(result)
                self._used -= e.bytes
                self._outstanding -= 1
        return running

    def idle(self):
        """No request submitted and unfinished (queued, being assembled,
        waiting for a batch slot or solving)."""
        with self._lock:
            return self._outstanding == 0

    def drain(self, timeout=600.0):
        t_end = time.perf_counter() + timeout
        while not self.idle():
            self.poll()
            if time.perf_counter() > t_end:
                raise RuntimeError("server did not drain")

    def close(self):
        self._stop = True
        self._asm_thread.join()
        self._mg_thread.join()
        self.batch.close()
        self.ctx.close()
        self.ctx_asm.close()
        self.ctx_mg.close()


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
