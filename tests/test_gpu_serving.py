"""The continuous-batching server (gismo_amd/serving.py) on the GPU: every
request's solutions are bitwise the same when served alone, in a mix that
arrives at once, and under a memory budget that admits one request at a
time; the solutions reach the FP64 oracle's."""
import time

import numpy as np
import pytest

import oracle as O
from gismo_amd import gsx, serving

pytestmark = pytest.mark.gpu


def _requests(count=8, seed=99):
    rs = serving.make_requests(count, seed=seed, size_range=(1e3, 3e4), nrhs_range=(1, 3))
    for r in rs:
        r.tol = 1e-12
        r.keep = True
    return rs


def _serve(rs, budget=160 << 30, precision=gsx.MIXED, stagger=0.0, mg_min_dofs=20000):
    srv = serving.Server(precision=precision, budget_bytes=budget, iters_per_step=8, mg_min_dofs=mg_min_dofs)
    for r in rs:
        r.arrival = srv.now()
        srv.submit(r)
        if stagger:
            time.sleep(stagger)
            srv.poll()
    srv.drain(timeout=120)
    res = {r.rid: r for r in srv.results}
    srv.close()
    return res


def test_serving_alone_equals_mix():
    rs = _requests()
    mix = _serve(rs)
    assert sorted(mix) == [r.rid for r in rs]
    for r in rs[:4]:
        alone = _serve([r])[r.rid]
        for a, b in zip(alone.solutions, mix[r.rid].solutions):
            assert np.array_equal(a, b), r.rid
    # staggered arrivals and a budget that admits one request at a time
    one = _serve(rs, budget=1, stagger=0.002)
    for r in rs:
        res = one[r.rid]
        assert all(s == 1 for s in res.status)
        for a, b in zip(res.solutions, mix[r.rid].solutions):
            assert np.array_equal(a, b), r.rid


def test_serving_solutions_match_oracle():
    rs = _requests(count=4, seed=7)
    res = _serve(rs)
    for r in rs:
        P = r.problem
        dofs, nfree = O.mapper(P.patches, P.dir_sides, P.ifaces)
        Ao, bo = O.assemble(P.patches, dofs, nfree, 0)
        for k, seed in enumerate(r.rhs_seeds):
            b = bo if k == 0 else serving.rhs_vector(nfree, seed)
            xo, _, _ = O.cg_solve(Ao, b, tol=1e-14, maxit=20000)
            x = res[r.rid].solutions[k]
            err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
            print(P.name, k, res[r.rid].iterations[k], err, flush=True)
            assert err <= 1e-8
        assert res[r.rid].latency > 0


def test_serving_multigrid_lane():
    """Requests that coarsen dyadically go to the multigrid lane (hierarchy
    built once, every right-hand side solved by MG-PCG with it resident);
    their solutions are bitwise the same alone and in a mix with batch-lane
    requests, and reach the FP64 oracle's."""
    rs = _requests(count=6, seed=11)
    mix = _serve(rs, mg_min_dofs=400)
    lanes = {rid: res.lane for rid, res in mix.items()}
    print(lanes, flush=True)
    assert "mg" in lanes.values()
    mg_reqs = [r for r in rs if lanes[r.rid] == "mg"]
    for r in mg_reqs[:2]:
        alone = _serve([r], mg_min_dofs=400)[r.rid]
        assert alone.lane == "mg"
        for a, b in zip(alone.solutions, mix[r.rid].solutions):
            assert np.array_equal(a, b), r.rid
    for r in mg_reqs[:2]:
        P = r.problem
        dofs, nfree = O.mapper(P.patches, P.dir_sides, P.ifaces)
        Ao, bo = O.assemble(P.patches, dofs, nfree, 0)
        for k, seed in enumerate(r.rhs_seeds):
            b = bo if k == 0 else serving.rhs_vector(nfree, seed)
            xo, _, _ = O.cg_solve(Ao, b, tol=1e-14, maxit=20000)
            x = mix[r.rid].solutions[k]
            err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
            print(P.name, k, mix[r.rid].iterations[k], err, flush=True)
            assert mix[r.rid].status[k] == 1 and err <= 1e-8
