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


def _serve(rs, budget=160 << 30, precision=gsx.MIXED, stagger=0.0, mg_min_dofs=20000, group_max=1):
    # group_rows 1: the small test operators group too
    srv = serving.Server(precision=precision, budget_bytes=budget, iters_per_step=8, mg_min_dofs=mg_min_dofs,
                         group_max=group_max, group_rows=1)
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


def _fold(req):
    """Move one interior control point far across its neighbours: the map
    folds and assembly raises GSX_ERR_JACOBIAN."""
    P = req.problem.patches[0]
    nb = P.nbasis_dir()
    i = 2 + nb[0] * (2 + (nb[1] * 2 if P.d == 3 else 0))
    P.ctrl = P.ctrl.copy()
    P.ctrl[i, 0] += 3.0
    return req


@pytest.mark.parametrize("mg_min_dofs", [20000, 400])
def test_serving_survives_a_failing_request(mg_min_dofs):
    """A request that raises (a folded patch) is reported as failed with the
    library's message and its budget returned; the requests around it, in
    the batch lane and in the multigrid lane, still complete with the same
    solutions as without it."""
    rs = _requests(count=6, seed=11)
    ref = _serve(rs, mg_min_dofs=mg_min_dofs)
    bad = [_fold(r) for r in _requests(count=6, seed=11)]
    mixed = [rs[0], bad[1], rs[2], bad[3], rs[4], rs[5]]
    srv = serving.Server(precision=gsx.MIXED, iters_per_step=8, mg_min_dofs=mg_min_dofs)
    for r in mixed:
        r.arrival = srv.now()
        srv.submit(r)
    srv.drain(timeout=120)
    got = {r.rid: r for r in srv.results}
    assert srv._used == 0
    srv.close()
    assert sorted(got) == list(range(6))
    for rid in (1, 3):
        assert got[rid].status == [serving.FAILED] * got[rid].nrhs and "Jacobian" in got[rid].error
    for rid in (0, 2, 4, 5):
        assert got[rid].error is None and all(s == 1 for s in got[rid].status)
        for a, b in zip(got[rid].solutions, ref[rid].solutions):
            assert np.array_equal(a, b), rid


def test_serving_shared_spmv_groups_give_the_same_solutions():
    """With the right-hand sides of a request sharing SpMV launches
    (group_max 4) every solution is bitwise the one served without groups."""
    rs = _requests(count=6, seed=5)
    for r in rs:
        r.rhs_seeds = (r.rhs_seeds + [r.rhs_seeds[0] * 7 + k for k in range(1, 4)])[:4]
    plain = _serve(rs)
    grouped = _serve(rs, group_max=4)
    for r in rs:
        assert len(grouped[r.rid].solutions) == 4
        for a, b in zip(grouped[r.rid].solutions, plain[r.rid].solutions):
            assert np.array_equal(a, b), r.rid
