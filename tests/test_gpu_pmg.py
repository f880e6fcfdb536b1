"""p-multigrid on the device (gsx_problem_create_geo, gsx_pmg_create)
against the oracle (oracle/pmg.py, pinned on the CPU by
tests/test_oracle_pmg.py): level operators assembled on a level basis over
the fine geometry, mixed-mass prolongations P = L_f^-1 M and restrictions
R = L_c^-1 M^T, knot-insertion transfers and Galerkin operators of h
transitions, and MG-PCG iterations and solutions.  Values agree with the
oracle to rounding (1e-12 of the largest entry); the sums run in another
order."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
import oracle.pmg as PM
from gismo_amd import gsx, problems

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _rel(A, B):
    A = A.toarray() if sp.issparse(A) else np.asarray(A)
    B = B.toarray() if sp.issparse(B) else np.asarray(B)
    assert A.shape == B.shape, (A.shape, B.shape)
    return np.abs(A - B).max() / np.abs(B).max()


def _noisy(d, p, n, seed, kappa=True):
    P = problems.box_patch(d, p, n)
    problems.perturb_interior(P, 0.15 / n, seed)
    if kappa:
        P.kappa = problems.lognormal_kappa(P.nelem, 0.5, seed)
    P.fkind = 1
    return problems.Problem([P], problems.all_sides(d), [], name=f"noisy{d}d-p{p}-n{n}")


@pytest.mark.parametrize("d,pg,ng,q,nq", [(2, 3, 6, 2, 6), (2, 3, 6, 1, 3), (2, 4, 4, 2, 2), (3, 2, 4, 1, 2),
                                         (3, 3, 4, 2, 4)])
def test_level_assembly_matches_oracle(ctx, d, pg, ng, q, nq):
    """A level basis (degree q, nq elements) over a curved fine patch with
    kappa (degree pg, ng elements): matrix and load as the oracle's."""
    prob = _noisy(d, pg, ng, 40 + q)
    print(f"[case] geo d={d} pg={pg} ng={ng} q={q} nq={nq}", flush=True)
    G = prob.patches[0]
    B = PM.Basis(d, [q] * d, [PM.open_knots(problems.knots_uniform(nq, 1), q) for _ in range(d)])
    dofs, nfree = O.mapper([B], prob.dir_sides, [])
    Ko, Fo = PM.assemble_geo([B], [G], dofs, nfree)
    gp = gsx.Problem(ctx, prob)
    lp = gsx.Problem.with_geometry(gp, [B])
    assert lp.nfree == nfree and np.array_equal(lp.dofs(), dofs)
    A, b = lp.assemble(gsx.PATTERN_FULL | gsx.ASM_CSR_ONLY, gsx.FP64)
    K = A.to_scipy()
    assert _rel(K, Ko) <= TOL
    assert _rel(b.download()[:nfree], Fo) <= TOL
    assert (K != K.T).nnz == 0
    # the same through slabs of two element layers (a budget of q + 2
    # layers of element matrices: the slab recomputes its q shared layers)
    nb = (q + 1) ** d
    ctx.set_elem_budget(nq ** (d - 1) * (nb * nb + nb) * 8 * (q + 2))
    A2, b2 = lp.assemble(gsx.PATTERN_FULL | gsx.ASM_CSR_ONLY, gsx.FP64)
    ctx.set_elem_budget(16 << 30)
    assert np.array_equal(A2.to_scipy().toarray(), K.toarray())
    for o in (A, b, A2, b2, lp, gp):
        o.close()


def _check_hierarchy(ctx, prob, nlevels, coarsening, projection, coarse_operator):
    gp = gsx.Problem(ctx, prob)
    A, b = gp.assemble(gsx.PATTERN_FULL, gsx.FP64)
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    Ao, bo = O.assemble(prob.patches, dofs, nfree, 1)
    As, Ps, Rs, levels, maps = PM.pmg_setup(prob.patches, prob.dir_sides, prob.ifaces, Ao, nlevels, coarsening,
                                            projection, coarse_operator)
    mg = gsx.Multigrid.pmg(gp, A, nlevels, coarsening, projection, coarse_operator, damping=0.6)
    for l in range(nlevels - 1):
        Al, Pl = mg.level(l)
        Rl = mg.restriction(l)
        print(f"  level {l}: n={Al.shape[0]} A {_rel(Al, As[l]):.2e} P {_rel(Pl, Ps[l]):.2e} "
              f"R {_rel(Rl, Rs[l]):.2e}", flush=True)
        assert _rel(Al, As[l]) <= TOL, l
        assert _rel(Pl, Ps[l]) <= TOL, l
        assert _rel(Rl, Rs[l]) <= TOL, l
    return gp, A, b, mg, (As, Ps, Rs), Ao, bo


HIER = [
    ("noisy2d-zz", lambda: _noisy(2, 3, 8, 7), 3, "zz", 2, 2),
    ("noisy2d-hz-galerkin", lambda: _noisy(2, 3, 8, 7), 3, "hz", 2, 2),
    ("noisy2d-hz-rediscretised", lambda: _noisy(2, 3, 8, 7), 3, "hz", 2, 1),
    ("noisy2d-p-to-1", lambda: _noisy(2, 3, 6, 8), 2, "p", 1, 2),
    ("noisy3d-z", lambda: _noisy(3, 2, 4, 9), 2, "z", 2, 2),
    ("grid2x2-z", lambda: problems.square_grid(2, 2, n=4, p=2), 2, "z", 2, 2),
]


@pytest.mark.parametrize("name,make,nlevels,coarsening,projection,coarse_operator", HIER,
                         ids=[h[0] for h in HIER])
def test_pmg_hierarchy_matches_oracle(ctx, name, make, nlevels, coarsening, projection, coarse_operator):
    print(f"[case] pmg {name}", flush=True)
    gp, A, b, mg, _, _, _ = _check_hierarchy(ctx, make(), nlevels, coarsening, projection, coarse_operator)
    for o in (mg, A, b, gp):
        o.close()


@pytest.mark.parametrize("precision", [gsx.FP64, gsx.MIXED])
def test_pmg_pcg_matches_oracle(ctx, precision):
    """MG-PCG with the p-multigrid cycle (degree 3 to 1 by z coarsening,
    damped Jacobi): the oracle's iterations (FP64: to one; mixed: within
    two) and solution."""
    prob = _noisy(2, 3, 16, 11)
    print(f"[case] pmg pcg precision={precision}", flush=True)
    gp = gsx.Problem(ctx, prob)
    A, b = gp.assemble(gsx.PATTERN_FULL, precision)
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    Ao, bo = O.assemble(prob.patches, dofs, nfree, 1)
    As, Ps, Rs, _, _ = PM.pmg_setup(prob.patches, prob.dir_sides, [], Ao, 3, "zz")
    cyc = PM.vcycle(As, Ps, Rs, PM.jacobi_smoothers(As, 0.6))
    xo, ito, erro = O.pcg(Ao, bo, cyc, tol=1e-10, maxit=200)
    mg = gsx.Multigrid.pmg(gp, A, 3, "zz", damping=0.6, precision=precision)
    x = ctx.array(nfree).zero()
    it, rel, ok = gsx.mg_cg_solve(A, mg, b, x, tol=1e-10, maxit=200)
    print(f"  device {it} iterations ({rel:.2e}), oracle {ito} ({erro:.2e})", flush=True)
    assert ok
    assert abs(it - ito) <= (1 if precision == gsx.FP64 else 2), (it, ito)
    xs = x.download()
    assert np.linalg.norm(xs - xo) <= 1e-8 * np.linalg.norm(xo)
    for o in (x, mg, A, b, gp):
        o.close()
