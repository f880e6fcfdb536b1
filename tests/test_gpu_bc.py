"""Boundary data on the GPU (gsx_problem_set_bc, gsx_assemble_bdr): the
Dirichlet values by L2 projection, the right-hand side with the fixed
columns moved over (single-patch structured gather and multipatch gather)
and the Neumann load against the oracle; the affine patch test and the
manufactured convergence order through the device path."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as O
from gismo_amd import gsx, problems

pytestmark = pytest.mark.gpu

AFFINE = (0.3, 1.1, -0.7, 0.5)


def _curved_patch(d, p, n, seed, lo=None, hi=None, fkind=0):
    P = problems.box_patch(d, p, n, lo, hi)
    rng = np.random.default_rng(seed)
    P.ctrl = P.ctrl + rng.uniform(-0.1 / n, 0.1 / n, P.ctrl.shape)
    P.fkind = fkind
    if fkind == 0:
        P.felem = rng.uniform(-1.0, 1.0, P.nelem)
    return P


def _cases():
    out = []
    # single patch, mixed Dirichlet (affine / manufactured) and Neumann
    P = _curved_patch(2, 3, 6, seed=1)
    out.append(("patch2d-p3", problems.Problem([P], [(0, 1), (0, 3)], []),
                [(0, 1, 1, 2, AFFINE), (0, 3, 1, 1, ()), (0, 2, 2, 1, ()), (0, 4, 2, 2, AFFINE)]))
    P = _curved_patch(3, 2, 4, seed=2)
    out.append(("patch3d-p2", problems.Problem([P], [(0, 1), (0, 4), (0, 6)], []),
                [(0, 1, 1, 2, AFFINE), (0, 4, 1, 1, ()), (0, 6, 1, 0, ()), (0, 2, 2, 1, ()), (0, 5, 2, 2, AFFINE)]))
    # multipatch (2 x 2 squares, glued): Dirichlet on the outer west/south
    # sides, Neumann on the outer east/north sides
    G = problems.square_grid(2, 2, n=4, p=2)
    dsides = [ps for ps in G.dir_sides if ps[1] in (1, 3)]
    nsides = [ps for ps in G.dir_sides if ps[1] in (2, 4)]
    prob = problems.Problem(G.patches, dsides, G.ifaces)
    bcs = [(pt, sd, 1, 2, AFFINE) for pt, sd in dsides] + [(pt, sd, 2, 1, ()) for pt, sd in nsides]
    out.append(("grid2x2-p2", prob, bcs))
    return out


CASES = _cases()


def _oracle(prob, bcs):
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    nglued = int(max(dofs)) + 1 if len(dofs) else 0
    nfix = nglued - nfree
    dbc = [b for b in bcs if b[2] == 1]
    # every eliminated side takes part in the projection (homogeneous ones too)
    allb = [(pt, sd, 1, 0, ()) for pt, sd in prob.dir_sides if not any(b[0] == pt and b[1] == sd for b in dbc)]
    order = {ps: i for i, ps in enumerate(prob.dir_sides)}
    drecs = sorted(dbc + allb, key=lambda b: order[(b[0], b[1])])
    fixed = O.dirichlet_values(prob.patches, dofs, nfree, nfix, drecs)
    A, b = O.assemble(prob.patches, dofs, nfree, 0, fixed=fixed)
    neu = O.neumann_load(prob.patches, dofs, nfree, bcs)
    return dofs, nfree, fixed, A, b, neu


@pytest.mark.parametrize("name,prob,bcs", CASES, ids=[c[0] for c in CASES])
def test_boundary_data_match_oracle(ctx, name, prob, bcs):
    print(f"[case] {name}", flush=True)
    dofs, nfree, fixed_o, Ao, bo, neu_o = _oracle(prob, bcs)
    gp = gsx.Problem(ctx, prob)
    gp.set_bc(bcs)
    fixed = gp.fixed_values()
    assert fixed.shape == fixed_o.shape
    assert np.abs(fixed - fixed_o).max() <= 1e-12 * max(1.0, np.abs(fixed_o).max())
    for prec in (gsx.FP64, gsx.MIXED):
        A, b = gp.assemble(gsx.PATTERN_EXPR, prec)
        rhs_vol = b.download()
        gp.assemble_bdr(b)
        rhs = b.download()
        scale = np.abs(bo).max()
        assert np.abs(rhs_vol - bo).max() <= 1e-12 * scale, prec
        assert np.abs((rhs - rhs_vol) - neu_o).max() <= 1e-12 * max(1.0, np.abs(neu_o).max()), prec
        As = A.to_scipy()
        assert np.array_equal(As.indptr, Ao.indptr) and np.array_equal(As.indices, Ao.indices)
        assert np.abs(As.data - Ao.data).max() <= 1e-12 * np.abs(Ao.data).max()
        A.close()
        b.close()
    gp.close()


@pytest.mark.parametrize("d,p,dir_sides", [(2, 3, [1, 3]), (3, 2, [2, 3, 5])])
def test_patch_test_on_device(ctx, d, p, dir_sides):
    """u = c + a . x with projected Dirichlet data and flux a . nv: the
    device solution is u's coefficients (c + a . C_m) to solver tolerance,
    on the structured and the CSR operator."""
    P = _curved_patch(d, p, 5, seed=9)
    P.felem = np.zeros(P.nelem)
    prob = problems.Problem([P], [(0, s) for s in dir_sides], [])
    neu = [s for s in range(1, 2 * d + 1) if s not in dir_sides]
    bcs = [(0, s, 1, 2, AFFINE) for s in dir_sides] + [(0, s, 2, 2, AFFINE) for s in neu]
    gp = gsx.Problem(ctx, prob)
    gp.set_bc(bcs)
    A, b = gp.assemble(gsx.PATTERN_EXPR, gsx.MIXED)
    gp.assemble_bdr(b)
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    vals = AFFINE[0] + P.ctrl @ np.array(AFFINE[1:1 + d])
    expect = np.zeros(nfree)
    for m, g in enumerate(dofs):
        if g < nfree:
            expect[g] = vals[m]
    for structured in (True, False):
        A.set_operator(structured)
        for prec in (gsx.FP64, gsx.MIXED):
            x = ctx.array(nfree).zero()
            it, rel, ok = A.cg_solve(b, x, tol=1e-13, maxit=5000, precision=prec)
            assert ok, (structured, prec, it, rel)
            assert np.abs(x.download() - expect).max() <= 1e-9 * np.abs(expect).max(), (structured, prec)
            x.close()
    A.close()
    b.close()
    gp.close()


def test_manufactured_mixed_bc_convergence_on_device(ctx):
    """sin(pi x) sin(pi y) on [0.25, 1.25]^2 with projected Dirichlet data on
    west/south and the flux on east/north: the device L2 error (fixed values
    included) equals the oracle's and falls at order p + 1 = 3."""
    errs = []
    for n in (4, 8, 16):
        P = problems.box_patch(2, 2, n, [0.25, 0.25], [1.25, 1.25])
        P.fkind = 1
        prob = problems.Problem([P], [(0, 1), (0, 3)], [])
        bcs = [(0, 1, 1, 1, ()), (0, 3, 1, 1, ()), (0, 2, 2, 1, ()), (0, 4, 2, 1, ())]
        gp = gsx.Problem(ctx, prob)
        gp.set_bc(bcs)
        A, b = gp.assemble()
        gp.assemble_bdr(b)
        x = ctx.array(gp.nfree).zero()
        A.cg_solve(b, x, tol=1e-13, maxit=5000)
        e = gp.l2_error(x)
        dofs, nfree, fixed_o, Ao, bo, neu_o = _oracle(prob, bcs)
        xo = spla.spsolve(Ao.tocsc(), bo + neu_o)
        eo = O.l2_error(prob.patches, dofs, nfree, xo, fixed=fixed_o)
        assert abs(e - eo) <= 1e-8 * eo, (n, e, eo)
        errs.append(e)
        for o in (x, A, b, gp):
            o.close()
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    print(errs, rates, flush=True)
    assert rates[-1] > 2.8, rates


def test_projection_of_inhomogeneous_sides_only(ctx):
    """DIR_L2PROJECTION_INHOM (gsAssembler::computeDirichletDofs): only the
    sides with nonzero data take part in the projection, so the corner
    functions shared with a homogeneous side keep the projected data of the
    other side, and functions on homogeneous sides alone stay 0.  Equals the
    oracle's projection over the inhomogeneous records; differs from the
    all-sides projection at the shared corners."""
    P = _curved_patch(2, 3, 6, seed=4)
    prob = problems.Problem([P], [(0, 1), (0, 2), (0, 3)], [])
    bcs = [(0, 1, 1, 2, AFFINE), (0, 2, 1, 1, ()), (0, 3, 1, 0, ()), (0, 4, 2, 1, ())]
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    nfix = int(max(dofs)) + 1 - nfree
    want = O.dirichlet_values(prob.patches, dofs, nfree, nfix, [b for b in bcs if b[2] == 1 and b[3] != 0])
    gp = gsx.Problem(ctx, prob)
    gp.set_bc(bcs, gsx.DIR_L2PROJECTION_INHOM)
    got = gp.fixed_values()
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    gp.set_bc(bcs, gsx.DIR_L2PROJECTION)
    every = gp.fixed_values()
    assert np.abs(every - want).max() > 1e-6
    gp.close()



@pytest.mark.parametrize("name,prob,bcs", CASES, ids=[c[0] for c in CASES])
def test_dirichlet_interpolation_matches_oracle(ctx, name, prob, bcs):
    """GSX_DIR_INTERPOLATION (gsAssembler::computeDirichletDofsIntpl): the
    fixed values equal the oracle's interpolation over the problem's
    Dirichlet sides in order (homogeneous sides write zeros)."""
    print(f"[case] interpolation {name}", flush=True)
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    nfix = int(max(dofs)) + 1 - nfree
    dbc = {(b[0], b[1]): b for b in bcs if b[2] == 1}
    recs = [dbc.get((pt, sd), (pt, sd, 1, 0, ())) for pt, sd in prob.dir_sides]
    want = O.dirichlet_interpolation(prob.patches, dofs, nfree, nfix, recs)
    gp = gsx.Problem(ctx, prob)
    gp.set_bc(bcs, gsx.DIR_INTERPOLATION)
    got = gp.fixed_values()
    assert np.abs(got - want).max() <= 1e-12 * max(1.0, np.abs(want).max())
    A, b = gp.assemble()
    Ao, bo = O.assemble(prob.patches, dofs, nfree, 0, fixed=want)
    assert np.abs(b.download() - bo).max() <= 1e-12 * np.abs(bo).max()
    for o in (A, b, gp):
        o.close()
