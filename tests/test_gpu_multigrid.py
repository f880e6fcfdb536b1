"""GPU multigrid against the oracle: transfers, Galerkin level operators,
one cycle, and MG-preconditioned CG (FP64 and FP32 cycles)."""
import numpy as np
import pytest

import oracle as O
from gismo_amd import gsx, problems

pytestmark = pytest.mark.gpu


def _hierarchy(make, ns):
    levels = [make(n) for n in ns]
    maps = [O.mapper(p.patches, p.dir_sides, p.ifaces) for p in levels]
    return levels, maps


def _oracle_mg(levels, maps, npre=2, npost=2, damping=0.6):
    Af, bf = O.assemble(levels[-1].patches, maps[-1][0], maps[-1][1], 0)
    Ps = [O.transfer_multipatch(levels[i].patches, maps[i][0], maps[i][1], maps[i + 1][0], maps[i + 1][1])
          for i in range(len(levels) - 1)]
    return Af, bf, Ps, O.Multigrid(Af, Ps, npre=npre, npost=npost, damping=damping)


def _rowscaled(A, B):
    A = A.tocsr()
    B = B.tocsr()
    assert np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)
    err = 0.0
    for i in range(A.shape[0]):
        s, e = A.indptr[i], A.indptr[i + 1]
        if e > s:
            err = max(err, np.abs(A.data[s:e] - B.data[s:e]).max() / np.abs(B.data[s:e]).max())
    return err


CASES = {
    "grid2x2-p3": (lambda n: problems.square_grid(2, 2, n, 3), (2, 4, 8)),
    "grid3x2-p2": (lambda n: problems.square_grid(3, 2, n, 2), (2, 4, 8)),
    "cube-p2": (lambda n: problems.poisson_noisy_cube(n=n, p=2, seed=4), (2, 4, 8)),
    # the multigrid workload's layout (16 degree-3 patches tiling [0,4]^2)
    # on three levels: 4, 8, 16 elements per patch and direction
    "grid4x4-p3": (lambda n: problems.square_grid(4, 4, n, 3), (4, 8, 16)),
}


@pytest.mark.parametrize("name", list(CASES))
def test_transfers_and_galerkin_levels(ctx, name):
    make, ns = CASES[name]
    print(f"[case] {name}", flush=True)
    levels, maps = _hierarchy(make, ns)
    Af, bf, Ps, omg = _oracle_mg(levels, maps)
    gp = gsx.Problem(ctx, levels[-1])
    A, b = gp.assemble()
    mg = gsx.Multigrid(gp, A, len(ns), 2, 2, 0.6)
    for lvl in range(len(ns)):
        Ag, Pg = mg.level(lvl)
        Ao = omg.matrix(lvl)
        assert _rowscaled(Ag, Ao) <= 1e-12, lvl
        if Pg is not None:
            Po = Ps[lvl]
            Po.sort_indices()
            assert np.array_equal(Pg.indptr, Po.indptr) and np.array_equal(Pg.indices, Po.indices)
            assert np.array_equal(Pg.data, Po.data)
    # one cycle
    rng = np.random.default_rng(1)
    r = rng.standard_normal(gp.nfree)
    rd = ctx.to_device(r)
    zd = ctx.array(gp.nfree)
    mg.apply(rd, zd)
    zo = omg.apply(r)
    assert np.linalg.norm(zd.download() - zo) <= 1e-10 * np.linalg.norm(zo)
    mg.close()


@pytest.mark.parametrize("name,make", [
    ("grid2x2-n8-p3", lambda: problems.square_grid(2, 2, 8, 3)),       # 289 DOFs, 5 tiles of 64
    ("cube-n6-p2", lambda: problems.poisson_noisy_cube(n=6, p=2, seed=8)),   # 216 DOFs, 4 tiles
    ("grid2x2-n16-p2", lambda: problems.square_grid(2, 2, 16, 2)),     # 1,089 DOFs, 18 tiles
])
def test_coarse_inverse_multi_tile(ctx, name, make):
    """A one-level multigrid is the exact coarse solve: its apply is the dense
    inverse (blocked symmetric sweep over 64x64 tiles, trailing update on
    the FP64 matrix cores), checked against a sparse direct solve."""
    import scipy.sparse.linalg as spla
    prob = make()
    print(f"[case] {name}", flush=True)
    gp = gsx.Problem(ctx, prob)
    A, b = gp.assemble(gsx.PATTERN_FULL)
    As = A.to_scipy().tocsc()
    assert As.shape[0] > 128
    mg = gsx.Multigrid(gp, A, 1, 1, 1, 1.0)
    rng = np.random.default_rng(3)
    r = rng.standard_normal(gp.nfree)
    rd = ctx.to_device(r)
    zd = ctx.array(gp.nfree)
    mg.apply(rd, zd)
    z = zd.download()
    zs = spla.spsolve(As, r)
    assert np.linalg.norm(z - zs) <= 1e-10 * np.linalg.norm(zs)
    for o in (mg, rd, zd, A, b):
        o.close()


@pytest.mark.parametrize("name", ["grid2x2-p3", "cube-p2"])
def test_galerkin_routes_agree(ctx, name):
    """Coarse levels formed from element matrices (recomputed by the setup,
    or kept from the assembly) are bitwise equal, and equal to the global
    product P^T (A P) (taken for a matrix handed over as CSR) to 1e-12."""
    make, ns = CASES[name]
    print(f"[case] {name}", flush=True)
    levels, maps = _hierarchy(make, ns)
    gp = gsx.Problem(ctx, levels[-1])
    A1, b1 = gp.assemble(gsx.PATTERN_EXPR | gsx.ASM_KEEP_ELEMENTS)
    mg1 = gsx.Multigrid(gp, A1, len(ns), 2, 2, 0.6)
    A2, b2 = gp.assemble(gsx.PATTERN_EXPR)
    mg2 = gsx.Multigrid(gp, A2, len(ns), 2, 2, 0.6)
    rp, ci, va = A2.download()
    A3 = gsx.Matrix.from_csr(ctx, rp, ci, va)
    mg3 = gsx.Multigrid(gp, A3, len(ns), 2, 2, 0.6)
    for lvl in range(len(ns) - 1):
        L1, _ = mg1.level(lvl)
        L2, _ = mg2.level(lvl)
        L3, _ = mg3.level(lvl)
        assert np.array_equal(L1.indptr, L2.indptr) and np.array_equal(L1.indices, L2.indices)
        assert np.array_equal(L1.data, L2.data), lvl
        assert _rowscaled(L1, L3) <= 1e-12, lvl
    for o in (mg1, mg2, mg3, A1, A2, A3, b1, b2):
        o.close()


@pytest.mark.parametrize("name", ["grid2x2-p3", "cube-p2", "grid4x4-p3"])
def test_mg_pcg_matches_oracle(ctx, name):
    make, ns = CASES[name]
    levels, maps = _hierarchy(make, ns)
    Af, bf, Ps, omg = _oracle_mg(levels, maps)
    xo, ito, erro = O.cg_solve(Af, bf, tol=1e-10, maxit=200, precond="mg", mg=omg)
    gp = gsx.Problem(ctx, levels[-1])
    A, b = gp.assemble()
    for prec in (gsx.FP64, gsx.MIXED):
        mg = gsx.Multigrid(gp, A, len(ns), 2, 2, 0.6, precision=prec)
        x = ctx.array(gp.nfree).zero()
        it, rel, ok = gsx.mg_cg_solve(A, mg, b, x, tol=1e-10, maxit=200)
        assert ok and rel < 1e-10
        if prec == gsx.FP64:
            assert abs(it - ito) <= 1
        else:
            assert it <= ito + 3
        xs = x.download()
        assert np.linalg.norm(xs - xo) <= 1e-8 * np.linalg.norm(xo)
        mg.close()


@pytest.mark.parametrize("name,smoother", [("grid2x2-p3", "chebyshev"), ("cube-p2", "chebyshev"),
                                           ("grid3x2-p2", "richardson"), ("grid2x2-p3", "gauss_seidel"),
                                           ("cube-p2", "gauss_seidel")])
def test_polynomial_smoothers_match_oracle(ctx, name, smoother):
    """Richardson, Chebyshev and Gauss-Seidel smoothing: the Chebyshev
    bounds (power iteration) agree with the oracle's, one cycle agrees with
    the oracle's (Gauss-Seidel by level sets: the sequential sweeps' values),
    and MG-PCG needs the oracle's iteration count (FP64; FP32 cycle within 3)."""
    make, ns = CASES[name]
    levels, maps = _hierarchy(make, ns)
    Af, bf = O.assemble(levels[-1].patches, maps[-1][0], maps[-1][1], 0)
    Ps = [O.transfer_multipatch(levels[i].patches, maps[i][0], maps[i][1], maps[i + 1][0], maps[i + 1][1])
          for i in range(len(levels) - 1)]
    damping = 0.6 if smoother == "chebyshev" else 0.3 / np.abs(Af.diagonal()).max()
    omg = O.Multigrid(Af, Ps, npre=2, npost=2, damping=damping, smoother=smoother, smoother_param=20.0)
    xo, ito, erro = O.cg_solve(Af, bf, tol=1e-10, maxit=300, precond="mg", mg=omg)
    gp = gsx.Problem(ctx, levels[-1])
    A, b = gp.assemble()
    for prec in (gsx.FP64, gsx.MIXED):
        mg = gsx.Multigrid(gp, A, len(ns), 2, 2, damping, precision=prec, smoother=smoother, smoother_param=20.0)
        if smoother == "chebyshev":
            for lvl in range(1, len(ns)):
                assert mg.lmax(lvl) == pytest.approx(omg.lmax(lvl), rel=1e-12)
        if prec == gsx.FP64:
            r = ctx.to_device(bf)
            z = ctx.array(gp.nfree)
            mg.apply(r, z)
            zo = omg.apply(bf)
            assert np.linalg.norm(z.download() - zo) <= 1e-11 * np.linalg.norm(zo)
            r.close()
            z.close()
        x = ctx.array(gp.nfree).zero()
        it, rel, ok = gsx.mg_cg_solve(A, mg, b, x, tol=1e-10, maxit=300)
        print(name, smoother, prec, it, ito, flush=True)
        assert ok and rel < 1e-10
        if prec == gsx.FP64:
            assert abs(it - ito) <= 1
        else:
            assert it <= ito + 3
        assert np.linalg.norm(x.download() - xo) <= 1e-8 * np.linalg.norm(xo)
        x.close()
        mg.close()
    A.close()
    b.close()
    gp.close()


SCMS_CASES = {
    "square-p3": (lambda n: problems.poisson_square(n=n, p=3), (4, 8, 16)),
    "square-p4": (lambda n: problems.poisson_square(n=n, p=4), (2, 4, 8)),   # level 1: no interior, re-estimated
    "cube-p2": (lambda n: problems.poisson_noisy_cube(n=n, p=2, seed=4), (2, 4, 8)),
}


@pytest.mark.parametrize("name", list(SCMS_CASES))
def test_scms_smoother_matches_oracle(ctx, name):
    """Subspace corrected mass smoothing (the reference multigrid example's
    scms smoother on one patch with Dirichlet sides): the per-level damping
    from the power iteration, one cycle, and MG-PCG iterations and solution
    against the oracle's dense smoother in its numpy V-cycle."""
    make, ns = SCMS_CASES[name]
    print(f"[case] scms smoother {name}", flush=True)
    levels, maps = _hierarchy(make, ns)
    Af, bf, Ps, omg = _oracle_mg(levels, maps)
    As = [omg.matrix(lvl) for lvl in range(len(ns))]
    lv = O.scms_smoother_levels([L.patches[0] for L in levels], As)

    def smoother(lvl):
        S, tau = lv[lvl]

        def f(b, x, post):
            x += tau * (S @ (b - As[lvl] @ x))
        return f
    cyc = O.mg_vcycle(As, Ps, [None] + [smoother(lvl) for lvl in range(1, len(ns))], npre=1, npost=1)
    xo, ito, _ = O.pcg(Af, bf, cyc, tol=1e-10, maxit=300)
    gp = gsx.Problem(ctx, levels[-1])
    A, b = gp.assemble()
    for prec in (gsx.FP64, gsx.MIXED):
        mg = gsx.Multigrid(gp, A, len(ns), 1, 1, 1.0, precision=prec, smoother="scms")
        for lvl in range(1, len(ns)):
            assert mg.lmax(lvl) == pytest.approx(lv[lvl][1], rel=1e-10), lvl
        if prec == gsx.FP64:
            r = ctx.to_device(bf)
            z = ctx.array(gp.nfree)
            mg.apply(r, z)
            zo = cyc(bf)
            assert np.linalg.norm(z.download() - zo) <= 1e-11 * np.linalg.norm(zo)
            r.close()
            z.close()
        x = ctx.array(gp.nfree).zero()
        it, rel, ok = gsx.mg_cg_solve(A, mg, b, x, tol=1e-10, maxit=300)
        print(name, prec, it, ito, flush=True)
        assert ok and rel < 1e-10
        if prec == gsx.FP64:
            assert abs(it - ito) <= 1
        else:
            assert it <= ito + 3
        assert np.linalg.norm(x.download() - xo) <= 1e-8 * np.linalg.norm(xo)
        x.close()
        mg.close()
    A.close()
    b.close()
    gp.close()


def test_scms_smoother_error_paths(ctx):
    """Several components (patches, interfaces, sides without a Dirichlet
    condition) run FP64 cycles only, and need a hierarchy built from a
    problem (not from the caller's transfers)."""
    grid = problems.square_grid(2, 1, n=4, p=2)
    gp = gsx.Problem(ctx, grid)
    A, b = gp.assemble(gsx.PATTERN_EXPR, gsx.MIXED)
    with pytest.raises(gsx.GsxError):
        gsx.Multigrid(gp, A, 2, precision=gsx.MIXED, smoother="scms")
    levels, maps = _hierarchy(lambda n: problems.square_grid(2, 1, n, 2), (2, 4))
    _, _, Ps, _ = _oracle_mg(levels, maps)
    mg = gsx.Multigrid.from_transfers(ctx, A, Ps)
    with pytest.raises(gsx.GsxError):
        mg.set_level_smoother(1, "scms")
    mg.close()
    A.close()
    b.close()
    gp.close()


SPLIT_CASES = {
    "grid2x2-p3": (lambda n: problems.square_grid(2, 2, n, 3), (2, 4, 8)),
    "square-neumann-p3": (lambda n: problems.Problem([problems.poisson_square(n=n, p=3).patches[0]],
                                                     [(0, 1), (0, 2), (0, 3)], []), (4, 8, 16)),
    "two-cubes-p2": (lambda n: problems.two_cubes(n, 2), (2, 4)),
}


@pytest.mark.parametrize("name", list(SPLIT_CASES))
def test_multipatch_scms_smoother_matches_oracle(ctx, name):
    """The subspace corrected mass smoother with several components (patch
    interiors and 3D faces with their box smoothers and power-iteration
    dampings, lower-dimensional components solved exactly): one cycle and
    MG-PCG iterations and solution against the oracle's dense split
    smoother in its numpy V-cycle."""
    make, ns = SPLIT_CASES[name]
    print(f"[case] split scms {name}", flush=True)
    levels, maps = _hierarchy(make, ns)
    Af, bf, Ps, omg = _oracle_mg(levels, maps)
    As = [omg.matrix(lvl) for lvl in range(len(ns))]
    split = O.scms_split_levels(levels, As)

    def smoother(lvl):
        def f(b, x, post):
            x += split[lvl] @ (b - As[lvl] @ x)
        return f
    cyc = O.mg_vcycle(As, Ps, [None] + [smoother(lvl) for lvl in range(1, len(ns))], npre=1, npost=1)
    xo, ito, _ = O.pcg(Af, bf, cyc, tol=1e-10, maxit=300)
    gp = gsx.Problem(ctx, levels[-1])
    A, b = gp.assemble()
    mg = gsx.Multigrid(gp, A, len(ns), 1, 1, 1.0, smoother="scms")
    r = ctx.to_device(bf)
    z = ctx.array(gp.nfree)
    mg.apply(r, z)
    zo = cyc(bf)
    print("  cycle", np.linalg.norm(z.download() - zo) / np.linalg.norm(zo), flush=True)
    assert np.linalg.norm(z.download() - zo) <= 1e-11 * np.linalg.norm(zo)
    x = ctx.array(gp.nfree).zero()
    it, rel, ok = gsx.mg_cg_solve(A, mg, b, x, tol=1e-10, maxit=300)
    print(f"  {it} iterations, oracle {ito}", flush=True)
    assert ok and abs(it - ito) <= 1
    assert np.linalg.norm(x.download() - xo) <= 1e-8 * np.linalg.norm(xo)
    for o in (r, z, x, mg, A, b, gp):
        o.close()


@pytest.mark.parametrize("name", ["grid2x2-p3", "cube-p2", "grid4x4-p3"])
def test_multigrid_from_caller_transfers(ctx, name):
    """gsMultiGridOp::make(fineMatrix, transferMatrices) (the reference's
    construction, gsMultiGrid.hpp:19-35,70-94): the oracle's transfers
    uploaded, P^T A P formed on the device.  The level operators equal the
    oracle's Galerkin products to 1e-12 (row-scaled), one V(2,2) cycle its
    cycle, and the MG-PCG iteration count its count."""
    make, ns = CASES[name]
    print(f"[case] {name}", flush=True)
    levels, maps = _hierarchy(make, ns)
    Af, bf, Ps, omg = _oracle_mg(levels, maps)
    gp = gsx.Problem(ctx, levels[-1])
    A, b = gp.assemble(gsx.PATTERN_EXPR, gsx.FP64)
    mg = gsx.Multigrid.from_transfers(ctx, A, Ps, 2, 2, 0.6)
    for lvl in range(len(ns)):
        Ag, Pg = mg.level(lvl)
        assert _rowscaled(Ag, omg.matrix(lvl)) <= 1e-12, lvl
        if Pg is not None:
            Po = Ps[lvl].tocsr()
            Po.sort_indices()
            assert np.array_equal(Pg.indptr, Po.indptr) and np.array_equal(Pg.indices, Po.indices)
            assert np.array_equal(Pg.data, Po.data)
    rng = np.random.default_rng(3)
    r = rng.standard_normal(gp.nfree)
    rd, zd = ctx.to_device(r), ctx.array(gp.nfree)
    mg.apply(rd, zd)
    zo = omg.apply(r)
    assert np.linalg.norm(zd.download() - zo) <= 1e-10 * np.linalg.norm(zo)
    x = ctx.array(gp.nfree).zero()
    it, rel, ok = gsx.mg_cg_solve(A, mg, b, x, tol=1e-8, maxit=200)
    xo, ito, erro = O.cg_solve(Af, bf, tol=1e-8, maxit=200, precond="mg", mg=omg)
    assert ok and it == ito, (it, ito)
    assert np.linalg.norm(x.download() - xo) <= 1e-8 * np.linalg.norm(xo)
    mg.close()
    A.close()
    gp.close()


def test_per_level_smoothers(ctx):
    """gsMultiGridOp::setSmoother(lvl, op): setting every level to the same
    smoother one level at a time gives bitwise the cycle of
    gsx_mg_set_smoother; levels may mix Gauss-Seidel, Richardson and damped
    Jacobi smoothing, and the mixed hierarchy still preconditions CG."""
    make, ns = CASES["grid2x2-p3"]
    levels, maps = _hierarchy(make, ns)
    Af, bf, Ps, omg = _oracle_mg(levels, maps)
    gp = gsx.Problem(ctx, levels[-1])
    A, b = gp.assemble(gsx.PATTERN_EXPR, gsx.FP64)
    rng = np.random.default_rng(5)
    rd = ctx.to_device(rng.standard_normal(gp.nfree))
    outs = []
    for per_level in (False, True):
        mg = gsx.Multigrid.from_transfers(ctx, A, Ps, 1, 1, 0.7)
        if per_level:
            for lvl in range(1, len(ns)):
                mg.set_level_smoother(lvl, "gauss_seidel")
        else:
            lib_ok = gsx.lib().gsx_mg_set_smoother(mg.h, gsx.Multigrid.SMOOTHERS["gauss_seidel"], 0.0)
            assert lib_ok == 0
        zd = ctx.array(gp.nfree)
        mg.apply(rd, zd)
        outs.append(zd.download())
        mg.close()
    assert np.array_equal(outs[0], outs[1])
    # mixed: coarse-most smoothed level Richardson, the finest damped Jacobi
    mg = gsx.Multigrid.from_transfers(ctx, A, Ps, 2, 2, 0.6)
    mg.set_level_smoother(1, "richardson", damping=0.05)
    mg.set_level_smoother(2, "jacobi", damping=0.6)
    x = ctx.array(gp.nfree).zero()
    it, rel, ok = gsx.mg_cg_solve(A, mg, b, x, tol=1e-8, maxit=300)
    xo, ito, erro = O.cg_solve(Af, bf, tol=1e-8, maxit=3000, precond="jacobi")
    assert ok and it < ito, (it, ito)
    assert np.linalg.norm(x.download() - xo) <= 1e-6 * np.linalg.norm(xo)
    mg.close()
    A.close()
    gp.close()
