"""Exact coarse solve by the banded FP64 Cholesky factor (gsx_band.hip, in
reverse Cuthill-McKee order), forced onto small coarsest levels through
gsx_ctx_set_coarse_dense_max.  A one-level multigrid is the coarse solve
alone: its apply agrees with scipy's sparse direct solve to rounding, on
coarse sizes that are and are not multiples of the 64-row panel, below one
panel, 2D and 3D, one patch and several.  With several levels the cycle is
the one of the dense inverse: the same MG-PCG iterations and solution, in
FP64 and mixed precision."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from gismo_amd import gsx, problems

pytestmark = pytest.mark.gpu

CASES = [
    ("square-n4-p2", lambda: problems.poisson_square(n=4, p=2)),
    ("square-n40-p2", lambda: problems.poisson_square(n=40, p=2)),
    ("grid2x2-n16-p2", lambda: problems.square_grid(2, 2, 16, 2)),
    ("grid4x4-n4-p3", lambda: problems.square_grid(4, 4, 4, 3)),
    ("cube-n8-p2", lambda: problems.poisson_noisy_cube(n=8, p=2, seed=4)),
]


@pytest.fixture
def band_ctx(ctx):
    ctx.set_coarse_dense_max(1)
    yield ctx
    ctx.set_coarse_dense_max(32768)


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_band_coarse_solve_is_exact(band_ctx, name, make):
    ctx = band_ctx
    print(f"[case] band coarse solve {name}", flush=True)
    gp = gsx.Problem(ctx, make())
    A, b = gp.assemble(gsx.PATTERN_FULL)
    As = A.to_scipy().tocsc()
    mg = gsx.Multigrid(gp, A, 1, 1, 1, 1.0)
    rng = np.random.default_rng(3)
    r = rng.standard_normal(gp.nfree)
    rd = ctx.to_device(r)
    zd = ctx.array(gp.nfree)
    mg.apply(rd, zd)
    z = zd.download()
    zs = spla.spsolve(As, r)
    rel = np.linalg.norm(z - zs) / np.linalg.norm(zs)
    print(f"  n={gp.nfree} rel err {rel:.2e}", flush=True)
    assert rel <= 1e-12
    for o in (mg, rd, zd, A, b, gp):
        o.close()


@pytest.mark.parametrize("precision", [gsx.FP64, gsx.MIXED])
def test_band_and_dense_cycles_agree(ctx, precision):
    prob = problems.square_grid(2, 2, 32, 3)
    print(f"[case] band vs dense MG-PCG precision={precision}", flush=True)
    gp = gsx.Problem(ctx, prob)
    A, b = gp.assemble(gsx.PATTERN_FULL, precision)
    out = []
    for dense_max in (32768, 1):
        ctx.set_coarse_dense_max(dense_max)
        try:
            mg = gsx.Multigrid(gp, A, 3, 2, 2, 0.6, precision=precision)
        finally:
            ctx.set_coarse_dense_max(32768)
        x = ctx.array(gp.nfree).zero()
        it, rel, ok = gsx.mg_cg_solve(A, mg, b, x, tol=1e-10, maxit=200)
        assert ok
        out.append((it, x.download()))
        for o in (x, mg):
            o.close()
    (itd, xd), (itb, xb) = out
    print(f"  dense {itd} iterations, band {itb}", flush=True)
    assert itd == itb
    assert np.linalg.norm(xb - xd) <= 1e-9 * np.linalg.norm(xd)
    for o in (A, b, gp):
        o.close()


def test_coarse_dense_max_rejects_out_of_range(ctx):
    for bad in (0, 32769):
        with pytest.raises(gsx.GsxError):
            ctx.set_coarse_dense_max(bad)
