"""Allocation-free CG (gsx_cg_solve_noalloc, workspace in a fixed device
array) against gsx_cg_solve's FP64 CSR CG on the same matrices: the same
iteration count, residual and solution bit for bit, on two calls in a row,
the second with a smaller system; a request beyond the workspace is refused
with an error before any kernel runs."""
import ctypes as C

import numpy as np
import pytest

from gismo_amd import gsx, problems

pytestmark = pytest.mark.gpu


def _noalloc(A, b, x, precond, tau, tol=1e-10, maxit=2000):
    it, rr = C.c_int32(), C.c_double()
    st = gsx.lib().gsx_cg_solve_noalloc(A.h, precond, tau, b.ptr, x.ptr, tol, maxit, C.byref(it), C.byref(rr))
    return st, it.value, rr.value


@pytest.mark.parametrize("precond,tau", [(gsx.PREC_JACOBI, 1.0), (gsx.PREC_NONE, 1.0), (gsx.PREC_JACOBI, 0.8)])
def test_noalloc_cg_matches_cg(ctx, precond, tau):
    # a 3D cube (1,331 rows), then a smaller 2D multipatch system (289 rows)
    for prob in (problems.poisson_noisy_cube(n=10, p=3, seed=1), problems.square_grid(2, 2, 8, 2)):
        print(f"[case] noalloc {prob.name} precond={precond} tau={tau}", flush=True)
        gp = gsx.Problem(ctx, prob)
        A, b = gp.assemble(gsx.PATTERN_EXPR, gsx.FP64)
        A.set_operator(False)   # the CSR CG that the allocation-free entry runs
        x1 = ctx.array(gp.nfree).zero()
        it1, r1, ok = A.cg_solve(b, x1, tol=1e-10, maxit=2000, precond=precond, tau=tau)
        assert ok
        x2 = ctx.array(gp.nfree).zero()
        st, it2, r2 = _noalloc(A, b, x2, precond, tau)
        assert st == 0, gsx.lib().gsx_last_error(ctx.h)
        assert it1 == it2 and r1 == r2, (it1, it2, r1, r2)
        assert np.array_equal(x1.download(), x2.download())
        for a in (A, b, x1, x2):
            a.close()
        gp.close()


def test_noalloc_cg_refuses_oversized_requests(ctx):
    gp = gsx.Problem(ctx, problems.poisson_square(8, 2))
    A, b = gp.assemble(gsx.PATTERN_EXPR, gsx.FP64)
    x = ctx.array(gp.nfree).zero()
    st, _, _ = _noalloc(A, b, x, gsx.PREC_JACOBI, 1.0, maxit=20_000_000)   # 160 MB of history
    assert st == gsx.GSX_ERR_ARG
    assert b"workspace" in gsx.lib().gsx_last_error(ctx.h)
    # the entry stays usable after the refusal
    st, it, rr = _noalloc(A, b, x, gsx.PREC_JACOBI, 1.0)
    assert st == 0 and rr < 1e-10
    for a in (A, b, x):
        a.close()
    gp.close()
