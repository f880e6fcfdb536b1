"""The C++ drop-in example drivers (gismo_amd/host/*_example.cpp, built by
build()) on the GPU, against the oracle and the Python mirror on the same
problems: per-level DOF counts, pattern sizes and CG iteration counts of
the poisson2_example call sequence (gsExprAssembler and gsPoissonAssembler
paths), the L2 convergence order, and the multigrid-preconditioned CG of the
multiGrid_example call sequence."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from gismo_amd import gsx, problems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "gismo_amd", "host", "bin")

pytestmark = pytest.mark.gpu


def _run(name, *args):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} missing: run __graft_entry__.build()")
    print(f"running {name} {' '.join(args)}", flush=True)
    out = subprocess.run([exe, *args], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    return out.stdout


def _levels(text):
    rows = []
    for m in re.finditer(r"level (\d+) dofs (\d+) nnz (\d+) iterations (\d+) relres (\S+) L2 (\S+)", text):
        rows.append((int(m[2]), int(m[3]), int(m[4]), float(m[5]), float(m[6])))
    return rows


def _h1(text):
    return np.array([float(m[1]) for m in re.finditer(r"level \d+ .* H1 (\S+)", text)])


def test_poisson2_example_square_p2():
    rows = _levels(_run("poisson2_example", "-r", "3", "-e", "1"))
    assert len(rows) == 4
    for lvl, (dofs, nnz, its, rel, l2) in enumerate(rows):
        n = 2 ** (lvl + 1)
        prob = problems.poisson_square(n=n, p=2)
        dmap, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
        Ao, bo = O.assemble(prob.patches, dmap, nfree, 0)
        _, ito, _ = O.cg_solve(Ao, bo, tol=1e-10, maxit=10000)
        assert (dofs, nnz) == (nfree, Ao.nnz)
        assert its == ito, (n, its, ito)
        assert rel < 1e-10
    l2 = np.array([r[4] for r in rows])
    rates = np.log2(l2[:-1] / l2[1:])
    assert rates[-1] > 2.7, rates


def test_poisson2_example_cube_full_pattern_mixed():
    rows = _levels(_run("poisson2_example", "-d", "3", "-e", "2", "-r", "1", "--full", "--mixed"))
    assert len(rows) == 2
    for lvl, (dofs, nnz, its, rel, l2) in enumerate(rows):
        n = 2 ** (lvl + 1)
        prob = problems.box_patch(3, 3, n)
        prob = problems.Problem([prob], problems.all_sides(3), [], name="cube")
        prob.patches[0].fkind = 1
        dmap, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
        Ao, bo = O.assemble(prob.patches, dmap, nfree, 1)
        assert (dofs, nnz) == (nfree, Ao.nnz)
        assert rel < 1e-10


def test_poisson2_example_grid_matches_python_mirror(ctx):
    rows = _levels(_run("poisson2_example", "--grid", "3", "2", "-e", "2", "-r", "2", "--last"))
    assert len(rows) == 1
    dofs, nnz, its, rel, l2 = rows[0]
    prob = problems.square_grid(3, 2, n=8, p=3)
    gp = gsx.Problem(ctx, prob)
    A, b = gp.assemble(gsx.PATTERN_EXPR, gsx.FP64)
    x = ctx.array(gp.nfree).zero()
    it, rel_py, ok = A.cg_solve(b, x, tol=1e-10, maxit=100000)
    assert (dofs, nnz, its) == (gp.nfree, A.nnz, it)
    assert abs(l2 - gp.l2_error(x)) <= 1e-12 * l2
    A.close()
    gp.close()


def test_multigrid_example_matches_python_mirror(ctx):
    text = _run("multiGrid_example", "--grid", "2", "2", "-r", "3", "-p", "2", "-l", "3", "-s", "j",
                "--MG.Damping", "0.8", "-t", "1e-8")
    m = re.search(r"dofs (\d+) nnz (\d+)", text)
    s = re.search(r"solver cg iterations (\d+) relres (\S+)", text)
    l2 = float(re.search(r"L2 (\S+)", text)[1])
    prob = problems.square_grid(2, 2, n=8, p=2)
    gp = gsx.Problem(ctx, prob)
    A, b = gp.assemble(gsx.PATTERN_FULL, gsx.FP64)
    mg = gsx.Multigrid(gp, A, 3, 1, 1, 0.8)
    x = ctx.array(gp.nfree).zero()
    it, rel, ok = gsx.mg_cg_solve(A, mg, b, x, tol=1e-8, maxit=100)
    assert (int(m[1]), int(m[2])) == (gp.nfree, A.nnz)
    assert int(s[1]) == it and ok
    assert float(s[2]) < 1e-8
    assert abs(l2 - gp.l2_error(x)) <= 1e-10 * l2
    # the reference's sequence ran: gsGridHierarchy::buildByCoarsening, then
    # gsMultiGridOp::make(matrix, transfers) with a Cholesky coarse solver
    # and one Jacobi smoother per level; same iterations as the oracle's
    # multigrid on its own transfers
    h = re.search(r"hierarchy levels (\d+) transfers (.*)", text)
    assert h and int(h[1]) == 3
    levels = [problems.square_grid(2, 2, n, 2) for n in (2, 4, 8)]
    maps = [O.mapper(q.patches, q.dir_sides, q.ifaces) for q in levels]
    Ps = [O.transfer_multipatch(levels[i].patches, maps[i][0], maps[i][1], maps[i + 1][0], maps[i + 1][1])
          for i in range(2)]
    assert h[2].split() == [f"{P.shape[0]}x{P.shape[1]}" for P in Ps]
    Af, bf = O.assemble(levels[-1].patches, maps[-1][0], maps[-1][1], 1)
    omg = O.Multigrid(Af, Ps, npre=1, npost=1, damping=0.8)
    _, ito, _ = O.cg_solve(Af, bf, tol=1e-8, maxit=100, precond="mg", mg=omg)
    assert int(s[1]) == ito
    # the stationary multigrid iteration converges too
    text = _run("multiGrid_example", "--grid", "2", "2", "-r", "3", "-p", "2", "-l", "3", "-i", "d", "-s", "j",
                "--MG.Damping", "0.8", "-t", "1e-8")
    s = re.search(r"solver d iterations (\d+) relres (\S+)", text)
    assert float(s[2]) < 1e-8 and int(s[1]) < 100
    mg.close()
    A.close()
    gp.close()


@pytest.mark.parametrize("fname,args,rate", [("ldomain.xml", ["-e", "1", "-r", "3"], 2.7),
                                            ("fichera.xml", ["-e", "1", "-r", "2"], 2.5)])
def test_poisson2_example_reads_reference_xml(fname, args, rate):
    """The example driver reads the reference's own multipatch XML inputs
    (L-shaped domain, Fichera corner; u = prod sin(pi x_i) vanishes on their
    boundaries) and converges at the order p+1 = 3 in L2."""
    path = os.path.join(ROOT, "tests", "golden", "xml", fname)
    rows = _levels(_run("poisson2_example", "-f", path, *args))
    l2 = np.array([r[4] for r in rows])
    assert all(r[3] < 1e-10 for r in rows)
    rates = np.log2(l2[:-1] / l2[1:])
    print(fname, l2, rates, flush=True)
    assert rates[-1] > rate, rates


def test_multigrid_example_chebyshev_smoother():
    """multiGrid_example with MG.Smoother = Chebyshev on the unit cube
    (degree 3): converges, in fewer iterations than damped Jacobi."""
    its = {}
    for sm in ("j", "c"):
        text = _run("multiGrid_example", "-d", "3", "-r", "3", "-p", "3", "-l", "3", "-s", sm,
                    "--MG.NumPreSmooth", "3", "--MG.NumPostSmooth", "3", "--MG.Damping", "0.6", "-t", "1e-10",
                    "--Solver.MaxIterations", "300")
        s = re.search(r"solver cg iterations (\d+) relres (\S+)", text)
        assert float(s[2]) < 1e-10
        its[sm] = int(s[1])
    print(its, flush=True)
    assert its["c"] < its["j"], its


def _bvp_oracle(d, p, n, dsides, nsides, mode, interp=False):
    """The --bvp problem of poisson2_example at n elements per direction:
    u = prod sin(pi x_i) on [0.25, 1.25]^d, projected Dirichlet data on
    dsides, the flux grad u on nsides."""
    P = problems.box_patch(d, p, n, [0.25] * d, [1.25] * d)
    P.fkind = 1
    prob = problems.Problem([P], [(0, s) for s in dsides], [])
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    nfix = int(max(dofs)) + 1 - nfree
    recs = [(0, s, 1, 1, ()) for s in dsides]
    fixed = (O.dirichlet_interpolation(prob.patches, dofs, nfree, nfix, recs) if interp
             else O.dirichlet_values(prob.patches, dofs, nfree, nfix, recs))
    A, b = O.assemble(prob.patches, dofs, nfree, mode, fixed=fixed)
    if nsides:
        b = b + O.neumann_load(prob.patches, dofs, nfree, [(0, s, 2, 1, ()) for s in nsides])
    xo, ito, _ = O.cg_solve(A, b, tol=1e-10, maxit=10000)
    eo = O.l2_error(prob.patches, dofs, nfree, xo, fixed=fixed)
    return nfree, A.nnz, ito, eo, eo + O.h1_error(prob.patches, dofs, nfree, xo, fixed=fixed)


@pytest.mark.parametrize("d,args,rate", [(2, ["-e", "1", "-r", "3"], 2.7),
                                        (3, ["-d", "3", "-e", "1", "-r", "2"], 2.5)])
def test_poisson2_example_bvp(d, args, rate):
    """poisson2_example --bvp: Dirichlet data u on the west/south(/front)
    sides by L2 projection (gsFeSpace::setup with dirichlet::l2Projection)
    and the flux on the others through gsExprAssembler::assembleBdr with
    u * g_N.tr() * nv(G).  Per level the DOFs, pattern, CG iterations and
    L2 error (fixed values included) agree with the oracle on the same
    problem; the error falls at order p + 1 = 3."""
    text = _run("poisson2_example", "--bvp", *args)
    rows = _levels(text)
    h1 = _h1(text)
    assert len(h1) == len(rows)
    lower = [1, 3, 5][:d]
    upper = [2, 4, 6][:d]
    for lvl, (dofs, nnz, its, rel, l2) in enumerate(rows):
        nfree, onnz, ito, eo, ho = _bvp_oracle(d, 2, 2 ** (lvl + 1), lower, upper, 0)
        assert (dofs, nnz) == (nfree, onnz)
        assert abs(its - ito) <= 2, (lvl, its, ito)
        assert rel < 1e-10
        assert abs(l2 - eo) <= 1e-4 * eo, (lvl, l2, eo)
        assert abs(h1[lvl] - ho) <= 1e-4 * ho, (lvl, h1[lvl], ho)
    l2 = np.array([r[4] for r in rows])
    rates = np.log2(l2[:-1] / l2[1:])
    h1rates = np.log2(h1[:-1] / h1[1:])
    print(d, l2, rates, h1, h1rates, flush=True)
    assert rates[-1] > rate, rates
    # H1 error falls at order p = 2
    assert h1rates[-1] > 1.8, h1rates


@pytest.mark.parametrize("dv", [None, "102"])
def test_poisson2_example_bvp_poisson_assembler(dv):
    """poisson2_example --bvp --full: gsPoissonAssembler with u on every side,
    its default DirichletValues = 101 (interpolation at the side anchors) or
    102 (L2 projection), on the full pattern, against the oracle; converges
    at order 3."""
    extra = [] if dv is None else ["--dv", dv]
    rows = _levels(_run("poisson2_example", "--bvp", "--full", "-e", "1", "-r", "3", *extra))
    for lvl, (dofs, nnz, its, rel, l2) in enumerate(rows):
        nfree, onnz, ito, eo, _ = _bvp_oracle(2, 2, 2 ** (lvl + 1), [1, 2, 3, 4], [], 1, interp=dv is None)
        assert (dofs, nnz) == (nfree, onnz)
        assert abs(its - ito) <= 2, (lvl, its, ito)
        assert abs(l2 - eo) <= 1e-4 * eo, (lvl, l2, eo)
    l2 = np.array([r[4] for r in rows])
    rates = np.log2(l2[:-1] / l2[1:])
    assert rates[-1] > 2.7, rates


def test_poisson2_example_bvp_from_xml():
    """poisson2_example -f square_pair_bvp.xml: source, boundary conditions
    (Dirichlet u, Neumann flux grad u) and exact solution read from the file
    in the layout of the reference's poisson2d_bvp.xml; two glued patches.
    Per level the DOFs, pattern and L2 error agree with the oracle."""
    path = os.path.join(ROOT, "tests", "golden", "xml", "square_pair_bvp.xml")
    text = _run("poisson2_example", "-f", path, "-e", "1", "-r", "3")
    assert "Boundary conditions from the file" in text
    rows = _levels(text)
    assert len(rows) == 4
    dsides = [(0, 1), (0, 3), (1, 3)]
    nsides = [(0, 4), (1, 4), (1, 2)]
    for lvl, (dofs, nnz, its, rel, l2) in enumerate(rows):
        n = 2 ** (lvl + 1)
        patches = []
        for i in range(2):
            P = problems.box_patch(2, 2, n, [0.25 + i, 0.25], [1.25 + i, 1.25])
            P.fkind = 1
            patches.append(P)
        prob = problems.Problem(patches, dsides, [(0, 2, 1, 1, 1, 0, 0, 0)])
        dmap, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
        nfix = int(max(dmap)) + 1 - nfree
        fixed = O.dirichlet_values(prob.patches, dmap, nfree, nfix, [(p, s, 1, 1, ()) for p, s in dsides])
        A, b = O.assemble(prob.patches, dmap, nfree, 0, fixed=fixed)
        b = b + O.neumann_load(prob.patches, dmap, nfree, [(p, s, 2, 1, ()) for p, s in nsides])
        xo, ito, _ = O.cg_solve(A, b, tol=1e-10, maxit=10000)
        eo = O.l2_error(prob.patches, dmap, nfree, xo, fixed=fixed)
        assert (dofs, nnz) == (nfree, A.nnz)
        assert abs(its - ito) <= 2, (lvl, its, ito)
        assert abs(l2 - eo) <= 1e-4 * eo, (lvl, l2, eo)
    l2 = np.array([r[4] for r in rows])
    rates = np.log2(l2[:-1] / l2[1:])
    print(l2, rates, flush=True)
    assert rates[-1] > 2.7, rates


@pytest.mark.parametrize("args", [["-d", "3", "-e", "2", "-r", "2"], ["--bvp", "-e", "1", "-r", "3"],
                                  ["--bvp", "--full", "-e", "1", "-r", "2"]])
def test_poisson2_example_fast_diagonalization(args):
    """poisson2_example --fd: gsConjugateGradient preconditioned by
    gsPatchPreconditionersCreator::fastDiagonalizationOp(basis, bc).  The
    example's domains are affine boxes with unit coefficient, where the
    stiffness matrix is the Kronecker sum the operator inverts (Neumann
    ends kept for --bvp): one iteration per level, and the same DOFs and
    errors as the diagonally preconditioned run."""
    jac = _levels(_run("poisson2_example", *args))
    fd = _levels(_run("poisson2_example", "--fd", *args))
    assert len(fd) == len(jac)
    for (dj, nj, ij, rj, lj), (df, nf, itf, rf, lf) in zip(jac, fd):
        assert (dj, nj) == (df, nf)
        assert itf == 1, (args, itf)
        assert rf < 1e-10
        assert abs(lf - lj) <= 1e-6 * lj


def test_multigrid_example_default_gauss_seidel(ctx):
    """multiGrid_example with its default smoother, Gauss-Seidel (the
    reference example's default): MG-PCG iterations and solution equal the
    Python mirror's with the same smoother (damped Jacobi printed beside)."""
    text = _run("multiGrid_example", "--grid", "2", "2", "-r", "3", "-p", "2", "-l", "3", "-t", "1e-8")
    s = re.search(r"solver cg iterations (\d+) relres (\S+)", text)
    l2 = float(re.search(r"L2 (\S+)", text)[1])
    prob = problems.square_grid(2, 2, n=8, p=2)
    gp = gsx.Problem(ctx, prob)
    A, b = gp.assemble(gsx.PATTERN_FULL, gsx.FP64)
    its = {}
    for sm in ("gauss_seidel", "jacobi"):
        mg = gsx.Multigrid(gp, A, 3, 1, 1, 0.8, smoother=sm)
        x = ctx.array(gp.nfree).zero()
        it, rel, ok = gsx.mg_cg_solve(A, mg, b, x, tol=1e-8, maxit=100)
        its[sm] = it
        if sm == "gauss_seidel":
            assert int(s[1]) == it and ok
            assert abs(l2 - gp.l2_error(x)) <= 1e-10 * l2
        mg.close()
        x.close()
    print(its, flush=True)
    A.close()
    gp.close()


def test_multigrid_example_scms_smoother(ctx):
    """multiGrid_example -s scms (the subspace corrected mass smoother,
    one patch with Dirichlet sides, degree 3): the Python mirror's
    iteration count, and fewer iterations than damped Jacobi smoothing."""
    args = ["-d", "2", "-r", "4", "-p", "3", "-l", "3", "-t", "1e-8"]
    text = _run("multiGrid_example", *args, "-s", "scms")
    s = re.search(r"solver cg iterations (\d+) relres (\S+)", text)
    assert float(s[2]) < 1e-8
    prob = problems.poisson_square(n=16, p=3)
    gp = gsx.Problem(ctx, prob)
    A, b = gp.assemble(gsx.PATTERN_FULL, gsx.FP64)
    mg = gsx.Multigrid(gp, A, 3, 1, 1, smoother="scms")
    x = ctx.array(gp.nfree).zero()
    it, rel, ok = gsx.mg_cg_solve(A, mg, b, x, tol=1e-8, maxit=100)
    assert ok and int(s[1]) == it, (s[1], it)
    tj = _run("multiGrid_example", *args, "-s", "j")
    sj = re.search(r"solver cg iterations (\d+) relres (\S+)", tj)
    print("scms", it, "jacobi", sj[1], flush=True)
    assert it < int(sj[1])
    for o in (x, mg, A, b, gp):
        o.close()


def test_poisson2_example_scms_matches_python_mirror(ctx):
    """poisson2_example --scms: gsConjugateGradient preconditioned by
    gsPatchPreconditionersCreator::subspaceCorrectedMassSmootherOp(basis,
    bc) through the C++ drop-in classes takes the Python mirror's CG
    iterations on the same square (16 x 16 elements, p = 2)."""
    rows = _levels(_run("poisson2_example", "--scms", "-r", "3", "-e", "1", "--last"))
    assert len(rows) == 1
    dofs, nnz, its, rel, l2 = rows[0]
    prob = problems.poisson_square(n=16, p=2)
    gp = gsx.Problem(ctx, prob)
    A, b = gp.assemble(gsx.PATTERN_EXPR, gsx.FP64)
    sm = gsx.SubspaceMassSmoother(gp)
    x = ctx.array(gp.nfree).zero()
    it, rel_py, ok = gsx.fd_cg_solve(A, sm, b, x, tol=1e-10, maxit=100000)
    assert ok and (dofs, nnz, its) == (gp.nfree, A.nnz, it), (dofs, its, gp.nfree, it)
    assert rel < 1e-10
    assert abs(l2 - gp.l2_error(x)) <= 1e-9 * l2
    for o in (x, sm, A, b, gp):
        o.close()
