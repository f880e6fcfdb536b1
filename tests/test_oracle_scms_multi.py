"""Pins of the oracle's component split of the subspace corrected mass
smoother (oracle.scms_components, oracle.scms_split_levels): against the
reference's golden component assignment of the refined Fichera corner
(unittests/gsBasisComponents_test.cpp:41-79), the exact partition of the
free DOFs with the component sizes a box grid fixes, the reduction to the
pinned single-patch smoother, and symmetry and convergence of the cycle."""
import json
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O
from gismo_amd import problems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def two_cubes(n, p):
    return problems.two_cubes(n, p)


def test_two_cubes_glue():
    prob = two_cubes(4, 2)
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    assert nfree == 9 * 4 * 4
    A, b = O.assemble(prob.patches, dofs, nfree, 0)
    assert abs(A - A.T).max() == 0.0
    assert np.all(np.linalg.eigvalsh(A.toarray()) > 0)


def test_components_partition_the_free_dofs_of_a_grid():
    """2 x 2 degree-2 patches with 4 elements each (6 functions per
    direction), Dirichlet outside: four interiors of 4 x 4, four interfaces
    of 4, the one free corner; the outer sides and corners are empty."""
    prob = problems.square_grid(2, 2, n=4, p=2)
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    comps = O.scms_components(prob.patches, dofs, nfree)
    sizes = sorted((dim, len(idx)) for dim, idx, _ in comps if len(idx))
    assert sizes == [(0, 1)] + [(1, 4)] * 4 + [(2, 16)] * 4
    allidx = np.concatenate([idx for _, idx, _ in comps])
    assert np.array_equal(np.sort(allidx), np.arange(nfree))
    # a side without a Dirichlet condition is a component of its own
    P = problems.box_patch(2, 3, 5)
    P.fkind = 1
    single = problems.Problem([P], [(0, 1), (0, 2), (0, 3)], [])
    dofs, nfree = O.mapper(single.patches, single.dir_sides, [])
    comps = O.scms_components(single.patches, dofs, nfree)
    sizes = sorted((dim, len(idx)) for dim, idx, _ in comps if len(idx))
    assert sizes == [(1, 6), (2, 6 * 6)]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_components_on_the_fichera_golden():
    """The refined Fichera corner (7 degree-1 patches, 2 elements per
    direction, no Dirichlet side): 26 corners, 51 edges, 33 faces and 7
    interiors of one DOF each, and the same grouping of the 117 DOFs as the
    reference's golden assignment."""
    from gismo_amd import gsx
    gsx.lib()
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "gsxml_check")
        libdir = os.path.join(ROOT, "gismo_amd")
        subprocess.check_call(["g++", "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                               "-I" + os.path.join(ROOT, "gismo_amd", "host"),
                               os.path.join(ROOT, "tests", "cpp", "gsxml_check.cpp"), "-o", exe, "-L" + libdir,
                               "-lgsx", "-Wl,-rpath," + libdir])
        out = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "xml")], capture_output=True, text=True,
                             timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    ifaces = [tuple(int(v) for v in ln.split()[1:]) for ln in out.stdout.splitlines()
              if ln.startswith("fichera_iface")]
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "fichera_components.json")))
    kn = problems.knots_uniform(2, 1)
    pats = [problems.Patch(d=3, p=[1, 1, 1], knots=[kn] * 3, ctrl=np.zeros((27, 3))) for _ in range(7)]
    dofs, nfree = O.mapper(pats, [], ifaces)
    comps = O.scms_components(pats, dofs, nfree, combine_corners=False)
    per_dim = [sum(1 for dim, _, _ in comps if dim == k) for k in range(4)]
    assert per_dim == gold["components_per_dim"]
    assert all(len(idx) == 1 for _, idx, _ in comps)
    label = np.full(nfree, -1)
    for c, (_, idx, _) in enumerate(comps):
        label[idx] = c
    g = np.array(gold["component_of_dof"])
    # the same partition: golden label <-> oracle component is one to one
    pairs = set(zip(g.tolist(), label.tolist()))
    assert len(pairs) == len(set(g.tolist())) == len(set(label.tolist())) == nfree
    combined = O.scms_components(pats, dofs, nfree)
    assert sorted(len(idx) for dim, idx, _ in combined if dim == 0) == [26]


def test_single_patch_split_is_the_pinned_smoother():
    """One patch with every side Dirichlet: the interior is the only
    non-empty component, and the split smoother is tau S of the pinned
    single-patch smoother."""
    levels = [problems.poisson_square(n=n, p=3) for n in (4, 8, 16)]
    maps = [O.mapper(L.patches, L.dir_sides, L.ifaces) for L in levels]
    Af, _ = O.assemble(levels[-1].patches, maps[-1][0], maps[-1][1], 0)
    Ps = [O.transfer_multipatch(levels[i].patches, maps[i][0], maps[i][1], maps[i + 1][0], maps[i + 1][1])
          for i in range(2)]
    omg = O.Multigrid(Af, Ps)
    As = [omg.matrix(l) for l in range(3)]
    split = O.scms_split_levels(levels, As)
    single = O.scms_smoother_levels([L.patches[0] for L in levels], As)
    for l in (1, 2):
        S, tau = single[l]
        assert np.abs(split[l] - tau * S).max() <= 1e-14 * np.abs(tau * S).max()


@pytest.mark.parametrize("name", ["grid2x2", "two-cubes"])
def test_split_smoother_is_symmetric_and_mg_pcg_converges(name):
    make, ns = {"grid2x2": (lambda n: problems.square_grid(2, 2, n, 3), (2, 4, 8)),
                "two-cubes": (lambda n: two_cubes(n, 2), (2, 4))}[name]
    levels = [make(n) for n in ns]
    maps = [O.mapper(L.patches, L.dir_sides, L.ifaces) for L in levels]
    Af, bf = O.assemble(levels[-1].patches, maps[-1][0], maps[-1][1], 0)
    Ps = [O.transfer_multipatch(levels[i].patches, maps[i][0], maps[i][1], maps[i + 1][0], maps[i + 1][1])
          for i in range(len(ns) - 1)]
    omg = O.Multigrid(Af, Ps)
    As = [omg.matrix(l) for l in range(len(ns))]
    split = O.scms_split_levels(levels, As)
    for S in split[1:]:
        assert np.abs(S - S.T).max() <= 1e-12 * np.abs(S).max()

    def sm(l):
        def f(b, x, post):
            x += split[l] @ (b - As[l] @ x)
        return f
    cyc = O.mg_vcycle(As, Ps, [None] + [sm(l) for l in range(1, len(ns))])
    x, it, err = O.pcg(Af, bf, cyc, tol=1e-10, maxit=200)
    assert err < 1e-10 and it <= 30, it
