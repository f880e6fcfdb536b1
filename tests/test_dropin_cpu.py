"""The C++ drop-in classes (gismo_amd/host/gismo.h) on the CPU: knot vector
construction and refinement against gsKnotVector's definitions, topology
detection against the interface records of the reference's square grid,
the geometry mapped into the discretisation basis against the Python
problem generator (bitwise), and the DOF mapper checked physically (shared
global index <=> coinciding control points), for glued, flipped, rotated
and permuted patches in 2D and 3D."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O
from gismo_amd import gsx, problems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def host_check(tmp_path_factory):
    gsx.lib()   # the library must be built
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("hc") / "gismo_host_check")
    libdir = os.path.join(ROOT, "gismo_amd")
    subprocess.check_call([cxx, "-O1", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-Wno-unused-parameter",
                           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "gismo_amd", "host"),
                           os.path.join(ROOT, "tests", "cpp", "gismo_host_check.cpp"), "-o", exe,
                           "-L" + libdir, "-lgsx", "-Wl,-rpath," + libdir])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = {}
    for ln in out.stdout.splitlines():
        key, _, rest = ln.partition(" ")
        lines[key] = rest.split()
    return lines


def _knots(tokens):
    return int(tokens[0]), np.array([float(t) for t in tokens[1:]])


def test_knot_vectors(host_check):
    deg, k = _knots(host_check["kv_uniform"])
    assert deg == 2 and np.array_equal(k, problems.knots_uniform(4, 2))
    deg, k = _knots(host_check["kv_refine1"])
    assert deg == 1 and np.array_equal(k, [0, 0, 0.5, 1, 1])
    # two knots per span: thirds of the halves
    deg, k = _knots(host_check["kv_refine2"])
    ref = [0, 0] + [0.5 * j / 3 for j in (1, 2)] + [0.5] + [0.5 + 0.5 * j / 3 for j in (1, 2)] + [1, 1]
    assert np.allclose(k, ref, rtol=0, atol=1e-16)
    deg, k = _knots(host_check["kv_elevate"])
    assert deg == 3 and np.array_equal(k, [0] * 4 + [1] * 4)
    # elevation keeps the continuity: interior multiplicity grows too
    deg, k = _knots(host_check["kv_elevate_interior"])
    assert deg == 2 and np.array_equal(k, [0, 0, 0, 0.5, 0.5, 1, 1, 1])


def _records(tokens):
    return sorted(tuple(int(v) for v in t.split(",")) for t in tokens)


def test_grid_topology_matches_reference_numbering(host_check):
    prob = problems.square_grid(3, 2, n=4, p=3)
    assert _records(host_check["grid3x2_ifaces"]) == sorted(prob.ifaces)
    assert _records(host_check["grid3x2_bdr"]) == sorted(prob.dir_sides)


def test_geometry_in_discretisation_basis_is_bitwise(host_check):
    prob = problems.square_grid(3, 2, n=4, p=3)
    ctrl = np.array([float(t) for t in host_check["grid3x2_ctrl4"]]).reshape(-1, 2)
    assert np.array_equal(ctrl, prob.patches[4].ctrl)


@pytest.mark.parametrize("name,nfree,nglued", [
    ("grid3x2", 17 * 11, 19 * 13),
    ("flip2d", None, None),
    ("perm3d", 4 * 4 * 9, 6 * 6 * 11),
    ("perm3d_p3", 3 * 3 * 7, 5 * 5 * 9),
    ("cubes2x2", 9 * 9 * 4, 11 * 11 * 6),
])
def test_mapper_glues_coinciding_functions(host_check, name, nfree, nglued):
    nf, ng, ok = (int(v) for v in host_check[name + "_mapper"])
    assert ok == 1
    if nfree is not None:
        assert (nf, ng) == (nfree, nglued)


def test_flipped_and_rotated_neighbours_found(host_check):
    recs = _records(host_check["flip2d_ifaces"])
    # patch 1 meets patch 0 with its east side (parameter 0 runs along -x);
    # patch 2 meets with its west side, tangential direction flipped
    assert recs == [(0, 2, 1, 2, 1, 0, 0, 0), (0, 4, 2, 1, 1, 1, 0, 0)]
    assert _records(host_check["perm3d_ifaces"]) == [(0, 2, 1, 4, 2, 0, 0, 1)]


def test_xml_reader_on_reference_inputs():
    """gsXml.h reads the reference's own XML inputs (tests/golden/xml, copied
    from its filedata/): their interface records agree with corner-matched
    topology, the refined Fichera corner has the 117 free DOFs asserted by
    the reference's gsBasisComponents_test and their golden component
    assignment (host mapper and oracle mapper), and matrices, sparse matrices and
    multipatches survive a write/read round trip."""
    gsx.lib()
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "gsxml_check")
        libdir = os.path.join(ROOT, "gismo_amd")
        subprocess.check_call([cxx, "-O1", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-Wno-unused-parameter",
                               "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "gismo_amd", "host"),
                               os.path.join(ROOT, "tests", "cpp", "gsxml_check.cpp"), "-o", exe,
                               "-L" + libdir, "-lgsx", "-Wl,-rpath," + libdir])
        out = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "xml")], capture_output=True, text=True,
                             timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "fichera_refined_free 117" in out.stdout
    assert out.stdout.count("topology_match 1") == 3
    assert "bvp dirichlet 3 neumann 3" in out.stdout
    # the reference's golden DOF -> component assignment of the refined
    # Fichera corner (unittests/gsBasisComponents_test.cpp:41-79)
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "fichera_components.json")))
    lines = {ln.split()[0]: ln.split()[1:] for ln in out.stdout.splitlines() if ln.startswith("fichera_")}
    assert [int(v) for v in lines["fichera_components_per_dim"]][::-1] == gold["components_per_dim"][::-1]
    assert [int(v) for v in lines["fichera_component_of_dof"]] == gold["component_of_dof"]
    # the oracle's mapper on the same interface records gives the same
    # numbering: each free DOF lands on the golden's component
    ifaces = [tuple(int(v) for v in ln.split()[1:]) for ln in out.stdout.splitlines()
              if ln.startswith("fichera_iface")]
    local = np.array([int(v) for v in lines["fichera_local_component"]]).reshape(7, 27)
    kn = problems.knots_uniform(2, 1)
    pats = [problems.Patch(d=3, p=[1, 1, 1], knots=[kn] * 3, ctrl=np.zeros((27, 3))) for _ in range(7)]
    dofs, nfree = O.mapper(pats, [], ifaces)
    assert nfree == 117
    comp = np.full(nfree, -1)
    for k in range(7):
        for l in range(27):
            g = dofs[27 * k + l]
            assert comp[g] in (-1, local[k, l])
            comp[g] = local[k, l]
    assert comp.tolist() == gold["component_of_dof"]


def _transfer(tokens):
    parts = " ".join(tokens).split("|")
    rows, cols = (int(t) for t in parts[0].split())
    rp = np.array([int(t) for t in parts[1].split()])
    ci = np.array([int(t) for t in parts[2].split()], dtype=np.int64)
    va = np.array([float(t) for t in parts[3].split()])
    import scipy.sparse as sp
    return sp.csr_matrix((va, ci, rp), shape=(rows, cols))


@pytest.mark.parametrize("name,grid,p,nfine,levels", [("hier2d", 2, 2, 8, 3), ("hier3d", 0, 3, 4, 2)])
def test_grid_hierarchy_transfers_match_oracle(host_check, name, grid, p, nfine, levels):
    """gsGridHierarchy::buildByCoarsening's transfers (the C++ host:
    coarsened knots, 1D knot insertion, tensor product, free DOFs, take_first
    at interfaces) are bitwise the oracle's combined transfers
    (oc_transfer_multipatch) between the same levels, coarsest first."""
    for lvl in range(levels - 1):
        nc = nfine >> (levels - 1 - lvl)
        mk = (lambda n: problems.square_grid(grid, grid, n, p)) if grid else \
             (lambda n: problems.Problem([problems.box_patch(3, p, n)], problems.all_sides(3), []))
        coarse, fine = mk(nc), mk(2 * nc)
        dc, ncf = O.mapper(coarse.patches, coarse.dir_sides, coarse.ifaces)
        df, nff = O.mapper(fine.patches, fine.dir_sides, fine.ifaces)
        Po = O.transfer_multipatch(coarse.patches, dc, ncf, df, nff)
        P = _transfer(host_check[f"{name}_T{lvl}"])
        assert P.shape == Po.shape
        assert np.array_equal(P.indptr, Po.indptr) and np.array_equal(P.indices, Po.indices)
        assert np.array_equal(P.data, Po.data)
        deg, k = _knots(host_check[f"{name}_kv{lvl}"])
        assert deg == p and np.array_equal(k, problems.knots_uniform(nc, p))
