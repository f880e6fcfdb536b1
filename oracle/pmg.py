"""p-multigrid restated on the CPU: assembly with the geometry in its own
basis, mixed and lumped mass, the level hierarchy of
examples/pMultiGrid_example.cpp and its V-cycle with a restriction other
than P^T.

TEST INFRASTRUCTURE ONLY: a checker for the device p-multigrid; only
tests/ may import it.  Element data are formed per element from the C
oracle's 1D routines (Gauss rule, span search, de Boor values and
derivatives) and summed with numpy: the sums run in another order than on
the device, so the two agree to rounding, not bitwise.

Reference (G+Smo):
  levels and transfers  examples/pMultiGrid_example.cpp:296-435
  operators             examples/pMultiGrid_example.cpp:436-470
                        (gsCDRAssembler(mp, bases[i], ...): the geometry mp is
                        kept, the basis changes; quadrature from the basis,
                        gsQuadrature::get, src/gsAssembler/gsQuadrature.h:24-60)
  lumped mass           examples/pMultiGrid_example.cpp:803-826
  mixed mass            examples/pMultiGrid_example.cpp:855-880
  cycle                 src/gsMultiGrid/gsMultiGrid.hpp:148-199
"""
import numpy as np
import scipy.sparse as sp

from . import bspline_ders, find_span, gauss_legendre, mapper, transfer_multipatch


class Basis:
    """A tensor B-spline basis of one patch (degree per direction, open knot
    vectors) without geometry; zero control points so that the C oracle's
    mapper, which reads only the knots, accepts it."""

    def __init__(self, d, p, knots):
        self.d = d
        self.p = [int(v) for v in p]
        self.knots = [np.asarray(k, dtype=np.float64) for k in knots]
        self.ctrl = np.zeros((int(np.prod(self.nb())), d))
        self.kappa = None
        self.felem = None
        self.fkind = 1

    def nb(self):
        return [len(self.knots[k]) - self.p[k] - 1 for k in range(self.d)]

    def nel(self):
        return [len(self.knots[k]) - 2 * self.p[k] - 1 for k in range(self.d)]


def open_knots(knots, q):
    """Open knot vector of degree q over the breakpoints of knots (simple
    interior knots)."""
    bp = np.unique(np.asarray(knots, dtype=np.float64))
    return np.concatenate([[bp[0]] * q, bp, [bp[-1]] * q])


def coarsen_knots(knots, p):
    """Every other interior breakpoint removed (gsKnotVector::coarsen(1, 1)
    on a vector with simple interior knots)."""
    bp = np.unique(np.asarray(knots, dtype=np.float64))
    nel = len(bp) - 1
    if nel % 2 or nel < 2:
        raise ValueError("multigrid needs an even number of elements per direction on every level")
    return open_knots(bp[::2], p)


def _nodes(knots, p, e, nq):
    """Gauss nodes and weights mapped to element e (knot span p + e): the
    mapping of gsQuadRule::mapTo (src/gsAssembler/gsQuadRule.hpp:24-117)."""
    x, w = gauss_legendre(nq)
    a, b = knots[p + e], knots[p + e + 1]
    h = (b - a) / 2.0
    return h * (x + 1.0) + a, (0.5 if h == 0.0 else h) * w


def _source(fkind, d, x, felem, gel):
    if fkind == 1:
        return d * np.pi ** 2 * np.prod(np.sin(np.pi * np.asarray(x)))
    return 0.0 if felem is None else felem[gel]


def _element(U, G, ei, V=None):
    """Points of element ei of basis U (its q + 1 Gauss nodes per direction,
    direction 0 fastest) with weights w * |det J|, J^-1, det, x, the geometry
    element, U's values and parameter gradients, and V's values (V's element
    containing ei) when V is given."""
    d = U.d
    nodes, wts, NU, DU, NV, geo = [], [], [], [], [], []
    for k in range(d):
        u, w = _nodes(U.knots[k], U.p[k], ei[k], U.p[k] + 1)
        nodes.append(u)
        wts.append(w)
        tab = [bspline_ders(U.knots[k], U.p[k], U.p[k] + ei[k], uu, 1) for uu in u]
        NU.append(np.array([t[0] for t in tab]))
        DU.append(np.array([t[1] for t in tab]))
        gk = []
        for uu in u:
            s = find_span(G.knots[k], G.p[k], uu)
            gk.append((s, bspline_ders(G.knots[k], G.p[k], s, uu, 1)))
        geo.append(gk)
        if V is not None:
            mid = 0.5 * (U.knots[k][U.p[k] + ei[k]] + U.knots[k][U.p[k] + ei[k] + 1])
            vs = find_span(V.knots[k], V.p[k], mid)
            NV.append(np.array([bspline_ders(V.knots[k], V.p[k], vs, uu, 0)[0] for uu in u]))
    nbg = [len(G.knots[k]) - G.p[k] - 1 for k in range(d)]
    nelg = [len(G.knots[k]) - 2 * G.p[k] - 1 for k in range(d)]
    nq = [U.p[k] + 1 for k in range(d)]
    pts = []
    for qi in np.ndindex(*nq[::-1]):
        qi = qi[::-1]                       # direction 0 fastest
        x = np.zeros(d)
        J = np.zeros((d, d))
        for ci in np.ndindex(*[G.p[k] + 1 for k in range(d)]):
            m, stride = 0, 1
            v = 1.0
            g = np.ones(d)
            for k in range(d):
                s, t = geo[k][qi[k]]
                m += (s - G.p[k] + ci[k]) * stride
                stride *= nbg[k]
                v *= t[0][ci[k]]
                for j in range(d):
                    g[j] *= t[1][ci[k]] if j == k else t[0][ci[k]]
            x += G.ctrl[m] * v
            J += np.outer(G.ctrl[m], g)
        gel, stride = 0, 1
        for k in range(d):
            gel += (geo[k][qi[k]][0] - G.p[k]) * stride
            stride *= nelg[k]
        det = np.linalg.det(J)
        w = np.prod([wts[k][qi[k]] for k in range(d)])
        # tensor values/gradients of U's functions (local index, direction 0 fastest)
        nbu = [U.p[k] + 1 for k in range(d)]
        phi, grad = [], []
        for ai in np.ndindex(*nbu[::-1]):
            ai = ai[::-1]
            v = np.prod([NU[k][qi[k]][ai[k]] for k in range(d)])
            gr = np.array([np.prod([(DU[k] if k == j else NU[k])[qi[k]][ai[k]] for k in range(d)])
                           for j in range(d)])
            phi.append(v)
            grad.append(gr)
        psi = None
        if V is not None:
            nbv = [V.p[k] + 1 for k in range(d)]
            psi = np.array([np.prod([NV[k][qi[k]][bi[::-1][k]] for k in range(d)])
                            for bi in np.ndindex(*nbv[::-1])])
        pts.append(dict(w=w, det=det, meas=abs(det), Jinv=np.linalg.inv(J), x=x, gel=gel, phi=np.array(phi),
                        grad=np.array(grad), psi=psi))
    return pts


def _nb(B):
    return [len(B.knots[k]) - B.p[k] - 1 for k in range(B.d)]


def _nel(B):
    return [len(B.knots[k]) - 2 * B.p[k] - 1 for k in range(B.d)]


def _local_dofs(B, dofs, off, ei):
    """Global indices of the functions of element ei of B (local order)."""
    nb = _nb(B)
    out = []
    for ai in np.ndindex(*[B.p[k] + 1 for k in range(B.d)][::-1]):
        ai = ai[::-1]
        m, stride = 0, 1
        for k in range(B.d):
            m += (ei[k] + ai[k]) * stride
            stride *= nb[k]
        out.append(dofs[off + m])
    return np.array(out)


def _elements(B):
    for ei in np.ndindex(*_nel(B)[::-1]):
        yield ei[::-1]


def assemble_geo(bases, geoms, dofs, nfree, unit_f=False, matrix=True):
    """Stiffness (kappa of the geometry element at each point) and load (the
    geometry patch's source, or 1 with unit_f: the lumped mass) of the bases
    over the geometry patches, restricted to free DOFs (eliminated columns
    dropped, homogeneous data).  Returns (K scipy CSR or None, F)."""
    rows, cols, vals = [], [], []
    F = np.zeros(nfree)
    off = 0
    for B, G in zip(bases, geoms):
        kap = getattr(G, "kappa", None)
        for ei in _elements(B):
            pts = _element(B, G, ei)
            gl = _local_dofs(B, dofs, off, ei)
            nbl = len(gl)
            Ke = np.zeros((nbl, nbl))
            Fe = np.zeros(nbl)
            for q in pts:
                gx = q["grad"] @ q["Jinv"]             # rows: grad_x phi_a
                k = 1.0 if kap is None else kap[q["gel"]]
                Ke += (q["w"] * q["meas"] * k) * (gx @ gx.T)
                f = 1.0 if unit_f else _source(G.fkind, B.d, q["x"], getattr(G, "felem", None), q["gel"])
                Fe += q["w"] * (f * q["meas"]) * q["phi"]
            for a in range(nbl):
                if gl[a] >= nfree:
                    continue
                F[gl[a]] += Fe[a]
                if matrix:
                    for b in range(nbl):
                        if gl[b] < nfree:
                            rows.append(gl[a])
                            cols.append(gl[b])
                            vals.append(Ke[a, b])
        off += int(np.prod(_nb(B)))
    K = sp.csr_matrix((vals, (rows, cols)), shape=(nfree, nfree)) if matrix else None
    return K, F


def lumped_mass(bases, geoms, dofs, nfree):
    """L(i) = int phi_i dx with the basis' own quadrature."""
    return assemble_geo(bases, geoms, dofs, nfree, unit_f=True, matrix=False)[1]


def mixed_mass(ubases, vbases, geoms, dofs_u, nfree_u, dofs_v, nfree_v):
    """M(i, j) = int phi_i psi_j dx on U's elements with U's quadrature (rows:
    U's free DOFs, columns: V's)."""
    rows, cols, vals = [], [], []
    offu = offv = 0
    for U, V, G in zip(ubases, vbases, geoms):
        for ei in _elements(U):
            pts = _element(U, G, ei, V)
            gu = _local_dofs(U, dofs_u, offu, ei)
            # V's element containing ei
            ev = []
            for k in range(U.d):
                mid = 0.5 * (U.knots[k][U.p[k] + ei[k]] + U.knots[k][U.p[k] + ei[k] + 1])
                ev.append(find_span(V.knots[k], V.p[k], mid) - V.p[k])
            gv = _local_dofs(V, dofs_v, offv, ev)
            Me = np.zeros((len(gu), len(gv)))
            for q in pts:
                Me += (q["w"] * q["meas"]) * np.outer(q["phi"], q["psi"])
            for a in range(len(gu)):
                if gu[a] >= nfree_u:
                    continue
                for b in range(len(gv)):
                    if gv[b] < nfree_v:
                        rows.append(gu[a])
                        cols.append(gv[b])
                        vals.append(Me[a, b])
        offu += int(np.prod(_nb(U)))
        offv += int(np.prod(_nb(V)))
    return sp.csr_matrix((vals, (rows, cols)), shape=(nfree_u, nfree_v))


def default_coarsening(nlevels, p, projection=2):
    """The --Coarsening default of pMultiGrid_example.cpp:267-288."""
    if projection == 1:
        return "h" * (nlevels - 2) + "z"
    if nlevels > p:
        return "h" * (nlevels - p) + "z" * (p - 1)
    return "z" * (nlevels - 1)


def level_bases(patches, nlevels, coarsening, projection=2):
    """Bases of every level (index 0 = coarsest), each a list of Basis per
    patch; level nlevels - 1 carries the patches' own bases."""
    levels = [None] * nlevels
    levels[-1] = [Basis(P.d, P.p, P.knots) for P in patches]
    for l in range(nlevels - 2, -1, -1):
        t = coarsening[l]
        cur = []
        for B in levels[l + 1]:
            q = B.p[0] if t == "h" else (1 if projection == 1 else B.p[0] - 1)
            if q < 1:
                raise ValueError("degree cannot be reduced further")
            kn = [coarsen_knots(K, B.p[k]) if t != "p" else K for k, K in enumerate(B.knots)]
            cur.append(Basis(B.d, [q] * B.d, [open_knots(K, q) for K in kn]))
        levels[l] = cur
    return levels


def pmg_setup(patches, dir_sides, ifaces, A, nlevels, coarsening, projection=2, coarse_operator=2):
    """Operators As, prolongations Ps and restrictions Rs (index l: between
    levels l and l + 1) of the p-multigrid hierarchy, level nlevels - 1
    being A on the patches' own bases."""
    levels = level_bases(patches, nlevels, coarsening, projection)
    maps = [mapper(L, dir_sides, ifaces) for L in levels]
    As = [None] * nlevels
    Ps = [None] * (nlevels - 1)
    Rs = [None] * (nlevels - 1)
    As[-1] = sp.csr_matrix(A)
    for l in range(nlevels - 2, -1, -1):
        t = coarsening[l]
        (dc, nc), (df, nf) = maps[l], maps[l + 1]
        if t == "h":
            P = transfer_multipatch(levels[l], dc, nc, df, nf)
            R = P.T.tocsr()
        else:
            M = mixed_mass(levels[l + 1], levels[l], patches, df, nf, dc, nc)
            Lf = lumped_mass(levels[l + 1], patches, df, nf)
            Lc = lumped_mass(levels[l], patches, dc, nc)
            P = sp.diags(1.0 / Lf) @ M
            R = sp.diags(1.0 / Lc) @ M.T
        Ps[l], Rs[l] = P.tocsr(), R.tocsr()
        if t == "h" and coarse_operator == 2:
            As[l] = (Rs[l] @ As[l + 1] @ Ps[l]).tocsr()
        else:
            As[l] = assemble_geo(levels[l], patches, dc, nc)[0]
    return As, Ps, Rs, levels, maps


def vcycle(As, Ps, Rs, smoothers, npre=1, npost=1, cgc=1.0):
    """gsMultiGridOp's V-cycle from a zero start with restrictions Rs
    (src/gsMultiGrid/gsMultiGrid.hpp:148-199): smoothers[l](b, x, post)
    update x in place; As[0] is solved exactly.  Returns the apply
    function."""
    import scipy.sparse.linalg as spla
    coarse = spla.factorized(sp.csc_matrix(As[0]))

    def step(l, b, x):
        if l == 0:
            x[:] = coarse(b)
            return
        for _ in range(npre):
            smoothers[l](b, x, False)
        cr = Rs[l - 1] @ (b - As[l] @ x)
        cc = np.zeros(As[l - 1].shape[0])
        step(l - 1, cr, cc)
        x += cgc * (Ps[l - 1] @ cc)
        for _ in range(npost):
            smoothers[l](b, x, True)

    def apply(r):
        x = np.zeros(As[-1].shape[0])
        step(len(As) - 1, np.asarray(r, dtype=np.float64), x)
        return x
    return apply


def jacobi_smoothers(As, damping=1.0):
    """Damped Jacobi x += tau D^-1 (b - A x) on every level above the
    coarsest (gsJacobiOp::step, src/gsSolver/gsSimplePreconditioners.h:178-202)."""
    out = [None]
    for A in As[1:]:
        dinv = 1.0 / A.diagonal()

        def sm(b, x, post, A=A, dinv=dinv):
            x += damping * dinv * (b - A @ x)
        out.append(sm)
    return out
