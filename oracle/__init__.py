"""ctypes bindings of the CPU FP64 oracle (oracle/libgsx_oracle.so).

TEST INFRASTRUCTURE ONLY: the oracle restates the reference algorithm on the
CPU and is the parity checker.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this module; the product path
(gismo_amd, libgsx) never does.
"""
import ctypes as C
import os
import subprocess
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libgsx_oracle.so")
_CPUREF_PATH = os.path.join(_HERE, "libgsx_cpuref.so")
_lib = None
_cpuref = None

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class OcPatch(C.Structure):
    _fields_ = [
        ("d", C.c_int),
        ("p", C.c_int * 3),
        ("nknots", C.c_int * 3),
        ("knots", C.POINTER(C.c_double) * 3),
        ("ctrl", C.POINTER(C.c_double)),
        ("kappa", C.POINTER(C.c_double)),
        ("felem", C.POINTER(C.c_double)),
        ("fkind", C.c_int),
    ]


class OcCsr(C.Structure):
    _fields_ = [
        ("n", C.c_int),
        ("nnz", C.c_int64),
        ("rowptr", C.POINTER(C.c_int)),
        ("col", C.POINTER(C.c_int)),
        ("val", C.POINTER(C.c_double)),
    ]


def build():
    """Compile the oracle (gcc) in place."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.oc_gauss_legendre.argtypes = [C.c_int, _f64p, _f64p]
        L.oc_knots_uniform.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _f64p]
        L.oc_find_span.argtypes = [_f64p, C.c_int, C.c_int, C.c_double]
        L.oc_find_span.restype = C.c_int
        L.oc_bspline_ders.argtypes = [_f64p, C.c_int, C.c_int, C.c_double, C.c_int, _f64p]
        L.oc_element_geometry.argtypes = [C.POINTER(OcPatch), C.c_int, _f64p, _f64p, _f64p, _f64p]
        L.oc_element_geometry.restype = C.c_int
        L.oc_mapper_build.argtypes = [C.c_int, C.POINTER(OcPatch), C.c_int, _i32p, C.c_int, _i32p, _i32p]
        L.oc_mapper_build.restype = C.c_int
        L.oc_assemble.argtypes = [C.c_int, C.POINTER(OcPatch), _i32p, C.c_int, C.c_int, _f64p]
        L.oc_assemble.restype = C.POINTER(OcCsr)
        L.oc_csr_free.argtypes = [C.POINTER(OcCsr)]
        L.oc_assemble_rows.argtypes = [C.POINTER(OcPatch), _i32p, C.c_int, C.c_int, _i32p, C.c_int,
                                       _i32p, _i32p, _f64p, _f64p]
        L.oc_cg_solve.argtypes = [C.POINTER(OcCsr), _f64p, _f64p, C.c_double, C.c_int, C.c_int,
                                  C.c_double, C.c_void_p, C.POINTER(C.c_double), C.c_void_p]
        L.oc_cg_solve.restype = C.c_int
        L.oc_transfer_1d.argtypes = [_f64p, C.c_int, C.c_int, _f64p]
        L.oc_transfer_1d.restype = C.c_int
        L.oc_transfer_multipatch.argtypes = [C.c_int, C.POINTER(OcPatch), _i32p, C.c_int, _i32p, C.c_int]
        L.oc_transfer_multipatch.restype = C.POINTER(OcCsr)
        L.oc_mg_create.argtypes = [C.POINTER(OcCsr), C.c_int, C.POINTER(C.POINTER(OcCsr)),
                                   C.c_int, C.c_int, C.c_double, C.c_double]
        L.oc_mg_create.restype = C.c_void_p
        L.oc_mg_apply.argtypes = [C.c_void_p, _f64p, _f64p]
        L.oc_mg_set_smoother.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.oc_mg_set_smoother.restype = None
        L.oc_gs_sweep.argtypes = [C.c_void_p, _f64p, _f64p, C.c_int]
        L.oc_gs_sweep.restype = None
        L.oc_mg_lmax.argtypes = [C.c_void_p, C.c_int]
        L.oc_mg_lmax.restype = C.c_double
        L.oc_mg_matrix.argtypes = [C.c_void_p, C.c_int]
        L.oc_mg_matrix.restype = C.POINTER(OcCsr)
        L.oc_mg_free.argtypes = [C.c_void_p]
        L.oc_assemble_bc.argtypes = [C.c_int, C.POINTER(OcPatch), _i32p, C.c_int, C.c_int, C.c_void_p, _f64p]
        L.oc_assemble_bc.restype = C.POINTER(OcCsr)
        L.oc_dirichlet_l2.argtypes = [C.c_int, C.POINTER(OcPatch), _i32p, C.c_int, C.c_int, C.c_int,
                                      C.c_void_p, _f64p]
        L.oc_dirichlet_l2.restype = C.c_int
        L.oc_neumann.argtypes = [C.c_int, C.POINTER(OcPatch), _i32p, C.c_int, C.c_int, C.c_void_p, _f64p]
        L.oc_neumann.restype = None
        L.oc_l2_error_bc.argtypes = [C.c_int, C.POINTER(OcPatch), _i32p, C.c_int, _f64p, C.c_void_p]
        L.oc_l2_error_bc.restype = C.c_double
        L.oc_h1_error_bc.argtypes = [C.c_int, C.POINTER(OcPatch), _i32p, C.c_int, _f64p, C.c_void_p]
        L.oc_h1_error_bc.restype = C.c_double
        L.oc_l2_error.argtypes = [C.c_int, C.POINTER(OcPatch), _i32p, C.c_int, _f64p]
        L.oc_l2_error.restype = C.c_double
        L.oc_h1_error.argtypes = [C.c_int, C.POINTER(OcPatch), _i32p, C.c_int, _f64p]
        L.oc_h1_error.restype = C.c_double
        _lib = L
    return _lib


# ---------------------------------------------------------------------------
# numpy-friendly wrappers
# ---------------------------------------------------------------------------

def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().oc_gauss_legendre(n, x, w)
    return x, w


def knots_uniform(nelem, p, a=0.0, b=1.0):
    k = np.zeros(nelem + 2 * p + 1)
    lib().oc_knots_uniform(nelem, p, a, b, k)
    return k


def find_span(knots, p, u):
    knots = np.ascontiguousarray(knots, dtype=np.float64)
    return lib().oc_find_span(knots, len(knots), p, float(u))


def bspline_ders(knots, p, span, u, nder=1):
    knots = np.ascontiguousarray(knots, dtype=np.float64)
    out = np.zeros((nder + 1) * (p + 1))
    lib().oc_bspline_ders(knots, p, span, float(u), nder, out)
    return out.reshape(nder + 1, p + 1)


class Patches:
    """Keeps the numpy arrays behind an array of OcPatch alive."""

    def __init__(self, patch_list):
        # patch_list: sequence of objects with attributes d, p, knots, ctrl,
        # kappa (or None), felem (or None), fkind
        self._keep = []
        n = len(patch_list)
        self.arr = (OcPatch * n)()
        for i, P in enumerate(patch_list):
            o = self.arr[i]
            o.d = P.d
            for k in range(3):
                if k < P.d:
                    kn = np.ascontiguousarray(P.knots[k], dtype=np.float64)
                    self._keep.append(kn)
                    o.p[k] = int(P.p[k])
                    o.nknots[k] = len(kn)
                    o.knots[k] = kn.ctypes.data_as(C.POINTER(C.c_double))
                else:
                    o.p[k] = 0
                    o.nknots[k] = 0
            ctrl = np.ascontiguousarray(P.ctrl, dtype=np.float64)
            self._keep.append(ctrl)
            o.ctrl = ctrl.ctypes.data_as(C.POINTER(C.c_double))
            for name in ("kappa", "felem"):
                a = getattr(P, name, None)
                if a is None:
                    setattr(o, name, C.POINTER(C.c_double)())
                else:
                    a = np.ascontiguousarray(a, dtype=np.float64)
                    self._keep.append(a)
                    setattr(o, name, a.ctypes.data_as(C.POINTER(C.c_double)))
            o.fkind = int(P.fkind)
        self.n = n


def element_geometry(patch, elem):
    P = Patches([patch])
    nq = 1
    for k in range(patch.d):
        nq *= patch.p[k] + 1
    d = patch.d
    xq = np.zeros(nq * d)
    jac = np.zeros(nq * d * d)
    meas = np.zeros(nq)
    wq = np.zeros(nq)
    lib().oc_element_geometry(P.arr, elem, xq, jac, meas, wq)
    return xq.reshape(nq, d), jac.reshape(nq, d, d), meas, wq


def cpuref_lib():
    """The oracle built for timing (-O3, AVX2/FMA, OpenMP): libgsx_cpuref.so."""
    global _cpuref
    if _cpuref is None:
        if not os.path.exists(_CPUREF_PATH):
            build()
        L = C.CDLL(_CPUREF_PATH)
        L.oc_mapper_build.argtypes = [C.c_int, C.POINTER(OcPatch), C.c_int, _i32p, C.c_int, _i32p, _i32p]
        L.oc_mapper_build.restype = C.c_int
        L.oc_assemble_par.argtypes = [C.c_int, C.POINTER(OcPatch), _i32p, C.c_int, C.c_int, _f64p, C.c_int]
        L.oc_assemble_par.restype = C.POINTER(OcCsr)
        L.oc_cg_solve.argtypes = [C.POINTER(OcCsr), _f64p, _f64p, C.c_double, C.c_int, C.c_int,
                                  C.c_double, C.c_void_p, C.POINTER(C.c_double), C.c_void_p]
        L.oc_cg_solve.restype = C.c_int
        L.oc_csr_free.argtypes = [C.POINTER(OcCsr)]
        _cpuref = L
    return _cpuref


def baseline_assemble_solve(patch_list, dir_sides, ifaces, tol, nthreads, maxit=5000):
    """CPU baseline run (timing only): DOF map, the reference-style threaded
    assembly on nthreads threads and Jacobi-CG to tol (serial, as the
    reference's column-major sparse product is), all from libgsx_cpuref.so.
    Returns (assembly seconds, solve seconds, CG iterations, free DOFs)."""
    L = cpuref_lib()
    P = Patches(patch_list)
    total = sum(int(np.prod([len(q.knots[k]) - q.p[k] - 1 for k in range(q.d)])) for q in patch_list)
    ds = np.ascontiguousarray(np.asarray(dir_sides, dtype=np.int32).reshape(-1))
    ifs = np.ascontiguousarray(np.asarray(ifaces, dtype=np.int32).reshape(-1))
    ds = ds if ds.size else np.zeros(2, np.int32)
    ifs = ifs if ifs.size else np.zeros(8, np.int32)
    dofs = np.zeros(total, np.int32)
    t0 = time.perf_counter()
    nfree = L.oc_mapper_build(P.n, P.arr, len(dir_sides), ds, len(ifaces), ifs, dofs)
    rhs = np.zeros(nfree)
    ptr = L.oc_assemble_par(P.n, P.arr, dofs, nfree, 0, rhs, nthreads)
    t1 = time.perf_counter()
    x = np.zeros(nfree)
    err = C.c_double(0)
    it = L.oc_cg_solve(ptr, rhs, x, tol, maxit, 1, 1.0, None, C.byref(err), None)
    t2 = time.perf_counter()
    L.oc_csr_free(ptr)
    return t1 - t0, t2 - t1, it, nfree


def mapper(patch_list, dir_sides, ifaces):
    """dir_sides: list of (patch, side); ifaces: list of 8-int records.
    Returns (dofs concatenated per patch, nfree)."""
    P = Patches(patch_list)
    total = 0
    for Pp in patch_list:
        s = 1
        for k in range(Pp.d):
            s *= len(Pp.knots[k]) - Pp.p[k] - 1
        total += s
    ds = np.ascontiguousarray(np.asarray(dir_sides, dtype=np.int32).reshape(-1))
    if ds.size == 0:
        ds = np.zeros(2, np.int32)
    ifs = np.ascontiguousarray(np.asarray(ifaces, dtype=np.int32).reshape(-1))
    if ifs.size == 0:
        ifs = np.zeros(8, np.int32)
    out = np.zeros(total, np.int32)
    nfree = lib().oc_mapper_build(P.n, P.arr, len(dir_sides), ds, len(ifaces), ifs, out)
    return out, nfree


def _csr_to_scipy(ptr, free=True):
    import scipy.sparse as sp
    A = ptr.contents
    n, nnz = A.n, A.nnz
    rp = np.ctypeslib.as_array(A.rowptr, shape=(n + 1,)).copy()
    if nnz:
        ci = np.ctypeslib.as_array(A.col, shape=(nnz,)).copy()
        va = np.ctypeslib.as_array(A.val, shape=(nnz,)).copy()
    else:
        ci = np.zeros(0, np.int32)
        va = np.zeros(0)
    ncols = n
    M = sp.csr_matrix((va, ci, rp), shape=(n, max(ncols, int(ci.max()) + 1 if nnz else n)))
    if free:
        lib().oc_csr_free(ptr)
    return M


def assemble(patch_list, dofs, nfree, mode=0, fixed=None):
    """Returns (scipy CSR matrix with sorted indices, rhs); `fixed` (values
    of the eliminated DOFs) moves K_ij g_j to the rhs."""
    P = Patches(patch_list)
    rhs = np.zeros(nfree)
    dofs = np.ascontiguousarray(dofs, dtype=np.int32)
    if fixed is None:
        ptr = lib().oc_assemble(P.n, P.arr, dofs, nfree, mode, rhs)
    else:
        fx = np.ascontiguousarray(fixed, np.float64)
        ptr = lib().oc_assemble_bc(P.n, P.arr, dofs, nfree, mode, fx.ctypes.data_as(C.c_void_p), rhs)
    return _csr_to_scipy(ptr), rhs


class OcBc(C.Structure):
    _fields_ = [("patch", C.c_int), ("side", C.c_int), ("type", C.c_int), ("kind", C.c_int),
                ("value", C.c_double * 4)]


def _bc_array(bcs):
    arr = (OcBc * max(1, len(bcs)))()
    for i, (patch, side, typ, kind, vals) in enumerate(bcs):
        arr[i].patch, arr[i].side, arr[i].type, arr[i].kind = patch, side, typ, kind
        for k, v in enumerate(list(vals)[:4]):
            arr[i].value[k] = v
    return arr


def dirichlet_values(patch_list, dofs, nfree, nfix, bcs):
    """Fixed values (nfix) by the L2 projection over the Dirichlet records
    bcs = [(patch, side, 1, kind, values), ...] (every Dirichlet side)."""
    P = Patches(patch_list)
    out = np.zeros(max(1, nfix))
    arr = _bc_array(bcs)
    lib().oc_dirichlet_l2(P.n, P.arr, np.ascontiguousarray(dofs, np.int32), nfree, nfix, len(bcs),
                          C.cast(arr, C.c_void_p), out)
    return out[:nfix]


def neumann_load(patch_list, dofs, nfree, bcs):
    """Neumann boundary load (nfree) of the records with type 2."""
    P = Patches(patch_list)
    out = np.zeros(max(1, nfree))
    arr = _bc_array(bcs)
    lib().oc_neumann(P.n, P.arr, np.ascontiguousarray(dofs, np.int32), nfree, len(bcs),
                     C.cast(arr, C.c_void_p), out)
    return out[:nfree]


def assemble_rows(patch, dofs, nfree, rows, maxc=1024):
    """Full rows (cols, vals) and rhs entries of a single-patch problem for the
    requested free DOFs, from the elements in their supports."""
    P = Patches([patch])
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    n = rows.size
    counts = np.zeros(n, np.int32)
    cols = np.zeros(n * maxc, np.int32)
    vals = np.zeros(n * maxc)
    rhs = np.zeros(n)
    lib().oc_assemble_rows(P.arr, np.ascontiguousarray(dofs, np.int32), nfree, n, rows, maxc,
                           counts, cols, vals, rhs)
    out = []
    for t in range(n):
        c = int(counts[t])
        out.append((cols[t * maxc:t * maxc + c].copy(), vals[t * maxc:t * maxc + c].copy()))
    return out, rhs


class _CsrHolder:
    def __init__(self, M):
        M = M.tocsr()
        M.sort_indices()
        self.rp = np.ascontiguousarray(M.indptr, dtype=np.int32)
        self.ci = np.ascontiguousarray(M.indices, dtype=np.int32)
        self.va = np.ascontiguousarray(M.data, dtype=np.float64)
        self.s = OcCsr()
        self.s.n = M.shape[0]
        self.s.nnz = M.nnz
        self.s.rowptr = self.rp.ctypes.data_as(C.POINTER(C.c_int))
        self.s.col = self.ci.ctypes.data_as(C.POINTER(C.c_int))
        self.s.val = self.va.ctypes.data_as(C.POINTER(C.c_double))


def cg_solve(A, b, x0=None, tol=1e-10, maxit=1000, precond="jacobi", tau=1.0, mg=None,
             history=False):
    H = _CsrHolder(A)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, dtype=np.float64).copy()
    err = C.c_double(0)
    pc = {"none": 0, "jacobi": 1, "mg": 2}[precond]
    hist = np.zeros(maxit + 1) if history else None
    it = lib().oc_cg_solve(C.byref(H.s), b, x, tol, maxit, pc, tau,
                           mg.handle if mg is not None else None, C.byref(err),
                           hist.ctypes.data_as(C.c_void_p) if history else None)
    if history:
        return x, it, err.value, hist[: it + 1]
    return x, it, err.value


def gs_sweep(A, b, x, reverse=False):
    """One Gauss-Seidel sweep on x (a copy is returned), rows ascending or
    descending (gsSimplePreconditioners.hpp:21-68)."""
    H = _CsrHolder(A)
    x = np.ascontiguousarray(x, dtype=np.float64).copy()
    lib().oc_gs_sweep(C.byref(H.s), np.ascontiguousarray(b, dtype=np.float64), x, 1 if reverse else 0)
    return x


def transfer_1d(knots, p):
    knots = np.ascontiguousarray(knots, dtype=np.float64)
    nc = len(knots) - p - 1
    T = np.zeros((2 * nc + 8) * nc)
    nf = lib().oc_transfer_1d(knots, len(knots), p, T)
    return T[: nf * nc].reshape(nf, nc)


def transfer_multipatch(coarse_patches, dofs_c, nfree_c, dofs_f, nfree_f):
    P = Patches(coarse_patches)
    ptr = lib().oc_transfer_multipatch(P.n, P.arr, np.ascontiguousarray(dofs_c, np.int32), nfree_c,
                                       np.ascontiguousarray(dofs_f, np.int32), nfree_f)
    import scipy.sparse as sp
    M = _csr_to_scipy(ptr)
    return sp.csr_matrix((M.data, M.indices, M.indptr), shape=(nfree_f, nfree_c))


class Multigrid:
    SMOOTHERS = {"jacobi": 0, "richardson": 1, "chebyshev": 2, "gauss_seidel": 3}

    def __init__(self, A_fine, transfers, npre=1, npost=1, damping=1.0, cgc_damping=1.0, smoother="jacobi",
                 smoother_param=30.0):
        self._A = _CsrHolder(A_fine)
        self._P = [_CsrHolder(P) for P in transfers]
        arr = (C.POINTER(OcCsr) * max(1, len(self._P)))()
        for i, h in enumerate(self._P):
            arr[i] = C.pointer(h.s)
        self._arr = arr
        self.nlev = len(transfers) + 1
        self.handle = lib().oc_mg_create(C.byref(self._A.s), self.nlev, arr, npre, npost,
                                         damping, cgc_damping)
        if smoother != "jacobi":
            lib().oc_mg_set_smoother(self.handle, self.SMOOTHERS[smoother], smoother_param)

    def lmax(self, lvl):
        """Chebyshev bound of level lvl (1.1 x power-iteration lambda_max(D^-1 A))."""
        return lib().oc_mg_lmax(self.handle, lvl)

    def apply(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros_like(r)
        lib().oc_mg_apply(self.handle, r, z)
        return z

    def matrix(self, lvl):
        return _csr_to_scipy(lib().oc_mg_matrix(self.handle, lvl), free=False)

    def __del__(self):
        try:
            lib().oc_mg_free(self.handle)
        except Exception:
            pass


def _fixed_ptr(fixed):
    if fixed is None:
        return None, None
    fx = np.ascontiguousarray(fixed, np.float64)
    return fx, fx.ctypes.data_as(C.c_void_p)


def l2_error(patch_list, dofs, nfree, sol, fixed=None):
    P = Patches(patch_list)
    fx, fp = _fixed_ptr(fixed)
    return lib().oc_l2_error_bc(P.n, P.arr, np.ascontiguousarray(dofs, np.int32), nfree,
                                np.ascontiguousarray(sol, np.float64), fp)


def h1_error(patch_list, dofs, nfree, sol, fixed=None):
    """H1 seminorm error of the discrete solution against prod sin(pi x_r)."""
    P = Patches(patch_list)
    fx, fp = _fixed_ptr(fixed)
    return lib().oc_h1_error_bc(P.n, P.arr, np.ascontiguousarray(dofs, np.int32), nfree,
                                np.ascontiguousarray(sol, np.float64), fp)


# ---------------------------------------------------------------------------
# Fast diagonalization (gsPatchPreconditionersCreator, src/gsSolver/
# gsPatchPreconditionersCreator.hpp) restated with numpy/scipy
# ---------------------------------------------------------------------------

def tensor_1d(knots, p, lo, hi, beta=1.0):
    """Univariate stiffness (times beta) and mass of one knot vector on the
    parameter line, p+1 Gauss points per span (assembleStiffness /
    assembleMass, gsPatchPreconditionersCreator.hpp:70-96), restricted to
    the functions lo..hi (eliminateDirichlet1D, :43-66)."""
    knots = np.ascontiguousarray(knots, dtype=np.float64)
    n = len(knots) - p - 1
    gx, gw = gauss_legendre(p + 1)
    K = np.zeros((n, n))
    M = np.zeros((n, n))
    for span in range(p, n):
        a, b = knots[span], knots[span + 1]
        if not b > a:
            continue
        h = 0.5 * (b - a)
        for x, w in zip(gx, gw):
            B = bspline_ders(knots, p, span, a + h * (x + 1.0), 1)
            sl = slice(span - p, span + 1)
            M[sl, sl] += h * w * np.outer(B[0], B[0])
            K[sl, sl] += h * w * np.outer(B[1], B[1])
    return beta * K[lo:hi + 1, lo:hi + 1], M[lo:hi + 1, lo:hi + 1]


def _free_range(patch, k, dir_sides):
    nb = len(patch.knots[k]) - patch.p[k] - 1
    lo = 1 if (2 * k + 1) in dir_sides else 0
    hi = nb - 2 if (2 * k + 2) in dir_sides else nb - 1
    return lo, hi


def kron_sum(patch, dir_sides, alpha=0.0, beta=1.0):
    """alpha M + beta sum_k M x .. x K_k x .. x M over the free DOFs of a
    single patch (direction 0 fastest: the last Kronecker factor), the
    matrix stiffnessMatrix(basis, bc, opt, alpha, beta) assembles
    (gsPatchPreconditionersCreator.hpp:184-220)."""
    import scipy.sparse as sp
    mats = [tensor_1d(patch.knots[k], patch.p[k], *_free_range(patch, k, dir_sides), beta)
            for k in range(patch.d)]
    d = patch.d

    def kron_all(pick):
        out = sp.csr_matrix(np.ones((1, 1)))
        for k in reversed(range(d)):
            out = sp.kron(out, sp.csr_matrix(pick(k)), format="csr")
        return out
    A = kron_all(lambda k: mats[k][1]) * alpha
    for j in range(d):
        A = A + kron_all(lambda k: mats[k][0] if k == j else mats[k][1])
    return A.tocsr()


def fd_operator(patch, dir_sides, alpha=0.0, beta=1.0):
    """fastDiagonalizationOp (gsPatchPreconditionersCreator.hpp:258-348):
    per direction K_k q = lambda M_k q with Q^T M Q = I (scipy's eigh), and
    r -> (Q x .. x Q) diag(1/(alpha + sum lambda)) (Q x .. x Q)^T r over the
    free DOFs (direction 0 fastest).  Returns the apply function."""
    import scipy.linalg as sla
    d = patch.d
    Qs, lams, ms = [], [], []
    for k in range(d):
        lo, hi = _free_range(patch, k, dir_sides)
        K, M = tensor_1d(patch.knots[k], patch.p[k], lo, hi, beta)
        lam, Q = sla.eigh(K, M)
        Qs.append(Q)
        lams.append(lam)
        ms.append(hi - lo + 1)
    shape = tuple(ms[::-1])               # axes: direction d-1 .. 0
    den = np.full(shape, float(alpha))
    for k in range(d):
        ax = d - 1 - k
        den = den + lams[k].reshape([-1 if a == ax else 1 for a in range(d)])

    def mode(X, M, k):
        ax = d - 1 - k
        return np.moveaxis(np.tensordot(M, X, axes=([1], [ax])), 0, ax)

    def apply(r):
        X = np.asarray(r, dtype=np.float64).reshape(shape)
        for k in range(d):
            X = mode(X, Qs[k].T, k)
        X = X / den
        for k in range(d):
            X = mode(X, Qs[k], k)
        return X.ravel()
    return apply


def tilde_oneside(knots, p, left, dirichlet):
    """tildeSpaceBasis_oneside (gsPatchPreconditionersCreator.hpp:368-414,
    odd derivatives): D holds h^i times the i-th derivatives (i = 1, 3, ..
    < p) at the end of the p - b functions next to it (the function that
    the Dirichlet condition removes, b = 1, and the (p+1)-th function from
    the end left out); the tilde basis is the right singular vectors of D
    for the (p+1)/2 - b smallest singular values (its null space), the
    complement the rest.  Returns (tilde, compl) as (p-b) x cols arrays."""
    b = int(bool(dirichlet))
    m = p - b
    if m <= 0:
        return np.zeros((0, 0)), np.zeros((0, 0))
    knots = np.asarray(knots, dtype=np.float64)
    h = np.max(np.diff(np.unique(knots)))
    u = knots[0] if left else knots[-1]
    ders = bspline_ders(knots, p, find_span(knots, p, u), u, p - 1)
    D = np.zeros((m, m))
    off = b if left else 1
    for i in range(1, p, 2):
        D[i - b, :] = h ** i * ders[i, off:off + m]
    V = np.linalg.svd(D)[2].T            # columns by descending singular value
    nt = (p + 1) // 2 - b
    return V[:, m - nt:], V[:, :m - nt]


def tilde_space(knots, p, dir_west, dir_east):
    """tildeSpaceBasis (:416-512): B_tilde (n x (n_L + n_I + n_R)) with the
    left block, the identity on the interior functions and the right block;
    B_l2compl (n x (n_cL + n_cR)) with the two complement blocks; with no
    interior functions left, B_tilde is empty and B_l2compl the identity."""
    nb = len(knots) - p - 1
    n = nb - int(bool(dir_west)) - int(bool(dir_east))
    bL, cL = tilde_oneside(knots, p, True, dir_west)
    bR, cR = tilde_oneside(knots, p, False, dir_east)
    mL, mR = bL.shape[0], bR.shape[0]
    nL, nR, ncL, ncR = bL.shape[1], bR.shape[1], cL.shape[1], cR.shape[1]
    nI = n - nL - nR - ncL - ncR
    if nI <= 0:
        return np.zeros((n, 0)), np.eye(n)
    Bt = np.zeros((n, nL + nI + nR))
    Bt[:mL, :nL] = bL
    Bt[mL:mL + nI, nL:nL + nI] = np.eye(nI)
    Bt[mL + nI:, nL + nI:] = bR
    Bc = np.zeros((n, ncL + ncR))
    Bc[:mL, :ncL] = cL
    Bc[mL + nI:, ncL:] = cR
    return Bt, Bc


def _kron_dirs(mats):
    """Kronecker product of per-direction matrices, direction 0 fastest."""
    out = np.ones((1, 1))
    for A in reversed(mats):
        out = np.kron(out, A)
    return out


def scms_matrix(patch, dir_sides, sigma=0.12, alpha=0.0, beta=1.0):
    """subspaceCorrectedMassSmootherOp (gsPatchPreconditionersCreator.hpp:
    574-758) as a dense matrix over the free DOFs of a single patch: per
    subspace type J (directions with the complement B_compl = M^-1
    B_l2compl, the others with B_tilde) the transfer T_J, the correction
    C_J = s (x)_k N_k + sum_{j in J} (x)_k (K_compl at j, N_k elsewhere)
    with N_k = B_tilde^T M B_tilde (interior directions) or B_compl^T M
    B_compl (directions in J) and s = alpha + beta (d - |J|) / (sigma h^2),
    h the smallest knot span; T_J C_J^-1 T_J^T summed over the types whose
    transfer has columns.  C_J is built and solved as it stands (no
    eigen-decomposition)."""
    d = patch.d
    h = min(np.min(np.diff(np.unique(np.asarray(patch.knots[k], dtype=np.float64)))) for k in range(d))
    Bt, Bc, Mt, Mc, Kc = [], [], [], [], []
    for k in range(d):
        lo, hi = _free_range(patch, k, dir_sides)
        nb = len(patch.knots[k]) - patch.p[k] - 1
        K, M = tensor_1d(patch.knots[k], patch.p[k], lo, hi, beta)
        bt, bl = tilde_space(patch.knots[k], patch.p[k], lo == 1, hi == nb - 2)
        bc = np.linalg.solve(M, bl)
        Bt.append(bt)
        Bc.append(bc)
        Mt.append(bt.T @ M @ bt)
        Mc.append(bc.T @ M @ bc)
        Kc.append(bc.T @ K @ bc)
    N = int(np.prod([B.shape[0] for B in Bt]))
    out = np.zeros((N, N))
    for J in range(1 << d):
        inJ = [bool((J >> k) & 1) for k in range(d)]
        T = [Bc[k] if inJ[k] else Bt[k] for k in range(d)]
        if any(t.shape[1] == 0 for t in T):
            continue
        s = alpha + beta * (d - sum(inJ)) / (sigma * h * h)
        Nk = [Mc[k] if inJ[k] else Mt[k] for k in range(d)]
        Cm = s * _kron_dirs(Nk)
        for j in range(d):
            if inJ[j]:
                Cm = Cm + _kron_dirs([Kc[k] if k == j else Nk[k] for k in range(d)])
        Tf = _kron_dirs(T)
        out += Tf @ np.linalg.solve(Cm, Tf.T)
    return out


def scms_damping(A, S, steps=10):
    """1 / lambda_max(S A) by estimateLargestEigenvalueOfPreconditionedSystem
    (src/gsSolver/gsPreconditioner.h:121-132) for the smoother x += S (b -
    A x): normalise, apply A, one step from zero (x = S (A x)), ten times;
    lambda = ||x||.  The reference starts from a random vector; here v0_i =
    1 / (1 + i mod 5)."""
    n = A.shape[0]
    x = 1.0 / (1.0 + (np.arange(n) % 5))
    for _ in range(steps):
        x = x / np.sqrt(x @ x)
        x = S @ (A @ x)
    return 1.0 / np.sqrt(x @ x)


def scms_smoother_levels(level_patches, As, sigma=0.12):
    """setupSubspaceCorrectedMassSmoother (examples/multiGrid_example.cpp:
    407-560) for one patch with every side Dirichlet, where the interior is
    the only component with free functions: per level l >= 1 the dense
    smoother S_l (scms_matrix, alpha = 0, beta = 1) and tau_l = scms_damping;
    the first level with l > nlev - 5 and more elements than the degree in
    every direction keeps its tau for the finer levels (:538-553).  Returns
    [None, (S_1, tau_1), ...]."""
    nlev = len(As)
    out = [None]
    saved = 0.0
    for l in range(1, nlev):
        P = level_patches[l]
        S = scms_matrix(P, list(range(1, 2 * P.d + 1)), sigma=sigma)
        if saved == 0.0:
            tau = scms_damping(As[l], S)
            if l > nlev - 5 and all(len(np.unique(P.knots[k])) - 1 > P.p[k] for k in range(P.d)):
                saved = tau
        else:
            tau = saved
        out.append((S, tau))
    return out


def scms_components(patch_list, dofs, nfree, combine_corners=True):
    """The components of setupSubspaceCorrectedMassSmoother
    (examples/multiGrid_example.cpp:407-567) with conforming interfaces: the
    3^d sub-boxes of every patch (gsBoxTopology::allComponents,
    src/gsCore/gsBoxTopology.cpp:294-340), those of glued patches with the
    same corner functions being one component, all corners one when
    combine_corners; a component's DOFs are its free functions interior to
    the sub-box in its open directions (componentBasis_withIndices with
    no_lower, src/gsCore/gsMultiBasis.hpp:262-328), first patch first.
    Returns [(dim, idx, box)] ordered by dimension (highest first), box =
    (patch, open directions) for patch interiors and the faces of 3D
    patches (the components with a smoother), else None; empty components
    are kept (idx of length 0)."""
    import itertools
    groups, order, offs, off = {}, [], [], 0
    for P in patch_list:
        offs.append(off)
        off += int(np.prod([len(P.knots[k]) - P.p[k] - 1 for k in range(P.d)]))
    for k, P in enumerate(patch_list):
        d = P.d
        nb = [len(P.knots[j]) - P.p[j] - 1 for j in range(d)]

        def flat(ii):
            m, stride = 0, 1
            for j in range(d):
                m += ii[j] * stride
                stride *= nb[j]
            return offs[k] + m
        for loc in itertools.product((0, 1, 2), repeat=d):      # first index, interior, last index
            dim = sum(v == 1 for v in loc)
            corners = frozenset(int(dofs[flat([nb[j] - 1 if cn[j] else 0 for j in range(d)])])
                                for cn in itertools.product((0, 1), repeat=d)
                                if not any((loc[j] == 0 and cn[j]) or (loc[j] == 2 and not cn[j]) for j in range(d)))
            if dim == d:
                key = (dim, ("patch", k))
            elif dim == 0 and combine_corners:
                key = (0, "corners")
            else:
                key = (dim, corners)
            if key not in groups:
                dirs = [j for j in range(d) if loc[j] == 1]
                box = (k, dirs) if (dim == d or (dim == d - 1 and dim >= 2)) else None
                groups[key] = dict(dim=dim, idx=[], seen=set(), box=box)
                order.append(key)
            g = groups[key]
            ranges = [range(0, 1) if loc[j] == 0 else range(nb[j] - 1, nb[j]) if loc[j] == 2 else
                      range(1, nb[j] - 1) for j in range(d)]
            for t in itertools.product(*reversed(ranges)):
                gd = int(dofs[flat(t[::-1])])
                if gd < nfree and gd not in g["seen"]:
                    g["seen"].add(gd)
                    g["idx"].append(gd)
    comps = sorted((groups[key] for key in order), key=lambda g: -g["dim"])
    return [(g["dim"], np.array(g["idx"], dtype=np.int64), g["box"]) for g in comps]


def scms_split_levels(level_problems, As, sigma=0.12):
    """setupSubspaceCorrectedMassSmoother on every level l >= 1 of a
    hierarchy (level_problems[l]: the problem of level l, As[l] its
    operator) as dense smoothers S_l = sum_c T_c op_c T_c^T: for patch
    interiors and 3D faces op_c = tau_c scms_matrix of the component's
    basis (Dirichlet on all its sides; faces with alpha = p/h, beta = h/p, h
    the patch's smallest knot span, p its degree) with tau_c = scms_damping
    of A_cc, kept per component under the example's rule (:536-554); for
    the others op_c = A_cc^-1.  Returns [None, S_1, ...]."""
    import types
    nlev = len(As)
    out = [None]
    saved = None
    for l in range(1, nlev):
        prob = level_problems[l]
        dofs, nfree = mapper(prob.patches, prob.dir_sides, prob.ifaces)
        comps = scms_components(prob.patches, dofs, nfree)
        if saved is None:
            saved = [0.0] * len(comps)
        A = As[l].toarray() if hasattr(As[l], "toarray") else np.asarray(As[l])
        S = np.zeros((nfree, nfree))
        for i, (dim, idx, box) in enumerate(comps):
            if len(idx) == 0:
                continue
            Acc = A[np.ix_(idx, idx)]
            if box is None:
                S[np.ix_(idx, idx)] += np.linalg.inv(Acc)
                continue
            k, dirs = box
            P = prob.patches[k]
            B = types.SimpleNamespace(d=len(dirs), p=[P.p[j] for j in dirs], knots=[P.knots[j] for j in dirs])
            if dim == P.d:
                Sc = scms_matrix(B, list(range(1, 2 * B.d + 1)), sigma=sigma)
            else:
                h = min(np.min(np.diff(np.unique(np.asarray(P.knots[j], dtype=np.float64)))) for j in range(P.d))
                pm = max(P.p[:P.d])
                Sc = scms_matrix(B, list(range(1, 2 * B.d + 1)), sigma=sigma, alpha=pm / h, beta=h / pm)
            if saved[i] == 0.0:
                tau = scms_damping(Acc, Sc)
                if l > nlev - 5 and all(len(np.unique(B.knots[j])) - 1 > B.p[j] for j in range(B.d)):
                    saved[i] = tau
            else:
                tau = saved[i]
            S[np.ix_(idx, idx)] += tau * Sc
        out.append(S)
    return out


def mg_vcycle(As, Ps, smoothers, npre=1, npost=1, cgc=1.0):
    """gsMultiGridOp's V-cycle from a zero start (oc_mg_apply / mg_step of
    the C oracle: src/gsSolver/gsMultiGrid.hpp:91-199) with smoothers given
    as functions smoothers[l](b, x, post) updating x in place; As[0] is
    solved exactly, Ps[l] maps level l to level l + 1.  Returns the apply
    function."""
    import scipy.sparse.linalg as spla
    coarse = spla.factorized(As[0].tocsc())

    def step(l, b, x):
        if l == 0:
            x[:] = coarse(b)
            return
        for _ in range(npre):
            smoothers[l](b, x, False)
        cr = Ps[l - 1].T @ (As[l] @ x - b)
        cc = np.zeros(As[l - 1].shape[0])
        step(l - 1, cr, cc)
        x -= cgc * (Ps[l - 1] @ cc)
        for _ in range(npost):
            smoothers[l](b, x, True)

    def apply(r):
        x = np.zeros(As[-1].shape[0])
        step(len(As) - 1, np.asarray(r, dtype=np.float64), x)
        return x
    return apply


def pcg(A, b, prec, tol=1e-10, maxit=1000):
    """gsConjugateGradient with a preconditioner given as a function
    (src/gsSolver/gsConjugateGradient.hpp:20-93 inside gsIterativeSolver::
    solve, gsIterativeSolver.h:114-124): the recurrences and stopping rule
    of oc_cg_solve.  Returns (x, iterations, relative residual)."""
    x = np.zeros_like(b)
    r = b.copy()
    bn = np.linalg.norm(b)
    if bn == 0.0:
        return x, 0, 0.0
    err = np.linalg.norm(r) / bn
    if err < tol:
        return x, 0, err
    p = prec(r)
    rz = r @ p
    it = 0
    while it < maxit:
        it += 1
        t = A @ p
        a = rz / (p @ t)
        x += a * p
        r -= a * t
        err = np.linalg.norm(r) / bn
        if err < tol:
            break
        z = prec(r)
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, it, err


# ---------------------------------------------------------------------------
# Dirichlet values by interpolation (gsAssembler::computeDirichletDofsIntpl)
# ---------------------------------------------------------------------------

def _greville(knots, p):
    n = len(knots) - p - 1
    return np.array([np.sum(knots[i + 1:i + p + 1]) / p for i in range(n)])


def geometry_at(patch, u):
    """x(u) = sum_m C_m N_m(u) of a patch (direction 0 fastest)."""
    d = patch.d
    nb = [len(patch.knots[k]) - patch.p[k] - 1 for k in range(d)]
    spans, vals = [], []
    for k in range(d):
        s = find_span(patch.knots[k], patch.p[k], u[k])
        spans.append(s)
        vals.append(bspline_ders(patch.knots[k], patch.p[k], s, u[k], 0)[0])
    x = np.zeros(d)
    for idx in np.ndindex(*[patch.p[k] + 1 for k in range(d)]):
        w = np.prod([vals[k][idx[k]] for k in range(d)])
        m, stride = 0, 1
        for k in range(d):
            m += (spans[k] - patch.p[k] + idx[k]) * stride
            stride *= nb[k]
        x += w * patch.ctrl[m]
    return x


def dirichlet_interpolation(patch_list, dofs, nfree, nfix, bcs):
    """Fixed values by interpolation (gsAssembler::computeDirichletDofsIntpl,
    src/gsAssembler/gsAssembler.hpp:314-390): the Dirichlet records
    bcs = [(patch, side, 1, kind, values)] in condition order; per side, g at
    the anchor grid (Greville abscissae of the tangential directions, the
    normal coordinate at the knot vector's end) mapped by the geometry,
    interpolated by the side's tensor basis (its collocation matrices); each
    side writes all of its functions, homogeneous sides zeros."""
    offs = np.cumsum([0] + [P.nbasis for P in patch_list])
    out = np.zeros(nfix)
    for pt, side, _, kind, vals in bcs:
        P = patch_list[pt]
        d = P.d
        nb = [len(P.knots[k]) - P.p[k] - 1 for k in range(d)]
        nd, par = (side - 1) // 2, (side - 1) % 2
        od = [k for k in range(d) if k != nd]
        ns = [nb[k] for k in od]

        def local(js):
            idx = [0] * 3
            idx[nd] = nb[nd] - 1 if par else 0
            for t, k in enumerate(od):
                idx[k] = js[t]
            return idx[0] + nb[0] * (idx[1] + (nb[1] if d > 1 else 1) * idx[2])
        if kind == 0:
            for js in np.ndindex(*ns[::-1]):
                out[dofs[offs[pt] + local(js[::-1])] - nfree] = 0.0
            continue
        grs = [_greville(P.knots[k], P.p[k]) for k in od]
        un = P.knots[nd][-1] if par else P.knots[nd][0]
        G = np.zeros(ns[::-1])   # axes: tangential directions, last one first
        for js in np.ndindex(*ns[::-1]):
            jj = js[::-1]
            u = [0.0] * d
            u[nd] = un
            for t, k in enumerate(od):
                u[k] = grs[t][jj[t]]
            x = geometry_at(P, u)
            if kind == 1:
                g = np.prod(np.sin(np.pi * x))
            else:
                g = vals[0] + np.dot(vals[1:1 + d], x)
            G[js] = g
        coef = G
        for t, k in enumerate(od):
            p = P.p[k]
            n = nb[k]
            Cm = np.zeros((n, n))
            for i, a in enumerate(grs[t]):
                s = find_span(P.knots[k], p, a)
                Cm[i, s - p:s + 1] = bspline_ders(P.knots[k], p, s, a, 0)[0]
            ax = len(od) - 1 - t
            coef = np.moveaxis(np.linalg.solve(Cm, np.moveaxis(coef, ax, 0).reshape(n, -1)).reshape(
                np.moveaxis(coef, ax, 0).shape), 0, ax)
        for js in np.ndindex(*ns[::-1]):
            out[dofs[offs[pt] + local(js[::-1])] - nfree] = coef[js]
    return out
