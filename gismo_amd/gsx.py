"""ctypes binding of libgsx (include/gsx.h) and a thin Python mirror of the
G+Smo call sequence for the isogeometric Poisson path:

    gsMultiPatch/gsMultiBasis + gsFeSpace::setup   -> Problem(...)
    gsExprAssembler::assemble / gsPoissonAssembler -> Problem.assemble()
    gsConjugateGradient(A, makeJacobiOp(A)).solve  -> Matrix.cg_solve()
    gsExprEvaluator::integral((u - u_ex)^2 meas)   -> Problem.l2_error()

The product path is libgsx only: if the shared library is missing this
module raises instead of computing anything on the CPU.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgsx.so")

GSX_OK = 0
GSX_ERR_ARG = 1
GSX_ERR_OOM = 2
GSX_ERR_JACOBIAN = 3
GSX_ERR_HIP = 4
GSX_ERR_NOT_CONVERGED = 5
GSX_ERR_NO_DEVICE = 6
GSX_ERR_BUDGET = 7

FP64 = 0
MIXED = 1
PATTERN_EXPR = 0
PATTERN_FULL = 1
ASM_REBUILD_PATTERN = 0x100
ASM_CSR_ONLY = 0x200
ASM_KEEP_ELEMENTS = 0x400
ASM_ELEMENTWISE = 0x800
ASM_ROWWISE = 0x1000
BC_DIRICHLET, BC_NEUMANN = 1, 2
BCF_ZERO, BCF_MANUFACTURED, BCF_AFFINE = 0, 1, 2
DIR_HOMOGENEOUS, DIR_L2PROJECTION, DIR_L2PROJECTION_INHOM, DIR_INTERPOLATION = 0, 1, 2, 3


class GsxBc(C.Structure):
    _fields_ = [("patch", C.c_int32), ("side", C.c_int32), ("type", C.c_int32), ("kind", C.c_int32),
                ("value", C.c_double * 4)]
PREC_NONE = 0
PREC_JACOBI = 1
PREC_MG = 2

_STATUS_NAMES = {
    0: "GSX_OK", 1: "GSX_ERR_ARG", 2: "GSX_ERR_OOM", 3: "GSX_ERR_JACOBIAN",
    4: "GSX_ERR_HIP", 5: "GSX_ERR_NOT_CONVERGED", 6: "GSX_ERR_NO_DEVICE", 7: "GSX_ERR_BUDGET",
}


class GsxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{_STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class PatchDesc(C.Structure):
    _fields_ = [
        ("d", C.c_int32),
        ("p", C.c_int32 * 3),
        ("nknots", C.c_int32 * 3),
        ("knots_host", C.POINTER(C.c_double) * 3),
        ("ctrl_host", C.POINTER(C.c_double)),
        ("kappa_host", C.POINTER(C.c_double)),
        ("felem_host", C.POINTER(C.c_double)),
        ("fkind", C.c_int32),
    ]




_lib = None
_vp = C.c_void_p
_pp = C.POINTER(C.c_void_p)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_dp = C.POINTER(C.c_double)


class ServerOpts(C.Structure):
    _fields_ = [("device", C.c_int32), ("precision", C.c_int32), ("budget_bytes", C.c_int64),
                ("iters_per_step", C.c_int32), ("capacity", C.c_int32), ("mg_min_free", C.c_int64),
                ("mg_max_coarse", C.c_int64), ("elem_budget", C.c_int64), ("group_max", C.c_int32),
                ("mg_workers", C.c_int32), ("reserve_bytes", C.c_int64), ("group_rows", C.c_int64)]


class ServerRequest(C.Structure):
    _fields_ = [("patch", C.POINTER(PatchDesc)), ("nrhs", C.c_int32), ("rhs_host", C.POINTER(_dp)),
                ("x_host", C.POINTER(_dp)), ("tol", C.c_double), ("maxit", C.c_int32),
                ("iters_host", _i32p), ("relres_host", _dp), ("status_host", _i32p)]


class ServerResult(C.Structure):
    _fields_ = [("rid", C.c_int64), ("status", C.c_int32), ("lane", C.c_int32), ("nrhs", C.c_int32),
                ("iters_max", C.c_int32), ("nbasis", C.c_int64), ("nfree", C.c_int64), ("relres_max", C.c_double),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_assembled", C.c_double),
                ("t_done", C.c_double), ("error", C.c_char * 160)]


SERVER_FAILED = 3

# (name, argtypes) of every exported function; restype is c_int (gsx_status)
# unless listed in _RESTYPES.
SIGNATURES = {
    "gsx_gauss_legendre": [C.c_int32, _dp, _dp],
    "gsx_mapper_build": [C.c_int32, C.c_int32, _i32p, C.c_int32, _i32p, C.c_int32, _i32p, _i32p,
                         _i64p, _i64p],
    "gsx_ctx_create": [C.c_int32, _pp],
    "gsx_ctx_destroy": [_vp],
    "gsx_last_error": [_vp],
    "gsx_ctx_synchronize": [_vp],
    "gsx_ctx_enable_timing": [_vp, C.c_int32],
    "gsx_ctx_reset_timing": [_vp],
    "gsx_ctx_timing_filter": [_vp, C.c_char_p],
    "gsx_ctx_kernel_stats": [_vp, C.c_char_p, _i64p, _dp],
    "gsx_device_alloc": [_vp, C.c_int64, _pp],
    "gsx_device_free": [_vp, _vp],
    "gsx_copy_to_device": [_vp, _vp, _vp, C.c_int64],
    "gsx_copy_to_host": [_vp, _vp, _vp, C.c_int64],
    "gsx_ctx_memory": [_vp, _i64p],
    "gsx_ctx_set_elem_budget": [_vp, C.c_int64],
    "gsx_ctx_set_coarse_dense_max": [_vp, C.c_int64],
    "gsx_problem_create": [_vp, C.c_int32, C.POINTER(PatchDesc), C.c_int32, _i32p, C.c_int32,
                           _i32p, _pp],
    "gsx_problem_create_geo": [_vp, C.c_int32, C.POINTER(PatchDesc), C.c_int32, _i32p, C.c_int32, _i32p, _pp],
    "gsx_problem_destroy": [_vp],
    "gsx_problem_info": [_vp, _i64p, _i64p, _i64p, _i64p],
    "gsx_problem_dofs": [_vp, _i32p],
    "gsx_problem_min_detj": [_vp, _dp],
    "gsx_assemble": [_vp, C.c_int32, C.c_int32, _pp, _vp],
    "gsx_matrix_destroy": [_vp],
    "gsx_matrix_from_csr": [_vp, C.c_int32, _i32p, _i32p, _dp, C.c_int32, _pp],
    "gsx_ctx_release_cache": [_vp],
    "gsx_matrix_info": [_vp, _i64p, _i64p],
    "gsx_matrix_download": [_vp, _i32p, _i32p, _dp],
    "gsx_matrix_device_ptrs": [_vp, _pp, _pp, _pp],
    "gsx_spmv": [_vp, _vp, _vp],
    "gsx_matrix_set_operator": [_vp, C.c_int32],
    "gsx_cg_solve": [_vp, C.c_int32, C.c_double, _vp, C.c_int32, _vp, _vp, C.c_double, C.c_int32,
                     _i32p, _dp],
    "gsx_cg_history": [_vp, _dp, C.c_int32, _i32p],
    "gsx_cg_solve_noalloc": [_vp, C.c_int32, C.c_double, _vp, _vp, C.c_double, C.c_int32, _i32p, _dp],
    "gsx_l2_error": [_vp, _vp, _dp],
    "gsx_problem_set_bc": [_vp, C.c_int32, _vp, C.c_int32],
    "gsx_problem_fixed_values": [_vp, _dp, C.c_int32],
    "gsx_assemble_bdr": [_vp, _vp],
    "gsx_error_norms": [_vp, _vp, _dp, _dp],
    "gsx_mg_create": [_vp, _vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double,
                      C.c_int32, _pp],
    "gsx_mg_destroy": [_vp],
    "gsx_mg_create_transfers": [_vp, C.c_int32, _i32p, C.POINTER(_i32p), C.POINTER(_i32p), C.POINTER(_dp),
                                C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_int32, _pp],
    "gsx_mg_set_level_smoother": [_vp, C.c_int32, C.c_int32, C.c_double, C.c_double],
    "gsx_mg_set_cycle": [_vp, C.c_int32, C.c_int32, C.c_double],
    "gsx_fd_create": [_vp, C.c_double, C.c_double, _pp],
    "gsx_fd_destroy": [_vp],
    "gsx_scms_create": [_vp, C.c_double, C.c_double, C.c_double, _vp],
    "gsx_fd_apply": [_vp, _vp, _vp],
    "gsx_cg_solve_fd": [_vp, _vp, C.c_int32, _vp, _vp, C.c_double, C.c_int32, _i32p, _dp],
    "gsx_mg_set_smoother": [_vp, C.c_int32, C.c_double],
    "gsx_mg_smoother_bound": [_vp, C.c_int32, _dp],
    "gsx_mg_apply": [_vp, _vp, _vp],
    "gsx_mg_level_info": [_vp, C.c_int32, _i64p, _i64p, _i64p],
    "gsx_mg_level_download": [_vp, C.c_int32, _i32p, _i32p, _dp, _i32p, _i32p, _dp],
    "gsx_pmg_create": [_vp, _vp, C.c_int32, C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                       C.c_double, C.c_int32, _pp],
    "gsx_mg_restriction_download": [_vp, C.c_int32, _i32p, _i32p, _dp],
    "gsx_batch_create": [_vp, C.c_int32, C.c_int32, _pp],
    "gsx_batch_destroy": [_vp],
    "gsx_batch_add": [_vp, _vp, C.c_int32, C.c_int32, C.c_int32, C.c_double, _vp, _vp, C.c_double,
                      C.c_int32, _i32p],
    "gsx_batch_add_box": [_vp, _vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                          C.c_int32, C.c_int32, C.c_int32, C.c_double, _vp, _vp, C.c_double, C.c_int32,
                          C.POINTER(C.c_int32)],
    "gsx_batch_set_tile": [_vp, C.c_int32, C.c_int32, C.c_int32],
    "gsx_batch_set_group": [_vp, C.c_int32],
    "gsx_batch_set_group_rows": [_vp, C.c_int64],
    "gsx_ctx_arch": [_vp, C.c_char_p, C.c_int32],
    "gsx_server_create": [C.POINTER(ServerOpts), _pp],
    "gsx_server_set_tile": [_vp, C.c_int32, C.c_int32, C.c_int64, C.c_int64, C.c_int32, C.c_int32],
    "gsx_server_submit": [_vp, C.POINTER(ServerRequest), _i64p],
    "gsx_server_poll": [_vp, C.c_int32, C.POINTER(ServerResult), _i32p],
    "gsx_server_status": [_vp, _i64p, _i64p, C.POINTER(C.c_double)],
    "gsx_server_drain": [_vp, C.c_double],
    "gsx_server_destroy": [_vp],
    "gsx_server_error": [_vp],
    "gsx_batch_step": [_vp, C.c_int32, _i32p],
    "gsx_batch_set_graph": [_vp, C.c_int32],
    "gsx_batch_graph_launches": [_vp, _i64p],
    "gsx_batch_result": [_vp, C.c_int32, _i32p, _i32p, _dp],
    "gsx_batch_history": [_vp, C.c_int32, _dp, C.c_int32, _i32p],
    "gsx_batch_remove": [_vp, C.c_int32],
}
_RESTYPES = {"gsx_last_error": C.c_char_p, "gsx_server_error": C.c_char_p}


def lib():
    """Load libgsx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libgsx not built: {LIB_PATH} is missing (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPES.get(name, C.c_int)
        _lib = L
    return _lib


def _check(ctx, status, allow=()):
    if status != GSX_OK and status not in allow:
        msg = lib().gsx_last_error(ctx).decode() if ctx else ""
        raise GsxError(status, msg)
    return status


def gauss_legendre(n):
    """Host-side Gauss-Legendre rule of libgsx (used by the kernels)."""
    x = np.zeros(n)
    w = np.zeros(n)
    _check(None, lib().gsx_gauss_legendre(n, x.ctypes.data_as(_dp), w.ctypes.data_as(_dp)))
    return x, w


def mapper(problem):
    """libgsx's DOF mapper (host): (dofs per (patch, local), nfree, nglued)."""
    d = problem.d
    nbd = np.ones((len(problem.patches), 3), np.int32)
    for k, P in enumerate(problem.patches):
        nbd[k, :d] = P.nbasis_dir()
    ds = np.ascontiguousarray(np.asarray(problem.dir_sides, dtype=np.int32).reshape(-1))
    ifs = np.ascontiguousarray(np.asarray(problem.ifaces, dtype=np.int32).reshape(-1))
    out = np.zeros(int(nbd.prod(axis=1).sum()), np.int32)
    nf, ng = C.c_int64(), C.c_int64()
    _check(None, lib().gsx_mapper_build(
        d, len(problem.patches), nbd.ctypes.data_as(_i32p), len(problem.dir_sides),
        ds.ctypes.data_as(_i32p) if ds.size else None, len(problem.ifaces),
        ifs.ctypes.data_as(_i32p) if ifs.size else None, out.ctypes.data_as(_i32p),
        C.byref(nf), C.byref(ng)))
    return out, nf.value, ng.value


class DeviceArray:
    """Device buffer owned by a Context (freed with it or on close())."""

    def __init__(self, ctx, n, dtype=np.float64):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        p = C.c_void_p()
        _check(ctx.h, lib().gsx_device_alloc(ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        _check(self.ctx.h, lib().gsx_copy_to_device(self.ctx.h, self.ptr, a.ctypes.data, self.nbytes))
        return self

    def download(self):
        out = np.empty(self.n, dtype=self.dtype)
        _check(self.ctx.h, lib().gsx_copy_to_host(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def zero(self):
        return self.upload(np.zeros(self.n, self.dtype))

    def view(self, offset, n):
        """Entries [offset, offset + n) (shares the memory; not freed on its own)."""
        assert 0 <= offset and offset + n <= self.n
        v = DeviceArray.__new__(DeviceArray)
        v.ctx, v.n, v.dtype = self.ctx, int(n), self.dtype
        v.ptr = C.c_void_p(self.ptr.value + offset * self.dtype.itemsize)
        v._base = self
        return v

    def close(self):
        if self.ptr and getattr(self, "_base", None) is None:
            lib().gsx_device_free(self.ctx.h, self.ptr)
        self.ptr = None


class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        st = lib().gsx_ctx_create(device, C.byref(h))
        if st != GSX_OK:
            raise GsxError(st, "gsx_ctx_create failed")
        self.h = h

    def array(self, n, dtype=np.float64):
        return DeviceArray(self, n, dtype)

    def to_device(self, a, dtype=np.float64):
        a = np.asarray(a, dtype=dtype)
        return DeviceArray(self, a.size, dtype).upload(a)

    def arch(self):
        """The device's architecture name (gfx950 on MI355X)."""
        buf = C.create_string_buffer(128)
        _check(self.h, lib().gsx_ctx_arch(self.h, buf, 128))
        return buf.value.decode()

    def synchronize(self):
        _check(self.h, lib().gsx_ctx_synchronize(self.h))

    def enable_timing(self, on=True):
        _check(self.h, lib().gsx_ctx_enable_timing(self.h, 1 if on else 0))

    def timing_filter(self, name=None):
        _check(self.h, lib().gsx_ctx_timing_filter(self.h, name.encode() if name else None))

    def reset_timing(self):
        _check(self.h, lib().gsx_ctx_reset_timing(self.h))

    def kernel_stats(self, name):
        n = C.c_int64()
        ms = C.c_double()
        _check(self.h, lib().gsx_ctx_kernel_stats(self.h, name.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def set_elem_budget(self, nbytes):
        """Bytes of element matrices per assembly pass (single patches larger
        than this are assembled in slabs)."""
        _check(self.h, lib().gsx_ctx_set_elem_budget(self.h, int(nbytes)))

    def set_coarse_dense_max(self, rows):
        """Largest coarsest multigrid level solved through its dense inverse
        (default 32768 rows); larger ones use a banded Cholesky factor."""
        _check(self.h, lib().gsx_ctx_set_coarse_dense_max(self.h, int(rows)))

    def memory(self):
        b = C.c_int64()
        _check(self.h, lib().gsx_ctx_memory(self.h, C.byref(b)))
        return b.value

    def close(self):
        if self.h:
            lib().gsx_ctx_destroy(self.h)
            self.h = None


def fill_patch_desc(o, P, keep):
    """Fill the PatchDesc o from a problems.Patch; the host arrays it points
    to are appended to keep (they must outlive the call that reads o)."""
    o.d = P.d
    for k in range(3):
        if k < P.d:
            kn = np.ascontiguousarray(P.knots[k], dtype=np.float64)
            keep.append(kn)
            o.p[k] = int(P.p[k])
            o.nknots[k] = kn.size
            o.knots_host[k] = kn.ctypes.data_as(_dp)
    ctrl = np.ascontiguousarray(P.ctrl, dtype=np.float64)
    keep.append(ctrl)
    o.ctrl_host = ctrl.ctypes.data_as(_dp)
    for name, attr in (("kappa_host", "kappa"), ("felem_host", "felem")):
        a = getattr(P, attr, None)
        if a is None:
            setattr(o, name, _dp())
        else:
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            setattr(o, name, a.ctypes.data_as(_dp))
    o.fkind = int(P.fkind)




class Problem:
    """Patches + DOF mapper on the device (gsMultiPatch, gsMultiBasis,
    gsFeSpace::setup with glued interfaces and eliminated Dirichlet sides)."""

    def __init__(self, ctx, problem):
        self.ctx = ctx
        self.src = problem
        patches = problem.patches
        n = len(patches)
        arr = (PatchDesc * n)()
        self._keep = []
        for i, P in enumerate(patches):
            fill_patch_desc(arr[i], P, self._keep)
        ds = np.ascontiguousarray(np.asarray(problem.dir_sides, dtype=np.int32).reshape(-1))
        ifs = np.ascontiguousarray(np.asarray(problem.ifaces, dtype=np.int32).reshape(-1))
        self._keep += [ds, ifs]
        h = C.c_void_p()
        _check(ctx.h, lib().gsx_problem_create(
            ctx.h, n, arr, len(problem.dir_sides), ds.ctypes.data_as(_i32p) if ds.size else None,
            len(problem.ifaces), ifs.ctypes.data_as(_i32p) if ifs.size else None, C.byref(h)))
        self.h = h
        nb, nf, ne, ng = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        _check(ctx.h, lib().gsx_problem_info(h, C.byref(nb), C.byref(nf), C.byref(ne), C.byref(ng)))
        self.nbasis, self.nfree, self.nelem, self.nglued = nb.value, nf.value, ne.value, ng.value

    @classmethod
    def with_geometry(cls, geo, bases, dir_sides=None, ifaces=None):
        """gsx_problem_create_geo: the bases (objects with d, p, knots, one per
        patch of geo) over the geometry, kappa and source of the Problem geo;
        Dirichlet sides and interfaces default to geo's."""
        self = cls.__new__(cls)
        self.ctx = geo.ctx
        self.src = None
        self.geo = geo
        n = len(bases)
        arr = (PatchDesc * n)()
        self._keep = []
        for i, B in enumerate(bases):
            o = arr[i]
            o.d = B.d
            for k in range(B.d):
                kn = np.ascontiguousarray(B.knots[k], dtype=np.float64)
                self._keep.append(kn)
                o.p[k] = int(B.p[k])
                o.nknots[k] = kn.size
                o.knots_host[k] = kn.ctypes.data_as(_dp)
            o.fkind = 1
        src = geo.src
        ds = np.ascontiguousarray(np.asarray(src.dir_sides if dir_sides is None else dir_sides,
                                             dtype=np.int32).reshape(-1))
        ifs = np.ascontiguousarray(np.asarray(src.ifaces if ifaces is None else ifaces, dtype=np.int32).reshape(-1))
        self._keep += [ds, ifs]
        h = C.c_void_p()
        _check(self.ctx.h, lib().gsx_problem_create_geo(
            geo.h, n, arr, ds.size // 2, ds.ctypes.data_as(_i32p) if ds.size else None, ifs.size // 8,
            ifs.ctypes.data_as(_i32p) if ifs.size else None, C.byref(h)))
        self.h = h
        nb, nf, ne, ng = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        _check(self.ctx.h, lib().gsx_problem_info(h, C.byref(nb), C.byref(nf), C.byref(ne), C.byref(ng)))
        self.nbasis, self.nfree, self.nelem, self.nglued = nb.value, nf.value, ne.value, ng.value
        return self

    def dofs(self):
        out = np.zeros(self.nbasis, np.int32)
        _check(self.ctx.h, lib().gsx_problem_dofs(self.h, out.ctypes.data_as(_i32p)))
        return out

    def min_detj(self):
        v = C.c_double()
        _check(self.ctx.h, lib().gsx_problem_min_detj(self.h, C.byref(v)))
        return v.value

    def assemble(self, mode=PATTERN_EXPR, precision=FP64, rhs=None):
        """Returns (Matrix, rhs DeviceArray)."""
        if rhs is None:
            rhs = self.ctx.array(max(1, self.nfree))
        h = C.c_void_p()
        _check(self.ctx.h, lib().gsx_assemble(self.h, mode, precision, C.byref(h), rhs.ptr))
        return Matrix(self.ctx, h), rhs

    def set_bc(self, bcs, dir_values=1):
        """Boundary data: bcs = [(patch, side, type, kind, values)] with type
        BC_DIRICHLET / BC_NEUMANN, kind BCF_ZERO / BCF_MANUFACTURED /
        BCF_AFFINE (values: c, a_0, a_1, a_2); dir_values DIR_L2PROJECTION
        (default: every Dirichlet side projected, gsFeSpace::setup),
        DIR_L2PROJECTION_INHOM (sides with nonzero data only,
        gsAssembler::computeDirichletDofs), DIR_INTERPOLATION (interpolation
        at the side anchors, gsAssembler::computeDirichletDofsIntpl) or
        DIR_HOMOGENEOUS."""
        arr = (GsxBc * max(1, len(bcs)))()
        for i, (patch, side, typ, kind, vals) in enumerate(bcs):
            arr[i].patch, arr[i].side, arr[i].type, arr[i].kind = patch, side, typ, kind
            for k, v in enumerate(list(vals)[:4]):
                arr[i].value[k] = v
        _check(self.ctx.h, lib().gsx_problem_set_bc(self.h, len(bcs), C.cast(arr, C.c_void_p), dir_values))

    def fixed_values(self):
        n = self.nglued - self.nfree
        out = np.zeros(max(1, n))
        _check(self.ctx.h, lib().gsx_problem_fixed_values(self.h, out.ctypes.data_as(_dp), n))
        return out[:n]

    def assemble_bdr(self, rhs):
        """rhs += Neumann boundary load (gsExprAssembler::assembleBdr)."""
        _check(self.ctx.h, lib().gsx_assemble_bdr(self.h, rhs.ptr))

    def l2_error(self, sol):
        v = C.c_double()
        _check(self.ctx.h, lib().gsx_l2_error(self.h, sol.ptr, C.byref(v)))
        return v.value

    def error_norms(self, sol):
        """(L2 error, H1 seminorm error) against prod sin(pi x_r)."""
        a, b = C.c_double(), C.c_double()
        _check(self.ctx.h, lib().gsx_error_norms(self.h, sol.ptr, C.byref(a), C.byref(b)))
        return a.value, b.value

    def close(self):
        if self.h:
            lib().gsx_problem_destroy(self.h)
            self.h = None


class Matrix:
    """Device CSR matrix (symmetric: the CSR arrays are also the CSC arrays
    of gsSparseMatrix)."""

    def __init__(self, ctx, h):
        self.ctx = ctx
        self.h = h
        n, nnz = C.c_int64(), C.c_int64()
        _check(ctx.h, lib().gsx_matrix_info(h, C.byref(n), C.byref(nnz)))
        self.n, self.nnz = n.value, nnz.value

    @staticmethod
    def from_csr(ctx, rowptr, col, val, with_fp32=False):
        """Device matrix from host CSR arrays (sorted columns)."""
        rp = np.ascontiguousarray(rowptr, np.int32)
        ci = np.ascontiguousarray(col, np.int32)
        va = np.ascontiguousarray(val, np.float64)
        h = C.c_void_p()
        _check(ctx.h, lib().gsx_matrix_from_csr(ctx.h, len(rp) - 1, rp.ctypes.data_as(_i32p),
                                                ci.ctypes.data_as(_i32p), va.ctypes.data_as(_dp),
                                                1 if with_fp32 else 0, C.byref(h)))
        return Matrix(ctx, h)

    def download(self):
        rp = np.zeros(self.n + 1, np.int32)
        ci = np.zeros(max(1, self.nnz), np.int32)
        va = np.zeros(max(1, self.nnz), np.float64)
        _check(self.ctx.h, lib().gsx_matrix_download(self.h, rp.ctypes.data_as(_i32p),
                                                     ci.ctypes.data_as(_i32p), va.ctypes.data_as(_dp)))
        return rp, ci[: self.nnz], va[: self.nnz]

    def to_scipy(self):
        import scipy.sparse as sp
        rp, ci, va = self.download()
        return sp.csr_matrix((va, ci, rp), shape=(self.n, self.n))

    def spmv(self, x, y):
        _check(self.ctx.h, lib().gsx_spmv(self.h, x.ptr, y.ptr))

    def set_operator(self, structured=True):
        """Solver operator format: structured stencil (single patch) or CSR."""
        _check(self.ctx.h, lib().gsx_matrix_set_operator(self.h, 1 if structured else 0))

    def cg_solve(self, b, x, tol=1e-10, maxit=1000, precond=PREC_JACOBI, tau=1.0,
                 precision=FP64, raise_on_nonconvergence=False):
        it = C.c_int32()
        rr = C.c_double()
        st = lib().gsx_cg_solve(self.h, precond, tau, None, precision, b.ptr, x.ptr, tol, maxit,
                                C.byref(it), C.byref(rr))
        allow = () if raise_on_nonconvergence else (GSX_ERR_NOT_CONVERGED,)
        _check(self.ctx.h, st, allow)
        return it.value, rr.value, st == GSX_OK

    def history(self):
        cnt = C.c_int32()
        _check(self.ctx.h, lib().gsx_cg_history(self.h, None, 0, C.byref(cnt)))
        h = np.zeros(max(1, cnt.value))
        _check(self.ctx.h, lib().gsx_cg_history(self.h, h.ctypes.data_as(_dp), cnt.value, C.byref(cnt)))
        return h[: cnt.value]

    def close(self):
        if self.h:
            lib().gsx_matrix_destroy(self.h)
            self.h = None


class Multigrid:
    """Geometric multigrid preconditioner (gsMultiGridOp) on the device."""

    SMOOTHERS = {"jacobi": 0, "richardson": 1, "chebyshev": 2, "gauss_seidel": 3, "scms": 4}

    def __init__(self, problem, matrix, nlevels, npre=1, npost=1, damping=1.0, cgc_damping=1.0,
                 precision=FP64, smoother="jacobi", smoother_param=None):
        """smoother_param: the Chebyshev interval ratio (default 30) or the
        scaling sigma of the subspace corrected mass smoother ("scms",
        default 0.12)."""
        if smoother_param is None:
            smoother_param = 0.12 if smoother == "scms" else 30.0
        self.ctx = problem.ctx
        h = C.c_void_p()
        _check(self.ctx.h, lib().gsx_mg_create(problem.h, matrix.h, nlevels, npre, npost, 1, damping,
                                               cgc_damping, precision, C.byref(h)))
        self.h = h
        self.nlevels = nlevels
        if smoother != "jacobi":
            _check(self.ctx.h, lib().gsx_mg_set_smoother(h, self.SMOOTHERS[smoother], smoother_param))

    @classmethod
    def from_transfers(cls, ctx, matrix, transfers, npre=1, npost=1, damping=1.0, cgc_damping=1.0,
                       precision=FP64):
        """gsMultiGridOp::make(fineMatrix, transferMatrices): transfers[l]
        (scipy CSR, n_{l+1} x n_l) prolongates level l to level l + 1, the
        last one to the matrix's level."""
        self = cls.__new__(cls)
        self.ctx = ctx
        nt = len(transfers)
        keep = []
        rps = (_i32p * max(1, nt))()
        cis = (_i32p * max(1, nt))()
        vas = (_dp * max(1, nt))()
        nc = np.zeros(max(1, nt), np.int32)
        for l, P in enumerate(transfers):
            P = P.tocsr()
            P.sort_indices()
            rp = np.ascontiguousarray(P.indptr, np.int32)
            ci = np.ascontiguousarray(P.indices, np.int32)
            va = np.ascontiguousarray(P.data, np.float64)
            keep += [rp, ci, va]
            rps[l] = rp.ctypes.data_as(_i32p)
            cis[l] = ci.ctypes.data_as(_i32p)
            vas[l] = va.ctypes.data_as(_dp)
            nc[l] = P.shape[1]
        h = C.c_void_p()
        _check(ctx.h, lib().gsx_mg_create_transfers(matrix.h, nt, nc.ctypes.data_as(_i32p), rps, cis, vas, npre,
                                                    npost, 1, damping, cgc_damping, precision, C.byref(h)))
        self.h = h
        self.nlevels = nt + 1
        return self

    @classmethod
    def pmg(cls, problem, matrix, nlevels, coarsening=None, projection=2, coarse_operator=2, npre=1, npost=1,
            damping=1.0, cgc_damping=1.0, precision=FP64):
        """p-multigrid (gsx_pmg_create) with the transitions of
        examples/pMultiGrid_example.cpp: coarsening[l] in 'h', 'p', 'z' from
        level l to l + 1; the default is the example's (projection 2:
        'h' * (nlevels - p) + 'z' * (p - 1), or 'z' * (nlevels - 1) when
        nlevels <= p; projection 1: 'h' * (nlevels - 2) + 'z')."""
        if coarsening is None:
            p = int(problem.src.patches[0].p[0])
            if projection == 1:
                coarsening = "h" * (nlevels - 2) + "z"
            elif nlevels > p:
                coarsening = "h" * (nlevels - p) + "z" * (p - 1)
            else:
                coarsening = "z" * (nlevels - 1)
        self = cls.__new__(cls)
        self.ctx = problem.ctx
        h = C.c_void_p()
        _check(self.ctx.h, lib().gsx_pmg_create(problem.h, matrix.h, nlevels, coarsening.encode(), projection,
                                                coarse_operator, npre, npost, damping, cgc_damping, precision,
                                                C.byref(h)))
        self.h = h
        self.nlevels = nlevels
        self.coarsening = coarsening
        return self

    def restriction(self, lvl):
        """The restriction from level lvl + 1 to level lvl (scipy CSR)."""
        import scipy.sparse as sp
        n, pnnz, nf = C.c_int64(), C.c_int64(), C.c_int64()
        _check(self.ctx.h, lib().gsx_mg_level_info(self.h, lvl, C.byref(n), None, C.byref(pnnz)))
        _check(self.ctx.h, lib().gsx_mg_level_info(self.h, lvl + 1, C.byref(nf), None, None))
        rp = np.zeros(n.value + 1, np.int32)
        ci = np.zeros(max(1, pnnz.value), np.int32)
        va = np.zeros(max(1, pnnz.value))
        _check(self.ctx.h, lib().gsx_mg_restriction_download(self.h, lvl, rp.ctypes.data_as(_i32p),
                                                             ci.ctypes.data_as(_i32p), va.ctypes.data_as(_dp)))
        return sp.csr_matrix((va[:pnnz.value], ci[:pnnz.value], rp), shape=(n.value, nf.value))

    def set_level_smoother(self, lvl, smoother, damping=1.0, param=0.0):
        """gsMultiGridOp::setSmoother(lvl, op): one level's smoother."""
        _check(self.ctx.h, lib().gsx_mg_set_level_smoother(self.h, lvl, self.SMOOTHERS[smoother], damping, param))

    def set_cycle(self, npre, npost, cgc_damping=1.0):
        _check(self.ctx.h, lib().gsx_mg_set_cycle(self.h, npre, npost, cgc_damping))

    def lmax(self, lvl):
        v = C.c_double()
        _check(self.ctx.h, lib().gsx_mg_smoother_bound(self.h, lvl, C.byref(v)))
        return v.value

    def apply(self, r, z):
        _check(self.ctx.h, lib().gsx_mg_apply(self.h, r.ptr, z.ptr))

    def level(self, lvl):
        """(operator CSR as scipy, transfer to the next finer level or None)."""
        import scipy.sparse as sp
        n, nnz, pnnz = C.c_int64(), C.c_int64(), C.c_int64()
        _check(self.ctx.h, lib().gsx_mg_level_info(self.h, lvl, C.byref(n), C.byref(nnz), C.byref(pnnz)))
        rp = np.zeros(n.value + 1, np.int32)
        ci = np.zeros(max(1, nnz.value), np.int32)
        va = np.zeros(max(1, nnz.value))
        nf = None
        if lvl < self.nlevels - 1:
            nf_, _, _ = C.c_int64(), None, None
            _check(self.ctx.h, lib().gsx_mg_level_info(self.h, lvl + 1, C.byref(nf_), None, None))
            nf = nf_.value
            prp = np.zeros(nf + 1, np.int32)
            pci = np.zeros(max(1, pnnz.value), np.int32)
            pva = np.zeros(max(1, pnnz.value))
            _check(self.ctx.h, lib().gsx_mg_level_download(
                self.h, lvl, rp.ctypes.data_as(_i32p), ci.ctypes.data_as(_i32p), va.ctypes.data_as(_dp),
                prp.ctypes.data_as(_i32p), pci.ctypes.data_as(_i32p), pva.ctypes.data_as(_dp)))
            P = sp.csr_matrix((pva[:pnnz.value], pci[:pnnz.value], prp), shape=(nf, n.value))
        else:
            _check(self.ctx.h, lib().gsx_mg_level_download(
                self.h, lvl, rp.ctypes.data_as(_i32p), ci.ctypes.data_as(_i32p), va.ctypes.data_as(_dp),
                None, None, None))
            P = None
        A = sp.csr_matrix((va[:nnz.value], ci[:nnz.value], rp), shape=(n.value, n.value))
        return A, P

    def close(self):
        if self.h:
            lib().gsx_mg_destroy(self.h)
            self.h = None


def mg_cg_solve(matrix, mg, b, x, tol=1e-8, maxit=500):
    """CG preconditioned by the multigrid cycle (gsConjugateGradient with a
    gsMultiGridOp preconditioner)."""
    it = C.c_int32()
    rr = C.c_double()
    st = lib().gsx_cg_solve(matrix.h, PREC_MG, 1.0, mg.h, FP64, b.ptr, x.ptr, tol, maxit, C.byref(it),
                            C.byref(rr))
    _check(matrix.ctx.h, st, (GSX_ERR_NOT_CONVERGED,))
    return it.value, rr.value, st == GSX_OK


class FastDiag:
    """Fast diagonalization preconditioner of a single-patch problem
    (gsPatchPreconditionersCreator::fastDiagonalizationOp): the exact inverse
    of alpha M + beta sum_k M x .. x K_k x .. x M on the parameter domain."""

    def __init__(self, problem, alpha=0.0, beta=1.0):
        self.ctx = problem.ctx
        h = C.c_void_p()
        _check(self.ctx.h, lib().gsx_fd_create(problem.h, alpha, beta, C.byref(h)))
        self.h = h

    def apply(self, r, z):
        _check(self.ctx.h, lib().gsx_fd_apply(self.h, r.ptr, z.ptr))

    def close(self):
        if self.h:
            lib().gsx_fd_destroy(self.h)
            self.h = None


class SubspaceMassSmoother(FastDiag):
    """Subspace corrected mass smoother of a single-patch problem
    (gsPatchPreconditionersCreator::subspaceCorrectedMassSmootherOp(basis,
    bc, opt, sigma, alpha, beta)): applied and used in CG like FastDiag."""

    def __init__(self, problem, sigma=0.12, alpha=0.0, beta=1.0):
        self.ctx = problem.ctx
        h = C.c_void_p()
        _check(self.ctx.h, lib().gsx_scms_create(problem.h, sigma, alpha, beta, C.byref(h)))
        self.h = h


def fd_cg_solve(matrix, fd, b, x, tol=1e-8, maxit=500, precision=FP64):
    """CG preconditioned by the fast diagonalization, on the structured
    operator (FP32 iterations with FP64 reliable updates for MIXED)."""
    it = C.c_int32()
    rr = C.c_double()
    st = lib().gsx_cg_solve_fd(matrix.h, fd.h, precision, b.ptr, x.ptr, tol, maxit, C.byref(it), C.byref(rr))
    _check(matrix.ctx.h, st, (GSX_ERR_NOT_CONVERGED,))
    return it.value, rr.value, st == GSX_OK


class Batch:
    """Independent CG solves advanced together on the device (one launch per
    CG stage for all members; a member's result does not depend on the
    others)."""

    def __init__(self, ctx, precision=FP64, capacity=1024):
        self.ctx = ctx
        h = C.c_void_p()
        _check(ctx.h, lib().gsx_batch_create(ctx.h, precision, capacity, C.byref(h)))
        self.h = h

    def add(self, matrix, b, x, row0=0, nrows=None, tol=1e-10, maxit=1000, precond=PREC_JACOBI, tau=1.0,
            box=None, tile=None):
        """box = (d, p, m0, m1, m2): the rows are the free functions of one
        tensor-product patch (gsx_batch_add_box): the member iterates on a
        structured copy of its block.  tile = (lanes, half): the SpMV tile
        shape (gsx_batch_set_tile)."""
        nrows = matrix.n - row0 if nrows is None else nrows
        i = C.c_int32()
        if box is None:
            _check(self.ctx.h, lib().gsx_batch_add(self.h, matrix.h, row0, nrows, precond, tau, b.ptr, x.ptr, tol,
                                                   maxit, C.byref(i)))
            if tile is not None:
                self.set_tile(i.value, *tile)
        else:
            d, p, m0, m1, m2 = box
            lanes, half = (1, 1) if tile is None else tile
            _check(self.ctx.h, lib().gsx_batch_add_box(self.h, matrix.h, row0, nrows, d, p, m0, m1, m2, lanes,
                                                       half, precond, tau, b.ptr, x.ptr, tol, maxit, C.byref(i)))
        return i.value

    def set_tile(self, i, lanes, half=1):
        _check(self.ctx.h, lib().gsx_batch_set_tile(self.h, i, lanes, half))

    def step(self, iters=8):
        run = C.c_int32()
        _check(self.ctx.h, lib().gsx_batch_step(self.h, iters, C.byref(run)))
        return run.value

    def run(self, iters=8, max_steps=100000):
        for _ in range(max_steps):
            if self.step(iters) == 0:
                return
        raise RuntimeError("batch did not finish")

    def result(self, i):
        """(status, iterations, relative residual); status 1 converged,
        2 iteration limit, 3 zero rhs, 0 running."""
        st, it, rr = C.c_int32(), C.c_int32(), C.c_double()
        _check(self.ctx.h, lib().gsx_batch_result(self.h, i, C.byref(st), C.byref(it), C.byref(rr)))
        return st.value, it.value, rr.value

    def history(self, i):
        n = C.c_int32()
        _check(self.ctx.h, lib().gsx_batch_history(self.h, i, None, 0, C.byref(n)))
        out = np.zeros(n.value)
        _check(self.ctx.h, lib().gsx_batch_history(self.h, i, out.ctypes.data_as(_dp), n.value, C.byref(n)))
        return out

    def set_group(self, group_max, rows=None):
        """Members on one operator share SpMV launches in groups of up to
        group_max (gsx_batch_set_group); members of at least rows rows
        (gsx_batch_set_group_rows; None: the default 131072)."""
        _check(self.ctx.h, lib().gsx_batch_set_group(self.h, group_max))
        if rows is not None:
            _check(self.ctx.h, lib().gsx_batch_set_group_rows(self.h, int(rows)))

    def set_graph(self, enable):
        """Capture each step's launches as a HIP graph (default) or not."""
        _check(self.ctx.h, lib().gsx_batch_set_graph(self.h, 1 if enable else 0))

    def graph_launches(self):
        n = C.c_int64()
        _check(self.ctx.h, lib().gsx_batch_graph_launches(self.h, C.byref(n)))
        return n.value

    def remove(self, i):
        _check(self.ctx.h, lib().gsx_batch_remove(self.h, i))

    def close(self):
        if self.h:
            lib().gsx_batch_destroy(self.h)
            self.h = None
