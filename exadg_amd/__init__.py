"""MI355X-native SIPG Laplace hot path of ExaDG: Python binding of libexadg_amd.so.

The compute path is the C ABI declared in include/exadg_amd.h (hand-written HIP for gfx950
behind host C++). This module only binds it with ctypes for tests and benchmarks; vectors are
device tensors (torch on ROCm is used purely for device memory and streams). There is no CPU
fallback: if the library is missing, loading fails loudly.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libexadg_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "exadg_amd.h")

EXADG_OK = 0
FP64, FP32 = 0, 1
DIRICHLET, NEUMANN = 0, 1

_lib = None


class ExaDGError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"exadg status {code}: {msg}")
        self.code = code


class MeshDesc(ctypes.Structure):
    _fields_ = [
        ("n_cells", ctypes.c_int32 * 3),
        ("lower", ctypes.c_double * 3),
        ("upper", ctypes.c_double * 3),
        ("deformation", ctypes.c_double),
        ("frequency", ctypes.c_int32),
        ("mapping_degree", ctypes.c_int32),
        ("periodic", ctypes.c_int32 * 3),
    ]


class BCDesc(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32 * 6)]


class MGDesc(ctypes.Structure):
    _fields_ = [
        ("n_dg_levels", ctypes.c_int32),
        ("dg_degrees", ctypes.c_int32 * 16),
        ("coarse_q1", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("smoother_degree", ctypes.c_int32),
        ("smoothing_range", ctypes.c_double),
        ("eig_cg_iterations", ctypes.c_int32),
        ("coarse_solver", ctypes.c_int32),
        ("coarse_rel_tol", ctypes.c_double),
        ("coarse_max_iter", ctypes.c_int32),
        ("operator_is_singular", ctypes.c_int32),
    ]


class SolverData(ctypes.Structure):
    _fields_ = [("max_iter", ctypes.c_int32), ("abs_tol", ctypes.c_double),
                ("rel_tol", ctypes.c_double), ("rel_mode", ctypes.c_int32)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("converged", ctypes.c_int32),
                ("not_finite", ctypes.c_int32), ("r0", ctypes.c_double), ("r_final", ctypes.c_double),
                ("b_norm", ctypes.c_double), ("rho", ctypes.c_double), ("n10", ctypes.c_double)]


PRECOND_NONE, PRECOND_POINT_JACOBI, PRECOND_MULTIGRID = 0, 1, 2
COARSE_CG, COARSE_CHEBYSHEV = 0, 1


def mg_desc(degree, degrees=None, precision=FP32, coarse_q1=True, coarse_solver=COARSE_CG,
            coarse_rel_tol=1e-3, smoother_degree=5, smoothing_range=20.0, eig_cg_iterations=20,
            singular=False):
    """MultigridData; default p-sequence Bisect (k, k/2, ..., 1) + continuous Q1 level."""
    d = MGDesc()
    check(lib().exadg_mg_desc_default(degree, ctypes.byref(d)))
    if degrees is not None:
        d.n_dg_levels = len(degrees)
        for i, p in enumerate(degrees):
            d.dg_degrees[i] = p
    d.precision = precision
    d.coarse_q1 = int(coarse_q1)
    d.coarse_solver = coarse_solver
    d.coarse_rel_tol = coarse_rel_tol
    d.smoother_degree = smoother_degree
    d.smoothing_range = smoothing_range
    d.eig_cg_iterations = eig_cg_iterations
    d.operator_is_singular = int(singular)
    return d


def mesh_desc(n_cells, lower=0.0, upper=1.0, deformation=0.0, frequency=1, mapping_degree=1,
              periodic=(0, 0, 0)):
    if isinstance(n_cells, int):
        n_cells = (n_cells,) * 3
    m = MeshDesc()
    m.n_cells[:] = list(n_cells)
    m.lower[:] = [float(lower)] * 3 if isinstance(lower, (int, float)) else list(lower)
    m.upper[:] = [float(upper)] * 3 if isinstance(upper, (int, float)) else list(upper)
    m.deformation = deformation
    m.frequency = frequency
    m.mapping_degree = mapping_degree
    m.periodic[:] = list(periodic)
    return m


def bc_desc(types=(DIRICHLET,) * 6):
    b = BCDesc()
    b.type[:] = list(types)
    return b


def lib():
    """Load libexadg_amd.so; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m exadg_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64p = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)
        ip = ctypes.POINTER(ctypes.c_int)
        L.exadg_last_error.restype = ctypes.c_char_p
        L.exadg_last_error.argtypes = []
        sig = {
            "exadg_laplace_create": [ctypes.POINTER(MeshDesc), ctypes.POINTER(BCDesc), ctypes.c_int,
                                     ctypes.c_int, ctypes.c_double, ctypes.POINTER(vp)],
            "exadg_laplace_destroy": [vp],
            "exadg_laplace_n_dofs": [vp, i64p],
            "exadg_laplace_info": [vp, i64p, ip, ip, ip],
            "exadg_laplace_vmult": [vp, vp, vp, vp],
            "exadg_laplace_vmult_add": [vp, vp, vp, vp],
            "exadg_laplace_calculate_inverse_diagonal": [vp, vp, vp],
            "exadg_laplace_calculate_diagonal": [vp, vp, vp],
            "exadg_laplace_rhs_sine": [vp, ctypes.c_double, vp, vp],
            "exadg_mg_desc_default": [ctypes.c_int, ctypes.POINTER(MGDesc)],
            "exadg_preconditioner_create": [vp, ctypes.c_int, ctypes.POINTER(MGDesc),
                                            ctypes.POINTER(vp), vp],
            "exadg_preconditioner_vmult": [vp, vp, vp, vp],
            "exadg_preconditioner_update": [vp, vp],
            "exadg_preconditioner_destroy": [vp],
            "exadg_preconditioner_mg_info": [vp, ip, ctypes.POINTER(ctypes.c_double), ip, ip],
            "exadg_cg_solve": [vp, vp, vp, vp, ctypes.POINTER(SolverData), ctypes.POINTER(SolveInfo),
                               vp],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.restype = ctypes.c_int
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != EXADG_OK:
        raise ExaDGError(status, lib().exadg_last_error().decode())


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


class LaplaceOperator:
    """Poisson::LaplaceOperator on the GPU (homogeneous SIPG operator, weak Dirichlet/Neumann)."""

    def __init__(self, mesh, bc, degree, precision=FP64, ip_factor=1.0):
        self._h = ctypes.c_void_p()
        check(lib().exadg_laplace_create(ctypes.byref(mesh), ctypes.byref(bc), degree, precision,
                                         ip_factor, ctypes.byref(self._h)))
        n = ctypes.c_int64()
        check(lib().exadg_laplace_n_dofs(self._h, ctypes.byref(n)))
        self.n = n.value
        nc, deg, prec, aff = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().exadg_laplace_info(self._h, ctypes.byref(nc), ctypes.byref(deg),
                                       ctypes.byref(prec), ctypes.byref(aff)))
        self.n_cells, self.degree, self.precision = nc.value, deg.value, prec.value
        self.is_affine = bool(aff.value)

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().exadg_laplace_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def dtype(self):
        import torch
        return torch.float64 if self.precision == FP64 else torch.float32

    def _check_vec(self, t):
        assert t.is_cuda and t.dtype == self.dtype and t.numel() == self.n and t.is_contiguous()

    def vmult(self, dst, src, stream=None):
        self._check_vec(dst)
        self._check_vec(src)
        check(lib().exadg_laplace_vmult(self._h, _ptr(dst), _ptr(src), _stream(stream)))

    def vmult_add(self, dst, src, stream=None):
        self._check_vec(dst)
        self._check_vec(src)
        check(lib().exadg_laplace_vmult_add(self._h, _ptr(dst), _ptr(src), _stream(stream)))

    def calculate_diagonal(self, diag, stream=None):
        self._check_vec(diag)
        check(lib().exadg_laplace_calculate_diagonal(self._h, _ptr(diag), _stream(stream)))

    def calculate_inverse_diagonal(self, diag, stream=None):
        self._check_vec(diag)
        check(lib().exadg_laplace_calculate_inverse_diagonal(self._h, _ptr(diag), _stream(stream)))

    def rhs_sine(self, omega, b, stream=None):
        import torch
        assert b.is_cuda and b.dtype == torch.float64 and b.numel() == self.n
        check(lib().exadg_laplace_rhs_sine(self._h, omega, _ptr(b), _stream(stream)))


class Preconditioner:
    """PreconditionerBase: None / PointJacobi / Multigrid (FP32 or FP64 V-cycle)."""

    def __init__(self, op, kind, mg=None, stream=None):
        self.kind = kind
        self._h = ctypes.c_void_p()
        check(lib().exadg_preconditioner_create(op.handle, kind,
                                                ctypes.byref(mg) if mg is not None else None,
                                                ctypes.byref(self._h), _stream(stream)))

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().exadg_preconditioner_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def vmult(self, dst, src, stream=None):
        check(lib().exadg_preconditioner_vmult(self._h, _ptr(dst), _ptr(src), _stream(stream)))

    def update(self, stream=None):
        check(lib().exadg_preconditioner_update(self._h, _stream(stream)))

    def mg_info(self):
        nl = ctypes.c_int()
        lmax = (ctypes.c_double * 32)()
        deg = (ctypes.c_int * 32)()
        cits = ctypes.c_int()
        check(lib().exadg_preconditioner_mg_info(self._h, ctypes.byref(nl), lmax, deg,
                                                 ctypes.byref(cits)))
        L = nl.value
        return dict(n_levels=L, lambda_max=[lmax[i] for i in range(L)],
                    chebyshev_degree=[deg[i] for i in range(L)],
                    last_coarse_iterations=cits.value)


def cg_solve(op, b, x, prec=None, rel_tol=1e-10, abs_tol=1e-20, max_iter=10000, rel_mode=0,
             stream=None, raise_on_error=True):
    """Krylov::KrylovSolver::solve (CG) on the FP64 operator; x holds the initial guess."""
    d = SolverData(max_iter, abs_tol, rel_tol, rel_mode)
    info = SolveInfo()
    rc = lib().exadg_cg_solve(op.handle, prec.handle if prec is not None else None, _ptr(b), _ptr(x),
                              ctypes.byref(d), ctypes.byref(info), _stream(stream))
    out = dict(iterations=info.iterations, converged=bool(info.converged),
               not_finite=bool(info.not_finite), r0=info.r0, r_final=info.r_final,
               b_norm=info.b_norm, rho=info.rho, n10=info.n10, status=rc)
    if rc != EXADG_OK and raise_on_error:
        raise ExaDGError(rc, lib().exadg_last_error().decode())
    return out
