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
EXADG_ERR_ARGUMENT, EXADG_ERR_HIP, EXADG_ERR_NOT_FINITE, EXADG_ERR_NO_CONVERGE = 1, 2, 3, 4
FP64, FP32 = 0, 1
DIRICHLET, NEUMANN = 0, 1
GUESS_LOSS_L2, GUESS_LOSS_ENERGY, GUESS_LOSS_RESIDUAL = 0, 1, 2

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
        ("capture_graph", ctypes.c_int32),
        ("mg_type", ctypes.c_int32),
        ("n_h_levels", ctypes.c_int32),
        ("smoother_preconditioner", ctypes.c_int32),
    ]


# MultigridType (exadg_mg_type), numbered as in the reference; "legacy" = pcMG/pMG by coarse_q1
MG_TYPES = ["legacy", "hMG", "chMG", "hcMG", "cMG", "pMG", "cpMG", "pcMG", "hpMG", "chpMG",
            "hcpMG", "hpcMG", "phMG", "cphMG", "pchMG", "phcMG"]


class SolverData(ctypes.Structure):
    _fields_ = [("max_iter", ctypes.c_int32), ("abs_tol", ctypes.c_double),
                ("rel_tol", ctypes.c_double), ("rel_mode", ctypes.c_int32)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("converged", ctypes.c_int32),
                ("not_finite", ctypes.c_int32), ("r0", ctypes.c_double), ("r_final", ctypes.c_double),
                ("b_norm", ctypes.c_double), ("rho", ctypes.c_double), ("n10", ctypes.c_double),
                ("r_true", ctypes.c_double), ("tolerance", ctypes.c_double)]


PRECOND_NONE, PRECOND_POINT_JACOBI, PRECOND_MULTIGRID = 0, 1, 2
PRECOND_BLOCK_JACOBI, PRECOND_ADDITIVE_SCHWARZ = 3, 4
SMOOTHER_POINT_JACOBI, SMOOTHER_BLOCK_JACOBI, SMOOTHER_ADDITIVE_SCHWARZ = 0, 1, 2
COARSE_CG, COARSE_CHEBYSHEV = 0, 1


def mg_desc(degree, degrees=None, precision=FP32, coarse_q1=True, coarse_solver=COARSE_CG,
            coarse_rel_tol=1e-3, smoother_degree=5, smoothing_range=20.0, eig_cg_iterations=20,
            singular=False, capture_graph=False, mg_type="legacy", n_h_levels=0,
            smoother_preconditioner=0):
    """MultigridData; default p-sequence Bisect (k, k/2, ..., 1) + continuous Q1 level.
    mg_type: a name of MG_TYPES (or its index); n_h_levels: meshes of the h-levels (0 =
    automatic)."""
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
    d.capture_graph = int(capture_graph)
    d.mg_type = MG_TYPES.index(mg_type) if isinstance(mg_type, str) else int(mg_type)
    d.n_h_levels = int(n_h_levels)
    d.smoother_preconditioner = int(smoother_preconditioner)
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
        dp = ctypes.POINTER(ctypes.c_double)
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
            "exadg_laplace_rhs_n_points": [vp, i64p, i64p],
            "exadg_laplace_rhs_points": [vp, vp, vp, vp],
            "exadg_laplace_rhs": [vp, vp, vp, vp, vp],
            "exadg_mg_desc_default": [ctypes.c_int, ctypes.POINTER(MGDesc)],
            "exadg_preconditioner_create": [vp, ctypes.c_int, ctypes.POINTER(MGDesc),
                                            ctypes.POINTER(vp), vp],
            "exadg_preconditioner_vmult": [vp, vp, vp, vp],
            "exadg_preconditioner_update": [vp, vp],
            "exadg_preconditioner_smooth": [vp, vp, vp, vp, vp],
            "exadg_preconditioner_destroy": [vp],
            "exadg_preconditioner_mg_info": [vp, ip, ctypes.POINTER(ctypes.c_double), ip, ip],
            "exadg_preconditioner_mg_levels": [vp, ctypes.c_int, ip, ip, ip, ip, i64p],
            "exadg_cg_solve": [vp, vp, vp, vp, ctypes.POINTER(SolverData), ctypes.POINTER(SolveInfo),
                               vp],
            "exadg_cg_solve_noalloc": [vp, vp, vp, vp, ctypes.POINTER(SolverData),
                                       ctypes.POINTER(SolveInfo), vp],
            "exadg_laplace_error_sine": [vp, ctypes.c_double, vp, dp, vp],
            "exadg_fourier_modes": [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ip],
            "exadg_laplace_rhs_divergence_fourier": [vp, dp, ctypes.c_int, ctypes.c_double, vp, vp],
            "exadg_vector_subtract_mean": [vp, ctypes.c_int64, vp],
            "exadg_vector_random": [vp, ctypes.c_int64, ctypes.c_int, ctypes.c_uint64, vp],
            "exadg_extrapolate": [ctypes.c_int, ctypes.POINTER(vp), vp, ctypes.c_int64, vp],
            "exadg_guess_create": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                   ctypes.POINTER(vp)],
            "exadg_guess_destroy": [vp],
            "exadg_guess_n_params": [vp, ip, ip, ip],
            "exadg_guess_get_params": [vp, vp, vp],
            "exadg_guess_set_params": [vp, vp, vp],
            "exadg_guess_set_scales": [vp, dp, ctypes.c_int],
            "exadg_guess_get_scales": [vp, dp, ctypes.c_int],
            "exadg_guess_loss_and_grad": [vp, vp, vp, ctypes.c_int, vp, dp, vp],
            "exadg_guess_train": [vp, vp, ctypes.c_int, vp, vp, ctypes.c_int64, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_uint64, dp, vp],
            "exadg_guess_evaluate": [vp, vp, ctypes.c_int, vp, vp, ctypes.c_int64, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, dp, vp],
            "exadg_guess_predict": [vp, ctypes.POINTER(vp), vp, vp, ctypes.c_int64, ip, vp],
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


def _require(cond, msg):
    if not cond:
        raise ValueError(msg)


def check_vector(t, n, dtype, what="vector"):
    """The C entry points take bare device pointers, so every wrapper checks a tensor before it
    is passed: n elements of `dtype`, contiguous, on the GPU. Raises ValueError otherwise (before
    any kernel launch)."""
    _require(t is not None, f"{what}: missing")
    _require(t.numel() == n, f"{what}: {t.numel()} entries, expected {n}")
    _require(t.dtype == dtype, f"{what}: dtype {t.dtype}, expected {dtype}")
    _require(t.is_contiguous(), f"{what}: must be contiguous")
    _require(t.is_cuda, f"{what}: must be a device tensor")


def check_sequence(P, Bv, t_begin, t_end, history=0):
    """A stored solution/right-hand-side sequence (steps x n_dofs, FP64, device) and the step
    range [t_begin, t_end) that a GuessNet call reads (history steps before t_begin + history
    included): both tensors 2-D of the same shape, 0 <= t_begin < t_end <= steps."""
    import torch
    _require(P.dim() == 2 and Bv.dim() == 2, "P and Bv must be 2-D (steps x n_dofs)")
    _require(tuple(P.shape) == tuple(Bv.shape), "P and Bv must have the same shape")
    _require(0 <= t_begin < t_end <= P.shape[0],
             f"step range [{t_begin}, {t_end}) outside the {P.shape[0]} stored steps")
    _require(t_end - t_begin > history, "step range shorter than the history")
    for t, what in ((P, "P"), (Bv, "Bv")):
        check_vector(t, t.numel(), torch.float64, what)


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

    def _check_vec(self, t, what="vector"):
        check_vector(t, self.n, self.dtype, what)

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
        check_vector(b, self.n, torch.float64, "b")
        check(lib().exadg_laplace_rhs_sine(self._h, omega, _ptr(b), _stream(stream)))

    def rhs_n_points(self):
        """(cell points, boundary points) at which rhs() takes its data."""
        nc, nb = ctypes.c_int64(), ctypes.c_int64()
        check(lib().exadg_laplace_rhs_n_points(self._h, ctypes.byref(nc), ctypes.byref(nb)))
        return nc.value, nb.value

    def rhs_points(self, stream=None):
        """Coordinates of the right-hand-side points: device tensors (3, n_cell_points) and
        (3, n_boundary_points), FP64."""
        import torch
        nc, nb = self.rhs_n_points()
        xc = torch.empty((3, nc), dtype=torch.float64, device="cuda")
        xb = torch.empty((3, max(nb, 0)), dtype=torch.float64, device="cuda")
        check(lib().exadg_laplace_rhs_points(self._h, _ptr(xc), _ptr(xb) if nb > 0 else None,
                                             _stream(stream)))
        return xc, xb

    def rhs(self, b, f=None, boundary=None, stream=None):
        """b = int f v + Dirichlet (-g dv/dn + 2 tau g v) + Neumann h v from point values (f at the
        cell points, g / h at the boundary points; None = zero)."""
        import torch
        nc, nb = self.rhs_n_points()
        check_vector(b, self.n, torch.float64, "b")
        if f is not None:
            check_vector(f, nc, torch.float64, "f")
        if boundary is not None:
            check_vector(boundary, nb, torch.float64, "boundary")
        check(lib().exadg_laplace_rhs(self._h, _ptr(f), _ptr(boundary), _ptr(b), _stream(stream)))

    def error_sine(self, omega, u, stream=None):
        """Relative L2 error against prod sin(omega x_d), QGauss(k+3) (FP64 operator)."""
        import torch
        check_vector(u, self.n, torch.float64, "u")
        e = ctypes.c_double()
        check(lib().exadg_laplace_error_sine(self._h, omega, _ptr(u), ctypes.byref(e),
                                             _stream(stream)))
        return e.value

    def rhs_divergence_fourier(self, modes, t, b, stream=None):
        """Weak DG divergence (central flux, no-slip walls) of the Fourier field `modes`
        (numpy float64, 10 per mode) at time t; mean removed for a singular operator."""
        import numpy as np
        import torch
        check_vector(b, self.n, torch.float64, "b")
        m = np.ascontiguousarray(modes, dtype=np.float64).reshape(-1)
        _require(m.size % 10 == 0 and m.size > 0, "modes: 10 doubles per mode expected")
        check(lib().exadg_laplace_rhs_divergence_fourier(
            self._h, m.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), m.size // 10, t, _ptr(b),
            _stream(stream)))


class Preconditioner:
    """PreconditionerBase: None / PointJacobi / Multigrid (FP32 or FP64 V-cycle)."""

    def __init__(self, op, kind, mg=None, stream=None):
        self.kind = kind
        # the C preconditioner refers to the operator (point Jacobi) and was built for its size:
        # keep the operator alive as long as this object
        self._op = op
        self.n = op.n
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
        import torch
        check_vector(dst, self.n, torch.float64, "dst")
        check_vector(src, self.n, torch.float64, "src")
        check(lib().exadg_preconditioner_vmult(self._h, _ptr(dst), _ptr(src), _stream(stream)))

    def update(self, stream=None):
        check(lib().exadg_preconditioner_update(self._h, _stream(stream)))

    def smooth(self, op, b, x, stream=None):
        """x <- x + S (b - A x), S the finest-level smoother (multigrid only)."""
        import torch
        _require(op.n == self.n, "operator and preconditioner sizes differ")
        check_vector(b, self.n, torch.float64, "b")
        check_vector(x, self.n, torch.float64, "x")
        check(lib().exadg_preconditioner_smooth(op.handle, self._h, _ptr(b), _ptr(x), _stream(stream)))

    def mg_levels(self):
        """[(h, degree, dg, n)] per level, coarse to fine (h = 0: the coarsest mesh)."""
        nl = ctypes.c_int()
        h, deg, dg = (ctypes.c_int * 64)(), (ctypes.c_int * 64)(), (ctypes.c_int * 64)()
        n = (ctypes.c_int64 * 64)()
        check(lib().exadg_preconditioner_mg_levels(self._h, 64, ctypes.byref(nl), h, deg, dg, n))
        return [(h[i], deg[i], bool(dg[i]), n[i]) for i in range(nl.value)]

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


NOALLOC_MAX_DOFS = 1 << 22  # EXADG_NOALLOC_MAX_DOFS


def cg_solve(op, b, x, prec=None, rel_tol=1e-10, abs_tol=1e-20, max_iter=10000, rel_mode=0,
             stream=None, raise_on_error=True, noalloc=False):
    """Krylov::KrylovSolver::solve (CG) on the FP64 operator; x holds the initial guess.
    The result carries the final FP64 true residual r_true = ||b - A x||. noalloc=True calls
    exadg_cg_solve_noalloc (work vectors in the library's static device workspace; at most
    NOALLOC_MAX_DOFS DoF)."""
    import torch
    _require(prec is None or prec.n == op.n, "operator and preconditioner sizes differ")
    check_vector(b, op.n, torch.float64, "b")
    check_vector(x, op.n, torch.float64, "x")
    d = SolverData(max_iter, abs_tol, rel_tol, rel_mode)
    info = SolveInfo()
    fn = lib().exadg_cg_solve_noalloc if noalloc else lib().exadg_cg_solve
    rc = fn(op.handle, prec.handle if prec is not None else None, _ptr(b), _ptr(x),
            ctypes.byref(d), ctypes.byref(info), _stream(stream))
    out = dict(iterations=info.iterations, converged=bool(info.converged),
               not_finite=bool(info.not_finite), r0=info.r0, r_final=info.r_final,
               b_norm=info.b_norm, rho=info.rho, n10=info.n10, r_true=info.r_true,
               tolerance=info.tolerance, status=rc)
    if rc != EXADG_OK and raise_on_error:
        raise ExaDGError(rc, lib().exadg_last_error().decode())
    return out


def fourier_modes(seed, k_min=1, k_max=8):
    """Seeded solenoidal Fourier field over the shell k_min <= |n| <= k_max (numpy, shape
    (n_modes, 10)): n(3), d(3), A, phi, omega, 0 per mode."""
    import numpy as np
    cnt = ctypes.c_int()
    check(lib().exadg_fourier_modes(seed, k_min, k_max, 0, None, ctypes.byref(cnt)))
    out = np.zeros((cnt.value, 10), dtype=np.float64)
    check(lib().exadg_fourier_modes(seed, k_min, k_max, cnt.value,
                                    out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                    ctypes.byref(cnt)))
    return out


DEFAULT_SEED = 0x5EED2026


def random_vector(n, seed=DEFAULT_SEED, precision=None, stream=None):
    """Device vector of n uniform[-1,1] values, SplitMix64 on (seed, index) (exadg_vector_random);
    FP64 unless precision is FP32."""
    import torch
    precision = FP64 if precision is None else precision
    v = torch.empty(n, dtype=torch.float64 if precision == FP64 else torch.float32, device="cuda")
    check(lib().exadg_vector_random(_ptr(v), n, precision, seed, _stream(stream)))
    return v


def subtract_mean(v, stream=None):
    import torch
    assert v.is_cuda and v.dtype == torch.float64 and v.is_contiguous()
    check(lib().exadg_vector_subtract_mean(_ptr(v), v.numel(), _stream(stream)))


def _history(vectors):
    arr = (ctypes.c_void_p * max(1, len(vectors)))()
    for i, t in enumerate(vectors):
        arr[i] = t.data_ptr()
    return arr


def extrapolate(history, x0, stream=None):
    """x0 = sum_i beta_i history[i] (history newest first, FP64 device vectors)."""
    import torch
    _require(1 <= len(history) <= 4, "extrapolation order must be in [1,4]")
    for t in history:
        check_vector(t, x0.numel(), torch.float64, "history vector")
    check_vector(x0, x0.numel(), torch.float64, "x0")
    check(lib().exadg_extrapolate(len(history), _history(history), _ptr(x0), x0.numel(),
                                  _stream(stream)))


class GuessNet:
    """Learned initial guess: per-cell MLP correcting the order-`history` extrapolation."""

    def __init__(self, degree, history=3, width=128, seed=1):
        self._h = ctypes.c_void_p()
        check(lib().exadg_guess_create(degree, history, width, seed, ctypes.byref(self._h)))
        npar, nf, no = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().exadg_guess_n_params(self._h, ctypes.byref(npar), ctypes.byref(nf),
                                         ctypes.byref(no)))
        self.degree, self.history, self.width = degree, history, width
        self.n_params, self.n_features, self.n_outputs = npar.value, nf.value, no.value

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().exadg_guess_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        """Copy (FP32, device) of the flat parameters [W1|b1|W2|b2|W3|b3]."""
        import torch
        out = torch.empty(self.n_params, dtype=torch.float32, device="cuda")
        check(lib().exadg_guess_get_params(self._h, _ptr(out), _stream(None)))
        return out

    def set_params(self, values):
        import torch
        v = values.to(device="cuda", dtype=torch.float32).contiguous()
        assert v.numel() == self.n_params
        torch.cuda.current_stream().synchronize()
        check(lib().exadg_guess_set_params(self._h, _ptr(v), _stream(None)))

    def scales(self):
        """[s_0 .. s_{H-1}, s_b, s_t]: RMS of the history differences, the rhs, the targets."""
        n = self.history + 2
        arr = (ctypes.c_double * n)()
        check(lib().exadg_guess_get_scales(self._h, arr, n))
        return list(arr)

    def set_scales(self, scales):
        n = self.history + 2
        assert len(scales) == n
        arr = (ctypes.c_double * n)(*scales)
        check(lib().exadg_guess_set_scales(self._h, arr, n))

    def loss_and_grad(self, X, T, stream=None):
        import torch
        _require(X.dim() == 2 and X.shape[1] == self.n_features and X.shape[0] > 0,
                 f"X must be B x {self.n_features}")
        B = X.shape[0]
        _require(tuple(T.shape) == (B, self.n_outputs), f"T must be {B} x {self.n_outputs}")
        check_vector(X, X.numel(), torch.float32, "X")
        check_vector(T, T.numel(), torch.float32, "T")
        g = torch.empty(self.n_params, dtype=torch.float32, device=X.device)
        loss = ctypes.c_double()
        check(lib().exadg_guess_loss_and_grad(self._h, _ptr(X), _ptr(T), B, _ptr(g),
                                              ctypes.byref(loss), _stream(stream)))
        torch.cuda.synchronize()
        return loss.value, g

    def train(self, P, Bv, t_begin, t_end, epochs, op=None, loss=None, batch_cells=4096, lr=1e-3,
              seed=7, stream=None):
        """Adam on the FP64 sequence P, Bv (steps x n_dofs, device). loss: GUESS_LOSS_* (default
        ENERGY with an operator, else L2). Returns the last epoch's normalized loss."""
        check_sequence(P, Bv, t_begin, t_end, self.history)
        if loss is None:
            loss = GUESS_LOSS_ENERGY if op is not None else GUESS_LOSS_L2
        out = ctypes.c_double()
        check(lib().exadg_guess_train(self._h, op.handle if op is not None else None, loss, _ptr(P),
                                      _ptr(Bv), P.shape[1], t_begin, t_end, epochs, batch_cells, lr,
                                      seed, ctypes.byref(out), _stream(stream)))
        return out.value

    def evaluate(self, P, Bv, t_begin, t_end, op=None, loss=None, batch_cells=4096, stream=None):
        check_sequence(P, Bv, t_begin, t_end, self.history)
        if loss is None:
            loss = GUESS_LOSS_ENERGY if op is not None else GUESS_LOSS_L2
        out = ctypes.c_double()
        check(lib().exadg_guess_evaluate(self._h, op.handle if op is not None else None, loss,
                                         _ptr(P), _ptr(Bv), P.shape[1], t_begin, t_end, batch_cells,
                                         ctypes.byref(out), _stream(stream)))
        return out.value

    def predict(self, history, rhs, x0, stream=None):
        """x0 from the FP64 history (newest first) and rhs; returns True if the network output was
        non-finite and the plain extrapolation was used."""
        import torch
        _require(len(history) == self.history, f"{self.history} history vectors expected")
        n3 = self.n_outputs
        _require(x0.numel() % n3 == 0 and x0.numel() > 0,
                 f"x0: {x0.numel()} entries is not a whole number of {n3}-DoF cells")
        check_vector(x0, x0.numel(), torch.float64, "x0")
        check_vector(rhs, x0.numel(), torch.float64, "rhs")
        for t in history:
            check_vector(t, x0.numel(), torch.float64, "history vector")
        fb = ctypes.c_int()
        check(lib().exadg_guess_predict(self._h, _history(history), _ptr(rhs), _ptr(x0),
                                        x0.numel() // n3, ctypes.byref(fb), _stream(stream)))
        return bool(fb.value)


# ---------------------------------------------------------------- snapshots (restart files)

class SnapshotInfo(ctypes.Structure):
    _fields_ = [("version", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("n_vectors", ctypes.c_int32), ("n", ctypes.c_int64), ("degree", ctypes.c_int32),
                ("mesh", MeshDesc), ("time", ctypes.c_double), ("step", ctypes.c_int64),
                ("user", ctypes.c_double * 16)]


def _snapshot_lib():
    L = lib()
    if not getattr(L, "_snapshot_sig", False):
        vp = ctypes.c_void_p
        L.exadg_snapshot_write.restype = ctypes.c_int
        L.exadg_snapshot_write.argtypes = [ctypes.c_char_p, ctypes.POINTER(SnapshotInfo),
                                           ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, vp]
        L.exadg_snapshot_read_info.restype = ctypes.c_int
        L.exadg_snapshot_read_info.argtypes = [ctypes.c_char_p, ctypes.POINTER(SnapshotInfo)]
        L.exadg_snapshot_read.restype = ctypes.c_int
        L.exadg_snapshot_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(SnapshotInfo),
                                          ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int64,
                                          ctypes.c_int, ctypes.c_int, vp]
        L._snapshot_sig = True
    return L


def _snapshot_vectors(vectors):
    """(pointer array, n, precision, on_device) of torch device tensors or numpy host arrays."""
    import numpy as np
    if len(vectors) == 0:
        return (ctypes.c_void_p * 1)(), 0, FP64, 0
    first = vectors[0]
    on_device = 0 if isinstance(first, np.ndarray) else 1
    n = first.size if on_device == 0 else first.numel()
    if on_device:
        import torch
        dtype = first.dtype
        _require(dtype in (torch.float64, torch.float32), "snapshot vectors must be FP64 or FP32")
        precision = FP64 if dtype == torch.float64 else FP32
        for v in vectors:
            check_vector(v, n, dtype, "snapshot vector")
        ptrs = [v.data_ptr() for v in vectors]
    else:
        dtype = first.dtype
        _require(dtype in (np.float64, np.float32), "snapshot vectors must be FP64 or FP32")
        precision = FP64 if dtype == np.float64 else FP32
        for v in vectors:
            _require(isinstance(v, np.ndarray) and v.dtype == dtype and v.size == n and
                     v.flags["C_CONTIGUOUS"], "host snapshot vectors: same size, dtype, contiguous")
        ptrs = [v.ctypes.data for v in vectors]
    arr = (ctypes.c_void_p * len(vectors))(*ptrs)
    return arr, n, precision, on_device


def _info_dict(s):
    return dict(version=s.version, precision=s.precision, n_vectors=s.n_vectors, n=s.n,
                degree=s.degree, time=s.time, step=s.step, user=list(s.user),
                mesh=dict(n_cells=list(s.mesh.n_cells), lower=list(s.mesh.lower),
                          upper=list(s.mesh.upper), deformation=s.mesh.deformation,
                          frequency=s.mesh.frequency, mapping_degree=s.mesh.mapping_degree,
                          periodic=list(s.mesh.periodic)))


def snapshot_write(path, vectors, time=0.0, step=0, degree=0, mesh=None, user=(), keep=1,
                   stream=None):
    """Write vectors (torch device tensors or numpy host arrays, all of one length and dtype) with
    the discretization parameters, time, step and up to 16 user values; older snapshots are kept
    as path.old.1 ... up to `keep` files."""
    arr, n, precision, on_device = _snapshot_vectors(vectors)
    s = SnapshotInfo()
    s.precision, s.n_vectors, s.n, s.degree = precision, len(vectors), n, degree
    if mesh is not None:
        s.mesh = mesh
    s.time, s.step = time, step
    _require(len(user) <= 16, "at most 16 user values")
    for i, u in enumerate(user):
        s.user[i] = u
    check(_snapshot_lib().exadg_snapshot_write(os.fsencode(path), ctypes.byref(s), arr, on_device,
                                               keep, _stream(stream) if on_device else None))


def snapshot_read_info(path):
    s = SnapshotInfo()
    check(_snapshot_lib().exadg_snapshot_read_info(os.fsencode(path), ctypes.byref(s)))
    return _info_dict(s)


def snapshot_read(path, vectors, stream=None):
    """Read a snapshot into the given vectors (their count, length and dtype must match)."""
    arr, n, precision, on_device = _snapshot_vectors(vectors)
    s = SnapshotInfo()
    check(_snapshot_lib().exadg_snapshot_read(os.fsencode(path), ctypes.byref(s), arr, len(vectors),
                                              n, precision, on_device,
                                              _stream(stream) if on_device else None))
    return _info_dict(s)


# ---------------------------------------------------------------- pressure step (INS dual splitting)

def _divergence_lib():
    L = lib()
    if not getattr(L, "_divergence_sig", False):
        vp, i64p = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)
        dp = ctypes.POINTER(ctypes.c_double)
        for name, args in {
            "exadg_divergence_create": [vp, ctypes.c_int, ctypes.POINTER(vp)],
            "exadg_divergence_destroy": [vp],
            "exadg_divergence_n_dofs": [vp, i64p, i64p],
            "exadg_divergence_apply": [vp, vp, vp, vp],
            "exadg_divergence_velocity_points": [vp, vp, vp],
            "exadg_divergence_interpolate_fourier": [vp, dp, ctypes.c_int, ctypes.c_double, vp, vp],
            "exadg_pressure_step": [vp, vp, vp, vp, ctypes.c_double, ctypes.c_int,
                                    ctypes.POINTER(vp), vp, vp, ctypes.POINTER(SolverData),
                                    ctypes.POINTER(SolveInfo), vp],
        }.items():
            f = getattr(L, name)
            f.restype = ctypes.c_int
            f.argtypes = args
        L._divergence_sig = True
    return L


class DivergenceOperator:
    """IncNS::DivergenceOperator (weak form, central flux) from FE_DGQ(degree_u)^3 velocities to
    the pressure space of a (FP64) LaplaceOperator; velocity layout: per cell three component
    blocks of (degree_u+1)^3 nodal values."""

    def __init__(self, op, degree_u):
        self._op = op
        self._h = ctypes.c_void_p()
        L = _divergence_lib()
        check(L.exadg_divergence_create(op.handle, degree_u, ctypes.byref(self._h)))
        nv, npr = ctypes.c_int64(), ctypes.c_int64()
        check(L.exadg_divergence_n_dofs(self._h, ctypes.byref(nv), ctypes.byref(npr)))
        self.degree_u, self.n_velocity, self.n_pressure = degree_u, nv.value, npr.value

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _divergence_lib().exadg_divergence_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def apply(self, dst, src, stream=None):
        import torch
        check_vector(dst, self.n_pressure, torch.float64, "dst")
        check_vector(src, self.n_velocity, torch.float64, "src")
        check(_divergence_lib().exadg_divergence_apply(self._h, _ptr(dst), _ptr(src), _stream(stream)))

    def velocity_points(self, stream=None):
        """Positions of the velocity nodes, (3, n_velocity / 3) device tensor."""
        import torch
        x = torch.empty((3, self.n_velocity // 3), dtype=torch.float64, device="cuda")
        check(_divergence_lib().exadg_divergence_velocity_points(self._h, _ptr(x), _stream(stream)))
        return x

    def interpolate_fourier(self, modes, t, u, stream=None):
        import numpy as np
        import torch
        check_vector(u, self.n_velocity, torch.float64, "u")
        m = np.ascontiguousarray(modes, dtype=np.float64).reshape(-1)
        _require(m.size % 10 == 0 and m.size > 0, "modes: 10 doubles per mode expected")
        check(_divergence_lib().exadg_divergence_interpolate_fourier(
            self._h, m.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), m.size // 10, t, _ptr(u),
            _stream(stream)))


def pressure_step(op, prec, div, velocity, scale, history, p_np, rhs, rel_tol=1e-10, abs_tol=1e-20,
                  max_iter=10000, rel_mode=1, stream=None, raise_on_error=True):
    """TimeIntBDFDualSplitting::pressure_step: rhs = scale * div(velocity) (scale = -gamma0/dt),
    p_np = extrapolation of `history` (newest first; [] keeps p_np as the guess), CG solve."""
    import torch
    _require(prec is None or prec.n == op.n, "operator and preconditioner sizes differ")
    _require(len(history) <= 4, "at most 4 history vectors")
    check_vector(velocity, div.n_velocity, torch.float64, "velocity")
    check_vector(p_np, op.n, torch.float64, "p_np")
    check_vector(rhs, op.n, torch.float64, "rhs")
    for h in history:
        check_vector(h, op.n, torch.float64, "history vector")
    d = SolverData(max_iter, abs_tol, rel_tol, rel_mode)
    info = SolveInfo()
    rc = _divergence_lib().exadg_pressure_step(
        op.handle, prec.handle if prec is not None else None, div.handle, _ptr(velocity), scale,
        len(history), _history(history), _ptr(p_np), _ptr(rhs), ctypes.byref(d), ctypes.byref(info),
        _stream(stream))
    out = dict(iterations=info.iterations, converged=bool(info.converged), r0=info.r0,
               r_final=info.r_final, b_norm=info.b_norm, r_true=info.r_true,
               tolerance=info.tolerance, status=rc)
    if rc != EXADG_OK and raise_on_error:
        raise ExaDGError(rc, lib().exadg_last_error().decode())
    return out
