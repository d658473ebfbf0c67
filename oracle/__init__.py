"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU-baseline leg of
bench.py as the checker. The product package never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

DIRICHLET, NEUMANN = 0, 1


class MeshDesc(ctypes.Structure):
    """Structured box, optionally deformed (DeformedCubeManifold) with MappingQ geometry."""

    _fields_ = [
        ("n_cells", ctypes.c_int * 3),
        ("lower", ctypes.c_double * 3),
        ("upper", ctypes.c_double * 3),
        ("deformation", ctypes.c_double),
        ("frequency", ctypes.c_int),
        ("mapping_degree", ctypes.c_int),
        ("periodic", ctypes.c_int * 3),
        ("boundary_type", ctypes.c_int * 6),
    ]


def make_mesh(n_cells, lower=0.0, upper=1.0, deformation=0.0, frequency=1, mapping_degree=1,
              periodic=(0, 0, 0), boundary_type=(DIRICHLET,) * 6):
    if isinstance(n_cells, int):
        n_cells = (n_cells,) * 3
    lo = (lower,) * 3 if np.isscalar(lower) else tuple(lower)
    hi = (upper,) * 3 if np.isscalar(upper) else tuple(upper)
    m = MeshDesc()
    m.n_cells[:] = list(n_cells)
    m.lower[:] = list(lo)
    m.upper[:] = list(hi)
    m.deformation = deformation
    m.frequency = frequency
    m.mapping_degree = mapping_degree
    m.periodic[:] = list(periodic)
    m.boundary_type[:] = list(boundary_type)
    return m


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def use_native_build(out_dir):
    """Build a copy of the oracle tuned for the host it runs on (-march=native) into out_dir and
    use it from now on: the CPU-baseline leg of bench.py, on the measuring host. Must be called
    before the library is first loaded. Returns the library path, or None if the build failed
    (the portable x86-64-v3 build is then used)."""
    global _LIB_PATH
    if _lib is not None:
        raise RuntimeError("the oracle library is already loaded")
    path = os.path.join(out_dir, "liboracle_native.so")
    r = subprocess.run(["make", "-s", "-B", "-C", _HERE, f"OUT={path}", "ARCH=-march=native"],
                       capture_output=True, text=True)
    if r.returncode != 0 or not os.path.exists(path):
        return None
    _LIB_PATH = path
    return path


def num_threads():
    """OpenMP threads of the oracle's parallel loops."""
    return lib().oracle_num_threads()


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.oracle_laplace_create.restype = ctypes.c_void_p
        L.oracle_laplace_create.argtypes = [ctypes.POINTER(MeshDesc), ctypes.c_int, ctypes.c_double]
        L.oracle_laplace_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_laplace_n_dofs.restype = ctypes.c_longlong
        L.oracle_laplace_n_dofs.argtypes = [ctypes.c_void_p]
        L.oracle_laplace_vmult.argtypes = [ctypes.c_void_p, dp, dp]
        L.oracle_laplace_rhs_sine.argtypes = [ctypes.c_void_p, ctypes.c_double, dp]
        L.oracle_divergence_apply.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp]
        L.oracle_laplace_rhs_divergence_fourier.argtypes = [ctypes.c_void_p, dp, ctypes.c_int,
                                                            ctypes.c_double, dp]
        L.oracle_laplace_error_sine.restype = ctypes.c_double
        L.oracle_laplace_error_sine.argtypes = [ctypes.c_void_p, ctypes.c_double, dp]
        L.oracle_laplace_diagonal.argtypes = [ctypes.c_void_p, dp]
        L.oracle_laplace_tau.argtypes = [ctypes.c_void_p, dp]
        L.oracle_laplace_cell_geometry.argtypes = [ctypes.c_void_p, ctypes.c_longlong, dp, dp]
        L.oracle_laplace_face_geometry.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int,
                                                   ctypes.c_int, dp, dp, dp, dp]
        L.oracle_tables.argtypes = [ctypes.c_int] + [dp] * 7
        L.oracle_random_vector.argtypes = [ctypes.c_longlong, ctypes.c_ulonglong, dp]
        L.oracle_num_threads.restype = ctypes.c_int
        L.oracle_num_threads.argtypes = []
        L.oracle_cg_solve.restype = ctypes.c_int
        L.oracle_cg_solve.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_int, dp]
        L.oracle_mg_create.restype = ctypes.c_void_p
        L.oracle_mg_create.argtypes = [ctypes.POINTER(MeshDesc), ctypes.POINTER(ctypes.c_int),
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int]
        L.oracle_laplace_block_diagonal.argtypes = [ctypes.c_void_p, dp]
        L.oracle_cg_solve_block_jacobi.restype = ctypes.c_int
        L.oracle_cg_solve_block_jacobi.argtypes = [ctypes.c_void_p, dp, dp, ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_int, dp]
        L.oracle_mg_level_list.restype = ctypes.c_int
        L.oracle_mg_level_list.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int] + \
            [ctypes.POINTER(ctypes.c_int)] * 3 + [ctypes.c_int]
        L.oracle_mg_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_mg_vmult.argtypes = [ctypes.c_void_p, dp, dp]
        L.oracle_mg_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                     ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_int),
                                     ctypes.POINTER(ctypes.c_int), dp, ctypes.POINTER(ctypes.c_int),
                                     ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.oracle_mg_level_vmult.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp]
        L.oracle_mg_prolongate_add.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp]
        L.oracle_mg_restrict_add.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp]
        L.oracle_mg_smoother_vmult.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp, dp]
        L.oracle_cg_solve_mg.restype = ctypes.c_int
        L.oracle_cg_solve_mg.argtypes = [ctypes.c_void_p, ctypes.c_void_p, dp, dp, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int, ctypes.c_int, dp]
        L.oracle_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


DEFAULT_SEED = 0x5EED2026


def random_vector(n, seed=DEFAULT_SEED):
    """SplitMix64(seed, index) -> uniform[-1,1], identical on CPU and GPU."""
    v = np.empty(n)
    lib().oracle_random_vector(n, seed, _p(v))
    return v


def tables(degree):
    n = degree + 1
    out = [np.empty(n), np.empty(n), np.empty(n), np.empty((n, n)), np.empty((n, n)), np.empty(n),
           np.empty(n)]
    lib().oracle_tables(degree, *[_p(a) for a in out])
    return dict(zip(["gl", "gp", "gw", "S", "Sd", "d0", "d1"], out))


class LaplaceOperator:
    """Oracle SIPG Laplace operator (homogeneous), mirrors Poisson::LaplaceOperator."""

    def __init__(self, mesh, degree, ip_factor=1.0):
        self.mesh = mesh
        self.degree = degree
        self._h = lib().oracle_laplace_create(ctypes.byref(mesh), degree, ip_factor)
        if not self._h:
            raise RuntimeError(lib().oracle_last_error().decode())
        self.n = lib().oracle_laplace_n_dofs(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_laplace_destroy(self._h)
            self._h = None

    def vmult(self, src):
        src = np.ascontiguousarray(src, dtype=np.float64)
        dst = np.empty(self.n)
        lib().oracle_laplace_vmult(self._h, _p(dst), _p(src))
        return dst

    def rhs_sine(self, omega):
        b = np.empty(self.n)
        lib().oracle_laplace_rhs_sine(self._h, omega, _p(b))
        return b

    def rhs_divergence_fourier(self, modes, t):
        """Weak DG divergence (central flux, zero wall flux) of the Fourier field `modes`
        (rows n(3) d(3) A phi omega pad) at time t."""
        m = np.ascontiguousarray(modes, dtype=np.float64).reshape(-1)
        b = np.empty(self.n)
        lib().oracle_laplace_rhs_divergence_fourier(self._h, _p(m), m.size // 10, t, _p(b))
        return b

    def divergence_apply(self, degree_u, u):
        """Weak DG divergence of the discrete velocity u (per cell three component blocks of
        (degree_u+1)^3 nodal values), tested with this operator's pressure space."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        assert u.size == (self.n // (self.degree + 1) ** 3) * 3 * (degree_u + 1) ** 3
        b = np.empty(self.n)
        lib().oracle_divergence_apply(self._h, degree_u, _p(u), _p(b))
        return b

    def error_sine(self, omega, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        return lib().oracle_laplace_error_sine(self._h, omega, _p(u))

    def diagonal(self):
        d = np.empty(self.n)
        lib().oracle_laplace_diagonal(self._h, _p(d))
        return d

    def block_diagonal(self):
        """the cells' diagonal blocks A_KK, shape (n_cells, n^3, n^3), [c, j, i] = (A_KK)_ij"""
        m = (self.degree + 1) ** 3
        blocks = np.empty((self.n // m, m, m))
        lib().oracle_laplace_block_diagonal(self._h, _p(blocks))
        return blocks

    def cg_solve_block_jacobi(self, b, x0=None, rel_tol=1e-10, abs_tol=1e-20, max_iter=100000):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        info = np.zeros(4)
        rc = lib().oracle_cg_solve_block_jacobi(self._h, _p(b), _p(x), rel_tol, abs_tol, max_iter,
                                                _p(info))
        if rc != 0:
            raise RuntimeError(lib().oracle_last_error().decode())
        return x, dict(iterations=int(info[0]), r0=info[1], r_final=info[2],
                       converged=bool(info[3]))

    def tau(self):
        nc = self.n // (self.degree + 1) ** 3
        t = np.empty(nc)
        lib().oracle_laplace_tau(self._h, _p(t))
        return t

    def cell_geometry(self, c):
        n3 = (self.degree + 1) ** 3
        jinv = np.empty((n3, 3, 3))
        jxw = np.empty(n3)
        lib().oracle_laplace_cell_geometry(self._h, c, _p(jinv), _p(jxw))
        return jinv, jxw

    def face_geometry(self, c, direction, side):
        n2 = (self.degree + 1) ** 2
        jinv, jxw, nrm, xq = np.empty((n2, 3, 3)), np.empty(n2), np.empty((n2, 3)), np.empty((n2, 3))
        lib().oracle_laplace_face_geometry(self._h, c, direction, side, _p(jinv), _p(jxw), _p(nrm),
                                           _p(xq))
        return jinv, jxw, nrm, xq

    def cg_solve(self, b, x0=None, rel_tol=1e-10, abs_tol=1e-20, max_iter=100000, jacobi=True):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        info = np.zeros(4)
        rc = lib().oracle_cg_solve(self._h, int(jacobi), _p(b), _p(x), rel_tol, abs_tol, max_iter,
                                   _p(info))
        if rc != 0:
            raise RuntimeError(lib().oracle_last_error().decode())
        return x, dict(iterations=int(info[0]), r0=info[1], r_final=info[2],
                       converged=bool(info[3]))


def bisect_degrees(k):
    """PSequenceType::Bisect: k, k/2, ..., 1 (multigrid_preconditioner_base.cpp:202-216)."""
    out = [k]
    while out[-1] > 1:
        out.append(max(out[-1] // 2, 1))
    return out


# MultigridType (multigrid_parameters.h:39-57), numbered as there; 0 = pcMG/pMG by coarse_q1
MG_TYPES = ["legacy", "hMG", "chMG", "hcMG", "cMG", "pMG", "cpMG", "pcMG", "hpMG", "chpMG",
            "hcpMG", "hpcMG", "phMG", "cphMG", "pchMG", "phcMG"]


def mg_level_list(mg_type, degrees, n_h_levels=1, coarse_q1=True):
    """Levels (coarse to fine) of a multigrid type as (h, degree, dg) tuples; mg_type is a name
    of MG_TYPES or its index."""
    t = MG_TYPES.index(mg_type) if isinstance(mg_type, str) else int(mg_type)
    arr = (ctypes.c_int * len(degrees))(*degrees)
    h, deg, dg = (ctypes.c_int * 64)(), (ctypes.c_int * 64)(), (ctypes.c_int * 64)()
    n = lib().oracle_mg_level_list(t, arr, len(degrees), n_h_levels, int(coarse_q1), h, deg, dg, 64)
    if n < 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return [(h[i], deg[i], bool(dg[i])) for i in range(n)]


class Multigrid:
    """Oracle hybrid multigrid V-cycle (FP64) over h-, p- and c-levels."""

    def __init__(self, mesh, degrees, coarse_q1=True, smoother_degree=5, smoothing_range=20.0,
                 eig_its=20, coarse_solver=0, coarse_rel_tol=1e-3, coarse_max_iter=10000,
                 singular=False, ip_factor=1.0, mg_type=0, n_h_levels=1,
                 smoother_preconditioner=0):
        t = MG_TYPES.index(mg_type) if isinstance(mg_type, str) else int(mg_type)
        arr = (ctypes.c_int * len(degrees))(*degrees)
        self._h = lib().oracle_mg_create(ctypes.byref(mesh), arr, len(degrees), int(coarse_q1),
                                         smoother_degree, smoothing_range, eig_its, coarse_solver,
                                         coarse_rel_tol, coarse_max_iter, int(singular), ip_factor,
                                         t, n_h_levels, int(smoother_preconditioner))
        if not self._h:
            raise RuntimeError(lib().oracle_last_error().decode())
        self.info()

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_mg_destroy(self._h)
            self._h = None

    def info(self):
        nl = ctypes.c_int()
        n = (ctypes.c_longlong * 32)()
        deg, dg, cd = (ctypes.c_int * 32)(), (ctypes.c_int * 32)(), (ctypes.c_int * 32)()
        hl = (ctypes.c_int * 32)()
        lmax = np.zeros(32)
        cits = ctypes.c_int()
        lib().oracle_mg_info(self._h, ctypes.byref(nl), n, deg, dg, _p(lmax), cd, ctypes.byref(cits),
                             hl)
        L = nl.value
        self.h_levels = [hl[i] for i in range(L)]
        self.n_levels = L
        self.sizes = [n[i] for i in range(L)]
        self.degrees = [deg[i] for i in range(L)]
        self.dg = [bool(dg[i]) for i in range(L)]
        self.lambda_max = lmax[:L].copy()
        self.cheb_degree = [cd[i] for i in range(L)]
        self.last_coarse_iterations = cits.value
        return self

    def vmult(self, src):
        src = np.ascontiguousarray(src, dtype=np.float64)
        dst = np.empty(self.sizes[-1])
        lib().oracle_mg_vmult(self._h, _p(dst), _p(src))
        return dst

    def smoother_vmult(self, level, src):
        src = np.ascontiguousarray(src, dtype=np.float64)
        dst = np.empty(self.sizes[level])
        td = np.zeros(2)
        lib().oracle_mg_smoother_vmult(self._h, level, _p(dst), _p(src), _p(td))
        return dst, td[0], td[1]

    def level_vmult(self, level, src):
        src = np.ascontiguousarray(src, dtype=np.float64)
        dst = np.empty(self.sizes[level])
        lib().oracle_mg_level_vmult(self._h, level, _p(dst), _p(src))
        return dst

    def prolongate(self, level, coarse):
        fine = np.zeros(self.sizes[level])
        lib().oracle_mg_prolongate_add(self._h, level, _p(fine),
                                       _p(np.ascontiguousarray(coarse, dtype=np.float64)))
        return fine

    def restrict(self, level, fine):
        coarse = np.zeros(self.sizes[level - 1])
        lib().oracle_mg_restrict_add(self._h, level, _p(coarse),
                                     _p(np.ascontiguousarray(fine, dtype=np.float64)))
        return coarse


def cg_solve_mg(op, mg, b, x0=None, rel_tol=1e-10, abs_tol=1e-20, max_iter=10000, rel_mode=0):
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(op.n) if x0 is None else np.array(x0, dtype=np.float64)
    info = np.zeros(4)
    rc = lib().oracle_cg_solve_mg(op._h, mg._h if mg is not None else None, _p(b), _p(x), rel_tol,
                                  abs_tol, max_iter, rel_mode, _p(info))
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return x, dict(iterations=int(info[0]), r0=info[1], r_final=info[2], converged=bool(info[3]))
