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
