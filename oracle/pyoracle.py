"""ctypes bindings of the CPU oracle (oracle/giem2g_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's CPU-baseline leg, as the checker.  The product package
(giem2g_amd) never imports this module.

Also binds oracle/_ref/liboracle_ref.so (the reference's own Fortran modules
compiled in this container) when it exists.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libgiem2g_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "liboracle_ref.so")
REF_QUAD_PATH = os.path.join(HERE, "_ref", "liboracle_ref_quad.so")
NTAPS = 501

_lib = None
_ref = None
_ref_quad = None

dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
cdp = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])


class GmresInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("stat", C.c_int), ("bea", C.c_double),
                ("be", C.c_double), ("matvecs", C.c_int64), ("dotprods", C.c_int64)]


MATVEC = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_model_new.restype = C.c_void_p
        L.or_model_new.argtypes = [C.c_int, dp, dp, C.c_double, C.c_int, C.c_int, C.c_int,
                                   C.c_double, C.c_double, dp]
        L.or_model_free.argtypes = [C.c_void_p]
        L.or_model_lnum.argtypes = [C.c_void_p, C.c_int]
        L.or_mu0.restype = C.c_double
        L.or_eps0.restype = C.c_double
        L.or_vbt_lms.argtypes = [C.c_double, dp]
        L.or_vbt_weights.argtypes = [C.c_double] * 5 + [dp]
        L.or_compute_weights.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, dp, dp]
        L.or_calc_apq.argtypes = [C.c_void_p, C.c_double, cdp, cdp, cdp, cdp, C.c_void_p]
        L.or_double_integral_U.argtypes = [C.c_void_p, C.c_double, dp, cdp]
        L.or_block_size.restype = C.c_int64
        L.or_block_size.argtypes = [C.c_int]
        L.or_tensor_element.argtypes = [C.c_void_p, C.c_int, C.c_int, cdp]
        L.or_tensor_element_abssum.argtypes = [C.c_void_p, C.c_int, C.c_int, dp]
        L.or_tensor_realspace.argtypes = [C.c_void_p, cdp, C.c_int]
        L.or_tensor_fft.argtypes = [C.c_void_p, cdp, cdp, C.c_int]
        L.or_apply.argtypes = [C.c_void_p, cdp, cdp, cdp, cdp, C.c_int]
        L.or_apply_bruteforce.argtypes = [C.c_void_p, cdp, cdp, cdp, cdp, C.c_int]
        L.or_apply_timings.argtypes = [dp]
        L.or_plane_wave_integral.argtypes = [C.c_void_p, C.c_int, cdp]
        L.or_fgmres.argtypes = [C.c_int64, C.c_int, C.c_int, C.c_int, C.c_double, MATVEC,
                                C.c_void_p, cdp, cdp, cdp, C.POINTER(GmresInfo), C.c_void_p, C.c_int]
        L.or_solve.argtypes = [C.c_void_p, cdp, cdp, cdp, C.c_void_p, C.c_int, C.c_int, C.c_int,
                               C.c_double, cdp, C.POINTER(GmresInfo), C.c_int]
        L.or_fft.argtypes = [cdp, C.c_int, C.c_int]
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        L.or_prepare_recvs.argtypes = [C.c_void_p, C.c_int, dp, dp, ip, ip]
        L.or_vbt_weights_single.argtypes = [C.c_double] * 5 + [dp]
        L.or_rc_integral_U.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_double, dp,
                                       cdp, cdp]
        L.or_rc_block_size.restype = C.c_int64
        L.or_rc_block_size.argtypes = [C.c_int, C.c_int]
        L.or_rc_tensor_element.argtypes = [C.c_void_p, C.c_int, dp, dp, dp, ip, ip, C.c_int,
                                           C.c_int, cdp]
        L.or_rc_tensor_realspace.argtypes = [C.c_void_p, C.c_int, dp, dp, dp, ip, ip, cdp, C.c_int]
        L.or_rc_tensor_fft.argtypes = [C.c_void_p, C.c_int, cdp, cdp, C.c_int]
        L.or_rc_apply.argtypes = [C.c_void_p, C.c_int, cdp, cdp, cdp, cdp, cdp, C.c_int]
        L.or_rc_apply_bruteforce.argtypes = [C.c_void_p, C.c_int, cdp, cdp, cdp, cdp, cdp, C.c_int]
        L.or_plane_wave.argtypes = [C.c_void_p, C.c_int, C.c_int, dp, cdp]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_PATH)


def ref(quad=False):
    """The reference's routines built from its Fortran sources (oracle/_ref);
    quad=True: the build with quad-precision Hankel-filter weights."""
    global _ref, _ref_quad
    if quad:
        if _ref_quad is None:
            _ref_quad = _bind(C.CDLL(REF_QUAD_PATH))
        return _ref_quad
    if _ref is None:
        _ref = _bind(C.CDLL(REF_PATH))
    return _ref


def _bind(R):
    """Argument types of the harness entry points."""
    R.ref_vbt_weights.argtypes = [C.c_double] * 5 + [dp, dp]
    R.ref_double_integral_u.argtypes = [C.c_int, dp, dp, C.c_double, C.c_int, dp,
                                        C.c_double, dp, cdp]
    R.ref_calc_apq.argtypes = [C.c_int, dp, dp, C.c_double, C.c_double, cdp, cdp, cdp, cdp, cdp]
    if hasattr(R, "ref_vbt_weights_single"):
        R.ref_vbt_weights_single.argtypes = [C.c_double] * 5 + [dp, dp]
        R.ref_rc_integral_u.argtypes = [C.c_int, dp, dp, C.c_double, C.c_int, dp, C.c_double,
                                        C.c_double, dp, cdp, cdp]
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        R.ref_prepare_recvs.argtypes = [C.c_int, dp, dp, C.c_double, C.c_int, dp, C.c_int, dp,
                                        dp, ip, ip]
    R.ref_fgmres_dense.argtypes = [C.c_int, cdp, cdp, cdp, C.c_int, C.c_int, C.c_int, C.c_double,
                                   cdp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double),
                                   C.POINTER(C.c_longlong)]
    return R


def nthreads_default():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


class Model:
    """One frequency of a layered background + anomaly geometry."""

    def __init__(self, csigma, thick, omega, Nx, Ny, Nz, dx, dy, z):
        self.csigma = np.ascontiguousarray(csigma, dtype=np.complex128)
        self.Nl = len(self.csigma) - 1
        self.thick = np.ascontiguousarray(np.r_[thick, 0.0], dtype=np.float64)
        self.omega = float(omega)
        self.Nx, self.Ny, self.Nz = int(Nx), int(Ny), int(Nz)
        self.dx, self.dy = float(dx), float(dy)
        self.z = np.ascontiguousarray(z, dtype=np.float64)
        cs = np.ascontiguousarray(self.csigma.view(np.float64))
        self.ptr = lib().or_model_new(self.Nl, cs, self.thick, self.omega, self.Nx, self.Ny,
                                      self.Nz, self.dx, self.dy, self.z)
        if not self.ptr:
            raise ValueError("bad model")

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.or_model_free(self.ptr)
            self.ptr = None

    @property
    def block_size(self):
        return int(lib().or_block_size(self.Nz))

    def lnum(self):
        return np.array([lib().or_model_lnum(self.ptr, i) for i in range(self.Nz + 1)])

    def tensor_element(self, lx, ly):
        out = np.zeros(self.block_size, np.complex128)
        lib().or_tensor_element(self.ptr, lx, ly, out)
        return out

    def tensor_element_abssum(self, lx, ly):
        """Per entry (real, block layout): sum over unmasked taps K and the
        six weights c of max|WT_c| |dG/dWT_c(K)|, with WT the unscaled filter
        weights.  Weight errors of delta times each vector's largest weight
        (how two FP64 weight codes differ) change the element by at most
        delta times this, however much the far-offset tap sums cancel."""
        out = np.zeros(self.block_size, np.float64)
        lib().or_tensor_element_abssum(self.ptr, lx, ly, out)
        return out

    def tensor_realspace(self, nthreads=None):
        G = np.zeros(self.Nx * self.Ny * self.block_size, np.complex128)
        lib().or_tensor_realspace(self.ptr, G, nthreads or nthreads_default())
        return G

    def tensor_fft(self, Greal, nthreads=None):
        G = np.zeros_like(Greal)
        lib().or_tensor_fft(self.ptr, np.ascontiguousarray(Greal), G, nthreads or nthreads_default())
        return G

    def apply(self, Gspec, csiga, v, nthreads=None):
        out = np.zeros_like(v)
        lib().or_apply(self.ptr, Gspec, np.ascontiguousarray(csiga, np.complex128),
                       np.ascontiguousarray(v, np.complex128), out, nthreads or nthreads_default())
        return out

    def apply_timings(self):
        t = np.zeros(3)
        lib().or_apply_timings(t)
        return t

    def apply_bruteforce(self, Greal, csiga, v, nthreads=None):
        out = np.zeros_like(v)
        lib().or_apply_bruteforce(self.ptr, Greal, np.ascontiguousarray(csiga, np.complex128),
                                  np.ascontiguousarray(v, np.complex128), out,
                                  nthreads or nthreads_default())
        return out

    def plane_wave(self, pol):
        E = np.zeros(3 * self.Nx * self.Ny * self.Nz, np.complex128)
        lib().or_plane_wave_integral(self.ptr, pol, E)
        return E

    def double_integral_U(self, lm, W):
        G = np.zeros(6 * self.Nz * self.Nz, np.complex128)
        lib().or_double_integral_U(self.ptr, lm, np.ascontiguousarray(W, np.float64), G)
        return G

    def calc_apq(self, lm, with_A=True):
        Nl = self.Nl
        Arr = np.zeros(2 * Nl, np.complex128)
        p = np.zeros_like(Arr)
        q = np.zeros_like(Arr)
        eta = np.zeros(Nl, np.complex128)
        A = np.zeros(Nl * Nl * 2, np.complex128) if with_A else None
        lib().or_calc_apq(self.ptr, lm, Arr, p, q, eta,
                          A.ctypes.data_as(C.c_void_p) if with_A else None)
        return Arr, p, q, eta, A

    def solve(self, Gspec, csiga, Ebkg, E0=None, m_out=30, m_in=30, maxit=100, tol=1e-8,
              nthreads=None):
        Esol = np.zeros_like(Ebkg)
        info = GmresInfo()
        e0 = None
        if E0 is not None:
            E0 = np.ascontiguousarray(E0, np.complex128)
            e0 = E0.ctypes.data_as(C.c_void_p)
        lib().or_solve(self.ptr, Gspec, np.ascontiguousarray(csiga, np.complex128),
                       np.ascontiguousarray(Ebkg, np.complex128), e0, m_out, m_in, maxit, tol,
                       Esol, C.byref(info), nthreads or nthreads_default())
        return Esol, info


class Receivers:
    """Receivers of the continuation (RECEIVER_TYPE, data_types_module.F90:43-47)
    after PrepareRecvs on a model: lateral shifts (m) from the cell centres,
    depths, background layer and anomaly cell index."""

    def __init__(self, model, xshift, yshift, depth):
        self.Nr = len(depth)
        self.xs = np.ascontiguousarray(xshift, np.float64)
        self.ys = np.ascontiguousarray(yshift, np.float64)
        self.depth = np.ascontiguousarray(depth, np.float64)
        self.zr = np.zeros(self.Nr)
        self.layer = np.zeros(self.Nr, np.int32)
        self.cell = np.zeros(self.Nr, np.int32)
        lib().or_prepare_recvs(model.ptr, self.Nr, self.depth, self.zr, self.layer, self.cell)
        self.model = model

    def args(self):
        return (self.Nr, self.xs, self.ys, self.zr, self.layer, self.cell)

    @property
    def block_size(self):
        return int(lib().or_rc_block_size(self.model.Nz, self.Nr))

    def tensor_element(self, lx, ly):
        out = np.zeros(self.block_size, np.complex128)
        lib().or_rc_tensor_element(self.model.ptr, *self.args(), lx, ly, out)
        return out

    def tensor_realspace(self, nthreads=None):
        m = self.model
        G = np.zeros(4 * m.Nx * m.Ny * self.block_size, np.complex128)
        lib().or_rc_tensor_realspace(m.ptr, *self.args(), G, nthreads or nthreads_default())
        return G

    def tensor_fft(self, Greal, nthreads=None):
        G = np.zeros_like(Greal)
        lib().or_rc_tensor_fft(self.model.ptr, self.Nr, np.ascontiguousarray(Greal), G,
                               nthreads or nthreads_default())
        return G

    def _fields(self):
        m = self.model
        n = self.Nr * 3 * m.Nx * m.Ny
        return np.zeros(n, np.complex128), np.zeros(n, np.complex128)

    def apply(self, Gspec, csiga, E, nthreads=None):
        Ea, Ha = self._fields()
        lib().or_rc_apply(self.model.ptr, self.Nr, Gspec, np.ascontiguousarray(csiga, np.complex128),
                          np.ascontiguousarray(E, np.complex128), Ea, Ha,
                          nthreads or nthreads_default())
        return Ea, Ha

    def apply_bruteforce(self, Greal, csiga, E, nthreads=None):
        Ea, Ha = self._fields()
        lib().or_rc_apply_bruteforce(self.model.ptr, self.Nr, Greal,
                                     np.ascontiguousarray(csiga, np.complex128),
                                     np.ascontiguousarray(E, np.complex128), Ea, Ha,
                                     nthreads or nthreads_default())
        return Ea, Ha

    def plane_wave(self, pol):
        """Normal field at the prepared receiver depths, as the reference
        driver evaluates it (giem2g.F90:281-283, 416)."""
        return plane_wave(self.model, pol, self.zr)


def plane_wave(model, pol, depths):
    """PlaneWave at the given depths: complex (Nr, 6) Ex,Ey,Ez,Hx,Hy,Hz."""
    depths = np.ascontiguousarray(depths, np.float64)
    F = np.zeros(6 * len(depths), np.complex128)
    lib().or_plane_wave(model.ptr, pol, len(depths), depths, F)
    return F.reshape(6, len(depths)).T.copy()


def rc_integral_U(model, zr, layer, cell, lm, W):
    GE = np.zeros(8 * model.Nz, np.complex128)
    GH = np.zeros(7 * model.Nz, np.complex128)
    lib().or_rc_integral_U(model.ptr, zr, layer, cell, lm, np.ascontiguousarray(W, np.float64),
                           GE, GH)
    return GE.reshape(model.Nz, 8).T.copy(), GH.reshape(model.Nz, 7).T.copy()


def vbt_weights_single(x, y, hx, hy, rmin):
    WT = np.zeros(NTAPS * 6)
    lib().or_vbt_weights_single(x, y, hx, hy, rmin, WT)
    return WT.reshape(6, NTAPS).T.copy()


def vbt_weights(x, y, hx, hy, rmin):
    WT = np.zeros(NTAPS * 6)
    lib().or_vbt_weights(x, y, hx, hy, rmin, WT)
    return WT.reshape(6, NTAPS).T.copy()


def compute_weights(lx, ly, dx, dy, rmin):
    WT = np.zeros(NTAPS * 6)
    W0 = np.zeros(6)
    lib().or_compute_weights(lx, ly, dx, dy, rmin, WT, W0)
    return WT.reshape(6, NTAPS).T.copy(), W0


def vbt_lms(rmin):
    out = np.zeros(NTAPS)
    lib().or_vbt_lms(rmin, out)
    return out


def fgmres(matvec, b, x0, m_out, m_in, maxit, tol, hist_len=0):
    """Nested FGMRES of the oracle with a Python matvec callback."""
    b = np.ascontiguousarray(b, np.complex128)
    x0 = np.ascontiguousarray(x0, np.complex128)
    N = b.size
    x = np.zeros_like(b)
    info = GmresInfo()

    def cb(ctx, pin, pout):
        vin = np.ctypeslib.as_array(pin, shape=(2 * N,)).view(np.complex128)
        vout = np.ctypeslib.as_array(pout, shape=(2 * N,)).view(np.complex128)
        vout[:] = matvec(vin)

    f = MATVEC(cb)
    hist = np.zeros(max(hist_len, 1))
    lib().or_fgmres(N, m_out, m_in, maxit, tol, f, None, x0, b, x, C.byref(info),
                    hist.ctypes.data_as(C.c_void_p) if hist_len else None, hist_len)
    return x, info, hist[: min(hist_len, info.iterations)] if hist_len else None
