"""Python mirror of the reference operator interface on the HIP library.

IEOperator follows the life cycle of giem2g_interfaces.F90:76-254
(calc_data_sizes -> prepare_ie_kernel -> calc_ie_kernel -> calc_fft_of_ie_kernel
-> set_anomaly_conductivity -> apply -> print_stats -> delete) and adds the
device-pointer and solver entry points.  Arrays: numpy (host) or torch
tensors on the GPU (device).  Everything goes through the C ABI.
"""
import ctypes as C

import numpy as np

from . import lib as L

ORDER_XYZ = 0
ORDER_YXZ = 1


def _ptr(x):
    """Raw address of a numpy array or torch tensor (contiguous)."""
    if isinstance(x, np.ndarray):
        if not x.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return x.ctypes.data
    if hasattr(x, "data_ptr"):
        if not x.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return x.data_ptr()
    if x is None:
        return None
    return int(x)


_DTYPE_NAMES = {"c128": "complex128", "c64": "complex64"}


def dev_ptrs(n, dtype, *tensors, allow_none=()):
    """Addresses of torch tensors handed to a device entry point, after the
    checks that keep a kernel from reading or writing out of bounds: each is
    a CUDA (HIP) tensor, contiguous, of the given dtype ("c128" or "c64")
    and with exactly n elements.  Positions listed in allow_none may be None
    (a nullable argument).  Raises ValueError (not assert, which python -O
    removes)."""
    import torch
    want = {"c128": torch.complex128, "c64": torch.complex64}[dtype]
    out = []
    for k, t in enumerate(tensors):
        if t is None:
            if k in allow_none:
                out.append(None)
                continue
            raise ValueError("argument %d must be a tensor" % k)
        if not isinstance(t, torch.Tensor):
            raise ValueError("argument %d must be a torch tensor on the GPU, got %s" % (k, type(t).__name__))
        if t.dtype != want:
            raise ValueError("argument %d must be %s, got %s" % (k, _DTYPE_NAMES[dtype], t.dtype))
        if t.numel() != n:
            raise ValueError("argument %d must have %d elements, got %d" % (k, n, t.numel()))
        if not t.is_contiguous():
            raise ValueError("argument %d must be contiguous" % k)
        if not t.is_cuda:
            raise ValueError("argument %d must be on the GPU (device %s)" % (k, t.device))
        out.append(t.data_ptr())
    return out


def _prec_of(t):
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError("expected a torch tensor on the GPU, got %s" % type(t).__name__)
    if t.dtype == torch.complex128:
        return 0, "c128"
    if t.dtype == torch.complex64:
        return 1, "c64"
    raise ValueError("vectors must be complex128 or complex64, got %s" % t.dtype)


class IEOperator:
    def __init__(self, Nx, Ny, Nz, dx, dy, z, order=ORDER_XYZ):
        self.lib = L.load()
        self.Nx, self.Ny, self.Nz = int(Nx), int(Ny), int(Nz)
        self.dx, self.dy = float(dx), float(dy)
        self.z = np.ascontiguousarray(z, np.float64)
        h = C.c_void_p()
        L.check(self.lib.giem2g_op_create(self.Nx, self.Ny, self.Nz, self.dx, self.dy,
                                          self.z.ctypes.data, C.byref(h)))
        self.h = h
        if order != ORDER_XYZ:
            L.check(self.lib.giem2g_set_component_order(self.h, order))

    @classmethod
    def from_model(cls, m, order=ORDER_XYZ):
        return cls(m.Nx, m.Ny, m.Nz, m.dx, m.dy, m.z, order)

    def close(self):
        if getattr(self, "h", None):
            L.check(self.lib.giem2g_op_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def block_size(self):
        return 4 * self.Nz * (self.Nz + 1) // 2 + 2 * self.Nz * self.Nz

    @property
    def ncell(self):
        return self.Nx * self.Ny * self.Nz

    @property
    def nvec(self):
        return 3 * self.Nx * self.Ny * self.Nz

    # --- frequency / tensor -------------------------------------------------
    def set_background(self, csigb, thick, omega):
        cs = np.ascontiguousarray(np.asarray(csigb, np.complex128))
        th = np.ascontiguousarray(np.r_[np.asarray(thick, np.float64), 0.0])
        L.check(self.lib.giem2g_op_set_background(self.h, len(cs) - 1, cs.ctypes.data,
                                                  th.ctypes.data, float(omega)))

    def build_tensor(self):
        L.check(self.lib.giem2g_op_build_tensor(self.h))

    def tensor_stage(self, kind):
        L.check(self.lib.giem2g_op_tensor_stage(self.h, kind))

    def get_tensor(self):
        G = np.empty(self.Nx * self.Ny * self.block_size, np.complex128)
        L.check(self.lib.giem2g_op_get_tensor(self.h, G.ctypes.data))
        return G

    def put_tensor(self, G, spectral=True):
        G = np.ascontiguousarray(G, np.complex128)
        if G.size != self.Nx * self.Ny * self.block_size:
            raise ValueError("tensor must have Nx*Ny*block_size = %d elements"
                             % (self.Nx * self.Ny * self.block_size))
        L.check(self.lib.giem2g_op_put_tensor(self.h, G.ctypes.data, 1 if spectral else 0))

    # --- conductivity --------------------------------------------------------
    def set_sigma(self, csiga):
        if isinstance(csiga, np.ndarray):
            c = np.ascontiguousarray(csiga, np.complex128)
            if c.size != self.ncell:
                raise ValueError("conductivity must have Nx*Ny*Nz = %d elements" % self.ncell)
            L.check(self.lib.giem2g_op_set_sigma(self.h, c.ctypes.data))
        else:
            (p,) = dev_ptrs(self.ncell, "c128", csiga)
            L.check(self.lib.giem2g_op_set_sigma_dev(self.h, p))

    def set_geometry(self, dx, dy, z):
        """New cell sizes and layer borders (same Nx, Ny, Nz); the tensor must
        be built again."""
        z = np.ascontiguousarray(z, np.float64)
        if z.size != self.Nz + 1:
            raise ValueError("z must hold Nz + 1 borders")
        L.check(self.lib.giem2g_op_set_geometry(self.h, float(dx), float(dy), z.ctypes.data))
        self.dx, self.dy, self.z = float(dx), float(dy), z

    # --- apply ---------------------------------------------------------------
    def apply_host(self, v):
        v = np.ascontiguousarray(v, np.complex128)
        if v.size != self.nvec:
            raise ValueError("vector must have 3*Nx*Ny*Nz = %d elements" % self.nvec)
        out = np.empty_like(v)
        L.check(self.lib.giem2g_op_apply_host(self.h, v.ctypes.data, out.ctypes.data))
        return out

    def apply_dev(self, vin, vout):
        """FP64 matvec on HBM (asynchronous on the operator stream)."""
        a, b = dev_ptrs(self.nvec, "c128", vin, vout)
        L.check(self.lib.giem2g_op_apply_dev(self.h, a, b))

    def apply_adjoint_dev(self, vin, vout):
        """vout = A^H vin (complex128 torch tensors on the GPU)."""
        a, b = dev_ptrs(self.nvec, "c128", vin, vout)
        L.check(self.lib.giem2g_op_apply_adjoint_dev(self.h, a, b))

    def apply_dev_f32(self, vin, vout):
        a, b = dev_ptrs(self.nvec, "c64", vin, vout)
        L.check(self.lib.giem2g_op_apply_dev_f32(self.h, a, b))

    def apply_batch_dev(self, vin, vout):
        """vout[r] = A vin[r] for the rows of two (nrhs, 3 Nx Ny Nz) torch
        tensors on the GPU (complex128 or complex64); one tensor read serves up
        to 8 complex64 or 4 complex128 rows (Nz 8 to 32) or 2 (Nz 33 to 64)."""
        prec, dt = _prec_of(vin)
        if vin.dim() != 2 or vin.shape[1] != self.nvec or vin.shape[0] < 1:
            raise ValueError("vin must be (nrhs, %d)" % self.nvec)
        nrhs = int(vin.shape[0])
        a, b = dev_ptrs(nrhs * self.nvec, dt, vin, vout)
        L.check(self.lib.giem2g_op_apply_batch_dev(self.h, prec, nrhs, a, b))

    def apply_batch_dev_noalloc(self, vin, vout):
        """As apply_batch_dev, with the FFT intermediate and the spectrum slots
        in the library's fixed 256 MiB device array instead of allocations
        (no hipMalloc in the call).  Grids whose one vector does not fit are
        rejected; complex64 needs the complex64 operator made beforehand."""
        prec, dt = _prec_of(vin)
        if vin.dim() != 2 or vin.shape[1] != self.nvec or vin.shape[0] < 1:
            raise ValueError("vin must be (nrhs, %d)" % self.nvec)
        nrhs = int(vin.shape[0])
        a, b = dev_ptrs(nrhs * self.nvec, dt, vin, vout)
        L.check(self.lib.giem2g_op_apply_batch_dev_noalloc(self.h, prec, nrhs, a, b))

    def apply_pair_dev(self, in0, in1, out0, out1):
        """out0 = A in0, out1 = A in1 with one tensor read (torch tensors on
        the GPU, all complex128 or all complex64; Nz from 8 to 64)."""
        prec, dt = _prec_of(in0)
        a, b, c, d = dev_ptrs(self.nvec, dt, in0, in1, out0, out1)
        L.check(self.lib.giem2g_op_apply_pair_dev(self.h, prec, a, b, c, d))

    def sync(self):
        L.check(self.lib.giem2g_op_sync(self.h))

    def time_apply(self, vin, vout, reps=10, precision=0):
        ms = np.zeros(3)
        a, b = dev_ptrs(self.nvec, "c64" if precision == 1 else "c128", vin, vout)
        L.check(self.lib.giem2g_op_time_apply(self.h, precision, a, b, reps, ms.ctypes.data))
        return ms

    def stats(self):
        o = np.zeros(11)
        L.check(self.lib.giem2g_op_stats(self.h, o.ctypes.data))
        keys = ["matvecs", "matvecs32", "dotprods", "t_apply", "t_fwd", "t_contr", "t_bwd", "t_dot",
                "t_tensor", "t_tensor_fft", "device_bytes"]
        return dict(zip(keys, o.tolist()))

    COUNTER_KEYS = ["matvecs", "matvecs32", "dotprods", "t_apply", "t_fwd", "t_contr", "t_bwd", "t_dot",
                    "t_tensor", "t_tensor_fft", "device_bytes", "gs_passes", "krylov_bytes",
                    "contr_launches_fp64", "contr_rhs_fp64", "t_contr_fp64",
                    "contr_launches_c64", "contr_rhs_c64", "t_contr_c64", "profiling"]

    def counters(self):
        """All operator counters (giem2g_op_counters); times in seconds."""
        o = np.zeros(len(self.COUNTER_KEYS))
        L.check(self.lib.giem2g_op_counters(self.h, o.ctypes.data, len(o)))
        return dict(zip(self.COUNTER_KEYS, o.tolist()))

    def reset_counters(self):
        L.check(self.lib.giem2g_op_reset_counters(self.h))

    def set_profiling(self, on):
        """Record HIP events around every apply stage and Gram-Schmidt pass
        (their times appear in counters())."""
        L.check(self.lib.giem2g_op_set_profiling(self.h, 1 if on else 0))

    # --- sources / solve -----------------------------------------------------
    def plane_wave(self, pol):
        E = np.empty(self.nvec, np.complex128)
        L.check(self.lib.giem2g_plane_wave(self.h, pol, E.ctypes.data))
        return E

    def solve(self, Ebkg, E0=None, outer_m=30, inner_m=30, max_restart=100, tol=1e-8,
              inner_precision=0, precond=0, verbose=0, learned_patience=0, inner_tol=0.0):
        """Nested FGMRES solve of the scaled system for the background field
        Ebkg (numpy: on the host; torch: on the GPU).  learned_patience:
        rejected learned steps in a row after which the network is not
        called again in this solve (0: 2, negative: never stop).  inner_tol:
        relative residual at which an inner solve stops; negative: relaxed,
        |inner_tol| tol / (the outer's current relative residual); 0: tol
        for FP64 inner solves (as in the reference), -1 for complex64 ones."""
        p = L.SolverParams(outer_m, inner_m, max_restart, tol, inner_precision, precond, verbose,
                           learned_patience, inner_tol)
        info = L.SolveInfo()
        if isinstance(Ebkg, np.ndarray):
            Ebkg = np.ascontiguousarray(Ebkg, np.complex128)
            out = np.empty_like(Ebkg)
            e0 = None if E0 is None else np.ascontiguousarray(E0, np.complex128)
            if Ebkg.size != self.nvec or (e0 is not None and e0.size != self.nvec):
                raise ValueError("fields must have 3*Nx*Ny*Nz = %d elements" % self.nvec)
            L.check(self.lib.giem2g_solve(self.h, Ebkg.ctypes.data,
                                          None if e0 is None else e0.ctypes.data, C.byref(p),
                                          out.ctypes.data, C.byref(info)))
        else:
            import torch
            out = torch.empty_like(Ebkg)
            a, b, c = dev_ptrs(self.nvec, "c128", Ebkg, E0, out, allow_none=(1,))
            L.check(self.lib.giem2g_solve_dev(self.h, a, b, C.byref(p), c, C.byref(info)))
        return out, info.as_dict()

    def scale(self, mode, vin, vout):
        """mode 0: vout = sqrt(Re sb) vin (rhs), 1: vin / a (field -> x),
        2: a vin (x -> field); torch complex128 tensors on the GPU."""
        a, b = dev_ptrs(self.nvec, "c128", vin, vout)
        L.check(self.lib.giem2g_op_scale(self.h, mode, a, b))

    def set_contraction(self, mode):
        """0 automatic (4 for FP64 at Nz from 8 to 32 and complex64 at Nz
        from 24 to 32, else 2 for Nz from 8 to 64, otherwise 1; complex64 at
        Nz = 17-23, 33-35, 49-63 uses 1), 1 vector ALU
        (two components in flight), 2 matrix cores (MFMA), 3 vector ALU (one
        component in flight), 4 small-block matrix cores (Nz from 8 to 32:
        every operand column a point), 5 the same with the x-y mirror (one
        tensor read for the quads (kx, ky) and (ky, kx); Nx = Ny, dx = dy
        and a tensor built by the library; automatic where it applies), 6
        the 16x16x4 tiles with the x-y mirror (Nz from 8 to 64)."""
        L.check(self.lib.giem2g_op_set_contraction(self.h, mode))

    def set_learned(self, net, accept=0.9):
        """Attach a learned.UNet as the right preconditioner of precond modes
        1 and 2 (None detaches)."""
        self._net = net  # keep the network alive while attached
        L.check(self.lib.giem2g_op_set_learned(self.h, None if net is None else net.h, float(accept)))

    def learned_features(self, r, e, x, t=None):
        """Network input x [vox][8] of residual r and target t [vox][6] of
        correction e (torch tensors on the GPU)."""
        import torch
        a, b = dev_ptrs(self.nvec, "c128", r, e, allow_none=(1,))
        for k, (arr, ch) in enumerate(((x, 8), (t, 6))):
            if arr is None and k == 1:
                continue
            if (not isinstance(arr, torch.Tensor) or not arr.is_cuda or arr.dtype != torch.float32
                    or not arr.is_contiguous() or arr.numel() != ch * self.ncell):
                raise ValueError("feature arrays must be contiguous float32 GPU tensors of %d x %d"
                                 % (self.ncell, ch))
        L.check(self.lib.giem2g_op_learned_features(self.h, a, b, x.data_ptr(),
                                                    None if t is None else t.data_ptr()))

    def learned_apply(self, v, z):
        a, b = dev_ptrs(self.nvec, "c128", v, z)
        L.check(self.lib.giem2g_op_learned_apply(self.h, a, b))

    def solve_pair(self, Eb0, Eb1, E00=None, E01=None, outer_m=30, inner_m=30, max_restart=100,
                   tol=1e-8, inner_precision=0, inner_tol=0.0):
        """Two solves (torch complex128 tensors on the GPU) with paired
        operator products; returns (E0, E1, info0, info1, [paired, alone])."""
        import torch
        p = L.SolverParams(outer_m, inner_m, max_restart, tol, inner_precision, 0, 0, 0, inner_tol)
        infos = (L.SolveInfo * 2)()
        cnt = (C.c_longlong * 2)()
        o0, o1 = torch.empty_like(Eb0), torch.empty_like(Eb1)
        a, b, c, d, e, f = dev_ptrs(self.nvec, "c128", Eb0, Eb1, E00, E01, o0, o1, allow_none=(2, 3))
        torch.cuda.synchronize()
        L.check(self.lib.giem2g_solve_pair_dev(self.h, a, b, c, d, C.byref(p), e, f, infos, cnt))
        return o0, o1, infos[0].as_dict(), infos[1].as_dict(), [cnt[0], cnt[1]]

    def history(self, maxlen=100000):
        buf = np.zeros(maxlen)
        n = C.c_int()
        L.check(self.lib.giem2g_last_history(self.h, buf.ctypes.data, maxlen, C.byref(n)))
        return buf[: min(n.value, maxlen)]


def operator_for(model, order=ORDER_XYZ, build=True):
    """IEOperator with background, tensor and conductivity of a models.Model."""
    op = IEOperator.from_model(model, order)
    op.set_background(model.csigb(), model.thick, model.omega)
    if build:
        op.build_tensor()
    op.set_sigma(model.csiga())
    return op


class ReceiverOperator:
    """Continuation of the solution to receivers (the post-solve field
    recalculation; Continuation_Function_Module.F90).  Follows the
    reference's life cycle PrepareContinuationOperator ->
    CalcRecalculationGreenTensor -> CalcFFTofRCTensor -> FinalizeRCOperator
    -> ReCalculation per frequency, on the background, grid and stream of an
    IEOperator (which must outlive it).  Receivers sit at every cell centre
    shifted by (xshift, yshift) metres, at the given depths; the fields are
    complex (Nr, 3, Nx, Ny) in Fortran order, returned as flat arrays."""

    def __init__(self, ie_op, xshift, yshift, depth):
        self.lib = L.load()
        self.ie = ie_op
        self.xs = np.ascontiguousarray(xshift, np.float64)
        self.ys = np.ascontiguousarray(yshift, np.float64)
        self.depth = np.ascontiguousarray(depth, np.float64)
        self.Nr = len(self.depth)
        h = C.c_void_p()
        L.check(self.lib.giem2g_rc_create(ie_op.h, self.Nr, self.xs.ctypes.data, self.ys.ctypes.data,
                                          self.depth.ctypes.data, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            L.check(self.lib.giem2g_rc_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def block_size(self):
        return 15 * self.Nr * self.ie.Nz

    @property
    def nfield(self):
        return self.Nr * 3 * self.ie.Nx * self.ie.Ny

    def calc_kernel(self):
        L.check(self.lib.giem2g_rc_calc_kernel(self.h))

    def calc_fft_of_kernel(self):
        L.check(self.lib.giem2g_rc_calc_fft_of_kernel(self.h))

    def build(self):
        """Tensor and tensor spectrum for the IE operator's current frequency."""
        self.calc_kernel()
        self.calc_fft_of_kernel()

    def get_tensor(self):
        G = np.zeros(4 * self.ie.Nx * self.ie.Ny * self.block_size, np.complex128)
        L.check(self.lib.giem2g_rc_get_tensor(self.h, G.ctypes.data))
        return G

    def put_tensor(self, G, spectral=True):
        G = np.ascontiguousarray(G, np.complex128)
        L.check(self.lib.giem2g_rc_put_tensor(self.h, G.ctypes.data, int(bool(spectral))))

    def receivers(self):
        zr = np.zeros(self.Nr)
        layer = np.zeros(self.Nr, np.int32)
        cell = np.zeros(self.Nr, np.int32)
        L.check(self.lib.giem2g_rc_receivers(self.h, zr.ctypes.data, layer.ctypes.data, cell.ctypes.data))
        return zr, layer, cell

    def set_sigma(self, csiga):
        if hasattr(csiga, "data_ptr"):
            (p,) = dev_ptrs(self.ie.ncell, "c128", csiga)
            L.check(self.lib.giem2g_rc_set_anomaly_conductivity_dev(self.h, p))
        else:
            cs = np.ascontiguousarray(csiga, np.complex128)
            L.check(self.lib.giem2g_rc_set_anomaly_conductivity(self.h, cs.ctypes.data))

    def recalculate(self, E):
        """(Ea, Ha) on the host from the total field E (host array)."""
        E = np.ascontiguousarray(E, np.complex128)
        Ea = np.zeros(self.nfield, np.complex128)
        Ha = np.zeros(self.nfield, np.complex128)
        L.check(self.lib.giem2g_rc_recalculate(self.h, E.ctypes.data, Ea.ctypes.data, Ha.ctypes.data))
        return Ea, Ha

    def recalculate_dev(self, E, Ea, Ha):
        """HBM tensors, ordered on the IE operator's stream."""
        (e,) = dev_ptrs(self.ie.nvec, "c128", E)
        a, b = dev_ptrs(self.nfield, "c128", Ea, Ha)
        L.check(self.lib.giem2g_rc_recalculate_dev(self.h, e, a, b))

    def plane_wave(self, pol):
        """Normal field at the prepared receiver depths: complex (Nr, 6)."""
        F = np.zeros(6 * self.Nr, np.complex128)
        L.check(self.lib.giem2g_rc_plane_wave(self.h, pol, F.ctypes.data))
        return F.reshape(6, self.Nr).T.copy()

    def stats(self):
        o = np.zeros(8)
        L.check(self.lib.giem2g_rc_stats(self.h, o.ctypes.data))
        keys = ["t_tensor", "t_tensor_fft", "ms_last", "ms_fwd", "ms_contract", "ms_inv", "recalcs",
                "bytes"]
        return dict(zip(keys, o))
