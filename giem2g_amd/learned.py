"""Learned right preconditioner: the U-Net handle, training pairs from the
solver's own FP64 runs, and training on the GPU with the library's kernels.

There is no counterpart in the reference, whose right preconditioner is the
inner GMRES (GFGMRES/seq_fgmres.F90:142-152).  The network maps the
residual r of the scaled system A x = b (ie_solver_module.F90:64-66) to the
correction e = A^-1 r, which the solver uses as z = M(v) (precond mode 2)
or as the start of the inner GMRES (mode 1).  Input channels: r / rms(r)
re/im of x, y, z, log10(Re sa / Re sb) and log10(|k_b| dx) (the frequency
enters through the background wavenumber k_b); output: the field correction
a * e in units of rms(r), so that z = rms * out / a.  Training targets are
normalized per sample to unit RMS: the solver uses only the direction of
M(v) (the outer FGMRES is invariant to its scale, and mode 1 rescales it by
the least-squares factor).

Training pairs (r_j, e_j) come from the outer FGMRES iterates x_j of the
reference algorithm: x_0 = b, and x_j is the iterate after j outer Arnoldi
steps (an FGMRES(j) cycle with the inner GMRES preconditioner); e_j = x* - x_j
with x* the FP64 solution at a tight tolerance.
"""
import ctypes as C
import os
import time

import numpy as np

from . import lib as L
from .operator import _ptr

# (cin, cout, taps) of the 11 layers, as in csrc/g2g_unet.hip
ARCH = [(8, 16, 27), (16, 16, 27), (16, 32, 27), (32, 32, 27), (32, 64, 27), (64, 64, 27),
        (96, 32, 27), (32, 32, 27), (48, 16, 27), (16, 16, 27), (16, 6, 1)]
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "data", "unet_c4.f32")


def n_params():
    return sum(co * t * ci + co for ci, co, t in ARCH)


def init_params(seed=0):
    """He-normal weights (std sqrt(2 / fan_in)), zero biases; the output layer
    is scaled by 0.1 so that the untrained network starts near zero."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for li, (ci, co, t) in enumerate(ARCH):
        std = np.sqrt(2.0 / (ci * t)) * (0.1 if li == len(ARCH) - 1 else 1.0)
        out.append(rng.standard_normal(co * t * ci) * std)
        out.append(np.zeros(co))
    return np.concatenate(out).astype(np.float32)


class UNet:
    """Handle of the HIP U-Net (giem2g_unet_* entry points)."""

    def __init__(self, params=None, seed=0):
        self.lib = L.load()
        h = C.c_void_p()
        L.check(self.lib.giem2g_unet_create(C.byref(h)))
        self.h = h
        n = C.c_longlong()
        L.check(self.lib.giem2g_unet_nparams(C.byref(n)))
        self.nparams = n.value
        assert self.nparams == n_params()
        self.set_params(init_params(seed) if params is None else params)

    def close(self):
        if getattr(self, "h", None):
            L.check(self.lib.giem2g_unet_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        L.check(self.lib.giem2g_unet_set_params(self.h, p.ctypes.data))

    def params(self):
        p = np.empty(self.nparams, np.float32)
        L.check(self.lib.giem2g_unet_get_params(self.h, p.ctypes.data))
        return p

    def grads(self):
        g = np.empty(self.nparams, np.float32)
        L.check(self.lib.giem2g_unet_get_grads(self.h, g.ctypes.data))
        return g

    def forward(self, x, keep=False):
        """x: float32 CUDA tensor (B, Nz, Ny, Nx, 8) -> (B, Nz, Ny, Nx, 6)."""
        import torch
        B, Z, Y, X, c = x.shape
        assert c == 8 and x.dtype == torch.float32 and x.is_contiguous()
        y = torch.empty((B, Z, Y, X, 6), dtype=torch.float32, device=x.device)
        torch.cuda.synchronize()
        L.check(self.lib.giem2g_unet_forward_dev(self.h, _ptr(x), _ptr(y), B, X, Y, Z, int(keep)))
        return y

    def backward(self, dy):
        """Parameter gradient of sum(dy * y) for the last forward(keep=True)."""
        import torch
        torch.cuda.synchronize()
        L.check(self.lib.giem2g_unet_backward_dev(self.h, _ptr(dy.contiguous())))
        return self.grads()

    def adam(self, lr):
        """One Adam step on the gradient of the last backward."""
        L.check(self.lib.giem2g_unet_adam(self.h, C.c_float(lr)))

    def train_step(self, x, t, lr):
        B, Z, Y, X, _ = x.shape
        loss = C.c_double()
        L.check(self.lib.giem2g_unet_train_step(self.h, _ptr(x), _ptr(t), B, X, Y, Z, C.c_float(lr),
                                                C.byref(loss)))
        return loss.value

    def save(self, path):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        self.params().tofile(path)

    @classmethod
    def load(cls, path=DEFAULT_WEIGHTS):
        p = np.fromfile(path, dtype=np.float32)
        return cls(params=p)


def training_pairs(op, m, pol, max_steps=5, inner_m=30, tol=1e-9, stop=1e-8):
    """(x, t) feature pairs of one model: residuals and corrections of the
    outer FGMRES iterates x_j, j = 0 .. max_steps-1, on the GPU.  op must be
    the model's operator with a learned network attached (for the feature
    context).  Returns float32 tensors (n, Nz, Ny, Nx, 8) and (n, ..., 6)."""
    import torch
    dev = "cuda"
    Eb = torch.from_numpy(op.plane_wave(pol)).to(dev)
    N = Eb.numel()
    b = torch.empty_like(Eb)
    op.scale(0, Eb, b)
    Es, info = op.solve(Eb, outer_m=30, inner_m=inner_m, max_restart=100, tol=tol)
    xs = torch.empty_like(Eb)
    op.scale(1, Es, xs)
    xj = b.clone()
    r = torch.empty_like(Eb)
    feats, targs = [], []
    shape = (m.Nz, m.Ny, m.Nx)
    bn = float(torch.linalg.vector_norm(b))
    for j in range(max_steps):
        if j > 0:
            Ej, _ = op.solve(Eb, outer_m=j, inner_m=inner_m, max_restart=1, tol=1e-14)
            op.scale(1, Ej, xj)
        op.apply_dev(xj, r)
        torch.cuda.synchronize()
        r = b - r
        if float(torch.linalg.vector_norm(r)) < stop * bn:
            break
        e = (xs - xj).contiguous()
        x = torch.empty(shape + (8,), dtype=torch.float32, device=dev)
        t = torch.empty(shape + (6,), dtype=torch.float32, device=dev)
        op.learned_features(r.contiguous(), e, x, t)
        t /= torch.sqrt(torch.mean(t * t))
        feats.append(x)
        targs.append(t)
    del N
    return feats, targs, info


def make_dataset(n_models, seed0=10000, max_steps=5, log=print):
    """Training set from n_models C4 random-box models (64x64x16, f
    log-uniform in [1e-2, 1e2] Hz), both polarizations alternating by seed."""
    import torch
    from . import models, operator as gop
    X, T = [], []
    t0 = time.perf_counter()
    dummy = UNet(seed=0)
    for k in range(n_models):
        m = models.c4_train(seed0 + k)
        op = gop.operator_for(m)
        op.set_learned(dummy)
        f, t, info = training_pairs(op, m, pol=k % 2, max_steps=max_steps)
        X += f
        T += t
        op.set_learned(None)
        op.close()
        if log and (k % 100 == 0 or k == n_models - 1):
            log("dataset: model %d/%d, %d pairs, %.1f s" % (k + 1, n_models, len(X),
                                                            time.perf_counter() - t0))
    dummy.close()
    return torch.stack(X), torch.stack(T)


def random_phase(x, t, gen):
    """Multiply residual and correction by a common random phase e^{i theta}
    (A is complex-linear, so the pair stays consistent)."""
    import torch
    B = x.shape[0]
    th = torch.rand(B, generator=gen, device=x.device) * (2 * np.pi)
    c, s = torch.cos(th).view(B, 1, 1, 1, 1), torch.sin(th).view(B, 1, 1, 1, 1)
    x = x.clone()
    t = t.clone()
    for k in range(3):
        xr, xi = x[..., 2 * k:2 * k + 1].clone(), x[..., 2 * k + 1:2 * k + 2].clone()
        x[..., 2 * k:2 * k + 1] = c * xr - s * xi
        x[..., 2 * k + 1:2 * k + 2] = s * xr + c * xi
        tr, ti = t[..., 2 * k:2 * k + 1].clone(), t[..., 2 * k + 1:2 * k + 2].clone()
        t[..., 2 * k:2 * k + 1] = c * tr - s * ti
        t[..., 2 * k + 1:2 * k + 2] = s * tr + c * ti
    return x, t


def train(net, X, T, steps=2000, batch=8, lr=1e-3, seed=0, log=print, log_every=200):
    """Adam on the MSE of net(X) vs T with the library's forward/backward
    kernels; lr decays by cosine to lr/20.  Returns the loss history."""
    import torch
    gen = torch.Generator(device=X.device)
    gen.manual_seed(seed)
    hist = []
    n = X.shape[0]
    t0 = time.perf_counter()
    for it in range(steps):
        idx = torch.randint(0, n, (batch,), generator=gen, device=X.device)
        xb, tb = random_phase(X[idx], T[idx], gen)
        xb, tb = xb.contiguous(), tb.contiguous()
        torch.cuda.synchronize()
        lr_t = lr * (0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * it / steps)))
        loss = net.train_step(xb, tb, lr_t)
        hist.append(loss)
        if log and (it % log_every == 0 or it == steps - 1):
            log("train: step %d loss %.4e (%.1f s)" % (it, np.mean(hist[-log_every:]),
                                                      time.perf_counter() - t0))
    return np.array(hist)


# --------------------------------------------------------------------------
# Training on the residual: loss ||r - A M(r)||^2 / ||r||^2 with the
# gradient through the operator's adjoint (giem2g_op_apply_adjoint_dev),
# so that the network is judged by what the solver sees.

class ResidualSample:
    __slots__ = ("op", "r", "x", "rms", "rn2")

    def __init__(self, op, r, x):
        import torch
        self.op, self.r, self.x = op, r, x
        self.rn2 = float(torch.vdot(r, r).real)
        self.rms = (self.rn2 / r.numel()) ** 0.5


def residual_dataset(n_models, seed0=10000, max_steps=5, log=print):
    """Residuals of the outer FGMRES iterates of n_models C4 models, with the
    models' operators kept on the GPU (about 0.1 GB each).  Returns (ops,
    samples); ops[k] = (operator, model, a) with a = the per-cell scaling
    a = 2 sqrt(Re sb) / (sa + conj sb) on every vector slot."""
    import torch
    from . import models, operator as gop
    ops, samples = [], []
    t0 = time.perf_counter()
    feat = UNet(seed=0)
    for k in range(n_models):
        m = models.c4_train(seed0 + k)
        op = gop.operator_for(m)
        op.set_learned(feat)
        n = 3 * m.ncell
        ones = torch.ones(n, dtype=torch.complex128, device="cuda")
        a = torch.empty_like(ones)
        op.scale(2, ones, a)
        Eb = torch.from_numpy(op.plane_wave(k % 2)).cuda()
        b = torch.empty_like(Eb)
        op.scale(0, Eb, b)
        xj = b.clone()
        r = torch.empty_like(Eb)
        bn = float(torch.linalg.vector_norm(b))
        for j in range(max_steps):
            if j > 0:
                Ej, _ = op.solve(Eb, outer_m=j, inner_m=30, max_restart=1, tol=1e-14)
                op.scale(1, Ej, xj)
            op.apply_dev(xj, r)
            torch.cuda.synchronize()
            rj = (b - r).contiguous()
            if float(torch.linalg.vector_norm(rj)) < 1e-8 * bn:
                break
            x = torch.empty((m.Nz, m.Ny, m.Nx, 8), dtype=torch.float32, device="cuda")
            op.learned_features(rj, None, x, None)
            samples.append(ResidualSample(len(ops), rj, x))
        ops.append((op, m, a))
        if log and (k % 50 == 0 or k == n_models - 1):
            log("residual dataset: model %d/%d, %d samples, %.1f s" % (k + 1, n_models, len(samples),
                                                                     time.perf_counter() - t0))
    return ops, samples, feat


def _to_vec(y, m):
    """(Nz, Ny, Nx, 6) real channels -> complex (3*Nz*Ny*Nx) operator vector."""
    import torch
    w = torch.complex(y[..., 0::2].double(), y[..., 1::2].double())  # (Nz, Ny, Nx, 3)
    return w.permute(3, 0, 1, 2).reshape(-1).contiguous()


def _to_chan(g, m):
    """complex operator vector -> (Nz, Ny, Nx, 6) real channels."""
    import torch
    w = g.reshape(3, m.Nz, m.Ny, m.Nx).permute(1, 2, 3, 0)
    out = torch.empty(w.shape[:3] + (6,), dtype=torch.float32, device=g.device)
    out[..., 0::2] = w.real.float()
    out[..., 1::2] = w.imag.float()
    return out


def residual_loss_step(net, ops, batch, lr):
    """One Adam step on mean_b ||r_b - A_b z_b||^2 / ||r_b||^2 with
    z_b = rms_b * net(x_b) / a_b.  Returns the mean loss."""
    import torch
    X = torch.stack([s.x for s in batch]).contiguous()
    Y = net.forward(X, keep=True)
    dY = torch.empty_like(Y)
    loss = 0.0
    B = len(batch)
    for i, s in enumerate(batch):
        op, m, a = ops[s.op]
        z = (s.rms * _to_vec(Y[i], m) / a).contiguous()
        az = torch.empty_like(z)
        op.apply_dev(z, az)
        torch.cuda.synchronize()
        res = (s.r - az).contiguous()
        loss += float(torch.vdot(res, res).real) / s.rn2 / B
        g = torch.empty_like(res)
        op.apply_adjoint_dev(res, g)
        torch.cuda.synchronize()
        # dL/d conj(z) = -A^H res / ||r||^2; dL/d conj(w) = conj(rms / a) dL/d conj(z)
        gw = torch.conj(s.rms / a) * (-g / s.rn2)
        dY[i] = 2.0 * _to_chan(gw, m) / B
    net.backward(dY.contiguous())
    net.adam(lr)
    return loss


def train_residual(net, ops, samples, steps=2000, batch=8, lr=3e-4, seed=0, log=print, log_every=100):
    import torch
    g = torch.Generator()
    g.manual_seed(seed)
    hist = []
    t0 = time.perf_counter()
    for it in range(steps):
        idx = torch.randint(0, len(samples), (batch,), generator=g).tolist()
        lr_t = lr * (0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * it / steps)))
        hist.append(residual_loss_step(net, ops, [samples[i] for i in idx], lr_t))
        if log and (it % log_every == 0 or it == steps - 1):
            log("train (residual loss): step %d loss %.4f (%.1f s)" % (it, np.mean(hist[-log_every:]),
                                                                    time.perf_counter() - t0))
    return np.array(hist)
