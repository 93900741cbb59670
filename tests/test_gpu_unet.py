"""GPU parity of the learned preconditioner's HIP kernels (U-Net forward,
backward and training on MFMA, feature packing) against the FP64 oracle
(oracle/unet_ref.py), and the solver's learned modes.

Bars: forward output within 2e-2 relative (2-norm) of the oracle with the
same bfloat16 roundings; per-layer parameter gradients within 5e-2
relative; features within 1e-6; learned solves reach the FP64 1e-8 true
residual and agree with the plain FP64 solve within 1e-6.
"""
import numpy as np
import pytest

from tests.conftest import rel

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU test without a GPU"
    return torch


def _params(seed):
    from oracle import unet_ref as U
    rng = np.random.default_rng(seed)
    p = U.init_params(seed)
    # nonzero biases and a larger output layer so every path carries signal
    out = []
    for (W, b), (name, *_rest) in zip(U.split(p.astype(np.float64)), U.LAYERS):
        s = 10.0 if name == "out" else 1.0
        out.append((W * s).ravel())
        out.append(0.05 * rng.standard_normal(b.size))
    return np.concatenate(out).astype(np.float32)


@pytest.mark.parametrize("B,Z,Y,X", [(1, 4, 4, 4), (2, 8, 8, 12), (1, 8, 16, 8)])
def test_unet_forward_backward_match_oracle(B, Z, Y, X):
    torch = _torch()
    from giem2g_amd import learned
    from oracle import unet_ref as U
    print("case", B, Z, Y, X, flush=True)
    rng = np.random.default_rng(B * 100 + Z + Y + X)
    p = _params(7)
    net = learned.UNet(params=p)
    x = rng.standard_normal((B, Z, Y, X, 8)).astype(np.float32)
    x[..., 7] = 0
    y = net.forward(torch.from_numpy(x).cuda(), keep=True)
    torch.cuda.synchronize()
    y = y.cpu().numpy().astype(np.float64)
    dy = rng.standard_normal(y.shape).astype(np.float32)
    g = net.backward(torch.from_numpy(dy).cuda()).astype(np.float64)
    gref = np.zeros(p.size)
    for b in range(B):
        o, cache = U.forward(p.astype(np.float64), x[b].astype(np.float64), round_bf16=True)
        e = rel(y[b], o)
        assert e < 2e-2, ("forward", b, e)
        gref += U.backward(p, cache, dy[b].astype(np.float64), round_bf16=True)
    off = 0
    for li, (W, bb) in enumerate(U.split(gref)):
        n = W.size
        ew = rel(g[off:off + n], W.ravel())
        eb = rel(g[off + n:off + n + bb.size], bb)
        off += n + bb.size
        assert ew < 5e-2 and eb < 5e-2, ("grad layer", li, ew, eb)
    net.close()


def test_unet_training_reduces_loss():
    torch = _torch()
    from giem2g_amd import learned
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0)
    x = torch.randn((4, 8, 8, 8, 8), generator=gen, device="cuda")
    x[..., 7] = 0
    # a smooth, learnable target: a fixed linear map of the input channels
    M = torch.randn((8, 6), generator=gen, device="cuda") * 0.5
    t = (x @ M).contiguous()
    net = learned.UNet(seed=1)
    losses = [net.train_step(x, t, 2e-3) for _ in range(150)]
    print("loss", losses[0], losses[-1], flush=True)
    assert losses[-1] < 0.3 * losses[0]
    net.close()


def test_learned_features(oracle):
    torch = _torch()
    from giem2g_amd import learned, models, operator as gop
    m = models.small(8, 8, 4, seed=3, freq=3.0)
    op = gop.operator_for(m)
    net = learned.UNet(seed=0)
    op.set_learned(net)
    rng = np.random.default_rng(5)
    n = 3 * m.ncell
    r = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    e = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x = torch.empty((m.Nz, m.Ny, m.Nx, 8), dtype=torch.float32, device="cuda")
    t = torch.empty((m.Nz, m.Ny, m.Nx, 6), dtype=torch.float32, device="cuda")
    op.learned_features(torch.from_numpy(r).cuda(), torch.from_numpy(e).cuda(), x, t)
    x, t = x.cpu().numpy(), t.cpu().numpy()
    rms = np.linalg.norm(r) / np.sqrt(n)
    rr = r.reshape(3, m.Nz, m.Ny, m.Nx)
    ee = e.reshape(3, m.Nz, m.Ny, m.Nx)
    om = oracle.Model(m.csigb(), m.thick, m.omega, m.Nx, m.Ny, m.Nz, m.dx, m.dy, m.z)
    ln = om.lnum()
    cb = m.csigb()
    sb = np.array([cb[ln[i + 1]] for i in range(m.Nz)])[:, None, None]
    sa = m.csiga().reshape(m.Nz, m.Ny, m.Nx)
    a = 2 * np.sqrt(sb.real) / (sa + np.conj(sb))
    for c in range(3):
        assert rel(x[..., 2 * c] + 1j * x[..., 2 * c + 1], rr[c] / rms) < 1e-6
        assert rel(t[..., 2 * c] + 1j * t[..., 2 * c + 1], a * ee[c] / rms) < 1e-6
    assert np.allclose(x[..., 6], np.log10(sa.real / sb.real), atol=1e-5)
    k = np.sqrt(m.omega * 4e-7 * np.pi * np.abs(sb))  # background wavenumber |k_b|
    assert np.allclose(x[..., 7], np.broadcast_to(np.log10(k * m.dx), x[..., 7].shape), atol=1e-5)
    op.set_learned(None)
    op.close()
    net.close()


@pytest.mark.parametrize("mode,prec", [(1, 0), (1, 1), (2, 0), (2, 1)])
def test_learned_modes_converge_with_untrained_net(mode, prec):
    """An untrained network must not break the solve: the fallback keeps the
    FP64 true residual below 1e-8 and the solution equal to plain FP64."""
    from giem2g_amd import learned, models, operator as gop
    print("mode", mode, "precision", prec, flush=True)
    m = models.small(16, 16, 8, seed=11, freq=1.0)
    op = gop.operator_for(m)
    E = op.plane_wave(1)
    E0, i0 = op.solve(E, outer_m=10, inner_m=10, max_restart=50, tol=1e-8)
    net = learned.UNet(seed=2)
    op.set_learned(net, accept=0.9)
    E1, i1 = op.solve(E, outer_m=10, inner_m=10, max_restart=50, tol=1e-8, inner_precision=prec,
                      precond=mode)
    print(i0, i1, flush=True)
    assert i1["status"] == 1 and i1["true_residual"] < 1e-8
    assert i1["precond_calls"] > 0
    assert rel(E1, E0) < 1e-6
    op.set_learned(None)
    op.close()
    net.close()
