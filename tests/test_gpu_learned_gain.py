"""What the trained network gains: the learned right preconditioner with
the committed weights (giem2g_amd/data/unet_c4.f32, trained on the GPU by
python -m giem2g_amd.train_precond on 2000 C4 models) against the same
solver without it, on held-out random-box models (seeds outside the
training range 10000..11999).

Measured with these weights: on the residuals of the outer iterates the
network's step reduces the residual by less than 1% (least-squares
rho = ||r - alpha A M(r)|| / ||r|| > 0.99), so the acceptance test rejects
it and the fallback runs the plain inner GMRES: the gain is nil, and after
two rejected steps in a row the solve stops calling the network, so the cost
is at most two extra operator applications and two network applications.  This test asserts what the fallback guarantees -- the same
convergence, the same field within 1e-6, at most that overhead -- and
prints the measured gain and rho.
"""
import os

import pytest

from tests.conftest import rel

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(os.path.dirname(HERE), "giem2g_amd", "data", "unet_c4.f32")


def _rho(op, torch, r):
    z = torch.empty_like(r)
    az = torch.empty_like(r)
    op.learned_apply(r, z)
    op.apply_dev(z, az)
    torch.cuda.synchronize()
    alpha = torch.vdot(az, r) / torch.vdot(az, az)
    return float(torch.linalg.vector_norm(r - alpha * az) / torch.linalg.vector_norm(r))


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="no trained weights committed")
@pytest.mark.parametrize("seed,mode", [(12100, 1), (12101, 2)])
def test_learned_preconditioner_gain(seed, mode):
    import torch
    from giem2g_amd import learned, models, operator as gop
    print("model seed", seed, "mode", mode, flush=True)
    m = models.c4_train(seed)
    op = gop.operator_for(m)
    net = learned.UNet.load(WEIGHTS)
    op.set_learned(net, accept=0.9)
    E = torch.from_numpy(op.plane_wave(seed % 2)).cuda()
    kw = dict(outer_m=30, inner_m=30, max_restart=100, tol=1e-8, inner_precision=1)
    E0, i0 = op.solve(E, precond=0, **kw)
    E1, i1 = op.solve(E, precond=mode, **kw)
    b = torch.empty_like(E)
    op.scale(0, E, b)
    r = torch.empty_like(E)
    op.apply_dev(b, r)
    torch.cuda.synchronize()
    rho = _rho(op, torch, (b - r).contiguous())
    cost0 = i0["matvecs_fp32"] + i0["matvecs_fp64"]
    cost1 = i1["matvecs_fp32"] + i1["matvecs_fp64"]
    print("plain", i0, "\nlearned", i1, flush=True)
    print("gain in operator applications %.3f, rho %.4f" % (cost0 / cost1, rho), flush=True)
    for info in (i0, i1):
        assert info["status"] == 1 and info["true_residual"] < 1e-8
    assert rel(E1, E0) < 1e-6
    # after two rejected learned steps in a row the solve stops calling the
    # network (learned_patience, default 2)
    rejected = i1["precond_fallbacks"]
    assert 1 <= i1["precond_calls"] <= i1["outer_iterations"]
    if rejected == i1["precond_calls"]:
        assert i1["precond_calls"] == min(2, i1["outer_iterations"])
    assert i1["outer_iterations"] <= i0["outer_iterations"] + 1
    assert cost1 <= cost0 + 2 * i1["precond_calls"] + 2
    op.set_learned(None)
    op.close()
    net.close()
