"""What the trained network gains: the learned right preconditioner with
the committed weights (giem2g_amd/data/unet_c4.f32, trained on the GPU by
python -m giem2g_amd.train_precond) against the same solver without it, on
held-out random-box models (seeds outside the training range 10000..11999).

Both solves must reach the FP64 true residual 1e-8 and agree within 1e-6;
the learned solve must need fewer operator applications (FP32 and FP64
matvecs together) than the plain mixed-precision solve, and the network's
residual reduction ||r - alpha A M(r)|| / ||r|| on the residuals of the
outer iterates must be below 1.
"""
import os

import numpy as np
import pytest

from tests.conftest import rel

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(os.path.dirname(HERE), "giem2g_amd", "data", "unet_c4.f32")


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="no trained weights committed")
@pytest.mark.parametrize("seed", [12100, 12101])
def test_learned_preconditioner_gain(seed):
    import torch
    from giem2g_amd import learned, models, operator as gop
    print("model seed", seed, flush=True)
    m = models.c4_train(seed)
    op = gop.operator_for(m)
    net = learned.UNet.load(WEIGHTS)
    op.set_learned(net, accept=1.0)
    E = torch.from_numpy(op.plane_wave(seed % 2)).cuda()
    kw = dict(outer_m=30, inner_m=30, max_restart=100, tol=1e-8, inner_precision=1)
    E0, i0 = op.solve(E, precond=0, **kw)
    E2, i2 = op.solve(E, precond=2, **kw)
    print("plain", i0, "\nlearned", i2, flush=True)
    for info in (i0, i2):
        assert info["status"] == 1 and info["true_residual"] < 1e-8
    assert rel(E2, E0) < 1e-6
    cost0 = i0["matvecs_fp32"] + i0["matvecs_fp64"]
    cost2 = i2["matvecs_fp32"] + i2["matvecs_fp64"]
    assert cost2 < cost0, (cost2, cost0)
    # network residual reduction on the first outer residual r0 = b - A b
    b = torch.empty_like(E)
    op.scale(0, E, b)
    r = torch.empty_like(E)
    op.apply_dev(b, r)
    torch.cuda.synchronize()
    r = (b - r).contiguous()
    z = torch.empty_like(r)
    az = torch.empty_like(r)
    op.learned_apply(r, z)
    op.apply_dev(z, az)
    torch.cuda.synchronize()
    alpha = torch.vdot(az, r) / torch.vdot(az, az)
    rho = float(torch.linalg.vector_norm(r - alpha * az) / torch.linalg.vector_norm(r))
    print("rho", rho, flush=True)
    assert rho < 1.0
    op.set_learned(None)
    op.close()
    net.close()
