"""The allocation-free batched apply (giem2g_op_apply_batch_dev_noalloc)
against the allocating one (giem2g_op_apply_batch_dev), through the C ABI.

Each case makes two calls in a row on one GPU, the second with a smaller
input (fewer vectors, then a smaller grid), so that what the first call left
in the shared scratch array cannot leak into the second.  FP64 results agree
within 1e-14 relative and complex64 within 2e-6 (the complex64 operator's
bar against FP64); where both entry points batch the same vectors per
tensor read they are bitwise equal.  Grids whose one vector does not fit in
the array, and complex64 before the complex64 operator exists, are errors.
"""
import numpy as np
import pytest

from tests.conftest import rel

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available()
    return torch


def make_op(Nx, Ny, Nz, seed):
    from giem2g_amd import models, operator as gop
    m = models.small(Nx, Ny, Nz, seed=seed, freq=0.7, dz=25.0)
    op = gop.IEOperator.from_model(m)
    op.set_background(m.csigb(), m.thick, m.omega)
    op.build_tensor()
    op.set_sigma(m.csiga())
    return m, op


def rand_rows(torch, nrhs, n, seed, dt):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((nrhs, n)) + 1j * rng.standard_normal((nrhs, n))
    return torch.from_numpy(v).cuda().to(dt).contiguous()


def compare(torch, op, nrhs, dt, seed):
    vin = rand_rows(torch, nrhs, op.nvec, seed, dt)
    ref = torch.empty_like(vin)
    got = torch.full_like(vin, float("nan"))
    op.apply_batch_dev(vin, ref)
    op.apply_batch_dev_noalloc(vin, got)
    op.sync()
    tol = 1e-14 if dt == torch.complex128 else 2e-6
    for r in range(nrhs):
        e = rel(got[r], ref[r])
        assert e < tol, (dt, nrhs, r, e)
    return got


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_noalloc_batch_matches_batch_two_calls(prec):
    torch = _torch()
    dt = torch.complex128 if prec == "c128" else torch.complex64
    print("noalloc 64x64x16 nrhs 5 then 3, then 32x32x8 nrhs 2,", prec, flush=True)
    _, op = make_op(64, 64, 16, 5)
    compare(torch, op, 5, dt, 1)
    compare(torch, op, 3, dt, 2)
    op.close()
    _, op = make_op(32, 32, 8, 6)
    compare(torch, op, 2, dt, 3)
    op.close()


def test_noalloc_batch_fewer_per_read_at_128x128x32():
    """At 128x128x32 FP64 the array holds two spectrum slots (one tensor read
    per pair, against four for the allocating apply); the products agree."""
    torch = _torch()
    print("noalloc 128x128x32 FP64 nrhs 4 then 3", flush=True)
    _, op = make_op(128, 128, 32, 7)
    compare(torch, op, 4, torch.complex128, 4)
    compare(torch, op, 3, torch.complex128, 5)
    op.close()


def test_noalloc_rejects_what_it_cannot_serve():
    torch = _torch()
    from giem2g_amd import operator as gop
    print("noalloc error paths", flush=True)
    _, op = make_op(32, 32, 8, 8)
    vin = rand_rows(torch, 2, op.nvec, 6, torch.complex64)
    with pytest.raises(Exception, match="complex64 operator not made"):
        op.apply_batch_dev_noalloc(vin, torch.empty_like(vin))
    op.close()
    # 256x256x32 FP64: one spectrum alone is 403 MB, more than the 256 MiB array
    big = gop.IEOperator(256, 256, 32, 100.0, 100.0, np.arange(33) * 25.0)
    v = torch.zeros((1, big.nvec), dtype=torch.complex128, device="cuda")
    with pytest.raises(Exception, match="too large for the allocation-free apply"):
        big.apply_batch_dev_noalloc(v, torch.empty_like(v))
    big.close()
