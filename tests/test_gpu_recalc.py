"""GPU parity of the continuation of the solution to receivers against the
oracle (Continuation_Function_Module.F90, calc_rc_tensor_module.F90,
rc_kernel_hankel_module.F90, sources_module.F90:26-66).

Bars: receiver preparation exact; plane wave at receivers within 1e-13;
real-space receiver tensor and its spectrum within 1e-10 (2-norm over the
tensor) of the oracle; the FP64 continuation on the oracle's spectrum within
1e-12 (2-norm) of the oracle; the whole GPU path (tensor + continuation)
within 1e-9.
"""
import numpy as np
import pytest

from tests.conftest import rel

pytestmark = pytest.mark.gpu

RECV = {  # xshift, yshift, depth: above, inside (moved to a border), on a border, below, layer 2
    3: ([0.0, 30.0, -45.0], [0.0, -40.0, 10.0], [0.0, 75.0, 350.0]),
    4: ([10.0, -20.0, 0.0, 35.0], [5.0, 0.0, -30.0, 45.0], [0.0, 100.0, 1200.0, 20.0]),
    1: ([12.5], [-7.5], [60.0]),
}


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU test without a GPU"
    return torch


def oracle_model(po, m):
    return po.Model(m.csigb(), m.thick, m.omega, m.Nx, m.Ny, m.Nz, m.dx, m.dy, m.z)


def rand_vec(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def make_ops(m, xs, ys, depth):
    from giem2g_amd import operator as gop
    op = gop.IEOperator.from_model(m)
    op.set_background(m.csigb(), m.thick, m.omega)
    rc = gop.ReceiverOperator(op, xs, ys, depth)
    return op, rc


def test_receiver_preparation_and_plane_wave(oracle):
    from giem2g_amd import models
    m = models.small(8, 8, 4, seed=3, freq=0.7)
    xs, ys = [0.0] * 6, [0.0] * 6
    depth = [0.0, 75.0, 100.0, 200.0, 350.0, 1500.0]
    op, rc = make_ops(m, xs, ys, depth)
    ro = oracle.Receivers(oracle_model(oracle, m), xs, ys, depth)
    zr, layer, cell = rc.receivers()
    assert np.array_equal(zr, ro.zr) and np.array_equal(layer, ro.layer) and np.array_equal(cell, ro.cell)
    for pol in (0, 1):
        assert rel(rc.plane_wave(pol), ro.plane_wave(pol)) < 1e-13, pol
    rc.close()
    op.close()


@pytest.mark.parametrize("Nx,Ny,Nz,Nr", [(8, 8, 4, 3), (16, 8, 3, 4), (8, 16, 5, 1),
                                          (10, 14, 5, 3), (8, 8, 72, 1)])
def test_receiver_tensor_matches_oracle(oracle, Nx, Ny, Nz, Nr):
    from giem2g_amd import models
    print("case", Nx, Ny, Nz, Nr, flush=True)
    m = models.small(Nx, Ny, Nz, seed=Nx + Nz, freq=1.3)
    xs, ys, depth = RECV[Nr]
    op, rc = make_ops(m, xs, ys, depth)
    ro = oracle.Receivers(oracle_model(oracle, m), xs, ys, depth)
    Gr = ro.tensor_realspace()
    rc.calc_kernel()
    Gg = rc.get_tensor()
    assert rel(Gg, Gr) < 1e-10
    blk = np.abs(Gg - Gr).reshape(4 * Nx * Ny, -1).max(1)
    assert blk.max() < 1e-9 * np.abs(Gr).max()
    rc.calc_fft_of_kernel()
    assert rel(rc.get_tensor(), ro.tensor_fft(Gr)) < 1e-10
    rc.close()
    op.close()


@pytest.mark.parametrize("Nx,Ny,Nz,Nr", [(8, 8, 4, 3), (16, 8, 3, 4), (8, 16, 5, 1), (8, 8, 1, 3),
                                          (16, 16, 32, 4), (12, 10, 6, 4), (8, 8, 70, 3)])
def test_continuation_matches_oracle_on_oracle_tensor(oracle, Nx, Ny, Nz, Nr):
    """The FFT continuation alone (forward FFTs, 17 block products, inverse
    FFTs) on the oracle's spectrum; host and HBM entry points agree exactly."""
    torch = _torch()
    from giem2g_amd import models
    print("case", Nx, Ny, Nz, Nr, flush=True)
    m = models.small(Nx, Ny, Nz, seed=7 * Nx + Nz, freq=0.4, dz=40.0)
    xs, ys, depth = RECV[Nr]
    op, rc = make_ops(m, xs, ys, depth)
    om = oracle_model(oracle, m)
    ro = oracle.Receivers(om, xs, ys, depth)
    Gs = ro.tensor_fft(ro.tensor_realspace())
    E = rand_vec(3 * m.ncell, Nz)
    Ea_r, Ha_r = ro.apply(Gs, m.csiga(), E)
    rc.put_tensor(Gs, spectral=True)
    rc.set_sigma(m.csiga())
    Ea, Ha = rc.recalculate(E)
    assert rel(Ea, Ea_r) < 1e-12 and rel(Ha, Ha_r) < 1e-12
    Et = torch.from_numpy(E).cuda()
    Ead = torch.empty(rc.nfield, dtype=torch.complex128, device="cuda")
    Had = torch.empty_like(Ead)
    rc.recalculate_dev(Et, Ead, Had)
    op.sync()
    assert np.array_equal(Ead.cpu().numpy(), Ea) and np.array_equal(Had.cpu().numpy(), Ha)
    rc.close()
    op.close()


def test_continuation_end_to_end_matches_oracle(oracle):
    """GPU tensor build + continuation against the oracle's, and the
    conductivity from the device entry point."""
    torch = _torch()
    from giem2g_amd import models
    m = models.small(16, 16, 4, seed=21, freq=3.0)
    xs, ys, depth = RECV[3]
    op, rc = make_ops(m, xs, ys, depth)
    ro = oracle.Receivers(oracle_model(oracle, m), xs, ys, depth)
    Gs = ro.tensor_fft(ro.tensor_realspace())
    E = rand_vec(3 * m.ncell, 5)
    Ea_r, Ha_r = ro.apply(Gs, m.csiga(), E)
    rc.build()
    rc.set_sigma(torch.from_numpy(m.csiga()).cuda())
    Ea, Ha = rc.recalculate(E)
    assert rel(Ea, Ea_r) < 1e-9 and rel(Ha, Ha_r) < 1e-9
    rc.close()
    op.close()


def test_full_size_reference_receivers(oracle):
    """128 x 128 x 32 (the C1 grid, 10 Hz) with the reference deck's receivers
    (test/giem2g_input.json: xshift 0, yshift -+20 km, depth 0) plus one in
    the anomaly: sampled tensor blocks against the oracle, the continuation
    on the GPU's spectrum against the oracle's FFT continuation, linearity,
    and zero field for sigma_a = sigma_b."""
    from giem2g_amd import models
    m = models.c1()
    xs, ys, depth = [0.0, 0.0, 25.0], [-20000.0, 20000.0, -15.0], [0.0, 0.0, 400.0]
    op, rc = make_ops(m, xs, ys, depth)
    om = oracle_model(oracle, m)
    ro = oracle.Receivers(om, xs, ys, depth)
    rc.calc_kernel()
    G = rc.get_tensor().reshape(4 * m.Nx * m.Ny, -1)
    Ny2 = 2 * m.Ny
    for lx, ly in [(0, 0), (1, 0), (-1, 3), (17, -40), (127, 127), (-127, -1)]:
        ix, iy = lx % (2 * m.Nx), ly % Ny2
        assert rel(G[ix * Ny2 + iy], ro.tensor_element(lx, ly)) < 1e-10, (lx, ly)
    rc.calc_fft_of_kernel()
    Gs = rc.get_tensor()
    rc.set_sigma(m.csiga())
    E1, E2 = rand_vec(3 * m.ncell, 1), rand_vec(3 * m.ncell, 2)
    Ea1, Ha1 = rc.recalculate(E1)
    Ea_r, Ha_r = ro.apply(Gs, m.csiga(), E1)
    assert rel(Ea1, Ea_r) < 1e-12 and rel(Ha1, Ha_r) < 1e-12
    Ea2, Ha2 = rc.recalculate(E2)
    a, b = 0.3 - 1.1j, 2.0 + 0.5j
    Ea3, Ha3 = rc.recalculate(a * E1 + b * E2)
    assert rel(Ea3, a * Ea1 + b * Ea2) < 1e-12 and rel(Ha3, a * Ha1 + b * Ha2) < 1e-12
    lay = om.lnum()[1:]
    rc.set_sigma(np.repeat(m.csigb()[lay], m.Nx * m.Ny))
    Ea0, Ha0 = rc.recalculate(E1)
    assert np.abs(Ea0).max() == 0 and np.abs(Ha0).max() == 0
    rc.close()
    op.close()


def test_receiver_errors():
    from giem2g_amd import lib as L, models, operator as gop
    m = models.small(8, 8, 2, seed=1)
    op = gop.IEOperator.from_model(m)
    with pytest.raises(L.G2GError, match="corner"):
        gop.ReceiverOperator(op, [50.0], [-50.0], [0.0])
    with pytest.raises(L.G2GError):
        gop.ReceiverOperator(op, [], [], [])
    op.set_background(m.csigb(), m.thick, m.omega)
    rc = gop.ReceiverOperator(op, [0.0], [0.0], [0.0])
    with pytest.raises(L.G2GError, match="not built"):
        rc.recalculate(np.zeros(3 * m.ncell, np.complex128))
    rc.build()
    with pytest.raises(L.G2GError, match="conductivity"):
        rc.recalculate(np.zeros(3 * m.ncell, np.complex128))
    rc.set_sigma(m.csiga())
    rc.recalculate(np.zeros(3 * m.ncell, np.complex128))
    op.set_background(m.csigb(2.0), m.thick, 2 * np.pi * 2.0)  # a new frequency
    with pytest.raises(L.G2GError, match="current frequency"):
        rc.recalculate(np.zeros(3 * m.ncell, np.complex128))
    with pytest.raises(L.G2GError, match="receiver operators"):
        op.close()  # the receiver operator still holds it
    rc.close()
    op.close()
