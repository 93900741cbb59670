"""Oracle parity at the benchmark configurations' own sizes.

C2 (256x256x32 x 25 m, 20 random boxes, 0.1 Hz), the headline solve:
  - real-space tensor blocks against the oracle's elements, each within its
    weight-error scale (tensor_block_tol: far offsets cancel by up to 1e10);
  - the whole tensor spectrum against the oracle's FFT of the same real-space
    tensor, 1e-10 (the 512-point transforms);
  - one FP64 apply against the oracle's operator, 1e-12;
  - the FP64 (EY) and mixed complex64-inner solves against the oracle's
    nested FGMRES solve, fields within 1e-6, the FP64 solve with the
    oracle's outer iteration and matvec counts, and both solutions' residual
    recomputed with the oracle's operator below 1e-8.
C3 (Nz = 64 layers of 20 m crossing the 1 km interface, random boxes):
  - at 128x128x64, tensor blocks at 1e-3, 1 and 1e3 Hz against the oracle's
    elements, and at 1 Hz the whole spectrum (the 256-point transforms);
  - at 32x32x64, the whole tensor and spectrum at 1e-3, 1 and 1e3 Hz, and
    paired complex64-inner solves of EY and EX against the oracle's solves
    (1e-6).
"""
import dataclasses

import numpy as np
import pytest

from tests.conftest import rel, tensor_block_tol

pytestmark = pytest.mark.gpu

C3_FREQS = (1e-3, 1.0, 1e3)


def oracle_model(po, m):
    return po.Model(m.csigb(), m.thick, m.omega, m.Nx, m.Ny, m.Nz, m.dx, m.dy, m.z)


def check_blocks(om, G, BS, Nx, offsets):
    """Real-space blocks against the oracle's elements.  Far offsets are sums
    that cancel by up to 1e10 (the diagonal ones most), so each block is held
    to its weight-error scale (tensor_block_tol), which the reference's own
    FP64 and quad-precision blocks calibrate."""
    for lx, ly in offsets:
        b = lx + Nx * ly
        blk = om.tensor_element(lx, ly)
        d = np.linalg.norm(G[b * BS:(b + 1) * BS] - blk)
        tol = tensor_block_tol(om, lx, ly)
        assert d < tol, (lx, ly, d, tol, np.linalg.norm(blk))


def scaled_residual(om, m, Gs, E, Ebkg):
    """||b - A x|| / ||b|| with the oracle's operator, x = E/a, b = r E_bkg
    (ie_solver_module.F90:64-66, 85-97)."""
    ln = om.lnum()
    sb = np.array([m.csigb()[ln[i + 1]] for i in range(m.Nz)])
    r = np.sqrt(sb.real)
    sa = m.csiga().reshape(m.Nz, -1)
    a = (2.0 * r[:, None] / (sa + np.conj(sb)[:, None])).reshape(-1)
    x = (E.reshape(3, -1) / a[None, :]).reshape(-1)
    b = (Ebkg.reshape(3, m.Nz, -1) * r[None, :, None]).reshape(-1)
    return np.linalg.norm(b - om.apply(Gs, m.csiga(), x)) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def c2_case(oracle):
    from giem2g_amd import models, operator as gop
    m = models.c2()
    om = oracle_model(oracle, m)
    op = gop.IEOperator.from_model(m)
    op.set_background(m.csigb(), m.thick, m.omega)
    st = {"m": m, "om": om, "op": op}
    yield st
    op.close()


def test_c2_tensor_and_spectrum(c2_case):
    m, om, op = c2_case["m"], c2_case["om"], c2_case["op"]
    print("C2 tensor build", flush=True)
    op.tensor_stage(0)
    Gr = op.get_tensor()
    check_blocks(om, Gr, op.block_size, m.Nx,
                 [(0, 0), (1, 0), (0, 1), (5, 3), (31, 200), (128, 128), (255, 255), (200, 7)])
    print("C2 oracle tensor FFT", flush=True)
    Gs_ref = om.tensor_fft(Gr)
    del Gr
    op.tensor_stage(1)
    e = rel(op.get_tensor(), Gs_ref)
    print("C2 spectrum rel %.2e" % e, flush=True)
    assert e < 1e-10
    c2_case["Gs"] = Gs_ref


def test_c2_fp64_apply(c2_case):
    m, om, op = c2_case["m"], c2_case["om"], c2_case["op"]
    Gs = c2_case["Gs"]
    op.set_sigma(m.csiga())
    rng = np.random.default_rng(22)
    v = rng.standard_normal(3 * m.ncell) + 1j * rng.standard_normal(3 * m.ncell)
    e = rel(op.apply_host(v), om.apply(Gs, m.csiga(), v))
    print("C2 apply rel %.2e" % e, flush=True)
    assert e < 1e-12


def test_c2_solves_match_oracle(c2_case):
    import torch
    m, om, op = c2_case["m"], c2_case["om"], c2_case["op"]
    Gs = c2_case["Gs"]
    Eb = op.plane_wave(1)  # EY, the reference's first polarization
    Ebt = torch.from_numpy(Eb).cuda()
    E64, i64 = op.solve(Ebt, outer_m=30, inner_m=30, max_restart=100, tol=1e-8)
    Emx, imx = op.solve(Ebt, outer_m=30, inner_m=30, max_restart=100, tol=1e-8, inner_precision=1)
    E64, Emx = E64.cpu().numpy(), Emx.cpu().numpy()
    print("C2 GPU solves", i64, imx, flush=True)
    print("C2 oracle solve", flush=True)
    Eo, oi = om.solve(Gs, m.csiga(), Eb, m_out=30, m_in=30, maxit=100, tol=1e-8)
    e64, emx = rel(E64, Eo), rel(Emx, Eo)
    r64, rmx = scaled_residual(om, m, Gs, E64, Eb), scaled_residual(om, m, Gs, Emx, Eb)
    print("C2 FP64 vs oracle %.2e (outer %d / %d, matvecs %d / %d), mixed %.2e; oracle-operator "
          "residuals %.2e %.2e" % (e64, i64["outer_iterations"], oi.iterations, i64["matvecs_fp64"],
                                   oi.matvecs, emx, r64, rmx), flush=True)
    assert i64["status"] == 1 and imx["status"] == 1 and oi.stat == 1
    assert e64 < 1e-6 and emx < 1e-6
    assert i64["outer_iterations"] == oi.iterations and i64["matvecs_fp64"] == oi.matvecs
    assert r64 < 1e-8 and rmx < 1e-8
    c2_case.pop("Gs")


def test_c3_tensor_at_frequency_extremes(oracle):
    from giem2g_amd import models, operator as gop
    m, _ = models.c3()
    op = gop.IEOperator.from_model(m)
    for f in C3_FREQS:
        mf = dataclasses.replace(m, freq=f)
        om = oracle_model(oracle, mf)
        print("C3 tensor at %g Hz" % f, flush=True)
        op.set_background(mf.csigb(), mf.thick, mf.omega)
        op.tensor_stage(0)
        Gr = op.get_tensor()
        check_blocks(om, Gr, op.block_size, m.Nx, [(0, 0), (1, 0), (0, 2), (9, 4), (64, 100), (127, 127)])
        if f == 1.0:
            Gs_ref = om.tensor_fft(Gr)
            del Gr
            op.tensor_stage(1)
            e = rel(op.get_tensor(), Gs_ref)
            print("C3 spectrum rel %.2e" % e, flush=True)
            assert e < 1e-10
            del Gs_ref
    op.close()


def c3_small(f):
    """32x32 cells of 100 m, the C3 layers (64 x 20 m, 0-1280 m), random boxes
    of 1e-3..10 S/m in a 0.01 S/m host, at frequency f."""
    from giem2g_amd import models
    rng = np.random.Generator(np.random.PCG64(33))
    siga = models.random_boxes(rng, 32, 32, 64, 6, 0.01, 1e-3, 10.0, size_lat=(2, 16), size_z=(2, 32))
    return models.Model("C3-32", 32, 32, 64, 100.0, 100.0, np.arange(65) * 20.0, models.BKG_SIGMA,
                        models.BKG_THICK, f, siga)


@pytest.mark.parametrize("f", C3_FREQS)
def test_c3_paired_mixed_solves_match_oracle(oracle, f):
    import torch
    from giem2g_amd import operator as gop
    print("C3 32x32x64 at %g Hz" % f, flush=True)
    m = c3_small(f)
    om = oracle_model(oracle, m)
    Gr = om.tensor_realspace()
    Gs = om.tensor_fft(Gr)
    op = gop.IEOperator.from_model(m)
    op.set_background(m.csigb(), m.thick, m.omega)
    op.tensor_stage(0)
    assert rel(op.get_tensor(), Gr) < 1e-10
    op.tensor_stage(1)
    assert rel(op.get_tensor(), Gs) < 1e-10
    op.set_sigma(m.csiga())
    Ey, Ex = op.plane_wave(1), op.plane_wave(0)
    oy, ox, iy, ix, cnt = op.solve_pair(torch.from_numpy(Ey).cuda(), torch.from_numpy(Ex).cuda(),
                                        inner_precision=1)
    assert cnt[0] > 0  # the two solves shared tensor reads
    for E, Eb, info, name in ((oy, Ey, iy, "EY"), (ox, Ex, ix, "EX")):
        Eo, oi = om.solve(Gs, m.csiga(), Eb, m_out=30, m_in=30, maxit=100, tol=1e-8)
        e = rel(E.cpu().numpy(), Eo)
        print("  %s: outer %d (oracle %d), mv64 %d mv32 %d, rel %.2e" %
              (name, info["outer_iterations"], oi.iterations, info["matvecs_fp64"], info["matvecs_fp32"], e),
              flush=True)
        assert info["status"] == 1 and info["true_residual"] < 1e-8
        assert e < 1e-6
    op.close()
