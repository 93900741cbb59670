"""CPU checks of the oracle's receiver continuation (the field recalculation
at receivers after the solve; Continuation_Function_Module.F90,
calc_rc_tensor_module.F90, rc_kernel_hankel_module.F90, sources_module.F90).

Pins:
  * the reference's own Fortran routines, compiled into oracle/_ref:
    single-integral Hankel weights (VBTransformWeightsAllSingleInt),
    PrepareRecvs and the receiver kernel Calc_Integral_U, through the fixtures
    of tests/golden/reference_vectors.npz, and tensor blocks composed from
    them;
  * physics that the reference's code does not state: the E tensor averaged
    over a receiver cell equals the integral-equation tensor of that cell
    pair divided by dz (itself pinned against the reference); the H tensor
    obeys Faraday's law curl E = i w mu0 H; the normal field at receivers
    averages to the cell-averaged plane wave and obeys the same law;
  * the FFT continuation equals the direct convolution sum.
"""
import numpy as np
import pytest

from tests.conftest import rel

MU0 = 4e-7 * np.pi
EPS0 = 1.0 / (MU0 * 299792458.0 ** 2)
BKG3 = ([0.01, 0.1, 1.0], [1000.0, 9000.0])


def csig(sig, f):
    w = 2 * np.pi * f
    return np.array([1e-12 - 1j * w * EPS0] + [s - 1j * w * EPS0 for s in sig])


def model(po, z, Nx=8, Ny=8, f=1.0, sig=BKG3[0], th=BKG3[1], dxy=100.0):
    z = np.asarray(z, np.float64)
    return po.Model(csig(sig, f), list(th), 2 * np.pi * f, Nx, Ny, len(z) - 1, dxy, dxy, z)


# ----------------------------------------------------------- golden pins --
def test_single_integral_weights_match_reference(oracle, golden):
    for (x, y, hx, hy), WTr in zip(golden["rc_weight_cases"], golden["rc_weights"]):
        WT = oracle.vbt_weights_single(x, y, hx, hy, np.hypot(hx, hy) / 2)
        assert np.abs(WT - WTr).max() <= 1e-12 * np.abs(WTr).max(), (x, y)


def test_prepare_receivers_matches_reference(oracle, golden):
    from tests.golden.make_golden import BKGS, ZS
    depths = golden["rc_depths"]
    for bi, (sig, th, f) in enumerate(BKGS):
        for zi, z in enumerate(ZS):
            m = oracle.Model(csig(sig, f), th, 2 * np.pi * f, 4, 4, len(z) - 1, 100.0, 100.0, z)
            rc = oracle.Receivers(m, np.zeros(len(depths)), np.zeros(len(depths)), depths)
            zr, layer, cell = golden["rc_recv_%d_%d" % (bi, zi)]
            assert np.array_equal(rc.zr, zr), (bi, zi)
            assert np.array_equal(rc.layer, layer.astype(np.int32)), (bi, zi)
            assert np.array_equal(rc.cell, cell.astype(np.int32)), (bi, zi)


def test_receiver_kernel_matches_reference(oracle, golden):
    """Calc_Integral_U for receivers above, inside (at a border, or moved up
    to one), and below the anomaly, in every background layer."""
    from tests.golden.make_golden import BKGS, LAMBDAS, W6, ZS
    depths = golden["rc_depths"]
    for bi, (sig, th, f) in enumerate(BKGS):
        for zi, z in enumerate(ZS):
            m = oracle.Model(csig(sig, f), th, 2 * np.pi * f, 4, 4, len(z) - 1, 100.0, 100.0, z)
            rc = oracle.Receivers(m, np.zeros(len(depths)), np.zeros(len(depths)), depths)
            U = golden["rcU_%d_%d" % (bi, zi)]
            for di in range(len(depths)):
                for li, lm in enumerate(LAMBDAS):
                    GE, GH = oracle.rc_integral_U(m, rc.zr[di], int(rc.layer[di]), int(rc.cell[di]), lm, W6)
                    G = np.concatenate([GE, GH])
                    Gr = U[di, li]
                    assert np.all(np.isfinite(G))
                    scale = np.abs(Gr).max()
                    if scale == 0:  # exp underflow far from the receiver: exactly zero in both
                        assert np.abs(G).max() == 0, (bi, zi, di, li)
                    else:
                        assert np.abs(G - Gr).max() <= 1e-12 * scale, (bi, zi, di, li)


def test_receiver_tensor_matches_reference_composition(oracle, golden):
    """Real-space receiver tensor blocks (receivers above, inside and below
    the anomaly, shifted laterally) against blocks composed from the
    reference's weights and kernel; observed <= 2.5e-11 per block."""
    from tests.golden.make_golden import TENSOR_MODELS
    for sig, th, f, z, tag in TENSOR_MODELS:
        m = oracle.Model(csig(sig, f), th, 2 * np.pi * f, 32, 32, len(z) - 1, 100.0, 100.0, z)
        R = golden["rc_recvs_%s" % tag]
        rc = oracle.Receivers(m, R[:, 0], R[:, 1], R[:, 2])
        for lx, ly in golden["rc_offsets"]:
            ref = golden["rctensor_%s_%d_%d" % (tag, lx, ly)]
            assert rel(rc.tensor_element(int(lx), int(ly)), ref) < 1e-9, (tag, lx, ly)


@pytest.mark.skipif("not __import__('oracle.pyoracle', fromlist=['x']).ref_available()")
def test_receiver_pieces_vs_reference_build_direct(oracle):
    """Fresh random receivers and offsets against oracle/_ref (only where it
    is built)."""
    R = oracle.ref()
    rng = np.random.default_rng(9)
    sig, th, f = [0.02, 0.3, 2.0], [700.0, 5000.0], 0.7
    cs = csig(sig, f)
    csd = np.ascontiguousarray(cs.view(np.float64))
    thd = np.array(th + [0.0])
    z = np.sort(np.r_[0.0, rng.uniform(10, 1500, 6)])
    m = oracle.Model(cs, th, 2 * np.pi * f, 4, 4, 6, 100.0, 100.0, z)
    for _ in range(6):
        dx, dy = rng.uniform(20, 500, 2)
        x, y = rng.uniform(-3e4, 3e4, 2)
        WTr = np.zeros(501 * 6)
        lms = np.zeros(501)
        R.ref_vbt_weights_single(np.hypot(dx, dy) / 2, x, y, dx, dy, WTr, lms)
        WT = oracle.vbt_weights_single(x, y, dx, dy, np.hypot(dx, dy) / 2)
        assert np.abs(WT - WTr.reshape(6, 501).T).max() <= 1e-12 * np.abs(WTr).max()
        depth = rng.uniform(0, 3000)
        lm = 10 ** rng.uniform(-5, -0.5)
        W = rng.standard_normal(6)
        GE = np.zeros(8 * 6, np.complex128)
        GH = np.zeros(7 * 6, np.complex128)
        R.ref_rc_integral_u(3, csd, thd, 2 * np.pi * f, 6, z, depth, lm, W, GE, GH)
        rc = oracle.Receivers(m, [0.0], [0.0], [depth])
        oE, oH = oracle.rc_integral_U(m, rc.zr[0], int(rc.layer[0]), int(rc.cell[0]), lm, W)
        ref = np.concatenate([GE.reshape(6, 8).T, GH.reshape(6, 7).T])
        assert np.abs(np.concatenate([oE, oH]) - ref).max() <= 1e-12 * np.abs(ref).max()


# ------------------------------------------------------- physics pins ----
def _ie_pair(po, zc, lx, ly):
    """Integral-equation tensor entries between the two cells of zc (three
    borders): receiver-cell-1 <- source-cell-0 and the reverse, per block
    name; zx and zy are the transposed, negated xz and yz blocks."""
    m = model(po, zc)
    blk = m.tensor_element(lx, ly)
    P, Nz = 3, 2
    xz = blk[4 * P:4 * P + 4]
    yz = blk[4 * P + 4:]

    def pair(i, j):  # receiver cell i, source cell j
        lo, hi = min(i, j), max(i, j)
        pi = lo + hi * (hi + 1) // 2
        return {"XX": blk[pi], "XY": blk[P + pi], "YY": blk[2 * P + pi], "ZZ": blk[3 * P + pi],
                "XZ": xz[i + Nz * j], "YZ": yz[i + Nz * j], "ZX": -xz[j + Nz * i],
                "ZY": -yz[j + Nz * i]}
    return pair(1, 0), pair(0, 1)


@pytest.mark.parametrize("lx,ly", [(3, 2), (1, 6), (5, 4)])
def test_receiver_tensor_cell_average_equals_ie_tensor(oracle, lx, ly):
    """The E part of the receiver tensor averaged (6^3 Gauss-Legendre
    points) over a receiver cell equals the integral-equation tensor between
    that cell and the source cell divided by the receiver cell's dz -- for
    receivers below the source (anomaly cell index Nz+1) and above it (0).
    Observed 1e-10 .. 1e-13."""
    zc = [0.0, 50.0, 100.0]
    below, above = _ie_pair(oracle, zc, lx, ly)
    gx, gw = np.polynomial.legendre.leggauss(6)
    X, Y, Z = np.meshgrid(gx, gx, gx, indexing="ij")
    Wt = (gw[:, None, None] * gw[None, :, None] * gw[None, None, :]).reshape(-1) / 8
    names = ["XX", "XY", "XZ", "YY", "YZ", "ZX", "ZY", "ZZ"]
    for src, (zlo, zhi), ie in (([0.0, 50.0], (50.0, 100.0), below), ([50.0, 100.0], (0.0, 50.0), above)):
        m = model(oracle, src)
        depth = (zlo + zhi) / 2 + Z.reshape(-1) * (zhi - zlo) / 2
        rc = oracle.Receivers(m, X.reshape(-1) * 50.0, Y.reshape(-1) * 50.0, depth)
        assert np.all(rc.cell == (2 if zlo >= 50.0 else 0))
        B = rc.tensor_element(lx, ly).reshape(15, rc.Nr)
        avg = B[:8] @ Wt
        for e, nm in enumerate(names):
            ref = ie[nm] / (zhi - zlo)
            assert abs(avg[e] - ref) <= 1e-8 * abs(ref), (lx, ly, zlo, nm)


def _curl_check(B, w, h):
    """Residual of curl E = i w mu0 H from receivers at (x0 + s, y0, z0),
    (x0, y0 + s, z0), (x0, y0, z0 + s), s = -2h, -h, h, 2h (entries 1..12),
    for unit currents along x, y and z; fourth-order central differences."""
    cf = np.array([1, -8, 8, -1]) / (12 * h)
    E = {0: (B[0], B[1], B[5]), 1: (B[1], B[3], B[6]), 2: (B[2], B[4], B[7])}
    H = {0: (B[8], B[11], B[13]), 1: (B[9], -B[8], B[14]), 2: (B[10], B[12], 0 * B[8])}
    worst = 0.0
    for c in range(3):
        Ex, Ey, Ez = E[c]
        d = lambda F, k: F[1 + 4 * k:5 + 4 * k] @ cf  # noqa: E731
        curl = np.array([d(Ez, 1) - d(Ey, 2), d(Ex, 2) - d(Ez, 0), d(Ey, 0) - d(Ex, 1)]) / (1j * w * MU0)
        Hc = np.array([H[c][0][0], H[c][1][0], H[c][2][0]])
        worst = max(worst, np.abs(curl - Hc).max() / np.abs(Hc).max())
    return worst


@pytest.mark.parametrize("lx,ly,z0", [(2, 1, 80.0), (0, 3, 80.0), (-2, 4, 20.0)])
def test_receiver_magnetic_tensor_obeys_faraday(oracle, lx, ly, z0):
    """H = curl E / (i w mu0) at receivers below (z0 = 80) and above (z0 = 20,
    source cell [40, 50]) the source cell; all seven H entries and the
    vanishing Hz of a vertical current.  Observed <= 1.2e-6 (the digital
    filter's finite-difference floor); a wrong sign or index is O(1)."""
    f = 1.0
    w = 2 * np.pi * f
    src = [0.0, 50.0] if z0 > 50 else [40.0, 50.0]
    m = model(oracle, src, f=f)
    h = 0.5
    x0, y0 = 13.0, -21.0
    st = np.array([-2, -1, 1, 2]) * h
    P = [(x0, y0, z0)] + [(x0 + s, y0, z0) for s in st] + [(x0, y0 + s, z0) for s in st] + \
        [(x0, y0, z0 + s) for s in st]
    P = np.array(P)
    rc = oracle.Receivers(m, P[:, 0], P[:, 1], P[:, 2])
    B = rc.tensor_element(lx, ly).reshape(15, rc.Nr)
    assert _curl_check(B, w, h) < 1e-5


def test_receiver_inside_anomaly_splits_into_above_and_below(oracle):
    """A receiver on an inner cell border (anomaly cell index 1..Nz-1) sees
    the cells above it as a receiver below a one-cell anomaly and the cells
    below it as a receiver above one."""
    z = [0.0, 50.0, 100.0]
    full = oracle.Receivers(model(oracle, z), [12.0], [-7.0], [50.0])
    assert full.cell[0] == 1
    up = oracle.Receivers(model(oracle, z[:2]), [12.0], [-7.0], [50.0])
    dn = oracle.Receivers(model(oracle, z[1:]), [12.0], [-7.0], [50.0])
    assert up.cell[0] == 2 and dn.cell[0] == 0
    for lx, ly in [(0, 0), (2, -1)]:
        F = full.tensor_element(lx, ly).reshape(15, 2)
        U = up.tensor_element(lx, ly)
        D = dn.tensor_element(lx, ly)
        assert rel(F[:, 0], U) < 1e-14 and rel(F[:, 1], D) < 1e-14


def test_plane_wave_at_receivers(oracle):
    """PlaneWave at receiver depths: its Gauss-Legendre average over each
    anomaly cell equals PlaneWaveIntegral (pinned by the half-space closed
    form), H = curl E / (i w mu0) by finite differences in depth, and the
    field is continuous across the 1 km layer border."""
    f = 0.5
    w = 2 * np.pi * f
    z = np.array([0.0, 300.0, 900.0, 1000.0, 1400.0])
    m = model(oracle, z, Nx=4, Ny=4, f=f)
    gx, gw = np.polynomial.legendre.leggauss(12)
    for pol in (0, 1):
        Eint = m.plane_wave(pol).reshape(3, len(z) - 1, 16)[pol, :, 0]
        for iz in range(len(z) - 1):
            d = (z[iz] + z[iz + 1]) / 2 + gx * (z[iz + 1] - z[iz]) / 2
            F = oracle.plane_wave(m, pol, d)
            assert abs(F[:, pol] @ gw / 2 - Eint[iz]) <= 1e-11 * abs(Eint[iz]), (pol, iz)
        h = 0.25
        d0 = np.array([150.0, 1200.0])
        D = np.concatenate([d0 - 2 * h, d0 - h, d0 + h, d0 + 2 * h, d0])
        F = oracle.plane_wave(m, pol, D).reshape(5, 2, 6)
        dE = (F[0, :, pol] - 8 * F[1, :, pol] + 8 * F[2, :, pol] - F[3, :, pol]) / (12 * h)
        Hc = 4 if pol == 0 else 3  # Hy for Ex, Hx for Ey
        want = dE / (1j * w * MU0) * (1 if pol == 0 else -1)
        assert np.abs(F[4, :, Hc] - want).max() <= 1e-8 * np.abs(want).max()
        others = [c for c in range(6) if c not in (pol, Hc)]
        assert np.all(F[:, :, others] == 0)
        J = oracle.plane_wave(m, pol, [1000.0 - 1e-7, 1000.0 + 1e-7])
        assert rel(J[0, [pol, Hc]], J[1, [pol, Hc]]) < 1e-8


# ---------------------------------------------------- convolution pins ----
@pytest.mark.parametrize("Nx,Ny,Nz,Nr", [(4, 4, 2, 2), (6, 4, 3, 1), (4, 8, 1, 3)])
def test_continuation_fft_equals_direct_sum(oracle, Nx, Ny, Nz, Nr):
    """or_rc_apply (2Nx x 2Ny FFTs, the 17 block products, inverse FFTs and
    1/(4 Nx Ny)) equals the direct sum over source cells of the real-space
    tensor; receivers above, inside and below the anomaly."""
    rng = np.random.default_rng(Nx * 100 + Nz * 10 + Nr)
    z = np.arange(Nz + 1) * 60.0 + 20.0
    m = oracle.Model(csig(BKG3[0], 2.0), BKG3[1], 2 * np.pi * 2.0, Nx, Ny, Nz, 100.0, 80.0, z)
    depths = np.array([0.0, z[-1] + 35.0, z[1] if Nz > 1 else 10.0])[:Nr]
    rc = oracle.Receivers(m, rng.uniform(-40, 40, Nr), rng.uniform(-40, 40, Nr), depths)
    Greal = rc.tensor_realspace(nthreads=4)
    Gspec = rc.tensor_fft(Greal, nthreads=4)
    cs = 10 ** rng.uniform(-3, 0, Nx * Ny * Nz) - 1j * 2 * np.pi * 2.0 * EPS0
    E = rng.standard_normal(3 * Nx * Ny * Nz) + 1j * rng.standard_normal(3 * Nx * Ny * Nz)
    Ea, Ha = rc.apply(Gspec, cs, E, nthreads=4)
    Eb, Hb = rc.apply_bruteforce(Greal, cs, E, nthreads=4)
    assert rel(Ea, Eb) < 1e-13 and rel(Ha, Hb) < 1e-13
    # sigma_a = sigma_b carries no current: no anomalous field
    lay = np.array([m.lnum()[i + 1] for i in range(Nz)])
    cb = csig(BKG3[0], 2.0)[lay]
    E0, H0 = rc.apply(Gspec, np.repeat(cb, Nx * Ny), E, nthreads=4)
    assert np.abs(E0).max() == 0 and np.abs(H0).max() == 0
