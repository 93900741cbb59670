"""Write the golden fixtures of tests/golden/ from the reference's own code.

Run in the build container (needs oracle/_ref/liboracle_ref.so, which
`make -C oracle ref` compiles from the reference's Fortran sources):

    python tests/golden/make_golden.py

Every array is produced by the reference's routines (VolumeBesselTransforms,
Calc_Double_Integral_U, Calc_Integral_U of the receiver continuation,
PrepareRecvs, Calc_APQ, GFGMRES) through oracle/ref_harness.F90;
the inputs are small synthetic cases.  With oracle/_ref/liboracle_ref_quad.so
(`make -C oracle refquad`, the reference's VOLUME_WEIGHT_QUAD build) the
quad-precision weights and the tensor elements built from them are added.
Output: tests/golden/reference_vectors.npz.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import pyoracle as po  # noqa: E402

MU0 = 4e-7 * np.pi
EPS0 = 1.0 / (MU0 * 299792458.0 ** 2)

# (x, y, hx, hy) offsets for the Hankel weights: diagonal cell, on-axis, far
WEIGHT_CASES = np.array([[50.0, 50.0, 100.0, 100.0], [150.0, 250.0, 100.0, 100.0],
                         [3150.0, 50.0, 100.0, 50.0], [50.0, 1250.0, 100.0, 100.0],
                         [12550.0, 9050.0, 100.0, 100.0]])
# backgrounds: (sigma list, thick list, f)
BKGS = [([0.01, 0.1, 1.0], [1000.0, 9000.0], 1.0), ([0.1, 0.01, 10.0], [10000.0, 20000.0], 0.1),
        ([0.05], [], 10.0)]
ZS = [np.linspace(0, 400, 9), np.linspace(900, 1100, 9), np.linspace(0, 1280, 17)]
LAMBDAS = [1e-5, 3e-4, 1e-2, 0.3]
W6 = np.array([0.3, -0.2, 0.5, 1.1, -0.7, 0.9])


TENSOR_OFFSETS = np.array([[0, 0], [1, 0], [0, 3], [5, 7], [31, 31], [12, 2]])
TENSOR_MODELS = [([0.01, 0.1, 1.0], [1000.0, 9000.0], 1.0, np.arange(9) * 50.0, "c0"),
                 ([0.01, 0.1, 1.0], [1000.0, 9000.0], 0.3, 900.0 + np.arange(9) * 25.0, "iface")]


def ref_tensor_element(W, sig, th, f, z, lx, ly, dx=100.0, dy=100.0, R=None):
    """Block of the real-space tensor from the reference's Fortran routines:
    Hankel weights from library W, layered kernel from R (default W)."""
    R = W if R is None else R
    Nz = len(z) - 1
    P = Nz * (Nz + 1) // 2
    cs = np.ascontiguousarray(csig(sig, f).view(np.float64))
    thd = np.array(list(th) + [0.0])
    rmin = np.hypot(dx, dy) / 2
    WT0 = np.zeros(501 * 6)
    lms = np.zeros(501)
    W.ref_vbt_weights(rmin, lx * dx + dx / 2, ly * dy + dy / 2, dx, dy, WT0, lms)
    WT0 = WT0.reshape(6, 501).T  # D0, DXX, DXY, DX, DYY, DY
    WT = np.stack([WT0[:, 1], WT0[:, 2], WT0[:, 4], WT0[:, 3], WT0[:, 5], WT0[:, 0]], 1)
    W0 = np.zeros(6)
    W0[0] = np.abs(WT[:, 0]).max()
    if lx != 0 and ly != 0:
        W0[1:5] = np.abs(WT[:, 1:5]).max()
    elif lx == 0:
        WT[:, 1] = 0
        W0[1] = -W0[0]
        W0[2] = np.abs(WT[:, 2]).max()
        WT[:, 3] = 0
        W0[3] = -W0[0]
        if ly != 0:
            W0[4] = np.abs(WT[:, 4]).max()
        else:
            WT[:, 4] = 0
            W0[4] = -W0[0]
    else:
        WT[:, 1] = 0
        W0[1] = -W0[0]
        W0[2] = np.abs(WT[:, 2]).max()
        W0[3] = np.abs(WT[:, 3]).max()
        WT[:, 4] = 0
        W0[4] = -W0[0]
    W0[5] = np.abs(WT[:, 5]).max()
    G = np.zeros(6 * Nz * Nz, np.complex128)
    for K in range(501):
        if np.max(np.abs(WT[K] / W0)) <= 1e-12:
            continue
        lm = lms[K]
        W = WT[K].copy()
        W[[0, 1, 2]] /= lm
        W[[3, 4, 5]] *= lm
        R.ref_double_integral_u(len(sig), cs, thd, 2 * np.pi * f, Nz, np.ascontiguousarray(z), lm,
                                W, G)
    G = G.reshape(Nz, Nz, 6) / np.pi / 4  # [j][i][comp]
    blk = np.zeros(4 * P + 2 * Nz * Nz, np.complex128)
    for j in range(Nz):
        for i in range(j + 1):
            pi = i + j * (j + 1) // 2
            blk[pi], blk[P + pi], blk[2 * P + pi], blk[3 * P + pi] = G[j, i, 0], G[j, i, 1], G[j, i, 3], G[j, i, 5]
    blk[4 * P:4 * P + Nz * Nz] = G[:, :, 2].reshape(-1)
    blk[4 * P + Nz * Nz:] = G[:, :, 4].reshape(-1)
    return blk


# receiver continuation (calc_rc_tensor_module.F90, rc_kernel_hankel_module.F90)
RC_WEIGHT_CASES = np.array([[-50.0, -50.0, 100.0, 100.0], [150.0, -250.0, 100.0, 100.0],
                            [3150.0, 50.0, 100.0, 50.0], [-1234.5, 20000.0, 312.5, 312.5],
                            [0.3, -50.0, 100.0, 100.0]])
RC_DEPTHS = np.array([0.0, 25.0, 50.0, 200.0, 430.0, 950.0, 1000.0, 1050.0, 1500.0, 12000.0,
                      50.000001, 899.0])
RC_OFFSETS = np.array([[0, 0], [1, 0], [-1, 2], [5, -7], [31, 31], [-12, 2]])
RC_RECVS = {"c0": np.array([[0.0, 0.0, 0.0], [30.0, -40.0, 130.0], [-45.0, 10.0, 600.0]]),
            "iface": np.array([[0.0, 0.0, 0.0], [30.0, -40.0, 950.0], [-45.0, 10.0, 1300.0]])}


def ref_rc_tensor_element(R, sig, th, f, z, lx, ly, recvs, dx=100.0, dy=100.0):
    """Block of the real-space receiver tensor from the reference's Fortran
    routines (single-integral weights and Calc_Integral_U), composed with the
    masking and tap sum of calc_rc_tensor_module.F90:191-224.  Layout
    [entry 0..14][receiver][z]."""
    Nz = len(z) - 1
    Nr = len(recvs)
    cs = np.ascontiguousarray(csig(sig, f).view(np.float64))
    thd = np.array(list(th) + [0.0])
    rmin = np.hypot(dx, dy) / 2
    blk = np.zeros((15, Nr, Nz), np.complex128)
    for r, (xs, ys, depth) in enumerate(recvs):
        WT0 = np.zeros(501 * 6)
        lms = np.zeros(501)
        R.ref_vbt_weights_single(rmin, lx * dx - dx / 2 + xs, ly * dy - dy / 2 + ys, dx, dy, WT0, lms)
        WT0 = WT0.reshape(6, 501).T  # D0, DXX, DXY, DX, DYY, DY
        WT = np.stack([WT0[:, 1], WT0[:, 2], WT0[:, 4], WT0[:, 3], WT0[:, 5], WT0[:, 0]], 1)
        W0 = np.abs(WT).max()
        GE = np.zeros(8 * Nz, np.complex128)
        GH = np.zeros(7 * Nz, np.complex128)
        for K in range(501):
            if np.max(np.abs(WT[K] / W0)) <= 1e-12:
                continue
            lm = lms[K]
            W = WT[K].copy()
            W[[0, 1, 2]] /= lm
            W[[3, 4, 5]] *= lm
            R.ref_rc_integral_u(len(sig), cs, thd, 2 * np.pi * f, Nz, np.ascontiguousarray(z), depth,
                                lm, W, GE, GH)
        blk[:8, r] = GE.reshape(Nz, 8).T / np.pi / 4
        blk[8:, r] = GH.reshape(Nz, 7).T / np.pi / 4
    return blk.reshape(-1)


def csig(sig, f):
    w = 2 * np.pi * f
    return np.array([1e-12 - 1j * w * EPS0] + [s - 1j * w * EPS0 for s in sig])


def main():
    R = po.ref()
    out = {}
    rmin = np.sqrt(100.0 ** 2 + 100.0 ** 2) / 2
    wts = []
    for x, y, hx, hy in WEIGHT_CASES:
        WT = np.zeros(501 * 6)
        lms = np.zeros(501)
        R.ref_vbt_weights(rmin, x, y, hx, hy, WT, lms)
        wts.append(WT.reshape(6, 501).T)
        out["lms"] = lms
    out["weight_cases"] = WEIGHT_CASES
    out["weight_rmin"] = np.array(rmin)
    out["weights"] = np.array(wts)
    k = 0
    for bi, (sig, th, f) in enumerate(BKGS):
        cs = csig(sig, f)
        csd = np.ascontiguousarray(cs.view(np.float64))
        thd = np.array(th + [0.0])
        Nl = len(sig)
        for zi, z in enumerate(ZS):
            Nz = len(z) - 1
            for lm in LAMBDAS:
                G = np.zeros(6 * Nz * Nz, np.complex128)
                R.ref_double_integral_u(Nl, csd, thd, 2 * np.pi * f, Nz, np.ascontiguousarray(z), lm,
                                        W6, G)
                out["U_%d_%d_%g" % (bi, zi, lm)] = G
                k += 1
        Arr = np.zeros(2 * Nl, np.complex128)
        p = np.zeros_like(Arr)
        q = np.zeros_like(Arr)
        eta = np.zeros(Nl, np.complex128)
        A = np.zeros(2 * Nl * Nl, np.complex128)
        R.ref_calc_apq(Nl, csd, thd, 2 * np.pi * f, 0.0, Arr, p, q, eta, A)
        out["apq_%d" % bi] = np.concatenate([Arr, p, q, eta, A])
    # receiver continuation: single-integral weights, PrepareRecvs and the
    # receiver kernel Calc_Integral_U
    wr = []
    for x, y, hx, hy in RC_WEIGHT_CASES:
        WT = np.zeros(501 * 6)
        lms = np.zeros(501)
        R.ref_vbt_weights_single(np.hypot(hx, hy) / 2, x, y, hx, hy, WT, lms)
        wr.append(WT.reshape(6, 501).T)
    out["rc_weight_cases"] = RC_WEIGHT_CASES
    out["rc_weights"] = np.array(wr)
    out["rc_depths"] = RC_DEPTHS
    for bi, (sig, th, f) in enumerate(BKGS):
        cs = csig(sig, f)
        csd = np.ascontiguousarray(cs.view(np.float64))
        thd = np.array(th + [0.0])
        Nl = len(sig)
        for zi, z in enumerate(ZS):
            Nz = len(z) - 1
            nd = len(RC_DEPTHS)
            zr = np.zeros(nd)
            lay = np.zeros(nd, np.int32)
            cell = np.zeros(nd, np.int32)
            R.ref_prepare_recvs(Nl, csd, thd, 2 * np.pi * f, Nz, np.ascontiguousarray(z), nd, RC_DEPTHS,
                                zr, lay, cell)
            out["rc_recv_%d_%d" % (bi, zi)] = np.stack([zr, lay, cell])
            U = np.zeros((nd, len(LAMBDAS), 15, Nz), np.complex128)
            for di, depth in enumerate(RC_DEPTHS):
                for li, lm in enumerate(LAMBDAS):
                    GE = np.zeros(8 * Nz, np.complex128)
                    GH = np.zeros(7 * Nz, np.complex128)
                    R.ref_rc_integral_u(Nl, csd, thd, 2 * np.pi * f, Nz, np.ascontiguousarray(z), depth,
                                        lm, W6, GE, GH)
                    U[di, li, :8] = GE.reshape(Nz, 8).T
                    U[di, li, 8:] = GH.reshape(Nz, 7).T
            out["rcU_%d_%d" % (bi, zi)] = U
    out["rc_offsets"] = RC_OFFSETS
    for sig, th, f, z, tag in TENSOR_MODELS:
        out["rc_recvs_%s" % tag] = RC_RECVS[tag]
        for lx, ly in RC_OFFSETS:
            out["rctensor_%s_%d_%d" % (tag, lx, ly)] = ref_rc_tensor_element(R, sig, th, f, z, lx, ly,
                                                                             RC_RECVS[tag])
    # GFGMRES on dense complex systems
    rng = np.random.Generator(np.random.PCG64(7))
    cases = [(60, 10, 5, 50, 1e-10), (120, 8, 4, 100, 1e-9), (50, 30, 30, 100, 1e-8),
             (80, 5, 3, 3, 1e-12)]
    for ci, (N, mo, mi, maxit, tol) in enumerate(cases):
        Am = np.eye(N) * 3 + (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) / np.sqrt(N)
        b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
        x0 = b.copy()
        xr = np.zeros(N, np.complex128)
        it, st, bea, be, nmv = C.c_int(), C.c_int(), C.c_double(), C.c_double(), C.c_longlong()
        R.ref_fgmres_dense(N, np.ascontiguousarray(Am.T.copy()), b, x0, mo, mi, maxit, tol, xr,
                           C.byref(it), C.byref(st), C.byref(bea), C.byref(be), C.byref(nmv))
        out["gm%d_A" % ci] = Am
        out["gm%d_b" % ci] = b
        out["gm%d_params" % ci] = np.array([N, mo, mi, maxit, tol])
        out["gm%d_x" % ci] = xr
        out["gm%d_info" % ci] = np.array([it.value, st.value, bea.value, be.value, nmv.value])
    # quad-precision weights (VOLUME_WEIGHT_QUAD build) and tensor elements
    # composed from them: the accuracy cross-check of the FP64 weights
    if os.path.exists(po.REF_QUAD_PATH):
        Q = po.ref(quad=True)
        wq = []
        for x, y, hx, hy in WEIGHT_CASES:
            WT = np.zeros(501 * 6)
            lms = np.zeros(501)
            Q.ref_vbt_weights(rmin, x, y, hx, hy, WT, lms)
            wq.append(WT.reshape(6, 501).T)
        out["weights_quad"] = np.array(wq)
        for sig, th, f, z, tag in TENSOR_MODELS:
            for lx, ly in TENSOR_OFFSETS:
                out["tensorq_%s_%d_%d" % (tag, lx, ly)] = ref_tensor_element(Q, sig, th, f, z, lx, ly, R=R)
    # tensor elements composed from the reference's own weights and kernel
    # (COMPUTE_WEIGHTS masking and the tap sum restated here, 20 lines of
    # calc_ie_tensor_module.F90:257-288)
    out["tensor_offsets"] = TENSOR_OFFSETS
    for sig, th, f, z, tag in TENSOR_MODELS:
        for lx, ly in TENSOR_OFFSETS:
            out["tensor_%s_%d_%d" % (tag, lx, ly)] = ref_tensor_element(R, sig, th, f, z, lx, ly)
    path = os.path.join(HERE, "reference_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, len(out), "arrays")


if __name__ == "__main__":
    main()
