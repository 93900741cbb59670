"""The Fortran driver giem2g_amd/_lib/giem2g_amd_host (ISO_C_BINDING over
the C ABI): deck parsing on the CPU, and on the GPU the multi-frequency,
two-polarization solve with warm starts and SOL_P{X,Y}_F*.bin checkpoints
(giem2g.F90:323-507, mpi_saveload_module.F90:568-675) against the Python
API and the oracle.

The deck fixture tests/golden/reference_deck_giem2g_input.json is the
reference's own test deck (test/giem2g_input.json), kept as data; its
conductivity file is not part of the reference, so the GPU test writes its
own deck and conductivity file.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import rel

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HOST = os.path.join(ROOT, "giem2g_amd", "_lib", "giem2g_amd_host")


def _host():
    if not os.path.exists(HOST):
        pytest.skip("Fortran driver not built (no Fortran compiler)")
    return HOST


def test_deck_parse_reference_deck():
    out = subprocess.run([_host(), "--check", os.path.join(HERE, "golden", "reference_deck_giem2g_input.json")],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = {l.split()[0]: l.split()[1:] for l in out.stdout.splitlines() if l.strip()}
    assert lines["Nl"] == ["3"]
    assert np.allclose([float(v) for v in lines["Conductivity"]], [0.1, 0.01, 10.0])
    assert np.allclose([float(v) for v in lines["Thickness"]], [10000, 20000])
    assert lines["Grid"][:3] == ["128", "128", "50"]
    assert np.isclose(float(lines["Grid"][3]), 312.5)
    assert np.allclose([float(v) for v in lines["zborders"]], [0.0, 10000.0])
    assert np.allclose([float(v) for v in lines["Frequencies"]], [1e-3, 1e-2, 1e-1, 1.0])
    assert lines["FGMRES"][:3] == ["15", "15", "300"] and float(lines["FGMRES"][3]) == 1e-8
    assert lines["SigmaFile"] == ["128_128_501.dat"]
    assert lines["SolveEquation"] == ["T"] and lines["RecalculateFields"] == ["T"]
    recv = [[float(v) for v in l.split()[1:]] for l in out.stdout.splitlines() if l.startswith("Reciever")]
    assert np.allclose(recv, [[0, -20000, 0], [0, 20000, 0]])


def test_deck_parse_errors(tmp_path):
    p = tmp_path / "bad.json"
    p.write_text('{"Background": {"Conductivity": [0.01, 0.1], "Thickness": []}}')
    out = subprocess.run([_host(), "--check", str(p)], capture_output=True, text=True, timeout=60)
    assert out.returncode != 0 and "Thickness" in out.stdout


def write_case(d, m, freqs, outer=10, inner=10, receivers=None, solve=True):
    # list-directed text in Fortran (Nz, Nx, Ny) order: z fastest, then x, then y
    sig = np.transpose(m.siga, (1, 2, 0)).reshape(-1)  # siga is (Nz, Ny, Nx), C order
    np.savetxt(os.path.join(d, "sigma.dat"), sig, fmt="%.17e")
    deck = {"Save Solution": True, "Solve Equation": solve, "Recalculate Fields": receivers is not None,
            "Background": {"Conductivity": list(m.sigma_bkg), "Permittivity": [],
                           "Thickness": list(m.thick)},
            "Anomaly": {"Nx": m.Nx, "Ny": m.Ny, "Nz": m.Nz, "dx": m.dx, "dy": m.dy,
                        "zborders": [float(v) for v in m.z], "Conductivity": "sigma.dat"},
            "Frequencies": list(freqs),
            "FGMRES": {"OuterBufferSize": outer, "InnerBufferSize": inner, "MaxRestart": 50,
                       "Tolerance": 1e-8}}
    if receivers is not None:
        deck["Recievers"] = [{"xshift": x, "yshift": y, "depth": z} for x, y, z in receivers]
    with open(os.path.join(d, "deck.json"), "w") as f:
        json.dump(deck, f, indent=1)


def read_sol(path, m):
    a = np.fromfile(path, dtype=np.complex128)
    f4 = a.reshape(m.Ny, m.Nx, 3, m.Nz)  # Fortran (z,c,x,y) as C (y,x,c,z)
    return np.transpose(f4, (2, 3, 0, 1)).reshape(-1)  # -> (c,z,y,x) = Fortran (x,y,z,c)


@pytest.mark.gpu
def test_fortran_host_multifrequency(tmp_path, oracle):
    from giem2g_amd import models, operator as gop
    host = _host()
    m = models.small(16, 16, 4, seed=21, freq=1.0)
    freqs = [0.5, 2.0]
    write_case(str(tmp_path), m, freqs)
    print("host run 1", flush=True)
    r = subprocess.run([host, "deck.json"], cwd=str(tmp_path), capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    runs = [l for l in r.stdout.splitlines() if l.startswith("freq")]
    assert len(runs) == 4 and all(" status   1" in l for l in runs), runs
    sols = {}
    for f in freqs:
        # WRITE (fnum, '(F22.10)') with blanks of the first 11 columns -> '0'
        fnum = "".join("0" if (i < 11 and c == " ") else c for i, c in enumerate("%22.10f" % f)).strip()
        for p in ("PY", "PX"):
            path = tmp_path / ("SOL_%s_F%s.bin" % (p, fnum))
            assert path.exists(), path
            sols[(f, p)] = read_sol(str(path), m)
    # first frequency: cold start, same as the Python API and the oracle
    m.freq = freqs[0]
    op = gop.operator_for(m)
    for p, pol in (("PY", 1), ("PX", 0)):
        E = op.plane_wave(pol)
        Es, info = op.solve(E, outer_m=10, inner_m=10, max_restart=50, tol=1e-8)
        assert rel(sols[(freqs[0], p)], Es) < 1e-12
        om = oracle.Model(m.csigb(), m.thick, m.omega, m.Nx, m.Ny, m.Nz, m.dx, m.dy, m.z)
        Gs = om.tensor_fft(om.tensor_realspace())
        Eo, _ = om.solve(Gs, m.csiga(), E, m_out=10, m_in=10, maxit=50, tol=1e-8)
        assert rel(sols[(freqs[0], p)], Eo) < 1e-6
    op.close()
    # second frequency: warm start from the first frequency's solution
    m.freq = freqs[1]
    op = gop.operator_for(m)
    for p, pol in (("PY", 1), ("PX", 0)):
        E = op.plane_wave(pol)
        Es, info = op.solve(E, E0=sols[(freqs[0], p)], outer_m=10, inner_m=10, max_restart=50, tol=1e-8)
        assert rel(sols[(freqs[1], p)], Es) < 1e-12
    op.close()
    # rerun: every solve starts from its saved checkpoint and converges at once
    print("host run 2", flush=True)
    r2 = subprocess.run([host, "deck.json"], cwd=str(tmp_path), capture_output=True, text=True, timeout=300)
    assert r2.returncode == 0, r2.stdout + r2.stderr
    assert r2.stdout.count("Load IE Solution from") == 4
    its = [int(l.split("outer")[1].split()[0]) for l in r2.stdout.splitlines() if l.startswith("freq")]
    its1 = [int(l.split("outer")[1].split()[0]) for l in runs]
    assert all(a <= 1 for a in its) and sum(its) < sum(its1), (its, its1)


def fnum_of(f):
    """WRITE (fnum, '(F22.10)') with blanks of the first 11 columns -> '0'."""
    return "".join("0" if (i < 11 and c == " ") else c for i, c in enumerate("%22.10f" % f)).strip()


def read_dat(path):
    """SaveOutputOneFile text (mpi_saveload_module.F90:391-483): title line
    and rows of x, y, z, f, then re/im of Ea, Ha, Et, Ht (x, y, z each)."""
    with open(path) as fh:
        title = fh.readline()
        rows = np.array([[float(v) for v in l.split()] for l in fh if l.strip()])
    return title, rows


@pytest.mark.gpu
def test_fortran_host_receivers(tmp_path, oracle):
    """Solve + continuation to receivers (giem2g.F90:412-428, 487-503): the
    P{Y,X}_F*.dat columns against the Python API (same solution, receiver
    operator, normal field) and the oracle's continuation; then a rerun
    with "Solve Equation": false recalculates from the saved solutions and
    writes the same files."""
    from giem2g_amd import models, operator as gop
    host = _host()
    m = models.small(16, 16, 4, seed=33, freq=1.0)
    freqs = [0.7]
    recv = [(0.0, 0.0, 0.0), (30.0, -40.0, 125.0), (-45.0, 10.0, 400.0)]
    write_case(str(tmp_path), m, freqs, receivers=recv)
    print("host run receivers", flush=True)
    r = subprocess.run([host, "deck.json"], cwd=str(tmp_path), capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    m.freq = freqs[0]
    op = gop.operator_for(m)
    rc = gop.ReceiverOperator(op, [v[0] for v in recv], [v[1] for v in recv], [v[2] for v in recv])
    rc.build()
    rc.set_sigma(m.csiga())
    zr, _, _ = rc.receivers()
    om = oracle.Model(m.csigb(), m.thick, m.omega, m.Nx, m.Ny, m.Nz, m.dx, m.dy, m.z)
    ro = oracle.Receivers(om, [v[0] for v in recv], [v[1] for v in recv], [v[2] for v in recv])
    Gs = ro.tensor_fft(ro.tensor_realspace())
    Nr = len(recv)
    texts = {}
    for p, pol in (("PY", 1), ("PX", 0)):
        path = tmp_path / ("%s_F%s.dat" % (p, fnum_of(freqs[0])))
        title, rows = read_dat(str(path))
        texts[p] = path.read_text()
        names = ["%s%s^%s" % (f, a, k) for k in "at" for f in "EH" for a in "xyz"]
        assert title.split() == ["%", "x_recv", "y_recv", "z_recv", "frequency"] + \
            [w for n in names for w in ("Re", n, "Im", n)]
        assert rows.shape == (Nr * m.Nx * m.Ny, 28)
        sol = read_sol(str(tmp_path / ("SOL_%s_F%s.bin" % (p, fnum_of(freqs[0])))), m)
        Ea, Ha = rc.recalculate(sol)
        F0 = rc.plane_wave(pol)
        A = lambda F: F.reshape(m.Ny, m.Nx, 3, Nr)  # Fortran (Nr,3,Nx,Ny) as C (y,x,c,r)  # noqa: E731
        Ea4, Ha4 = A(Ea), A(Ha)
        # rows: r outer, then y, then x
        want = np.zeros((Nr, m.Ny, m.Nx, 28))
        for ir in range(Nr):
            x = (np.arange(1, m.Nx + 1) * m.dx - m.dx / 2 + recv[ir][0])[None, :]
            y = (np.arange(1, m.Ny + 1) * m.dy - m.dy / 2 + recv[ir][1])[:, None]
            want[ir, :, :, 0] = x
            want[ir, :, :, 1] = y
            want[ir, :, :, 2] = zr[ir]
            want[ir, :, :, 3] = freqs[0]
            cols = [Ea4[:, :, k, ir] for k in range(3)] + [Ha4[:, :, k, ir] for k in range(3)] + \
                   [Ea4[:, :, k, ir] + F0[ir, k] for k in range(3)] + [Ha4[:, :, k, ir] + F0[ir, 3 + k] for k in range(3)]
            for j, c in enumerate(cols):
                want[ir, :, :, 4 + 2 * j] = c.real
                want[ir, :, :, 5 + 2 * j] = c.imag
        want = want.reshape(-1, 28)
        assert np.allclose(rows[:, :4], want[:, :4], rtol=1e-10, atol=1e-9)
        for j in range(4, 28, 2):  # printed with 11 significant digits
            got = rows[:, j] + 1j * rows[:, j + 1]
            ref = want[:, j] + 1j * want[:, j + 1]
            assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max(), (p, j)
        Eo, Ho = ro.apply(Gs, m.csiga(), sol)
        assert rel(Ea, Eo) < 1e-9 and rel(Ha, Ho) < 1e-9
    rc.close()
    op.close()
    # recalculation only, from the saved solutions
    write_case(str(tmp_path), m, freqs, receivers=recv, solve=False)
    for p in ("PY", "PX"):
        os.remove(str(tmp_path / ("%s_F%s.dat" % (p, fnum_of(freqs[0])))))
    r2 = subprocess.run([host, "deck.json"], cwd=str(tmp_path), capture_output=True, text=True, timeout=300)
    assert r2.returncode == 0, r2.stdout + r2.stderr
    assert not any(l.startswith("freq") for l in r2.stdout.splitlines())
    for p in ("PY", "PX"):
        assert (tmp_path / ("%s_F%s.dat" % (p, fnum_of(freqs[0])))).read_text() == texts[p]


@pytest.mark.gpu
@pytest.mark.timeout(300)
def test_fortran_host_reference_deck_geometry(tmp_path):
    """The reference's own deck (test/giem2g_input.json): 128 x 128 x 50 cells
    of 312.5 m on its graded z borders, its background, frequencies,
    receivers and FGMRES settings, with a synthetic random-box conductivity
    file in place of the deck's (not part of the reference).  Nz = 50 runs
    the zero-padded matrix-core contraction.  Every solve converges; each
    saved solution of f = 0.1 Hz satisfies the equation with the vector-ALU
    contraction too (restarted from it, the FP64 solver stops at once) and
    equals the Python API's solve warm-started the same way; the receiver
    files hold finite rows for both receivers.  The same deck with complex64
    inner solves gives every field within 1e-6 of the FP64 one."""
    import re
    from giem2g_amd import models, operator as gop
    host = _host()
    with open(os.path.join(HERE, "golden", "reference_deck_giem2g_input.json")) as fh:
        deck = json.loads(re.sub(r",\s*([\]}])", r"\1", fh.read()))  # the deck has trailing commas
    an, bg, fg = deck["Anomaly"], deck["Background"], deck["FGMRES"]
    assert (an["Nx"], an["Ny"], an["Nz"]) == (128, 128, 50)
    rng = np.random.Generator(np.random.PCG64(50))
    siga = models.random_boxes(rng, an["Nx"], an["Ny"], an["Nz"], 12, bg["Conductivity"][0], 0.01, 1.0)
    m = models.Model("deck", an["Nx"], an["Ny"], an["Nz"], an["dx"], an["dy"], np.array(an["zborders"]),
                     list(bg["Conductivity"]), list(bg["Thickness"]), 0.1, siga)
    sig = np.transpose(m.siga, (1, 2, 0)).reshape(-1)
    np.savetxt(str(tmp_path / "sigma.dat"), sig, fmt="%.17e")
    an["Conductivity"] = "sigma.dat"
    with open(str(tmp_path / "deck.json"), "w") as f:
        json.dump(deck, f, indent=1)
    print("host run reference deck", flush=True)
    r = subprocess.run([host, "deck.json"], cwd=str(tmp_path), capture_output=True, text=True, timeout=280)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    runs = [l for l in r.stdout.splitlines() if l.startswith("freq")]
    assert len(runs) == 2 * len(deck["Frequencies"]) and all(" status   1" in l for l in runs), runs
    recv = deck["Recievers"]
    for f in deck["Frequencies"]:
        for p in ("PY", "PX"):
            assert (tmp_path / ("SOL_%s_F%s.bin" % (p, fnum_of(f)))).exists()
            _, rows = read_dat(str(tmp_path / ("%s_F%s.dat" % (p, fnum_of(f)))))
            assert rows.shape == (len(recv) * m.Nx * m.Ny, 28) and np.isfinite(rows).all()
            assert np.abs(rows[:, 4:]).max() > 0
    f = 0.1
    m.freq = f
    op = gop.operator_for(m)
    prev = deck["Frequencies"][deck["Frequencies"].index(f) - 1]
    for p, pol in (("PY", 1), ("PX", 0)):
        sol = read_sol(str(tmp_path / ("SOL_%s_F%s.bin" % (p, fnum_of(f)))), m)
        warm = read_sol(str(tmp_path / ("SOL_%s_F%s.bin" % (p, fnum_of(prev)))), m)
        E = op.plane_wave(pol)
        op.set_contraction(0)
        Es, info = op.solve(E, E0=warm, outer_m=fg["OuterBufferSize"], inner_m=fg["InnerBufferSize"],
                            max_restart=fg["MaxRestart"], tol=fg["Tolerance"])
        assert info["status"] == 1 and rel(Es, sol) < 1e-12, (p, info)
        op.set_contraction(1)
        Ev, iv = op.solve(E, E0=sol, outer_m=fg["OuterBufferSize"], inner_m=fg["InnerBufferSize"],
                          max_restart=fg["MaxRestart"], tol=fg["Tolerance"])
        print(p, "Python solve", info, "restart on the vector ALU", iv, flush=True)
        assert iv["status"] == 1 and iv["outer_iterations"] <= 1 and iv["true_residual"] < fg["Tolerance"], iv
        assert rel(Ev, sol) < 1e-7
    op.close()
    # the same deck with complex64 inner solves (Solver.InnerPrecision = 1):
    # every solve converges in FP64 terms and its field is within 1e-6 of the
    # FP64 solve's
    mixed = tmp_path / "mixed"
    mixed.mkdir()
    deck["Solver"] = {"InnerPrecision": 1}
    deck["Recalculate Fields"] = False
    an["Conductivity"] = "../sigma.dat"
    with open(str(mixed / "deck.json"), "w") as f:
        json.dump(deck, f, indent=1)
    print("host run reference deck, complex64 inner", flush=True)
    r = subprocess.run([host, "deck.json"], cwd=str(mixed), capture_output=True, text=True, timeout=200)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    mruns = [l for l in r.stdout.splitlines() if l.startswith("freq")]
    assert len(mruns) == len(runs) and all(" status   1" in l for l in mruns), mruns
    assert all(int(l.split("mv32")[1].split()[0]) > 0 for l in mruns)
    t64 = sum(float(l.split()[-2]) for l in runs)
    t32 = sum(float(l.split()[-2]) for l in mruns)
    print("reference deck: FP64 solves %.3f s, complex64 inner %.3f s" % (t64, t32), flush=True)
    for f in deck["Frequencies"]:
        for p in ("PY", "PX"):
            a = read_sol(str(tmp_path / ("SOL_%s_F%s.bin" % (p, fnum_of(f)))), m)
            b = read_sol(str(mixed / ("SOL_%s_F%s.bin" % (p, fnum_of(f)))), m)
            assert rel(b, a) < 1e-6, (f, p, rel(b, a))
