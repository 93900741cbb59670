"""Benchmark of the MI355X IE solve (one GPU).

Headline workload: the time to a 1e-8 FP64 true residual on the 256x256x32
random-box model (25 m layers, 20 boxes of 1e-3..10 S/m in a 0.01 S/m host,
background 0.01/0.1/1 S/m, f = 0.1 Hz, EY polarization), solved by the
nested FGMRES(30) with a complex64 inner solver and the learned right
preconditioner, against the plain FP64 solve.  One step is one solve from
the plane-wave right-hand side, with the tensor and the conductivity
resident in HBM.  `value` is solves per second of the mixed+learned solve
over the K timed steps; the line also carries the FP64 and the mixed
(no network) solves timed the same way, their matvecs per second, iteration
counts and fractions of the HBM roofline, the roofline entry of the
dominant kernel (the complex64 tensor contraction), and the oracle's CPU
solve on the host cores.

Other records on the same line: the C1 FP64 operator apply (128x128x32,
matvecs/s, with its CPU baseline), eight right-hand sides through one
tensor read, the continuation to the reference deck's receivers, both
polarizations of the C2 model (one after the other and paired), and the
128x128x64 multi-frequency paired solves (16 frequencies, 1e-3..1e3 Hz).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-c3]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0     # MI355X HBM3E, GB/s
FP32_MFMA_TFLOPS = 157.3  # MI355X dense FP32 matrix peak
FP64_MFMA_TFLOPS = 78.6
CLOCK_WARMUP_S = 2.0      # applies run before any timed region (clocks settle)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def mirror_active(m, s):
    """The one-vector contraction reads one tensor block per mirrored pair of
    quads: a square grid with dx = dy, the library's own tensor, and the
    small-block kernel (Nz 8-32 in FP64, 24-32 in complex64)."""
    lo = 8 if s == 16 else 24
    return m.Nx == m.Ny and m.dx == m.dy and lo <= m.Nz <= 32


def algorithmic_bytes(Nx, Ny, Nz, s=16, mirror=False):
    """Bytes per matvec by kernel: n_c cells, n_v = 3 n_c, n_s = 12 Nx Ny Nz
    spectrum, n_G = Nx Ny 2Nz(2Nz+1) tensor entries; s = element size.  With
    the x-y mirror the contraction reads the Nx(Nx+1)/2 blocks with kx <= ky."""
    nc = Nx * Ny * Nz
    nv = 3 * nc
    ns = 4 * Nx * Ny * 3 * Nz
    nG = (Nx * (Nx + 1) // 2 if mirror else Nx * Ny) * 2 * Nz * (2 * Nz + 1)
    return {"fwd": s * (nv + nc + ns), "contraction": s * (nG + 2 * ns),
            "bwd": s * (ns + 2 * nv + nc), "total": s * (nG + 4 * ns + 3 * nv + 2 * nc)}


def read_traffic(path):
    """HBM bytes per contraction launch measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy-kernel calibration)."""
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch"):
                return float(line.split("=")[1])
    except OSError:
        pass
    return None


def host_info():
    """The host CPU the baseline runs on: model, logical CPUs, the CPUs this
    process may run on (pinning) and OMP_NUM_THREADS."""
    model = "?"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        allowed = sorted(os.sched_getaffinity(0))
    except AttributeError:
        allowed = list(range(os.cpu_count() or 1))
    return {"cpu_model": model, "nproc": os.cpu_count(), "allowed_cpus": len(allowed),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"),
            "pinning": "none (process affinity %d CPUs, threads not bound)" % len(allowed)}


def nthreads_default():
    return int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
                              else os.cpu_count() or 1))


# learned preconditioner in the benchmark: mode 2 (network step, fallback to
# inner GMRES) accepted when it cuts the residual by at least 10%
LEARNED_MODE = 2
LEARNED_ACCEPT = 0.9
SOLVER = dict(outer_m=30, inner_m=30, max_restart=100, tol=1e-8)


def clock_warmup(torch, op, v, out, seconds=CLOCK_WARMUP_S):
    """FP64 applies for `seconds` of wall time, outside every timed region."""
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < seconds:
        for _ in range(10):
            op.apply_dev(v, out)
        torch.cuda.synchronize()
        n += 10
    return n


def solve_record(m, dt, steps, info, c0, c1):
    """Per-solve figures from K timed solves and the operator counters before
    and after them: time, matvecs/s, and the fraction of the HBM roofline
    that the algorithmic bytes of its matvecs and Krylov kernels give."""
    b64 = algorithmic_bytes(m.Nx, m.Ny, m.Nz, 16, mirror_active(m, 16))["total"]
    b32 = algorithmic_bytes(m.Nx, m.Ny, m.Nz, 8, mirror_active(m, 8))["total"]
    mv64 = (c1["matvecs"] - c0["matvecs"]) / steps
    mv32 = (c1["matvecs32"] - c0["matvecs32"]) / steps
    kb = (c1["krylov_bytes"] - c0["krylov_bytes"]) / steps
    t = dt / steps
    byt = mv64 * b64 + mv32 * b32 + kb
    return {"time_to_1e-8_s": t, "solves_per_s": 1.0 / t, "outer_iterations": info["outer_iterations"],
            "matvecs_fp64": info["matvecs_fp64"], "matvecs_fp32": info["matvecs_fp32"],
            "matvecs_per_s": (mv64 + mv32) / t,
            "precond_calls": info["precond_calls"], "precond_fallbacks": info["precond_fallbacks"],
            "true_residual": info["true_residual"], "status": info["status"],
            "gs_passes": (c1["gs_passes"] - c0["gs_passes"]) / steps,
            "dot_products": (c1["dotprods"] - c0["dotprods"]) / steps,
            "algorithmic_bytes": {"matvec_fp64": mv64 * b64, "matvec_c64": mv32 * b32, "krylov": kb},
            "achieved_gbs": byt / t / 1e9, "roofline_frac": byt / t / 1e9 / HBM_PEAK_GBS,
            "kernel_s": {"fwd_fft": (c1["t_fwd"] - c0["t_fwd"]) / steps,
                         "contraction": (c1["t_contr"] - c0["t_contr"]) / steps,
                         "bwd_fft": (c1["t_bwd"] - c0["t_bwd"]) / steps,
                         "gram_schmidt": (c1["t_dot"] - c0["t_dot"]) / steps}}


def c2_headline(torch, gop, models, steps, warmup, clock_s=CLOCK_WARMUP_S):
    """The metric: K timed solves each of plain FP64, mixed (complex64 inner)
    and mixed+learned on the C2 model, EY polarization, profiling on (HIP
    events around every apply stage and Gram-Schmidt pass)."""
    from giem2g_amd import learned
    m = models.c2()
    t0 = time.perf_counter()
    op = gop.operator_for(m)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    st = op.stats()
    E = torch.from_numpy(op.plane_wave(1)).cuda()
    v = torch.randn(op.nvec, dtype=torch.complex128, device="cuda")
    out = torch.empty_like(v)
    nwarm = clock_warmup(torch, op, v, out, clock_s)
    del v, out
    runs = [("fp64", 0, 0), ("mixed", 1, 0)]
    net = None
    if os.path.exists(learned.DEFAULT_WEIGHTS):
        net = learned.UNet.load(learned.DEFAULT_WEIGHTS)
        op.set_learned(net, accept=LEARNED_ACCEPT)
        runs.append(("mixed+learned", 1, LEARNED_MODE))
    res = {"model": "256x256x32 x 25 m, 20 random boxes 1e-3..10 S/m in 0.01 S/m host, f=0.1 Hz, EY",
           "tensor_build_s": st["t_tensor"], "tensor_fft_s": st["t_tensor_fft"], "setup_s": t_setup,
           "clock_warmup_applies": nwarm}
    sols = {}
    for name, prec, mode in runs:
        for _ in range(warmup):
            op.solve(E, inner_precision=prec, precond=mode, **SOLVER)
    op.set_profiling(True)
    for name, prec, mode in runs:
        torch.cuda.synchronize()
        c0 = op.counters()
        t0 = time.perf_counter()
        for _ in range(steps):
            Es, info = op.solve(E, inner_precision=prec, precond=mode, **SOLVER)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        c1 = op.counters()
        sols[name] = Es
        rec = solve_record(m, dt, steps, info, c0, c1)
        rec["total_timed_s"] = dt
        for p, key in ((0, "fp64"), (1, "c64")):
            n = c1["contr_launches_" + key] - c0["contr_launches_" + key]
            if n > 0:
                rec["contraction_" + key] = {"launches": n,
                                             "avg_launch_ms": 1e3 * (c1["t_contr_" + key] - c0["t_contr_" + key]) / n,
                                             "rhs_per_launch": (c1["contr_rhs_" + key] - c0["contr_rhs_" + key]) / n}
        res[name] = rec
        log("solve", name, json.dumps(rec))
    op.set_profiling(False)
    for name in sols:
        if name != "fp64":
            d = torch.linalg.vector_norm(sols[name] - sols["fp64"]) / torch.linalg.vector_norm(sols["fp64"])
            res[name]["vs_fp64_rel"] = float(d)
            res[name]["speedup_vs_fp64"] = res["fp64"]["time_to_1e-8_s"] / res[name]["time_to_1e-8_s"]
    if net is not None:
        res["learned"] = {"weights": os.path.relpath(learned.DEFAULT_WEIGHTS, ROOT), "mode": LEARNED_MODE,
                          "accept": LEARNED_ACCEPT}
        op.set_learned(None)
        net.close()
    res["both_polarizations"] = both_polarizations(torch, op, E, res)
    return m, op, E, res


def both_polarizations(torch, op, Ey, res):
    """EX alone (with the EY solve above: one after the other), then EY and
    EX together with paired operator products (one tensor read per pair)."""
    Ex = torch.from_numpy(op.plane_wave(0)).cuda()
    both = {}
    for name, prec in (("fp64", 0), ("mixed", 1)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, ix = op.solve(Ex, inner_precision=prec, **SOLVER)
        torch.cuda.synchronize()
        t_x = time.perf_counter() - t0
        t0 = time.perf_counter()
        _, _, jy, jx, cnt = op.solve_pair(Ey, Ex, inner_precision=prec, **SOLVER)
        torch.cuda.synchronize()
        t_p = time.perf_counter() - t0
        both[name] = {"one_after_other_s": res[name]["time_to_1e-8_s"] + t_x, "paired_s": t_p,
                      "paired_applies": int(cnt[0]), "single_applies": int(cnt[1]),
                      "true_residual": max(jy["true_residual"], jx["true_residual"], ix["true_residual"])}
        log("solve both polarizations", name, both[name])
    both["mixed"]["speedup_vs_fp64_paired"] = both["fp64"]["paired_s"] / both["mixed"]["paired_s"]
    return both


def cpu_solve_baseline(torch, op, m, E, fp64_matvecs, nthreads):
    """The oracle's nested FGMRES (C restatement of the reference algorithm,
    OpenMP) on the C2 model, capped: tolerance 1e-2 stops it after its first
    outer iteration; its time per matvec is extrapolated to the matvec count
    of the full 1e-8 FP64 solve (the same count as the oracle's, by the
    parity tests).  The GPU FP64 solve with the same cap is compared with
    it.  Also: the oracle's FFT passes against torch's CPU FFT (MKL /
    pocketfft) of the same zero-padded 2Nx x 2Ny x 3Nz spectrum."""
    from oracle import pyoracle as po
    Gs = op.get_tensor()
    om = po.Model(m.csigb(), m.thick, m.omega, m.Nx, m.Ny, m.Nz, m.dx, m.dy, m.z)
    Eb = E.cpu().numpy()
    t0 = time.perf_counter()
    Eo, oi = om.solve(Gs, m.csiga(), Eb, m_out=30, m_in=30, maxit=100, tol=1e-2, nthreads=nthreads)
    t_cap = time.perf_counter() - t0
    split = om.apply_timings()
    Eg, ig = op.solve(E, outer_m=30, inner_m=30, max_restart=100, tol=1e-2)
    rel = float(np.linalg.norm(Eg.cpu().numpy() - Eo) / np.linalg.norm(Eo))
    t_full = t_cap * fp64_matvecs / oi.matvecs
    # the FFT cross-check: forward transforms of one vector
    torch.set_num_threads(nthreads)
    x = torch.zeros(3 * m.Nz, 2 * m.Ny, 2 * m.Nx, dtype=torch.complex128)
    x[:, :m.Ny, :m.Nx] = torch.randn(3 * m.Nz, m.Ny, m.Nx, dtype=torch.complex128)
    torch.fft.fft2(x)
    t0 = time.perf_counter()
    for _ in range(3):
        torch.fft.fft2(x)
    t_torch = (time.perf_counter() - t0) / 3
    del x
    fft_best = min(split[0], t_torch)
    return {"value": 1.0 / t_full, "unit": "solves/s", "cores": nthreads, "kind": "port",
            "sample": "oracle/giem2g_oracle.c nested FGMRES(30,30) on the C2 model (EY), OpenMP, stopped at "
                      "tolerance 1e-2 after %d outer iteration(s) and %d matvecs in %.2f s; extrapolated per "
                      "matvec to the %d matvecs of the full 1e-8 solve (%.1f s)"
                      % (oi.iterations, oi.matvecs, t_cap, fp64_matvecs, t_full),
            "capped": {"seconds": t_cap, "matvecs": oi.matvecs, "outer_iterations": oi.iterations,
                       "gpu_fp64_same_cap_rel": rel, "gpu_outer_iterations": ig["outer_iterations"],
                       "gpu_matvecs": ig["matvecs_fp64"]},
            "extrapolated_full_solve_s": t_full,
            "matvec_split_s": {"fwd_fft": split[0], "contraction": split[1], "bwd_fft": split[2]},
            "fft_cross_check": {"oracle_fwd_s": split[0], "torch_cpu_fft2_s": t_torch, "faster": "oracle" if split[0] <= t_torch else "torch",
                                "matvec_s_with_faster_fft": fft_best * (1 + split[2] / max(split[0], 1e-12)) + split[1]},
            "host": host_info()}


def c1_apply(torch, gop, models, steps, cpu):
    """C1: the FP64 operator apply, 128x128x32, f = 10 Hz, complex128
    currents N(0,1): matvecs/s, per-kernel times and rates, the contraction's
    roofline, and the oracle's operator on the host cores."""
    m = models.c1()
    op = gop.operator_for(m)
    v = torch.from_numpy(models.c1_currents(m)).cuda()
    out = torch.empty_like(v)
    clock_warmup(torch, op, v, out, 1.0)
    t0 = time.perf_counter()
    for _ in range(steps):
        op.apply_dev(v, out)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    stage_ms = op.time_apply(v, out, reps=max(10, min(100, steps)))
    v32 = v.to(torch.complex64)
    o32 = torch.empty_like(v32)
    stage32 = op.time_apply(v32, o32, reps=max(10, min(100, steps)), precision=1)
    b64 = algorithmic_bytes(m.Nx, m.Ny, m.Nz, 16, mirror_active(m, 16))
    b32 = algorithmic_bytes(m.Nx, m.Ny, m.Nz, 8, mirror_active(m, 8))
    ach = b64["contraction"] / (stage_ms[1] * 1e-3) / 1e9
    rec = {"workload": "C1 IE operator apply, FP64: 128x128x32 cells of 100x100x25 m, f=10 Hz",
           "matvec_s": steps / dt, "ms_per_matvec": 1e3 * dt / steps, "steps": steps,
           "kernels_ms": {"fwd_fft": float(stage_ms[0]), "contraction": float(stage_ms[1]),
                          "bwd_fft": float(stage_ms[2])},
           "kernels_gbs": {k: b64[k2] / (t * 1e-3) / 1e9 for k, k2, t in
                           (("fwd_fft", "fwd", stage_ms[0]), ("contraction", "contraction", stage_ms[1]),
                            ("bwd_fft", "bwd", stage_ms[2]))},
           "contraction_roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                    "frac": ach / HBM_PEAK_GBS,
                                    "traffic": read_traffic(os.path.join(ROOT, "profiles",
                                                                         "contraction_c1_fp64_traffic.txt")),
                                    "algorithmic_bytes_per_launch": b64["contraction"]},
           "matvec_total_gbs": b64["total"] / (dt / steps) / 1e9,
           "roofline_bound_matvec_s": HBM_PEAK_GBS * 1e9 / b64["total"],
           "complex64": {"kernels_ms": [float(x) for x in stage32], "matvec_s": 1e3 / float(np.sum(stage32)),
                         "contraction_gbs": b32["contraction"] / (stage32[1] * 1e-3) / 1e9}}
    if cpu:
        from oracle import pyoracle as po
        nth = nthreads_default()
        Gs = op.get_tensor()
        om = po.Model(m.csigb(), m.thick, m.omega, m.Nx, m.Ny, m.Nz, m.dx, m.dy, m.z)
        vh = models.c1_currents(m)
        t0 = time.perf_counter()
        reps = 20
        for _ in range(reps):
            ref = om.apply(Gs, m.csiga(), vh, nth)
        dtc = (time.perf_counter() - t0) / reps
        split = om.apply_timings()
        rec["cpu_baseline"] = {"value": 1.0 / dtc, "unit": "matvec/s", "cores": nth, "kind": "port",
                               "sample": "%d FP64 matvecs of the 128x128x32 operator (oracle, OpenMP), split "
                                         "fwd/contraction/bwd = %.3f/%.3f/%.3f s" % (reps, *split),
                               "gpu_vs_cpu_rel": float(np.linalg.norm(out.cpu().numpy() - ref) / np.linalg.norm(ref))}
    return m, op, v, stage_ms, stage32, rec


def batch_bench(torch, op, m, stage64, stage32, nrhs=8, reps=20):
    """Eight right-hand sides through giem2g_op_apply_batch_dev on the C1
    grid (complex64: one tensor read per contraction for all eight; FP64: two
    reads of four): time per batch and matvecs/s against one-vector applies."""
    out = {"nrhs": nrhs}
    g = torch.Generator(device="cuda").manual_seed(5)
    for name, dt, stage, reads in (("complex64", torch.complex64, stage32, 1), ("fp64", torch.complex128, stage64, 2)):
        vin = torch.randn(nrhs, 3 * m.ncell, dtype=dt, device="cuda", generator=g)
        vout = torch.empty_like(vin)
        for _ in range(3):
            op.apply_batch_dev(vin, vout)
        op.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            op.apply_batch_dev(vin, vout)
        op.sync()
        ms = 1e3 * (time.perf_counter() - t0) / reps
        # 9 sub-block uses x Nz/16 row tiles x Nz/4 layer steps x 2 MFMA x 2048
        # flop per (kx,ky) block, one 16-column operand per two right-hand sides
        flop = 9 * (m.Nz // 16) * (m.Nz // 4) * 2 * 2048 * (nrhs // 2) * m.Nx * m.Ny
        rec = {"ms_per_batch": ms, "matvec_s": nrhs / (ms * 1e-3),
               "single_matvec_s": 1e3 / float(stage[0] + stage[1] + stage[2]),
               "tensor_reads_per_batch": reads, "mfma_flop_per_batch": flop}
        log("batch", name, rec)
        out[name] = rec
    return out


# receivers of the reference's own deck (test/giem2g_input.json:32-35)
REF_DECK_RECEIVERS = [(0.0, -20000.0, 0.0), (0.0, 20000.0, 0.0)]


def recalc_bytes(Nx, Ny, Nz, Nr, s=16):
    """Algorithmic bytes of one continuation to Nr receivers."""
    nc = Nx * Ny * Nz
    npt = 4 * Nx * Ny
    ns = npt * 3 * Nz
    no = npt * 6 * Nr
    return {"fwd": s * (3 * nc + nc + ns), "contraction": s * (npt * 15 * Nr * Nz + ns + no),
            "inv": s * (no + no // 2 + no // 2 + Nx * Ny * 6 * Nr)}


def recalc_bench(torch, gop, op, m, E, cpu=True, steps=20):
    """Continuation of a C1 field to the reference deck's receivers: tensor
    build + spectrum per frequency, then recalculations from HBM."""
    xs, ys, zs = zip(*REF_DECK_RECEIVERS)
    rc = gop.ReceiverOperator(op, xs, ys, zs)
    rc.build()
    rc.set_sigma(m.csiga())
    Ea = torch.empty(rc.nfield, dtype=torch.complex128, device="cuda")
    Ha = torch.empty_like(Ea)
    for _ in range(3):
        rc.recalculate_dev(E, Ea, Ha)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    parts = np.zeros(4)
    for _ in range(steps):
        rc.recalculate_dev(E, Ea, Ha)
        st = rc.stats()
        parts += [st["ms_last"], st["ms_fwd"], st["ms_contract"], st["ms_inv"]]
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    parts /= steps
    b = recalc_bytes(m.Nx, m.Ny, m.Nz, rc.Nr)
    st = rc.stats()
    row = {"receivers": REF_DECK_RECEIVERS, "tensor_build_s": st["t_tensor"], "tensor_fft_s": st["t_tensor_fft"],
           "ms_per_recalc": 1e3 * dt, "kernels_ms": {"fwd_fft": parts[1], "contraction": parts[2],
                                                       "inv_fft": parts[3]},
           "contraction_gbs": b["contraction"] / (parts[2] * 1e-3) / 1e9,
           "contraction_frac": b["contraction"] / (parts[2] * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "recalc_gbs": sum(b.values()) / (parts[0] * 1e-3) / 1e9,
           "algorithmic_bytes": b, "tensor_bytes": 4 * m.Nx * m.Ny * 15 * rc.Nr * m.Nz * 16}
    if cpu:
        from oracle import pyoracle as po
        om = po.Model(m.csigb(), m.thick, m.omega, m.Nx, m.Ny, m.Nz, m.dx, m.dy, m.z)
        ro = po.Receivers(om, xs, ys, zs)
        Gs = rc.get_tensor()
        En = E.cpu().numpy()
        nth = nthreads_default()
        t0 = time.perf_counter()
        Eo, Ho = ro.apply(Gs, m.csiga(), En, nthreads=nth)
        tc = time.perf_counter() - t0
        rel = max(float(np.linalg.norm(Ea.cpu().numpy() - Eo) / np.linalg.norm(Eo)),
                  float(np.linalg.norm(Ha.cpu().numpy() - Ho) / np.linalg.norm(Ho)))
        row["cpu_baseline"] = {"recalc_s": tc, "cores": nth, "kind": "port", "gpu_vs_cpu_rel": rel}
    rc.close()
    log("recalc", json.dumps(row))
    return row


def c3_bench(torch, gop, models, nfreq=16, warm=True):
    """Multi-frequency run of the 128x128x64 random-box model: 16 frequencies
    log-spaced over 1e-3..1e3 Hz x 2 polarizations, each frequency with its
    own FP64 tensor; both polarizations of a frequency are solved together
    with paired operator products on the matrix cores.  Per frequency: the
    FP64 outer residual of each solve; overall: the paired complex64
    contraction's matrix-core rate (HIP events, profiling on).  warm: each
    frequency starts from the previous frequency's fields, as the reference
    driver does (giem2g.F90:366-373, 450-457)."""
    m, freqs = models.c3()
    freqs = freqs[:nfreq]
    op = gop.IEOperator.from_model(m)
    rows = []
    t_all = time.perf_counter()
    op.set_profiling(True)
    c0 = None
    prev = (None, None)
    for f in freqs:
        m.freq = float(f)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op.set_background(m.csigb(), m.thick, m.omega)
        op.build_tensor()
        op.set_sigma(m.csiga())
        torch.cuda.synchronize()
        t_build = time.perf_counter() - t0
        Ey = torch.from_numpy(op.plane_wave(1)).cuda()
        Ex = torch.from_numpy(op.plane_wave(0)).cuda()
        if c0 is None:
            c0 = op.counters()
        t0 = time.perf_counter()
        oy, ox, iy, ix, cnt = op.solve_pair(Ey, Ex, E00=prev[0], E01=prev[1], inner_precision=1, **SOLVER)
        t_solve = time.perf_counter() - t0
        if warm:
            prev = (oy, ox)
        row = {"freq": float(f), "build_s": t_build, "solve_pair_s": t_solve,
               "EY": {"outer": iy["outer_iterations"], "true_residual": iy["true_residual"],
                      "mv64": iy["matvecs_fp64"], "mv32": iy["matvecs_fp32"]},
               "EX": {"outer": ix["outer_iterations"], "true_residual": ix["true_residual"],
                      "mv64": ix["matvecs_fp64"], "mv32": ix["matvecs_fp32"]},
               "paired_applies": cnt[0], "single_applies": cnt[1]}
        log("c3", json.dumps(row))
        rows.append(row)
    c1 = op.counters()
    op.set_profiling(False)
    op.close()
    n32 = c1["contr_launches_c64"] - c0["contr_launches_c64"]
    rhs32 = c1["contr_rhs_c64"] - c0["contr_rhs_c64"]
    t32 = c1["t_contr_c64"] - c0["t_contr_c64"]
    # matrix-core work of one contraction of one or two right-hand sides: 9
    # sub-block uses x Nz/16 row tiles x Nz/4 layer steps x 2 MFMA x 2048 flop
    # per (kx,ky) block (one 16-column operand; a single vector repeats its
    # columns, so the work executed is the same)
    flop2 = 9 * (m.Nz // 16) * (m.Nz // 4) * 2 * 2048 * m.Nx * m.Ny
    mfma = None
    if n32 > 0:
        tf = flop2 * n32 / t32 / 1e12
        mfma = {"c64_contraction_launches": n32, "rhs_per_launch": rhs32 / n32,
                "avg_launch_ms": 1e3 * t32 / n32, "tflops": tf, "frac_fp32_mfma_peak": tf / FP32_MFMA_TFLOPS}
    return {"model": "128x128x64 x 20 m, 20 random boxes, %d frequencies 1e-3..1e3 Hz x EY, EX" % len(freqs),
            "total_s": time.perf_counter() - t_all, "solves": 2 * len(freqs),
            "max_true_residual": max(max(r["EY"]["true_residual"], r["EX"]["true_residual"]) for r in rows),
            "mfma_utilisation": mfma, "per_frequency": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed C2 solves of each kind")
    ap.add_argument("--warmup", type=int, default=2, help="untimed C2 solves of each kind")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c3", action="store_true", help="skip the 128x128x64 multi-frequency run")
    ap.add_argument("--c3-freqs", type=int, default=16, help="number of C3 frequencies")
    ap.add_argument("--c3-cold", action="store_true",
                    help="C3 frequencies without the warm start from the previous frequency")
    ap.add_argument("--headline-only", action="store_true", help="only the C2 solves (profiling runs)")
    ap.add_argument("--c3-only", action="store_true",
                    help="only the 128x128x64 multi-frequency paired solves (profiling runs)")
    ap.add_argument("--no-solve", action="store_true",
                    help="skip the C2 solves: only the C1 apply, continuation and batched records "
                         "(profiling runs of those kernels)")
    ap.add_argument("--clock-warmup", type=float, default=CLOCK_WARMUP_S,
                    help="seconds of FP64 applies before the timed C2 solves (0 for counter runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 1 GiB device copy (HBM counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path (the operator "
              "would shard by kx slabs with one all-to-all per FFT direction)", file=sys.stderr)
        return 2
    import torch
    from giem2g_amd import models, operator as gop
    assert torch.cuda.is_available(), "no GPU"
    if a.copy_calibration:
        src = torch.empty(1 << 27, dtype=torch.complex64, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)  # reads and writes exactly 1 GiB
        torch.cuda.synchronize()
        del src, dst
    if a.c3_only:
        line = {"metric": "C3 multi-frequency paired solves (profiling run)",
                "c3": c3_bench(torch, gop, models, nfreq=a.c3_freqs, warm=not a.c3_cold)}
        print(json.dumps(line), flush=True)
        return 0
    if a.no_solve:
        line = {"metric": "IE operator apply (profiling run, no solves)"}
        m1, op1, v1, st64, st32, c1rec = c1_apply(torch, gop, models, 200, not a.no_cpu)
        line["c1_apply"] = c1rec
        line["recalc"] = recalc_bench(torch, gop, op1, m1, v1, cpu=not a.no_cpu)
        line["batched"] = batch_bench(torch, op1, m1, st64, st32)
        op1.close()
        print(json.dumps(line), flush=True)
        return 0
    m2, op2, E2, head = c2_headline(torch, gop, models, a.steps, a.warmup, a.clock_warmup)
    top = head.get("mixed+learned", head["mixed"])
    top_name = "mixed+learned" if "mixed+learned" in head else "mixed"
    b32 = algorithmic_bytes(m2.Nx, m2.Ny, m2.Nz, 8, mirror_active(m2, 8))
    cc = top.get("contraction_c64", {})
    ach = b32["contraction"] / (cc["avg_launch_ms"] * 1e-3) / 1e9 if cc else None
    line = {
        "metric": "time-to-1e-8 residual + IE matvecs/s: mixed+learned vs plain FP64, 1 GPU",
        "value": top["solves_per_s"], "unit": "solves/s (C2 EY to a 1e-8 FP64 true residual)",
        "speedup_vs_fp64": top.get("speedup_vs_fp64"),
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * top["time_to_1e-8_s"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64 inner / f64 outer",
        "data": "synthetic (seeded PCG64 random-box model; trained network weights committed)",
        "config": {"workload": "C2 %s solve to 1e-8: 256x256x32 cells of 100x100x25 m, 20 random boxes "
                               "1e-3..10 S/m in 0.01 S/m, background 0.01/0.1/1 S/m, f=0.1 Hz, EY; nested "
                               "FGMRES(30) outer FP64, inner GMRES(30) complex64 stopping at the relaxed "
                               "tolerance tol/r (r: the outer's current relative residual); the FP64 solve "
                               "keeps the reference's shared tolerance" % top_name,
                   "grid": [m2.Nx, m2.Ny, m2.Nz], "solver": SOLVER, "headline_solve": top_name},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS if ach else None,
                     "traffic": read_traffic(os.path.join(ROOT, "profiles", "contraction_c2_c64_traffic.txt")),
                     "kernel": "k_contract1<float,32,false,4,true> (complex64 tensor contraction of the "
                               "inner solver, small-block matrix cores, x-y mirror: one tensor read per "
                               "(kx,ky) and (ky,kx) pair of quads)",
                     "algorithmic_bytes_per_launch": b32["contraction"],
                     "avg_launch_ms": cc.get("avg_launch_ms"), "launches": cc.get("launches")},
        "solve": head,
    }
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_solve_baseline(torch, op2, m2, E2, head["fp64"]["matvecs_fp64"],
                                                  nthreads_default())
        log("cpu baseline", json.dumps(line["cpu_baseline"]))
    else:
        line["cpu_baseline"] = None
    op2.close()
    del E2
    torch.cuda.empty_cache()
    if not a.headline_only:
        m1, op1, v1, st64, st32, c1rec = c1_apply(torch, gop, models, 200, not a.no_cpu)
        line["c1_apply"] = c1rec
        line["recalc"] = recalc_bench(torch, gop, op1, m1, v1, cpu=not a.no_cpu)
        line["batched"] = batch_bench(torch, op1, m1, st64, st32)
        op1.close()
        if not a.no_c3:
            line["c3"] = c3_bench(torch, gop, models, nfreq=a.c3_freqs, warm=not a.c3_cold)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
