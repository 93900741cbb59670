"""Benchmark of the MI355X IE solve (one GPU).

Headline workload (C1 of the benchmark set): the FP64 integral-equation
operator apply A*v on a 128x128x32 grid (100 m cells, 25 m layers, layered
background 0.01/0.1/1 S/m, f = 10 Hz), complex128 currents re/im ~ N(0,1).
`value` is matvecs per second over K timed applications with inputs
resident in HBM.  The JSON line also carries the roofline of the dominant
kernel (the tensor contraction), the oracle's CPU throughput on the same
workload, the continuation of the field to the reference deck's receivers
on the same grid, eight right-hand sides through one tensor read per
contraction (batched apply), and (unless --no-solve) the time to a 1e-8 FP64 true
residual of the 256x256x32 random-box model for the plain FP64 solve and
the mixed complex64-inner / FP64-outer solve.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-solve] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(Nx, Ny, Nz, s=16):
    """Bytes per matvec by kernel: n_c cells, n_v = 3 n_c, n_s = 12 Nx Ny Nz
    spectrum, n_G = Nx Ny 2Nz(2Nz+1) tensor entries; s = element size."""
    nc = Nx * Ny * Nz
    nv = 3 * nc
    ns = 4 * Nx * Ny * 3 * Nz
    nG = Nx * Ny * 2 * Nz * (2 * Nz + 1)
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


def cpu_baseline(m, Gspec, v, budget_s=12.0):
    """The oracle's FP64 operator (C restatement of the reference algorithm)
    on the host cores, same inputs, bounded sample of whole matvecs."""
    from oracle import pyoracle as po
    nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    om = po.Model(m.csigb(), m.thick, m.omega, m.Nx, m.Ny, m.Nz, m.dx, m.dy, m.z)
    t0 = time.perf_counter()
    om.apply(Gspec, m.csiga(), v, nthreads)
    t1 = time.perf_counter() - t0
    reps = max(1, min(50, int(budget_s / max(t1, 1e-3))))
    t0 = time.perf_counter()
    for _ in range(reps):
        out = om.apply(Gspec, m.csiga(), v, nthreads)
    dt = (time.perf_counter() - t0) / reps
    split = om.apply_timings()
    t0 = time.perf_counter()
    om.apply(Gspec, m.csiga(), v, 1)
    dt1 = time.perf_counter() - t0
    split1 = om.apply_timings()
    return {"value": 1.0 / dt, "unit": "matvec/s", "cores": nthreads, "kind": "port",
            "sample": "%d FP64 matvecs of the 128x128x32 operator (oracle/giem2g_oracle.c, "
                      "OpenMP, radix-2 FFT), split fwd/contraction/bwd = %.3f/%.3f/%.3f s"
                      % (reps, split[0], split[1], split[2]),
            "value_1thread": 1.0 / dt1,
            "sample_1thread": "1 matvec on one thread, split fwd/contraction/bwd = %.3f/%.3f/%.3f s"
                              % (split1[0], split1[1], split1[2])}, out


# learned preconditioner in the benchmark: mode 2 (network step, fallback to
# inner GMRES) accepted when it cuts the residual by at least 10%
LEARNED_MODE = 2
LEARNED_ACCEPT = 0.9


def solve_bench(torch, gop, models):
    """Time to a 1e-8 FP64 true residual on the 256x256x32 random-box model
    (0.1 Hz, EY polarization): plain FP64 nested FGMRES, complex64 inner, and
    complex64 inner with the learned preconditioner when its trained weights
    (giem2g_amd/data/unet_c4.f32) are present; then both polarizations, one
    after the other and paired, for the plain FP64 and complex64-inner solves."""
    from giem2g_amd import learned
    m = models.c2()
    t0 = time.perf_counter()
    op = gop.IEOperator.from_model(m)
    op.set_background(m.csigb(), m.thick, m.omega)
    op.build_tensor()
    op.set_sigma(m.csiga())
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    st = op.stats()
    E = torch.from_numpy(op.plane_wave(1)).cuda()
    res = {"model": "256x256x32 x 25 m, 20 random boxes 1e-3..10 S/m in 0.01 S/m host, f=0.1 Hz, EY",
           "tensor_build_s": st["t_tensor"], "tensor_fft_s": st["t_tensor_fft"], "setup_s": t_setup}
    sols = {}
    runs = [("fp64", 0, 0), ("mixed", 1, 0)]
    net = None
    if os.path.exists(learned.DEFAULT_WEIGHTS):
        net = learned.UNet.load(learned.DEFAULT_WEIGHTS)
        op.set_learned(net, accept=LEARNED_ACCEPT)
        runs += [("mixed+learned", 1, LEARNED_MODE)]
    for name, prec, mode in runs:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        Es, info = op.solve(E, outer_m=30, inner_m=30, max_restart=100, tol=1e-8,
                            inner_precision=prec, precond=mode)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        sols[name] = Es
        res[name] = {"time_to_1e-8_s": dt, "outer_iterations": info["outer_iterations"],
                     "matvecs_fp64": info["matvecs_fp64"], "matvecs_fp32": info["matvecs_fp32"],
                     "precond_calls": info["precond_calls"], "precond_fallbacks": info["precond_fallbacks"],
                     "true_residual": info["true_residual"], "status": info["status"]}
        log("solve", name, res[name])
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
    # both polarizations of the model: the EX solve alone (with the EY solve
    # above: one after the other), then EY and EX together with paired
    # operator products (one tensor read per pair of applies)
    Ex = torch.from_numpy(op.plane_wave(0)).cuda()
    both = {}
    for name, prec in (("fp64", 0), ("mixed", 1)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, ix = op.solve(Ex, outer_m=30, inner_m=30, max_restart=100, tol=1e-8, inner_precision=prec)
        torch.cuda.synchronize()
        t_x = time.perf_counter() - t0
        t0 = time.perf_counter()
        _, _, jy, jx, cnt = op.solve_pair(E, Ex, outer_m=30, inner_m=30, max_restart=100, tol=1e-8,
                                          inner_precision=prec)
        torch.cuda.synchronize()
        t_p = time.perf_counter() - t0
        both[name] = {"one_after_other_s": res[name]["time_to_1e-8_s"] + t_x, "paired_s": t_p,
                      "paired_applies": int(cnt[0]), "single_applies": int(cnt[1]),
                      "true_residual": max(jy["true_residual"], jx["true_residual"], ix["true_residual"])}
        log("solve both polarizations", name, both[name])
    both["mixed"]["speedup_vs_fp64_paired"] = both["fp64"]["paired_s"] / both["mixed"]["paired_s"]
    res["both_polarizations"] = both
    op.close()
    return res


def c3_bench(torch, gop, models, nfreq=16):
    """Multi-frequency run of the 128x128x64 random-box model: 16 frequencies
    log-spaced over 1e-3..1e3 Hz x 2 polarizations, each frequency with its
    own FP64 tensor.  Both polarizations of a frequency are solved together
    with paired operator products (the complex64 inner contraction reads the
    tensor once for both right-hand sides on the matrix cores); the FP64
    outer residual is reported per frequency and solve."""
    m, freqs = models.c3()
    freqs = freqs[:nfreq]
    op = gop.IEOperator.from_model(m)
    rows = []
    t_all = time.perf_counter()
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
        t0 = time.perf_counter()
        _, _, iy, ix, cnt = op.solve_pair(Ey, Ex, outer_m=30, inner_m=30, max_restart=100, tol=1e-8,
                                          inner_precision=1)
        t_solve = time.perf_counter() - t0
        row = {"freq": float(f), "build_s": t_build, "solve_pair_s": t_solve,
               "EY": {"outer": iy["outer_iterations"], "true_residual": iy["true_residual"],
                      "mv64": iy["matvecs_fp64"], "mv32": iy["matvecs_fp32"]},
               "EX": {"outer": ix["outer_iterations"], "true_residual": ix["true_residual"],
                      "mv64": ix["matvecs_fp64"], "mv32": ix["matvecs_fp32"]},
               "paired_applies": cnt[0], "single_applies": cnt[1]}
        log("c3", json.dumps(row))
        rows.append(row)
    op.close()
    return {"model": "128x128x64 x 20 m, 20 random boxes, %d frequencies 1e-3..1e3 Hz x EY, EX" % len(freqs),
            "total_s": time.perf_counter() - t_all, "solves": 2 * len(freqs),
            "max_true_residual": max(max(r["EY"]["true_residual"], r["EX"]["true_residual"]) for r in rows),
            "per_frequency": rows}


# receivers of the reference's own deck (test/giem2g_input.json:32-35)
REF_DECK_RECEIVERS = [(0.0, -20000.0, 0.0), (0.0, 20000.0, 0.0)]


def batch_bench(torch, op, m, stage64, stage32, nrhs=8, reps=20):
    """Eight right-hand sides through giem2g_op_apply_batch_dev on the C1
    grid (complex64: one tensor read per contraction for all eight; FP64: two
    reads of four): time per batch and matvecs/s against one-vector applies.
    The contraction kernel's own time, matrix-core rate and HBM rate come
    from the rocprofv3 trace of this run (kernel k_contract3<..., 32, 8 or 4,
    ...>); mfma_flop_per_batch is the matrix-core work of one batch."""
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


def recalc_bytes(Nx, Ny, Nz, Nr, s=16):
    """Algorithmic bytes of one continuation to Nr receivers: forward FFTs
    as in the IE apply (fwd), the contraction reading the receiver tensor
    (4NxNy points x 15 Nr Nz) and the current spectrum and writing 6 Nr
    output rows, and the inverse FFTs of those rows (y pass over the whole
    spectrum, x pass over the kept y rows) with the (Nr,3,Nx,Ny) fields."""
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
        nth = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        t0 = time.perf_counter()
        Eo, Ho = ro.apply(Gs, m.csiga(), En, nthreads=nth)
        tc = time.perf_counter() - t0
        rel = max(float(np.linalg.norm(Ea.cpu().numpy() - Eo) / np.linalg.norm(Eo)),
                  float(np.linalg.norm(Ha.cpu().numpy() - Ho) / np.linalg.norm(Ho)))
        row["cpu_baseline"] = {"recalc_s": tc, "cores": nth, "kind": "port", "gpu_vs_cpu_rel": rel}
    rc.close()
    log("recalc", json.dumps(row))
    return row


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--c3", action="store_true",
                    help="also run the 128x128x64 multi-frequency paired-polarization case")
    ap.add_argument("--c3-freqs", type=int, default=16, help="number of C3 frequencies (profiling runs)")
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
    m = models.c1()
    t0 = time.perf_counter()
    op = gop.operator_for(m)
    torch.cuda.synchronize()
    log("setup %.3f s" % (time.perf_counter() - t0), op.stats())
    v = torch.from_numpy(models.c1_currents(m)).cuda()
    out = torch.empty_like(v)
    for _ in range(a.warmup):
        op.apply_dev(v, out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        op.apply_dev(v, out)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per = 1e3 * dt / a.steps
    value = a.steps / dt
    # per-kernel times with HIP events on the operator's stream
    stage_ms = op.time_apply(v, out, reps=max(10, min(100, a.steps)))
    v32 = v.to(torch.complex64)
    o32 = torch.empty_like(v32)
    stage32 = op.time_apply(v32, o32, reps=max(10, min(100, a.steps)), precision=1)
    bytes64 = algorithmic_bytes(m.Nx, m.Ny, m.Nz, 16)
    bytes32 = algorithmic_bytes(m.Nx, m.Ny, m.Nz, 8)
    ach = bytes64["contraction"] / (stage_ms[1] * 1e-3) / 1e9
    traffic = read_traffic(os.path.join(ROOT, "profiles", "contraction_c1_fp64_traffic.txt"))
    line = {
        "metric": "time-to-1e-8 residual + IE matvecs/s: mixed+learned vs plain FP64, 1 GPU",
        "value": value, "unit": "matvec/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_per, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded PCG64 model and currents)",
        "config": {"workload": "C1 IE operator apply, FP64: 128x128x32 cells of 100x100x25 m, "
                               "background 0.01/0.1/1 S/m, f=10 Hz, complex128 currents N(0,1)",
                   "grid": [m.Nx, m.Ny, m.Nz], "precision": "complex128"},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_contract3<double,32,1,1,4> (tensor contraction on the FP64 matrix cores, "
                               "source layers split over four waves)",
                     "algorithmic_bytes_per_launch": bytes64["contraction"],
                     "avg_launch_ms": float(stage_ms[1])},
        "kernels_ms": {"fwd_fft": float(stage_ms[0]), "contraction": float(stage_ms[1]),
                       "bwd_fft": float(stage_ms[2])},
        "kernels_gbs": {k: bytes64[k2] / (t * 1e-3) / 1e9 for k, k2, t in
                        (("fwd_fft", "fwd", stage_ms[0]), ("contraction", "contraction", stage_ms[1]),
                         ("bwd_fft", "bwd", stage_ms[2]))},
        "matvec_total_gbs": bytes64["total"] / (ms_per * 1e-3) / 1e9,
        "roofline_bound_matvec_s": HBM_PEAK_GBS * 1e9 / bytes64["total"],
        "complex64": {"kernels_ms": [float(x) for x in stage32],
                      "matvec_s": 1e3 / float(np.sum(stage32)),
                      "contraction_gbs": bytes32["contraction"] / (stage32[1] * 1e-3) / 1e9},
    }
    if not a.no_cpu:
        Gs = op.get_tensor()
        cb, ref = cpu_baseline(m, Gs, models.c1_currents(m))
        res = out.cpu().numpy()
        cb["gpu_vs_cpu_rel"] = float(np.linalg.norm(res - ref) / np.linalg.norm(ref))
        line["cpu_baseline"] = cb
    else:
        line["cpu_baseline"] = None
    line["recalc"] = recalc_bench(torch, gop, op, m, v, cpu=not a.no_cpu)
    line["batched"] = batch_bench(torch, op, m, stage_ms, stage32)
    op.close()
    if not a.no_solve:
        line["solve"] = solve_bench(torch, gop, models)
    if a.c3:
        line["c3"] = c3_bench(torch, gop, models, nfreq=a.c3_freqs)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
