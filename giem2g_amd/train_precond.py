"""Train the learned right preconditioner on the GPU.

    python -m giem2g_amd.train_precond --models 300 --steps 3000 --out giem2g_amd/data/unet_c4.f32

1. Training data: for each C4 random-box model (64x64x16, f log-uniform in
   [1e-2, 1e2] Hz, seeds 10000+k, polarization EX/EY alternating) the FP64
   nested FGMRES is run and the residuals of its outer iterates are kept
   (with the correction from the FP64 solution for --loss correction).
2. Training: Adam with the library's BF16-MFMA forward/backward kernels,
   either on the residual the solver sees, ||r - A M(r)||^2 / ||r||^2, with
   the gradient through the operator adjoint (--loss residual, default), or
   on the MSE to the normalized FP64 correction (--loss correction, with a
   random complex phase per sample).
3. Evaluation on held-out models (seeds 12000+): the residual reduction
   ||r - A M(r)|| / ||r|| of the trained network on the held-out pairs, and
   the nested solve with and without the network.
Prints one JSON line with the figures; writes the FP32 parameter file.
"""
import argparse
import json
import os
import sys
import time

import numpy as np


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def residual_reduction(op, pols=(0, 1), max_steps=5):
    """rho = ||r - A M(r)|| / ||r|| for the outer-iterate residuals of one
    model (op has the network attached)."""
    import torch
    out = []
    for pol in pols:
        Eb = torch.from_numpy(op.plane_wave(pol)).cuda()
        b = torch.empty_like(Eb)
        op.scale(0, Eb, b)
        xj = b.clone()
        r = torch.empty_like(Eb)
        z = torch.empty_like(Eb)
        az = torch.empty_like(Eb)
        for j in range(max_steps):
            if j > 0:
                Ej, _ = op.solve(Eb, outer_m=j, inner_m=30, max_restart=1, tol=1e-14)
                op.scale(1, Ej, xj)
            op.apply_dev(xj, r)
            torch.cuda.synchronize()
            r = (b - r).contiguous()
            op.learned_apply(r, z)
            op.apply_dev(z, az)
            torch.cuda.synchronize()
            out.append(float(torch.linalg.vector_norm(r - az) / torch.linalg.vector_norm(r)))
    return out


def evaluate(net, n_eval=8, seed0=12000, accept=0.9):
    import torch
    from . import models, operator as gop
    rows = []
    for k in range(n_eval):
        m = models.c4_train(seed0 + k)
        op = gop.operator_for(m)
        op.set_learned(net, accept=accept)
        rho = residual_reduction(op, pols=(k % 2,), max_steps=3)
        E = op.plane_wave(k % 2)
        Et = torch.from_numpy(E).cuda()
        res = {"seed": seed0 + k, "freq": m.freq, "rho": rho}
        base = None
        for name, prec, mode in (("fp64", 0, 0), ("mixed", 1, 0), ("mixed+learned1", 1, 1),
                                 ("mixed+learned2", 1, 2), ("fp64+learned2", 0, 2)):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            Es, info = op.solve(Et, outer_m=30, inner_m=30, max_restart=100, tol=1e-8,
                                inner_precision=prec, precond=mode)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            if base is None:
                base = Es
            err = float(torch.linalg.vector_norm(Es - base) / torch.linalg.vector_norm(base))
            res[name] = {"s": dt, "it": info["outer_iterations"], "mv64": info["matvecs_fp64"],
                         "mv32": info["matvecs_fp32"], "calls": info["precond_calls"],
                         "fallbacks": info["precond_fallbacks"], "true": info["true_residual"],
                         "err_vs_fp64": err}
        rows.append(res)
        op.set_learned(None)
        op.close()
        log("eval", json.dumps(res))
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--models", type=int, default=2000)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--max-pairs", type=int, default=5, help="outer iterates per model")
    ap.add_argument("--eval", type=int, default=8)
    ap.add_argument("--init", default=None, help="start from this parameter file")
    ap.add_argument("--out", default=None)
    ap.add_argument("--loss", choices=["correction", "residual"], default="residual",
                    help="correction: MSE to the FP64 correction; residual: ||r - A M(r)||^2 "
                         "through the operator adjoint")
    a = ap.parse_args(argv)
    import torch
    from . import learned
    assert torch.cuda.is_available()
    net = learned.UNet.load(a.init) if a.init else learned.UNet(seed=0)
    t0 = time.perf_counter()
    if a.loss == "correction":
        X, T = learned.make_dataset(a.models, max_steps=a.max_pairs, log=log)
        t_data = time.perf_counter() - t0
        npairs = int(X.shape[0])
        log("dataset", tuple(X.shape), "%.1f s" % t_data)
        t0 = time.perf_counter()
        hist = learned.train(net, X, T, steps=a.steps, batch=a.batch, lr=a.lr, log=log)
        del X, T
    else:
        ops, samples, feat = learned.residual_dataset(a.models, max_steps=a.max_pairs, log=log)
        t_data = time.perf_counter() - t0
        npairs = len(samples)
        t0 = time.perf_counter()
        hist = learned.train_residual(net, ops, samples, steps=a.steps, batch=a.batch, lr=a.lr, log=log)
        for op, _, _ in ops:
            op.set_learned(None)
            op.close()
        del ops, samples
        feat.close()
    t_train = time.perf_counter() - t0
    if a.out:
        net.save(a.out)
    torch.cuda.empty_cache()
    rows = evaluate(net, a.eval) if a.eval else []
    summary = {"models": a.models, "pairs": npairs, "dataset_s": t_data,
               "train_steps": a.steps, "batch": a.batch, "train_s": t_train,
               "ms_per_step": 1e3 * t_train / max(1, a.steps),
               "loss_first": float(np.mean(hist[:50])), "loss_last": float(np.mean(hist[-50:])),
               "eval": rows}
    print(json.dumps(summary), flush=True)
    net.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
