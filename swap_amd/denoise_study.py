"""Learned-denoiser study: the moment error of the low-ppc run with the
denoiser against that of the plain high-ppc run, and the wall time per step
of each.

    python -m swap_amd.denoise_study --cells 128 128 128 --ppc 256 --stride 32

1. Pairs.  A high-ppc FP32 run (drifting Maxwellian plus a pickup-ion cloud,
   so the moments have structure) writes a pair every --pair-every steps:
   input x = moments of the particles with id % S == 0 (weights x S, a
   ppc/S subsample), target t = moments of the other particles (weights
   x S/(S-1)).  The two samples are disjoint, so their sampling noises are
   independent.  With s2 the per-value noise variance of the plain high-ppc
   moments (all particles) about the truth mu:
       var(x - mu) = S s2,  var(t - mu) = S/(S-1) s2,
       err_raw = E|x - t|^2 = S^2/(S-1) s2,
       err_net = E|f(x) - t|^2 = E|f(x) - mu|^2 + S/(S-1) s2.
   Hence s2 = err_raw (S-1)/S^2 and the net's error about the truth is
   e_net = err_net - err_raw/S, both from held-out pairs.  A plain run at
   ppc' has error s2 ppc/ppc', so the denoised low-ppc run matches the plain
   run at ppc_eq = ppc s2 / e_net.
2. Training: Adam on the first pairs with the library's own kernels; the
   error gate is set on the held-out pairs.
3. Wall time per step (FP32 split-precision low-ppc run with the denoiser
   inside the step, its gate opened on the held-out pairs copied into it)
   against the plain FP64 run at the high ppc and at ppc_eq.

Prints one JSON line; progress goes to stderr.
"""
import argparse
import json
import math
import sys
import time

from . import api, physics


def _log(*a):
    print("denoise_study:", *a, file=sys.stderr, flush=True)


def _capacity(p, cloud_steps):
    return p["npart"] + p["npart"] // 20 + 20000 * cloud_steps + 1000000


def step_time(s, steps, cloud, step0):
    s.sync()
    t0 = time.perf_counter()
    for k in range(steps):
        s.ionize_cloud(step0 + k, cloud)
        s.step(1)
    s.sync()
    return (time.perf_counter() - t0) * 1e3 / steps


def timed_run(nc, ppc, cl, precision, steps, warm, params=None, pairs=None, nscale=None,
              vscale=None):
    """ms/step of a fresh run at ppc; with params, the denoiser runs in the
    step after its gate is set on the given pairs"""
    p = physics.workload(*nc, ppc)
    s = physics.make_swap(p, precision=precision, capacity=_capacity(p, warm + steps),
                          flags=api.SPECIES)
    s.load_maxwellian(p["npart"], p["vsw"], p["vth"])
    gate = None
    if params is not None:
        s.dn_create(nscale, vscale, max_pairs=len(pairs), stride=1, seed=1)
        s.dn_set_params(params)
        for x, t in pairs:
            s.dn_put_pair(x, t)
        raw, net = s.dn_eval(0, len(pairs))
        gate = net < raw
        s.dn_enable(True)
    s.prime()
    step_time(s, warm, cl, 0)
    ms = step_time(s, steps, cl, warm)
    s.close()
    return ms, gate


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--cells", type=int, nargs=3, default=[128, 128, 128])
    ap.add_argument("--ppc", type=int, default=256)
    ap.add_argument("--stride", type=int, default=32)
    ap.add_argument("--spinup", type=int, default=20, help="steps before the first pair")
    ap.add_argument("--pairs", type=int, default=160)
    ap.add_argument("--pair-every", type=int, default=2)
    ap.add_argument("--test-pairs", type=int, default=16)
    ap.add_argument("--iters", type=int, default=2000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--time-steps", type=int, default=5)
    ap.add_argument("--no-timing", action="store_true")
    args = ap.parse_args(argv)
    nc = tuple(args.cells)
    S = args.stride
    p = physics.workload(*nc, args.ppc)
    cl = physics.cloud(p, N0_rate=4.0, r0_cells=3.0, H_cells=12.0)
    nsteps = args.spinup + args.pairs * args.pair_every

    # 1. pairs from the high-ppc FP32 run
    t0 = time.perf_counter()
    s = physics.make_swap(p, capacity=_capacity(p, nsteps),
                          flags=api.TRACK_IDS | api.SPECIES)
    s.load_maxwellian(p["npart"], p["vsw"], p["vth"])
    s.dn_create(p["n0"], p["vsw"], max_pairs=args.pairs, stride=S, seed=5)
    s.prime()
    k = 0
    for k in range(args.spinup):
        s.ionize_cloud(k, cl)
        s.step(1)
    tp = 0.0
    for j in range(args.pairs):
        for _ in range(args.pair_every):
            k += 1
            s.ionize_cloud(k, cl)
            s.step(1)
        s.sync()
        t1 = time.perf_counter()
        s.dn_add_pair(disjoint=True)
        tp += time.perf_counter() - t1
        if (j + 1) % 20 == 0:
            _log("pair %d/%d, %.1f s" % (j + 1, args.pairs, time.perf_counter() - t0))
    n_high = s.n
    gen_s = time.perf_counter() - t0

    # 2. training and held-out evaluation
    ntr = args.pairs - args.test_pairs
    raw0, net0 = s.dn_eval(ntr, args.pairs)
    t0 = time.perf_counter()
    done, loss = 0, float("nan")
    while done < args.iters:
        n = min(250, args.iters - done)
        loss = s.dn_train(n, args.lr, ntr)
        done += n
        _log("iter %d loss %.4g" % (done, loss))
    s.sync()
    train_s = time.perf_counter() - t0
    raw, net = s.dn_eval(ntr, args.pairs)
    held = [s.dn_get_pair(i) for i in range(ntr, args.pairs)]
    params = s.dn_params()
    s.close()
    s2 = raw * (S - 1) / S ** 2
    e_net = net - raw / S
    e_raw_low = S * s2
    ppc_eq = args.ppc * s2 / e_net if e_net > 0 else float("inf")
    out = {
        "study": "learned moment denoiser",
        "cells": list(nc), "ppc_high": args.ppc, "ppc_low": args.ppc // S,
        "pairs_train": ntr, "pairs_test": args.test_pairs, "pair_every": args.pair_every,
        "particles_high": n_high, "pair_gen_s": gen_s, "pair_deposit_s": tp,
        "iters": args.iters, "lr": args.lr, "train_s": train_s, "last_loss": loss,
        "err_raw": raw, "err_net": net, "err_net_untrained": net0,
        "mse_plain_high_ppc": s2, "mse_plain_low_ppc": e_raw_low, "mse_denoised_low_ppc": e_net,
        "variance_gain": e_raw_low / e_net if e_net > 0 else None,
        "ppc_equivalent": ppc_eq,
        "reaches_high_ppc_error": bool(e_net <= s2),
        "units": "mean squared error over (np/n0 - 1, up/vsw), per value",
    }
    _log(json.dumps(out))

    # 3. wall time per step
    if not args.no_timing:
        nscale, vscale = p["n0"], p["vsw"]
        T, W = args.time_steps, 2
        ms_dn, gate = timed_run(nc, args.ppc // S, cl, api.FP32, T, W, params, held, nscale,
                                vscale)
        _log("FP32 %d ppc + denoiser: %.2f ms/step (gate %s)" % (args.ppc // S, ms_dn, gate))
        ms_low, _ = timed_run(nc, args.ppc // S, cl, api.FP32, T, W)
        _log("FP32 %d ppc plain: %.2f ms/step" % (args.ppc // S, ms_low))
        ms64, _ = timed_run(nc, args.ppc, cl, api.FP64, T, W)
        _log("FP64 %d ppc plain: %.2f ms/step" % (args.ppc, ms64))
        out.update(ms_step_fp32_low_denoised=ms_dn, gate_open_in_low_run=gate,
                   ms_step_fp32_low_plain=ms_low, ms_step_fp64_high_plain=ms64,
                   speedup_vs_fp64_high=ms64 / ms_dn)
        if math.isfinite(ppc_eq) and args.ppc // S < ppc_eq < args.ppc:
            q = int(math.ceil(ppc_eq))
            ms_eq, _ = timed_run(nc, q, cl, api.FP64, T, W)
            _log("FP64 %d ppc plain: %.2f ms/step" % (q, ms_eq))
            out.update(ppc_eq_timed=q, ms_step_fp64_ppc_eq=ms_eq,
                       speedup_vs_fp64_equal_error=ms_eq / ms_dn)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
