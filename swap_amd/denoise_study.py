"""Learned-denoiser study: the moment error of the low-ppc run with the
denoiser against that of the plain high-ppc run, and the wall time per step
of each.

    python -m swap_amd.denoise_study --cells 128 128 128 --ppc 256 --stride 32

1. Pairs.  High-ppc FP32 runs (drifting Maxwellian plus a pickup-ion cloud,
   so the moments have structure; --segments seeded realisations) write a
   pair every --pair-every steps from the step's own final deposit:
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

import numpy as np

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


def gen_pairs(s, p, cl, pairs, segments, spinup, pair_every=1, seed0=20261203):
    """pairs from the high-ppc FP32 run s: `segments` seeded realisations,
    each a fresh load with B = B0, spun up for `spinup` steps (the cloud
    builds its structure), then a pair every `pair_every` steps; returns
    (particles at the end, generation s, pair-deposit s)"""
    t0 = time.perf_counter()
    per_seg = -(-pairs // segments)
    b0 = np.zeros((3, s.nz, s.ny, s.nx))
    b0[2] = p["b0z"]
    tp, j = 0.0, 0
    for seg in range(segments):
        s.set_seed(seed0 + 7919 * seg)
        s.load_maxwellian(p["npart"], p["vsw"], p["vth"])
        s.set_field("b1", b0)
        s.set_field("b12", b0)
        s.prime()
        k = 0
        for k in range(spinup):
            s.ionize_cloud(k, cl)
            s.step(1)
        for _ in range(per_seg):
            if j == pairs:
                break
            for _ in range(pair_every):
                k += 1
                s.ionize_cloud(k, cl)
                s.step(1)
            s.sync()
            t1 = time.perf_counter()
            s.dn_add_pair(api.DN_DISJOINT_STEP)
            tp += time.perf_counter() - t1
            j += 1
        _log("segment %d: %d pairs, %.1f s, %d particles, %d field subcycles" %
             (seg, j, time.perf_counter() - t0, s.n, s.check_time_step()[0]))
    return s.n, time.perf_counter() - t0, tp


def train(s, iters, lr, ntr):
    """Adam on the first ntr pairs: lr for the first half, 0.3 lr to 80 %,
    0.1 lr after; returns (last loss, seconds)"""
    t0 = time.perf_counter()
    done, loss = 0, float("nan")
    while done < iters:
        n = min(250, iters - done)
        f = 1.0 if done < 0.5 * iters else 0.3 if done < 0.8 * iters else 0.1
        loss = s.dn_train(n, lr * f, ntr)
        done += n
        _log("iter %d lr %.2g loss %.4g" % (done, lr * f, loss))
    s.sync()
    return loss, time.perf_counter() - t0


def decompose(raw, net, S, ppc):
    """errors about the truth from the held-out disjoint-pair errors (module
    docstring): plain high-ppc s2, denoised low-ppc e_net, and the ppc at
    which a plain run's error equals e_net"""
    s2 = raw * (S - 1) / S ** 2
    e_net = net - raw / S
    return s2, e_net, (ppc * s2 / e_net if e_net > 0 else float("inf"))


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


def equal_error(cells=(64, 64, 64), ppc=128, stride=16, pairs=2000, segments=20, spinup=20,
                iters=4000, lr=5e-4, n0_rate=1.0, time_steps=5, ppc_cap=1024):
    """The equal-error comparison of the headline metric on a structured
    case, small enough for a few minutes: a pickup-ion cloud (N0 rate
    n0_rate, r0 = 3 cells, H = 8 cells) in a drifting plasma at `cells`.
    Pairs come from FP32 runs at `ppc`, `segments` seeded realisations of
    pairs / segments pairs each; the last realisation's pairs are held out
    (their errors set the gate and every figure below), so training and
    evaluation never share a realisation.  The denoised FP32 run at
    ppc / stride is timed against the plain FP64 run (and the plain FP32
    run) at ppc_eq, the ppc whose moment error equals the denoised one
    (capped at ppc_cap: the ratio is then a lower bound).  For scale, the
    error of a predictor that knew each snapshot's mean but no structure is
    the truth's variance about its mean (signal_var), reported with its own
    equivalent ppc.  Short realisations (spinup + pairs / segments steps)
    keep the periodic domain from filling with pickup ions.  Returns a
    dict."""
    nc = tuple(cells)
    S = stride
    p = physics.workload(*nc, ppc)
    cl = physics.cloud(p, N0_rate=n0_rate, r0_cells=3.0, H_cells=8.0)
    per_seg = -(-pairs // segments)
    s = physics.make_swap(p, capacity=_capacity(p, spinup + per_seg),
                          flags=api.TRACK_IDS | api.SPECIES)
    s.dn_create(p["n0"], p["vsw"], max_pairs=pairs, stride=S, seed=5)
    n_high, gen_s, _ = gen_pairs(s, p, cl, pairs, segments, spinup)
    ntr = pairs - per_seg
    loss, train_s = train(s, iters, lr, ntr)
    raw, net = s.dn_eval(ntr, pairs)
    held = [s.dn_get_pair(i) for i in range(ntr, pairs)]
    tvar, tmax = 0.0, 0.0
    for _, t in held[:20]:
        tvar += float(np.mean(t.var(axis=(1, 2, 3)))) / min(len(held), 20)
        tmax = max(tmax, float(np.abs(t).max()))
    params = s.dn_params()
    s.close()
    s2, e_net, ppc_eq = decompose(raw, net, S, ppc)
    out = {
        "case": "pickup-ion cloud (N0 rate %g, r0 3 cells, H 8 cells) in a drifting plasma, "
                "%dx%dx%d cells" % ((n0_rate,) + nc),
        "cells": list(nc), "ppc_high": ppc, "ppc_low": ppc // S,
        "pairs_train": ntr, "pairs_test": pairs - ntr, "segments": segments,
        "spinup_steps": spinup, "iters": iters, "pair_gen_s": round(gen_s, 2),
        "train_s": round(train_s, 2), "last_loss": loss,
        "units": "mean squared error per value of (np/n0 - 1, up/vsw), about the truth",
        "mse_plain_ppc_high": s2, "mse_plain_ppc_low": S * s2, "mse_denoised_ppc_low": e_net,
        # variance of the truth about its mean, next to the plain high-ppc noise
        "signal_var": tvar - S / (S - 1) * s2, "target_absmax": tmax,
        "ppc_equivalent": ppc_eq,
    }
    sv = out["signal_var"]
    out["ppc_equivalent_mean_predictor"] = ppc * s2 / sv if sv > 0 else float("inf")
    nscale, vscale = p["n0"], p["vsw"]
    ms_dn, gate = timed_run(nc, ppc // S, cl, api.FP32, time_steps, 2, params, held, nscale,
                            vscale)
    q = int(math.ceil(min(ppc_eq, ppc_cap))) if math.isfinite(ppc_eq) else ppc_cap
    ms64, _ = timed_run(nc, q, cl, api.FP64, time_steps, 2)
    ms32, _ = timed_run(nc, q, cl, api.FP32, time_steps, 2)
    out.update({
        "gate_open": bool(gate), "ppc_timed": q, "lower_bound": bool(not ppc_eq <= ppc_cap),
        "ms_step_fp32_denoised_ppc_low": ms_dn, "ms_step_fp64_plain_ppc_eq": ms64,
        "ms_step_fp32_plain_ppc_eq": ms32,
        "speedup_vs_fp64_equal_error": ms64 / ms_dn,
        "speedup_vs_fp32_equal_error": ms32 / ms_dn,
    })
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--cells", type=int, nargs=3, default=[128, 128, 128])
    ap.add_argument("--ppc", type=int, default=256)
    ap.add_argument("--stride", type=int, default=32)
    ap.add_argument("--spinup", type=int, default=20, help="steps before the first pair")
    ap.add_argument("--pairs", type=int, default=500)
    ap.add_argument("--segments", type=int, default=10,
                    help="seeded realisations the pairs are drawn from")
    ap.add_argument("--pair-every", type=int, default=1)
    ap.add_argument("--test-pairs", type=int, default=50,
                    help="held out: the last pairs, from the last segment(s)")
    ap.add_argument("--iters", type=int, default=6000)
    ap.add_argument("--lr", type=float, default=1e-3,
                    help="Adam step for the first half; 0.3x for the next 30%%, 0.1x after")
    ap.add_argument("--n0-rate", type=float, default=1.0,
                    help="cloud ions born per node and step at r0 (structure in the moments)")
    ap.add_argument("--region", type=float, default=16.0,
                    help="radius (cells) of the cloud region reported on its own")
    ap.add_argument("--time-steps", type=int, default=5)
    ap.add_argument("--no-timing", action="store_true")
    args = ap.parse_args(argv)
    nc = tuple(args.cells)
    S = args.stride
    p = physics.workload(*nc, args.ppc)
    cl = physics.cloud(p, N0_rate=args.n0_rate, r0_cells=3.0, H_cells=12.0)

    # 1. pairs from the high-ppc FP32 run (short segments keep the cloud from
    # driving the plasma into the low-density, high-B state whose subcycling
    # blows up the step time)
    per_seg = -(-args.pairs // args.segments)
    s = physics.make_swap(p, capacity=_capacity(p, args.spinup + per_seg * args.pair_every),
                          flags=api.TRACK_IDS | api.SPECIES)
    s.dn_create(p["n0"], p["vsw"], max_pairs=args.pairs, stride=S, seed=5)
    n_high, gen_s, tp = gen_pairs(s, p, cl, args.pairs, args.segments, args.spinup,
                                  args.pair_every)

    # 2. training and held-out evaluation
    ntr = args.pairs - args.test_pairs
    raw0, net0 = s.dn_eval(ntr, args.pairs)
    loss, train_s = train(s, args.iters, args.lr, ntr)
    raw, net = s.dn_eval(ntr, args.pairs)
    held = [s.dn_get_pair(i) for i in range(ntr, args.pairs)]
    # where the structure is: the same decomposition restricted to the
    # voxels within --region cells of the cloud centre, and the variance of
    # the truth about its mean (signal) next to the noise, on held-out pairs
    zz, yy, xx = np.meshgrid(*(np.arange(n) + 1.0 for n in nc[::-1]), indexing="ij")
    rc = np.sqrt((xx - cl["cx"] / p["dx"]) ** 2 + (yy - cl["cy"] / p["dx"]) ** 2 +
                 (zz - cl["cz"] / p["dx"]) ** 2)
    inside = rc <= args.region
    acc = {"raw_in": 0.0, "net_in": 0.0, "tvar": 0.0, "tvar_in": 0.0}
    nh = min(len(held), 10)
    for x, t in held[:nh]:
        y = s.dn_infer(x)
        acc["raw_in"] += float(np.mean((x - t)[:, inside] ** 2)) / nh
        acc["net_in"] += float(np.mean((y - t)[:, inside] ** 2)) / nh
        acc["tvar"] += float(np.mean(t.var(axis=(1, 2, 3)))) / nh
        acc["tvar_in"] += float(np.mean(t[:, inside].var(axis=1))) / nh
    params = s.dn_params()
    s.close()
    s2, e_net, ppc_eq = decompose(raw, net, S, args.ppc)
    e_raw_low = S * s2
    out = {
        "study": "learned moment denoiser",
        "cells": list(nc), "ppc_high": args.ppc, "ppc_low": args.ppc // S,
        "cloud_n0_rate": args.n0_rate, "segments": args.segments,
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
    s2_in = acc["raw_in"] * (S - 1) / S ** 2
    e_in = acc["net_in"] - acc["raw_in"] / S
    out.update({
        "region_cells": args.region, "region_voxels": int(inside.sum()),
        "signal_var": acc["tvar"] - S / (S - 1) * s2,
        "signal_var_region": acc["tvar_in"] - S / (S - 1) * s2_in,
        "mse_plain_high_ppc_region": s2_in, "mse_denoised_low_ppc_region": e_in,
        "ppc_equivalent_region": args.ppc * s2_in / e_in if e_in > 0 else float("inf"),
        "reaches_high_ppc_error_region": bool(e_in <= s2_in),
    })
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
