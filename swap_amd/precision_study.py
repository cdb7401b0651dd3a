"""Precision-form study: the split-precision FP32 step (integer cell + FP32
offset positions, FP32 velocities, exact fixed-point moment sums, FP64 B
advance) against the FP64 build of the same kernels, from the same seeded
load, after the same number of steps; per-quantity errors and the time per
step of each.

    python -m swap_amd.precision_study --cells 512 256 256 --ppc 32 --steps 10
    python -m swap_amd.precision_study --cells 256 128 128 --ppc 16 --particles

The two runs go one after the other (each context is freed before the next),
so the precision form at 512x256x256 x 32 ppc (1.07e9 particles) and its
FP64 twin need not fit on the device together.  Errors (max norm over the
grid, relative to the max norm of the FP64 quantity; B as its deviation from
B0): np, up, E, dB; with --particles, per-particle position (relative to the
box) and velocity (relative to the drift) after matching particles by id.
Prints one JSON line; progress goes to stderr.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import api, physics


def _log(*a):
    print("precision_study:", *a, file=sys.stderr, flush=True)


def run(p, precision, steps, particles):
    flags = api.TRACK_IDS if particles else 0
    s = physics.make_swap(p, precision=precision, flags=flags)
    s.load_maxwellian(p["npart"], p["vsw"], p["vth"])
    s.prime()
    s.step(1)  # first step: includes the sort of the random-order load
    s.sync()
    t0 = time.perf_counter()
    s.step(steps - 1)
    s.sync()
    ms = (time.perf_counter() - t0) * 1e3 / max(steps - 1, 1)
    inner = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
    out = {"ms_per_step": ms, "diag": s.diagnostics()}
    for f in ("np", "up", "E", "b1"):
        a = s.get_field(f)
        out[f] = a[inner[1:]] if a.ndim == 3 else a[inner]
    if particles:
        ids = s.particle_array(api.P_ID)
        order = np.argsort(ids, kind="stable")
        out["x"] = s.xp()[:, order]
        out["v"] = s.vp()[:, order].astype(np.float64)
    s.close()
    return out


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--cells", type=int, nargs=3, default=[512, 256, 256])
    ap.add_argument("--ppc", type=int, default=32)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--particles", action="store_true",
                    help="also compare every particle (needs host memory for both runs)")
    args = ap.parse_args(argv)
    p = physics.workload(*args.cells, args.ppc)
    _log("FP32 run, %d particles" % p["npart"])
    r32 = run(p, api.FP32, args.steps, args.particles)
    _log("FP32 %.2f ms/step; FP64 run" % r32["ms_per_step"])
    r64 = run(p, api.FP64, args.steps, args.particles)
    _log("FP64 %.2f ms/step" % r64["ms_per_step"])
    b0 = np.zeros_like(r64["b1"])
    b0[2] = p["b0z"]
    err = {
        "np": rel(r32["np"], r64["np"]),
        "up": rel(r32["up"], r64["up"]),
        "E": rel(r32["E"], r64["E"]),
        "dB": rel(r32["b1"] - b0, r64["b1"] - b0),
    }
    rms = {k: float(np.sqrt(np.mean((r32[f] - r64[f]) ** 2)) / max(np.sqrt(np.mean(r64[f] ** 2)), 1e-300))
           for k, f in (("np", "np"), ("up", "up"), ("E", "E"))}
    d32, d64 = r32["diag"], r64["diag"]
    diag = {k: abs(d32[k] - d64[k]) / max(abs(d64[k]), 1e-300)
            for k in ("charge", "kinetic", "b_energy")}
    diag["momentum_x"] = abs(d32["momentum"][0] - d64["momentum"][0]) / abs(d64["momentum"][0])
    diag["divb_l2"] = {"fp32": d32["divb_l2"], "fp64": d64["divb_l2"]}
    out = {
        "study": "precision form vs FP64 build", "cells": args.cells, "ppc": args.ppc,
        "particles": p["npart"], "steps": args.steps,
        "ms_per_step_fp32": r32["ms_per_step"], "ms_per_step_fp64": r64["ms_per_step"],
        "max_rel_err": err, "rms_rel_err": rms, "diagnostics_rel_diff": diag,
    }
    if args.particles:
        L = np.array(args.cells, np.float64)[:, None] * p["dx"]
        dx = np.abs(r32["x"] - r64["x"])
        dx = np.minimum(dx, L - dx)
        out["particle_max_err"] = {"x_over_box": float(np.max(dx / L)),
                                   "x_over_cell": float(np.max(dx) / p["dx"]),
                                   "v_over_vsw": float(np.max(np.abs(r32["v"] - r64["v"])) /
                                                       p["vsw"])}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
