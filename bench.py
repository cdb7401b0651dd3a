"""Benchmark of the fused hybrid-PIC time step on one MI355X.

Workload: 256x128x128 cells x 64 ppc (268,435,456 protons), drifting
Maxwellian (vsw = 400 km/s, vth = 30 km/s), uniform B0 along z,
Omega_i dt = 0.05, 10 field subcycles per step, periodic boundaries, FP32
particle state (split-precision positions), FP64 B advance.  A "step" is one
full swap_step: the three particle passes, the counting sort when due, and
all field subcycles.  Particles are generated on the device before timing.

Prints one JSON line with the roofline of the dominant kernel, a CPU
baseline (the oracle with OpenMP on a bounded sample, FP32 and FP64) and the
metric's second half, "equal_error": the learned denoiser trained on 2000
pairs of a 64^3 pickup-cloud case, and the denoised 8-ppc FP32 step timed
against the plain FP64 (and FP32) step at the ppc of equal moment error
(swap_amd/denoise_study.py equal_error; about a minute).

Other workloads: --workload pickup (a neutral cloud injecting 28-amu ions
before every step), --precision 64 (the same kernels in FP64), --cells/--ppc
(e.g. 512 256 256 / 32 for the 1.07e9-particle precision form).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s incl. deposit, % HBM peak; s/step vs plain at equal error"

# algorithmic bytes of each particle pass: per particle (SoA state read and
# written) and per grid node (moments written, 6-component gather source read)
PASS_BYTES = {"P1": (40, 16), "P2": (28, 12 + 24), "P3": (68, 16 + 24)}
COPY_PROBE_BYTES = 1 << 30
# measured HBM bytes per launch of each pass (rocprofv3 FETCH_SIZE + WRITE_SIZE,
# scaled by the same counters on a plain copy), written from a profiled run
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_pass_traffic.json")


def measured_traffic(kernel, cfg):
    """HBM bytes per launch of `kernel` from the committed counter summary,
    if it was measured on this workload; else None"""
    try:
        t = json.load(open(TRAFFIC_FILE))
    except (OSError, ValueError):
        return None
    if t.get("config") != cfg:
        return None
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else k["fetch_bytes"] + k["write_bytes"]


def _host_cores():
    """cores this process may use on the host (its affinity set), at most 16
    (the GPU box's CPU share per GPU), and the CPU model"""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return min(n, 16), n, model


def cpu_baseline(steps=20):
    """The oracle (oracle/swap_oracle.c, OpenMP over the host's cores) on a
    bounded sample of the workload: the same 64 ppc drifting Maxwellian on
    64x32x32 cells (4.2M protons), `steps` full steps, in the reference's own
    width (FP32, makefile:1 -real-size 32) and in FP64."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    from swap_amd import physics
    cores, ncpu, model = _host_cores()
    p = physics.workload(64, 32, 32, 64)
    n = p["npart"]
    cell, frac, v = O.load_maxwellian(n, 20261203, (64, 32, 32), p["vsw"], p["vth"])
    i, j, k = cell & 2047, (cell >> 11) & 1023, cell >> 21
    xp = np.stack([(i + 1 + frac[0]) * p["dx"], (j + 1 + frac[1]) * p["dx"],
                   (k + 1 + frac[2]) * p["dx"]])
    rate = {}
    secs = {}
    for width in (32, 64):
        o = O.Oracle(66, 34, 34, p["dx"], p["dx"], p["dx"], p["dt"], 10, p["alpha"], p["beta"],
                     width=width)
        o.set_threads(cores)
        dt = o.dtype
        st = {nm: o.grid() for nm in ("b1", "b12", "b1p2", "bt", "btc", "up", "aj", "E")}
        for b in ("b1", "b12", "b1p2", "bt"):
            st[b][2] = p["b0z"]
        st.update(np=o.grid(1), nu=np.full((34, 34, 66), p["nu"], dt),
                  xp=np.ascontiguousarray(xp, dt), vp=np.ascontiguousarray(v, dt),
                  vp1=np.ascontiguousarray(v, dt), vplus=o.parts(n), vminus=o.parts(n),
                  Ep=o.parts(n), wght=o.parts(n, 8), ijkp=np.zeros((3, n), np.int32), Ni_tot=n)
        t0 = time.perf_counter()
        for _ in range(steps):
            o.step(st)
        secs[width] = time.perf_counter() - t0
        rate[width] = n * steps / secs[width]
    return {"value": rate[32], "unit": "particle-steps/s", "cores": cores, "kind": "port",
            "fp64_value": rate[64], "cpu_model": model, "host_cpus": ncpu,
            "sample": "oracle/swap_oracle.c with OpenMP on %d of the host's %d cores (%s), "
                      "64x32x32 cells x 64 ppc (%s protons), %d full steps incl. 10 field "
                      "subcycles each: FP32 %.1f s (value), FP64 %.1f s (fp64_value)"
                      % (cores, ncpu, model, format(n, ","), steps, secs[32], secs[64])}


def device_copy_gbs(probe=False):
    """Measured device-copy bandwidth: read + write of a 1 GiB device-to-device
    hipMemcpyAsync (a copy kernel), the best of 5 timed with HIP events on the
    null stream; the practical ceiling beside the 8 TB/s peak.  probe: one
    more copy, for the byte counters of a profiled run."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    ev = [C.c_void_p(), C.c_void_p()]

    def ok(rc, what):
        if rc != 0:
            raise RuntimeError("%s failed (%d)" % (what, rc))
    ok(hip.hipMalloc(C.byref(src), C.c_size_t(COPY_PROBE_BYTES)), "hipMalloc")
    ok(hip.hipMalloc(C.byref(dst), C.c_size_t(COPY_PROBE_BYTES)), "hipMalloc")
    ok(hip.hipMemset(src, 1, C.c_size_t(COPY_PROBE_BYTES)), "hipMemset")
    for e in ev:
        ok(hip.hipEventCreate(C.byref(e)), "hipEventCreate")
    d2d = 3  # hipMemcpyDeviceToDevice
    ok(hip.hipMemcpyAsync(dst, src, C.c_size_t(COPY_PROBE_BYTES), d2d, None), "copy")
    best = None
    for _ in range(5):
        ok(hip.hipEventRecord(ev[0], None), "hipEventRecord")
        ok(hip.hipMemcpyAsync(dst, src, C.c_size_t(COPY_PROBE_BYTES), d2d, None), "copy")
        ok(hip.hipEventRecord(ev[1], None), "hipEventRecord")
        ok(hip.hipEventSynchronize(ev[1]), "hipEventSynchronize")
        ms = C.c_float()
        ok(hip.hipEventElapsedTime(C.byref(ms), ev[0], ev[1]), "hipEventElapsedTime")
        best = ms.value if best is None else min(best, ms.value)
    if probe:
        ok(hip.hipMemcpyAsync(dst, src, C.c_size_t(COPY_PROBE_BYTES), d2d, None), "copy")
    ok(hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
    for e in ev:
        hip.hipEventDestroy(e)
    hip.hipFree(src)
    hip.hipFree(dst)
    return 2 * COPY_PROBE_BYTES / (best * 1e-3) / 1e9


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, nargs=3, default=[256, 128, 128])
    ap.add_argument("--ppc", type=int, default=64)
    ap.add_argument("--sort-every", type=int, default=8,
                    help="longest interval between counting sorts; the step sorts "
                         "earlier when a work item's box outgrows its deposit window")
    ap.add_argument("--precision", type=int, default=32)
    ap.add_argument("--workload", choices=("uniform", "pickup", "outflow"), default="uniform",
                    help="pickup: a neutral cloud at the domain centre injects 28-amu ions "
                         "(photoionization, charge exchange) before every step; outflow: "
                         "inflow/outflow x with the upstream plasma flowing in at x = 0")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-equal-error", action="store_true",
                    help="skip the equal-error block (the metric's second half: the "
                         "denoised low-ppc FP32 step against the plain FP64 step at the ppc "
                         "of equal moment error, on a 64^3 pickup-cloud case; about a minute)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run the 1 GiB device copy once more (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(multi-GPU sharding is a design note only)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    copy_gbs = device_copy_gbs(args.copy_probe)

    from swap_amd import api, physics
    ncx, ncy, ncz = args.cells
    p = physics.workload(ncx, ncy, ncz, args.ppc)
    npart = p["npart"]
    nn = (ncx + 2) * (ncy + 2) * (ncz + 2)
    pickup = args.workload == "pickup"
    outflow = args.workload == "outflow"
    cloud = physics.cloud(p) if pickup else None
    # the cloud adds a few tenths of a percent of ions per step; an open x
    # holds this step's entrants before the sort drops the ions that left
    cap = npart + (npart // 50 + 2000000) * pickup + (npart // 10) * outflow
    flags = (api.SPECIES if pickup else 0) | (api.OUTFLOW_X if outflow else 0)
    s = physics.make_swap(p, capacity=cap, precision=args.precision,
                          sort_every=args.sort_every, flags=flags)
    s.load_maxwellian(npart, p["vsw"], p["vth"])
    if outflow:
        s.set_inflow(args.ppc, p["vsw"], p["vth"])
    print("bench: loaded %d particles on %dx%dx%d cells" % (npart, ncx, ncy, ncz),
          file=sys.stderr, flush=True)

    def run(k, step0):
        """k steps; returns the particle-steps done"""
        if not pickup:
            # an open x keeps the ions that left as dead cells until its next
            # scheduled sort; reading the count would drop them with an extra
            # sort, so the count is taken before and after the k steps (the
            # inflow balances the outflow to a fraction of a percent)
            n0 = s.n
            s.step(k)
            return (n0 + s.n) * k // 2 if outflow else n0 * k
        done = 0
        for i in range(k):
            s.ionize_cloud(step0 + i, cloud)  # Ionize_pluto_mp before the step
            done += s.n
            s.step(1)
        return done

    run(args.warmup, 0)
    s.sync()
    s.profile(True)
    t0 = time.perf_counter()
    psteps = run(args.steps, args.warmup)
    s.sync()
    elapsed = time.perf_counter() - t0
    prof = s.profile_read()
    s.profile(False)
    diag = s.diagnostics()
    ms_step = elapsed * 1e3 / args.steps
    value = psteps / elapsed
    npart = psteps // args.steps  # mean particles per timed step

    # dominant particle pass and its roofline
    dom = max(PASS_BYTES, key=lambda k: prof[k][0])
    ms_tot, cnt = prof[dom]
    bp, bn = PASS_BYTES[dom]
    alg = bp * npart + bn * nn
    achieved = alg / (ms_tot / cnt * 1e-3) / 1e9
    step_alg = 136 * npart + 116 * nn
    kname = "k_pass<%s,%s>" % ("float" if args.precision == 32 else "double", dom[1])
    tcfg = {"cells": [ncx, ncy, ncz], "ppc": args.ppc, "sort_every": args.sort_every,
            "precision": args.precision}
    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32" if args.precision == 32 else "fp64",
        "data": "synthetic: seeded drifting Maxwellian (Philox), uniform B0" +
                (", neutral-cloud pickup ions" if pickup else "") +
                (", inflow at x = 0" if outflow else ""),
        "config": {"workload": "%dx%dx%d cells x %d ppc (%d %s), 10 B subcycles/step, "
                               "%s, %s particles, FP64 B advance"
                               % (ncx, ncy, ncz, args.ppc, npart,
                                  "ions incl. cloud pickup ions (mean)" if pickup else
                                  "protons (mean)" if outflow else "protons",
                                  "inflow/outflow x, periodic y z" if outflow else "periodic",
                                  "split-precision FP32" if args.precision == 32
                                  else "FP64"),
                   "cells": [ncx, ncy, ncz], "ppc": args.ppc, "n_particles": npart,
                   "sort_every": args.sort_every, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": kname,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(kname, tcfg) if args.workload == "uniform"
                     else None,
                     "alg_bytes_per_launch": alg,
                     "avg_launch_ms": ms_tot / cnt,
                     "copy_GBs": copy_gbs, "frac_of_copy": achieved / copy_gbs},
        "step_roofline": {"alg_bytes": step_alg, "achieved_GBs": step_alg / (ms_step * 1e-3) /
                          1e9, "frac": step_alg / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
                          "frac_of_copy": step_alg / (ms_step * 1e-3) / 1e9 / copy_gbs},
        "stages_ms_per_step": {k: v[0] / args.steps for k, v in prof.items()},
        "window_misses": diag["window_misses"],
    }
    if pickup:
        out["config"]["cloud"] = cloud
        out["config"]["injected"] = diag["injected"]
    s.close()
    if not args.no_cpu_baseline:
        print("bench: timing the CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline()
    if not args.no_equal_error and args.workload == "uniform" and args.precision == 32:
        from swap_amd import denoise_study
        print("bench: equal-error block (2000 pairs, training, timed runs)", file=sys.stderr,
              flush=True)
        out["equal_error"] = denoise_study.equal_error()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
