"""The plain reference method at its CPU-runnable size, as the benchmark's
first configuration states it: 64x32x32 cells, 8 ppc drifting Maxwellian
(vsw = 400 km/s, vth = 30 km/s), uniform B0 along z, Omega_i dt = 0.05,
10 field subcycles, against the FP64 oracle (maind.f:372-613 order).

At these parameters the step itself is unstable.  The cold beam drifting
across a grid of half an ion inertial length grows magnetic fluctuations
from the 8-ppc noise: the oracle alone (periodic) reaches |dB| ~ B0 by step
40-50 and 10-20 B0 by step 60-80, where check_time_step halves the subcycle
up to 160 times; with the open x the growth is faster and the run ends in
NaN by step 46.  The growth amplifies every rounding difference between two
codes, so per-ion parity is checked while the codes still track each other
(30 steps) and the rest of a 100-step periodic run through the growth is
compared by its statistics.

Open x with inflow (the configuration as written; the open x is
build-defined and pinned in test_oracle_open_x.py), 30 steps: every step the
ion count (entered, left, present), the charge sum 1/(beta beta_p), the
momentum and ||div B||; every 10 steps every ion's x and v matched by id and
np, up, E and B - B0 (max norm relative to the oracle's).  FP64 build: the
same ions, x and v within 1e-11 (of the box, of 400 km/s), fields within
1e-10 (measured at step 30: 5e-15, 1e-14 and at most 5e-13), momentum within
1e-10.  FP32 build: x within 1e-5, v 2e-5, np and up 1e-4, E 3e-3, B - B0
3e-2 (measured at step 30: 1.5e-6, 3.7e-6, 1.4e-5, 3.1e-6, 9.2e-4, 7.7e-3:
the FP32 differences ride on the growing fluctuations), momentum 1e-4, and
only ions on a face may leave or enter in one code alone.

Periodic (the survey's parity case), FP64, 100 steps: the same checks for the
first 30 steps; then every 10 steps the charge, div B and mean density, and
the growth curve: the rms of B - B0 and of E of the device within 25 % of the
oracle's, and the subcycle counts of check_time_step in the same range.
"""
import os

import numpy as np
import pytest

import oracle as O
from helpers import rel
from swap_amd import api, physics

pytestmark = pytest.mark.gpu

NC = (64, 32, 32)
PPC = 8
SEED = 20261203
TOL = {64: {"x": 1e-11, "v": 1e-11, "np": 1e-10, "up": 1e-10, "E": 1e-10, "dB": 1e-10,
            "momentum": 1e-10},
       32: {"x": 1e-5, "v": 2e-5, "np": 1e-4, "up": 1e-4, "E": 3e-3, "dB": 3e-2,
            "momentum": 1e-4}}


def _setup(precision, open_x=True):
    p = physics.workload(*NC, PPC)
    n = p["npart"]
    cap = 2 * n
    s = physics.make_swap(p, capacity=cap, precision=precision, sort_every=4,
                          flags=api.TRACK_IDS | (api.OUTFLOW_X if open_x else 0), seed=SEED)
    s.load_maxwellian(n, p["vsw"], p["vth"])
    o = O.Oracle(NC[0] + 2, NC[1] + 2, NC[2] + 2, p["dx"], p["dx"], p["dx"], p["dt"],
                 p["ntsub"], p["alpha"], p["beta"], width=64)
    o.set_threads(min(16, os.cpu_count() or 1))
    if open_x:
        s.set_inflow(PPC, p["vsw"], p["vth"])
        o.set_open_x(True)
        o.set_inflow(SEED, PPC, p["vsw"], p["vth"], next_id=n)
    st = {nm: o.grid() for nm in ("b1", "b12", "b1p2", "bt", "btc", "up", "aj", "E")}
    for b in ("b1", "b12", "b1p2", "bt"):
        st[b][2] = p["b0z"]
    xp, vp, vp1 = (np.zeros((3, cap)) for _ in range(3))
    xp[:, :n], vp[:, :n], vp1[:, :n] = s.xp(), s.vp(), s.vp1()
    st.update(np=o.grid(1), nu=np.full((NC[2] + 2, NC[1] + 2, NC[0] + 2), p["nu"]), xp=xp,
              vp=vp, vp1=vp1, vplus=o.parts(cap), vminus=o.parts(cap), Ep=o.parts(cap),
              wght=o.parts(cap, 8), ijkp=np.zeros((3, cap), np.int32), Ni_tot=n,
              id=np.arange(cap, dtype=np.uint32))
    return p, s, o, st


def _db(b, b0z):
    d = np.array(b, np.float64)
    d[2] -= b0z
    return d


def _compare(s, st, p, precision, step, note):
    """every ion by id and the grid fields against the oracle's state"""
    m = st["Ni_tot"]
    L = np.array(NC)[:, None] * p["dx"]
    ids = s.particle_array(api.P_ID)
    oid = st["id"][:m]
    common, ig, io = np.intersect1d(ids, oid, assume_unique=True, return_indices=True)
    xg, xo = s.xp()[:, ig], st["xp"][:, io]
    vg, vo = s.vp()[:, ig], st["vp"][:, io]
    dx = np.abs(xg - xo)
    dx[1:] = np.minimum(dx[1:], L[1:] - dx[1:])
    err = {"x": np.max(dx / L), "v": np.max(np.abs(vg - vo)) / 400.0}
    for f in ("np", "up", "E"):
        err[f] = rel(s.get_field(f), st[f])
    err["dB"] = rel(_db(s.get_field("b1"), p["b0z"]), _db(st["b1"], p["b0z"]))
    print("cfg1 FP%d step %3d: ions %d, " % (precision, step, s.n) +
          " ".join("%s %.2e" % kv for kv in err.items()), flush=True)
    for k, v in err.items():
        note(k, v)
        assert v < TOL[precision][k], (step, k, v)
    if precision == 64:
        assert common.size == m == ids.size, step
    else:
        assert common.size >= (1 - 1e-4) * m - 5, step


def _books(s, st, p, precision, step, n0, gone, added, note):
    """ion counts, charge, momentum and div B of one step"""
    m = st["Ni_tot"]
    d = s.diagnostics()
    assert d["nparticles"] == s.n == n0 - d["escaped"] + d["injected"], step
    if precision == 64:
        assert (d["escaped"], d["injected"], s.n) == (gone, added, m), step
    else:
        assert abs(s.n - m) <= 1e-4 * m + 5, step
    assert d["charge"] == pytest.approx(s.n / p["beta"], rel=1e-12), step
    mom_o = 1.67e-27 * st["vp"][:, :m].sum(axis=1) / p["beta"]
    em = np.max(np.abs(np.array(d["momentum"]) - mom_o)) / np.max(np.abs(mom_o))
    note("momentum", em)
    assert em < TOL[precision]["momentum"], (step, em)
    assert d["divb_l2"] < 1e-12 * p["b0z"] / p["dx"] * np.sqrt(np.prod(NC)), step
    return d


def _noter(worst):
    def note(k, v):
        worst[k] = max(worst.get(k, 0.0), float(v))
    return note


@pytest.mark.parametrize("precision", [64, 32])
def test_cfg1_open_x_against_the_oracle(precision):
    steps = 30
    print("cfg1: %dx%dx%d cells x %d ppc, open x with inflow, %d steps, FP%d"
          % (*NC, PPC, steps, precision), flush=True)
    p, s, o, st = _setup(precision, open_x=True)
    n0 = s.n
    gone = added = 0
    worst = {}
    note = _noter(worst)
    for step in range(1, steps + 1):
        s.step(1)
        o.step(st)
        gone += st["escaped"]
        added += st["injected"]
        _books(s, st, p, precision, step, n0, gone, added, note)
        if step % 10 == 0:
            _compare(s, st, p, precision, step, note)
    print("cfg1 open x FP%d worst over %d steps: %s" % (precision, steps,
          ", ".join("%s %.2e" % kv for kv in worst.items())), flush=True)
    assert gone > 0 and added > 0


def _rms(a):
    return float(np.sqrt(np.mean(np.square(a[:, 1:-1, 1:-1, 1:-1]))))


def test_cfg1_periodic_100_steps_through_the_growth():
    steps = 100
    print("cfg1: %dx%dx%d cells x %d ppc, periodic, %d steps, FP64" % (*NC, PPC, steps),
          flush=True)
    p, s, o, st = _setup(64, open_x=False)
    n0 = s.n
    worst = {}
    note = _noter(worst)
    b0z = p["b0z"]
    grew = False
    for step in range(1, steps + 1):
        s.step(1)
        o.step(st)
        if step <= 30:
            _books(s, st, p, 64, step, n0, 0, 0, note)
            if step % 10 == 0:
                _compare(s, st, p, 64, step, note)
            continue
        d = s.diagnostics()
        assert d["nparticles"] == n0
        assert d["charge"] == pytest.approx(n0 / p["beta"], rel=1e-12), step
        assert d["divb_l2"] < 1e-12 * b0z / p["dx"] * np.sqrt(np.prod(NC)) * 100, step
        if step % 10:
            continue
        npf = s.get_field("np")[1:-1, 1:-1, 1:-1]
        assert abs(npf.mean() / p["n0"] - 1) < 1e-9, step
        rb_g, rb_o = _rms(_db(s.get_field("b1"), b0z)), _rms(_db(st["b1"], b0z))
        re_g, re_o = _rms(s.get_field("E")), _rms(st["E"])
        ntf_g, _ = s.check_time_step()
        ntf_o = o.dtsub()[1]
        print("cfg1 periodic step %3d: rms dB/B0 %.3e (oracle %.3e), rms E %.3e (oracle %.3e), "
              "ntf %d (oracle %d)" % (step, rb_g / b0z, rb_o / b0z, re_g, re_o, ntf_g, ntf_o),
              flush=True)
        assert 0.8 < rb_g / rb_o < 1.25, step
        assert 0.8 < re_g / re_o < 1.25, step
        assert max(ntf_g, ntf_o) <= 4 * min(ntf_g, ntf_o), step
        grew = grew or rb_o > 0.3 * b0z
    assert grew  # the run went through the growth it is meant to test
