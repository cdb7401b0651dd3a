"""The fused time step (swap_step) against the FP64 oracle's maind.f-order step
and against the routine-level API on the device."""
import numpy as np
import pytest

from helpers import Pair, rel
from swap_amd import api

pytestmark = pytest.mark.gpu


def _by_id(s, arr):
    ids = s.particle_array(api.P_ID)
    out = np.empty_like(arr)
    out[..., ids] = arr
    return out


def _pdist(a, b, L):
    d = np.abs(a - b)
    return np.minimum(d, L - d)


@pytest.mark.parametrize("precision,xtol,vtol,gtol", [(32, 1e-5, 1e-5, 2e-3),
                                                       (64, 1e-11, 1e-11, 1e-8)])
def test_fused_step_matches_oracle(precision, xtol, vtol, gtol):
    pr = Pair(precision=precision)
    s, o, st, p = pr.s, pr.o, pr.st, pr.p
    L = np.array([p["ncx"], p["ncy"], p["ncz"]])[:, None] * p["dx"]
    for step in range(2):
        s.step(1)
        o.step(st)
        xg = _by_id(s, s.xp())
        vg = _by_id(s, s.vp())
        # per-particle x and v after each step
        assert np.max(_pdist(xg, st["xp"], L) / L) < xtol, step
        assert np.max(np.abs(vg - st["vp"])) / 400.0 < vtol, step
        # moments at n+1/2 and the advanced B
        assert rel(s.get_field("np"), st["np"]) < gtol * 0.1
        assert rel(s.get_field("up"), st["up"]) < gtol
        b0 = p["b0z"]
        db_g = s.get_field("b1")
        db_g[2] -= b0
        db_o = st["b1"].copy()
        db_o[2] -= b0
        assert rel(db_g, db_o) < gtol, step
        assert rel(s.get_field("E"), st["E"]) < gtol, step


def test_routine_level_step_matches_fused():
    """maind.f's call sequence through the routine-level API == swap_step"""
    a = Pair(load=True, sort_every=0)
    b = Pair(load=True, sort_every=0)
    s = a.s
    s.get_interp_weights()
    s.update_np()
    s.update_up()
    s.curlB("bt", "aj")
    s.edge_to_center("bt", "btc")
    s.extrapol_up()
    s.get_Ep()
    s.get_vplus_vminus()
    s.improve_up()
    s.get_Ep()
    s.get_vplus_vminus()
    s.get_vp_final()
    s.move_ion_half()
    s.get_interp_weights()
    s.update_np()
    s.update_up()
    ntf, dtsub = s.check_time_step()
    for _ in range(ntf):
        s.curlB("bt", "aj")
        s.predict_B()
        s.correct_B()
        s.f_update_tlev()
    s.move_ion_half()
    b.s.step(1)
    # swap_step may store particles in another order (the last pass keeps each
    # wavefront's particles in cell order): match them by id
    ia, ib = s.particle_array(api.P_ID), b.s.particle_array(api.P_ID)
    oa, ob = np.argsort(ia), np.argsort(ib)
    assert np.array_equal(ia[oa], ib[ob])
    L = np.array([a.p["ncx"], a.p["ncy"], a.p["ncz"]])[:, None] * a.p["dx"]
    assert np.max(_pdist(s.xp()[:, oa], b.s.xp()[:, ob], L) / L) < 1e-6
    assert np.max(np.abs(s.vp()[:, oa] - b.s.vp()[:, ob])) < 400 * 1e-6
    for f in ("np", "up", "E"):
        assert rel(s.get_field(f), b.s.get_field(f)) < 1e-5, f
    # B perturbation: both paths sum the deposits in a different order (FP32
    # atomics), which moves dB at the 1e-5 level after 10 subcycles
    b0 = np.zeros_like(s.get_field("b1"))
    b0[2] = a.p["b0z"]
    assert rel(s.get_field("b1") - b0, b.s.get_field("b1") - b0) < 1e-4


@pytest.mark.parametrize("precision", [32, 64])
def test_cooperative_field_launch_step_matches_two_launches(precision):
    """swap_step with the field subcycles in one cooperative launch (grid
    barriers between the phases) against two launches per subcycle, over
    steps with and without a sort, on a grid off the field tiles: bit for bit
    in the FP32 build (fixed-point moment sums); the FP64 build sums its
    moments with FP64 atomics, whose order varies from run to run, so there
    to 1e-11 with particles matched by id"""
    runs = []
    for extra in (0, api.PERSISTENT_FIELDS):
        a = Pair(40, 9, 33, ppc=4, precision=precision, sort_every=2,
                 flags=api.TRACK_IDS | extra)
        a.s.step(3)
        r = {f: a.s.get_field(f) for f in ("b1", "b12", "np", "up", "E", "aj")}
        ids = a.s.particle_array(api.P_ID)
        r.update(x=_by_id(a.s, a.s.xp()), v=_by_id(a.s, a.s.vp()), ids=np.sort(ids))
        if precision == 32:
            r.update(x_raw=a.s.xp(), id_raw=ids)
        runs.append(r)
        a.s.close()
    for k in runs[0]:
        if precision == 32:
            assert np.array_equal(runs[0][k], runs[1][k]), k
        elif k == "ids":
            assert np.array_equal(runs[0][k], runs[1][k]), k
        else:
            assert rel(runs[0][k], runs[1][k]) < 1e-11, k


def test_invariants_over_steps():
    pr = Pair(ncx=24, ncy=16, ncz=16, ppc=8)
    s, p = pr.s, pr.p
    d0 = s.diagnostics()
    assert d0["charge"] == pytest.approx(pr.n / p["beta"], rel=1e-12)
    s.step(5)
    d1 = s.diagnostics()
    assert d1["charge"] == pytest.approx(d0["charge"], rel=1e-12)
    assert d1["nparticles"] == pr.n
    # div B stays at rounding level (curlE updates are divergence free)
    assert d1["divb_l2"] < 1e-12 * p["b0z"] / p["dx"] * np.sqrt(24 * 16 * 16)
    # bulk drift is carried: total momentum changes by less than 1 %
    assert d1["momentum"][0] == pytest.approx(d0["momentum"][0], rel=1e-2)
    # the density stays near n0
    npf = s.get_field("np")[1:-1, 1:-1, 1:-1]
    assert abs(npf.mean() / p["n0"] - 1) < 1e-5
