"""Pickup-ion sources on the device against the oracle (chem_rates.f):
the cloud's photoionization and charge exchange bit for bit (counts, cells,
fractions, weights, the ions hit), Ionize_sw_mp's count and ring, and a step
after injection (charge = sum 1/(beta beta_p))."""
import math

import numpy as np
import pytest

import oracle as O
from swap_amd import api, physics

pytestmark = pytest.mark.gpu
NC = (16, 8, 8)
SEED = 20261203


def _setup(extra=60000):
    p = physics.workload(*NC, 4)
    s = physics.make_swap(p, capacity=p["npart"] + extra, flags=api.SPECIES | api.TRACK_IDS,
                          seed=SEED)
    s.load_maxwellian(p["npart"], p["vsw"], p["vth"])
    return p, s


def test_ionize_cloud_matches_oracle():
    p, s = _setup()
    n = p["npart"]
    cl = physics.cloud(p, N0_rate=3.0)
    # a cross-section that makes charge exchange frequent near the core
    nn_r0 = cl["N0"] * math.exp(-cl["r0"] / cl["H"])
    cl["sigma_chex"] = 0.05 / (p["dt"] * nn_r0 * 400.0)
    s.sort()  # ions no longer sit at their id: the chex draw follows the id
    ids0 = s.particle_array(api.P_ID)
    assert not np.array_equal(ids0, np.arange(n, dtype=np.uint32))
    cell0 = s.particle_array(api.P_CELL)
    frac0 = s.particle_array(api.P_FRAC)
    v0 = s.vp()
    mr0 = s.particle_array(api.P_MRAT)
    step = 3
    inj, nchex = s.ionize_cloud(step, cl)

    d = (p["dx"],) * 3
    hit = O.chex(cell0, frac0, v0, d, p["dt"], cl, SEED, step, ids=ids0).astype(bool)
    assert nchex == int(hit.sum()) and nchex > 0
    v1 = s.vp()
    mr1 = s.particle_array(api.P_MRAT)
    assert np.all(v1[:, :n][:, hit] == 0)
    assert np.array_equal(v1[:, :n][:, ~hit], v0[:, ~hit])
    assert np.all(mr1[:n][hit] == np.float32(cl["mrat_pu"]))
    assert np.array_equal(mr1[:n][~hit], mr0[~hit])

    cnt, bp = O.cloud_counts(NC, d, p["beta"], p["dt"], cl, SEED, step)
    assert inj == int(cnt.sum()) and inj > 0
    ocell, ofrac = O.cloud_ions(NC, cnt, SEED, step)
    assert s.n == n + inj
    assert np.array_equal(s.particle_array(api.P_CELL)[n:], ocell)
    assert np.array_equal(s.particle_array(api.P_FRAC)[:, n:], ofrac.astype(np.float32))
    assert np.array_equal(s.particle_array(api.P_BETA)[n:], np.repeat(bp, cnt))
    assert np.all(s.particle_array(api.P_MRAT)[n:] == np.float32(cl["mrat_pu"]))
    assert np.all(s.vp()[:, n:] == 0) and np.all(s.vp1()[:, n:] == 0)
    assert np.array_equal(s.particle_array(api.P_ID)[n:], np.arange(n, n + inj, dtype=np.uint32))

    # the step moves the new heavy ions with everything else; the charge
    # diagnostic counts each ion with its weight 1/(beta beta_p)
    s.prime()
    s.step(2)
    dg = s.diagnostics()
    want = (n + float(np.sum(cnt / bp.astype(np.float64)))) / p["beta"]
    assert dg["charge"] == pytest.approx(want, rel=1e-6)
    assert dg["injected"] == inj and dg["nparticles"] == n + inj
    np_g = s.get_field("np")
    assert np.isfinite(np_g).all()


def test_ionize_sw_count_and_ring():
    """Ionize_sw_mp (chem_rates.f:508-646): dNi = ndot dt beta beta_pu dx dy
    delz nz (:551), ring velocity vsw (1 + cos th, sin th, 0) (:567-571)"""
    p, s = _setup()
    n = p["npart"]
    beta_pu = 0.5
    per_ndot = p["dt"] * p["beta"] * beta_pu * p["dx"] ** 3 * (NC[2] + 2)
    ndot = 700.3 / per_ndot  # a rate that injects about 700 ions
    dni = ndot * per_ndot
    inj = s.ionize_sw(5, ndot, p["vsw"], beta_pu)
    assert inj == int(np.rint(dni))
    v = s.vp()[:, n:].astype(np.float64)
    ring = np.hypot(v[0] - p["vsw"], v[1])
    assert np.allclose(ring, p["vsw"], rtol=1e-5) and np.all(v[2] == 0)
    assert np.all(s.particle_array(api.P_BETA)[n:] == np.float32(beta_pu))
    cell = s.particle_array(api.P_CELL)[n:]
    assert np.all((cell & 2047) < NC[0]) and np.all(((cell >> 11) & 1023) < NC[1])


@pytest.mark.parametrize("dni", [700.4, 0.3, 0.7])
def test_ionize_sw_matches_oracle(dni):
    """Ionize_sw_mp on the device == the oracle's restatement: the count of
    every step (incl. the stochastic round-up below one ion), cells and
    fractions bit for bit, velocities to FP32 rounding of the FP64 ring
    (device and host cos/sin may differ in the last FP64 bit), vp1 = vp,
    beta_p = beta_pu, mrat = 1, ids continuing the counter"""
    p, s = _setup(extra=2000 if dni > 1 else 200)
    n = p["npart"]
    beta_pu = 0.5
    d = (p["dx"],) * 3
    per_ndot = p["dt"] * p["beta"] * beta_pu * p["dx"] ** 3 * (NC[2] + 2)
    ndot = dni / per_ndot
    steps = 1 if dni > 1 else 40
    tot = 0
    for step in range(steps):
        cell, frac, v = O.ionize_sw(NC, d, p["dt"], p["beta"], ndot, p["vsw"], beta_pu, SEED,
                                    step)
        m = s.ionize_sw(step, ndot, p["vsw"], beta_pu)
        assert m == cell.size, step
        a, b = n + tot, n + tot + m
        tot += m
        if m == 0:
            continue
        assert np.array_equal(s.particle_array(api.P_CELL)[a:b], cell)
        assert np.array_equal(s.particle_array(api.P_FRAC)[:, a:b], frac.astype(np.float32))
        vg = s.vp()[:, a:b]
        ulp = np.abs(vg.view(np.int32).astype(np.int64) -
                     v.astype(np.float32).view(np.int32).astype(np.int64))
        assert ulp.max() <= 1
        assert np.array_equal(s.vp1()[:, a:b], vg)
        assert np.all(s.particle_array(api.P_BETA)[a:b] == np.float32(beta_pu))
        assert np.all(s.particle_array(api.P_MRAT)[a:b] == 1.0)
        assert np.array_equal(s.particle_array(api.P_ID)[a:b],
                              np.arange(a, b, dtype=np.uint32))
    assert s.n == n + tot
    if dni < 0.5:
        assert 0 < tot < steps  # some steps round up, some do not
    elif dni < 1:
        assert tot == steps


class _OracleRun:
    """The oracle's state for a run with sources: particle arrays of a fixed
    capacity, the first Ni_tot live, with beta_p, mrat and ids."""

    def __init__(self, pr, cap):
        self.pr = pr
        o, st, dt = pr.o, pr.st, pr.o.dtype
        n = pr.n
        self.cap = cap
        for k, ncomp in (("xp", 3), ("vp", 3), ("vp1", 3), ("vplus", 3), ("vminus", 3),
                         ("Ep", 3), ("wght", 8)):
            a = np.zeros((ncomp, cap), dt)
            if k in st:
                a[:, :n] = st[k]
            st[k] = a
        st["ijkp"] = np.zeros((3, cap), np.int32)
        st["beta_p"] = np.ones(cap, dt)
        st["mrat"] = np.ones(cap, dt)
        self.ids = np.zeros(cap, np.uint32)
        self.ids[:n] = np.arange(n, dtype=np.uint32)
        st["Ni_tot"] = n

    def set_from_device(self, s):
        """the device's loaded/injected state (split positions exported as xp)"""
        st, dt = self.pr.st, self.pr.o.dtype
        m = s.n
        st["xp"][:, :m] = s.xp().astype(dt)
        st["vp"][:, :m] = s.vp().astype(dt)
        st["vp1"][:, :m] = s.vp1().astype(dt)
        st["beta_p"][:m] = s.particle_array(api.P_BETA).astype(dt)
        st["mrat"][:m] = s.particle_array(api.P_MRAT).astype(dt)
        self.ids[:m] = s.particle_array(api.P_ID)
        st["Ni_tot"] = m

    def append(self, cell, frac, v, beta_p, mrat, ids):
        st, dt, dx = self.pr.st, self.pr.o.dtype, self.pr.p["dx"]
        a, m = st["Ni_tot"], cell.size
        ijk = [cell & 2047, (cell >> 11) & 1023, cell >> 21]
        for ax in range(3):
            f = frac[ax]
            if self.pr.s.cfg.precision == 32:  # the fraction as the device stores it
                f = f.astype(np.float32).astype(np.float64)
            st["xp"][ax, a:a + m] = (ijk[ax] + 1 + f) * dx
            st["vp"][ax, a:a + m] = v[ax]
            st["vp1"][ax, a:a + m] = v[ax]
        st["beta_p"][a:a + m] = beta_p
        st["mrat"][a:a + m] = mrat
        self.ids[a:a + m] = ids
        st["Ni_tot"] = a + m

    def split(self):
        """cell + FP32 fraction of the live ions, as the device stores them"""
        st, dx = self.pr.st, self.pr.p["dx"]
        m = st["Ni_tot"]
        cell = np.zeros(m, np.uint32)
        frac = []
        for ax in range(3):
            t = st["xp"][ax, :m].astype(np.float64) / dx - 1.0
            c = np.ceil(t) - 1.0
            frac.append((t - c).astype(np.float32))
            cell |= (c.astype(np.uint32) << np.uint32((0, 11, 21)[ax]))
        return cell, frac


def _by_id(ids, arr, total):
    out = np.full(arr.shape[:-1] + (total,), np.nan)
    out[..., ids] = arr
    return out


def _pdist(a, b, L):
    d = np.abs(a - b)
    return np.minimum(d, L - d)


@pytest.mark.parametrize("precision,xtol,vtol,gtol", [(32, 1e-5, 1e-5, 2e-3),
                                                       (64, 1e-11, 1e-11, 1e-8)])
def test_pickup_step_matches_oracle(precision, xtol, vtol, gtol):
    """The mixed-species step: a solar-wind load plus Ionize_sw_mp ring
    protons (beta_p = 1/2) and cloud ions of 28 amu (mrat = 1/28, beta_p < 1
    where a node bears more than max_per_cell), charge exchange turning
    solar-wind ions into slow heavy ions; two steps on the device and in the
    FP64 oracle (maind.f:379-479 order: sources, then the step), the second
    with sources applied on both sides independently.  Per ion matched by id:
    x within xtol of the box, v within vtol x 400 km/s; np within gtol / 10,
    up, E and B - B0 within gtol (max norm relative to the oracle's), the
    tolerances of the uniform step.  The oracle's ions carry mrat and beta_p
    as the device stores them (FP32 arrays in both builds)."""
    from helpers import Pair, rel
    pr = Pair(ncx=NC[0], ncy=NC[1], ncz=NC[2], ppc=4, precision=precision,
              flags=api.SPECIES | api.TRACK_IDS, capacity=4096 + 8000)
    s, o, st, p = pr.s, pr.o, pr.st, pr.p
    d = (p["dx"],) * 3
    cl = physics.cloud(p, N0_rate=3.0)
    nn_r0 = cl["N0"] * math.exp(-cl["r0"] / cl["H"])
    cl["sigma_chex"] = 0.05 / (p["dt"] * nn_r0 * 400.0)  # frequent charge exchange
    beta_pu = 0.5
    ndot = 300.0 / (p["dt"] * p["beta"] * beta_pu * p["dx"] ** 3 * (NC[2] + 2))
    orun = _OracleRun(pr, 4096 + 8000)
    mrat_dev = float(np.float32(cl["mrat_pu"]))  # mrat as the device stores it (FP32)
    L = np.array(NC)[:, None] * p["dx"]

    # sources before the first step (checked bit for bit by the tests above);
    # the oracle starts from the device's state
    s.ionize_sw(0, ndot, p["vsw"], beta_pu)
    s.ionize_cloud(0, cl)
    orun.set_from_device(s)
    assert np.any(st["mrat"][:st["Ni_tot"]] != 1) and np.any(st["beta_p"][:st["Ni_tot"]] != 1)
    s.prime()
    for step in range(2):
        if step == 1:
            # sources on both sides in the order of the calls: the new ring
            # protons (Ionize_sw_mp), then the cloud: charge exchange over
            # every ion present (res_chex), then its new ions
            n_before = s.n
            m_sw = s.ionize_sw(step, ndot, p["vsw"], beta_pu)
            inj, nchex = s.ionize_cloud(step, cl)
            c1, f1, v1 = O.ionize_sw(NC, d, p["dt"], p["beta"], ndot, p["vsw"], beta_pu, SEED,
                                     step)
            assert c1.size == m_sw
            orun.append(c1, f1, v1.astype(np.float32).astype(np.float64)
                        if precision == 32 else v1, beta_pu, 1.0,
                        np.arange(n_before, n_before + m_sw, dtype=np.uint32))
            ocell, ofrac = orun.split()
            m0 = st["Ni_tot"]
            hit = O.chex(ocell, ofrac, [st["vp"][a, :m0] for a in range(3)], d, p["dt"], cl,
                         SEED, step, ids=orun.ids[:m0]).astype(bool)
            assert int(hit.sum()) == nchex and nchex > 0
            st["vp"][:, :m0][:, hit] = 0
            st["mrat"][:m0][hit] = mrat_dev
            cnt, bp = O.cloud_counts(NC, d, p["beta"], p["dt"], cl, SEED, step)
            assert int(cnt.sum()) == inj
            c2, f2 = O.cloud_ions(NC, cnt, SEED, step)
            orun.append(c2, f2, np.zeros((3, inj)), np.repeat(bp, cnt), mrat_dev,
                        np.arange(n_before + m_sw, n_before + m_sw + inj, dtype=np.uint32))
            assert st["Ni_tot"] == s.n
        s.step(1)
        o.step(st)
        m = st["Ni_tot"]
        ids_g = s.particle_array(api.P_ID)
        tot = s.n
        xg = _by_id(ids_g, s.xp(), tot)
        vg = _by_id(ids_g, s.vp(), tot)
        xo = _by_id(orun.ids[:m], st["xp"][:, :m].astype(np.float64), tot)
        vo = _by_id(orun.ids[:m], st["vp"][:, :m].astype(np.float64), tot)
        db_g = s.get_field("b1")
        db_g[2] -= p["b0z"]
        db_o = st["b1"].copy()
        db_o[2] -= p["b0z"]
        err = {"x": np.max(_pdist(xg, xo, L) / L), "v": np.max(np.abs(vg - vo)) / 400.0,
               "np": rel(s.get_field("np"), st["np"]), "up": rel(s.get_field("up"), st["up"]),
               "dB": rel(db_g, db_o), "E": rel(s.get_field("E"), st["E"])}
        print("pickup FP%d step %d: " % (precision, step) +
              " ".join("%s %.2e" % kv for kv in err.items()), flush=True)
        assert err["x"] < xtol and err["v"] < vtol, (step, err)
        assert err["np"] < gtol * 0.1, (step, err)
        for k in ("up", "dB", "E"):
            assert err[k] < gtol, (step, err)
    # the charge of the mixed load: sum 1/(beta beta_p) over every ion
    dg = s.diagnostics()
    want = float(np.sum(1.0 / st["beta_p"][:st["Ni_tot"]].astype(np.float64))) / p["beta"]
    assert dg["charge"] == pytest.approx(want, rel=1e-6)
