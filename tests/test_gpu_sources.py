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
