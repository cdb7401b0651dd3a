"""Pins of the oracle's Ionize_sw_mp restatement (chem_rates.f:508-646).

The count is fixed by the reference's three lines (:553-562): dNi = ndot dt
beta beta_pu dx dy delz nz with the reference's nz (cells + 2 ghosts), a
stochastic round-up to 1 below one ion, then nint.  Their consequence, which
the tests check by sampling, is E[dNi] = dNi for dNi < 1/2, exactly 1 for
1/2 <= dNi < 1 (nint rounds the kept value up), and nint(dNi) above.  The
ions (:567-610) lie on the ring |v - vsw x| = vsw in the x-y plane with a
uniform phase, at positions uniform in the domain.
"""
import math

import numpy as np
import pytest
from scipy import stats

import oracle as O

NC = (16, 8, 8)
D = (360.41, 360.41, 360.41)
DT = 2.609
SEED = 20261203
VSW = 400.0


def per_ndot(nc=NC, beta=1.0, beta_pu=1.0):
    return DT * beta * beta_pu * D[0] * D[1] * D[2] * (nc[2] + 2)


@pytest.mark.parametrize("dni,want", [(12.4, 12), (12.6, 13), (700.5, 701), (3.0, 3)])
def test_count_above_one_is_nint(dni, want):
    ndot = dni / per_ndot()
    for step in range(5):
        cell, _, _ = O.ionize_sw(NC, D, DT, 1.0, ndot, VSW, 1.0, SEED, step)
        assert cell.size == want


def test_count_uses_the_z_extent_with_ghosts():
    """nz in :553 is the storage extent (cells + 2): doubling ncx or ncy
    leaves the count alone, ncz = 6 -> 14 scales it by 16/8"""
    ndot = 100.0 / per_ndot()
    n0 = O.ionize_sw(NC, D, DT, 1.0, ndot, VSW, 1.0, SEED, 0)[0].size
    n_x = O.ionize_sw((32, 16, 8), D, DT, 1.0, ndot, VSW, 1.0, SEED, 0)[0].size
    n_z = O.ionize_sw((16, 8, 14), D, DT, 1.0, ndot, VSW, 1.0, SEED, 0)[0].size
    assert n0 == 100 and n_x == 100 and n_z == 160


def test_count_scales_with_beta_and_beta_pu():
    ndot = 50.0 / per_ndot()
    assert O.ionize_sw(NC, D, DT, 2.0, ndot, VSW, 1.0, SEED, 0)[0].size == 100
    assert O.ionize_sw(NC, D, DT, 1.0, ndot, VSW, 0.5, SEED, 0)[0].size == 25


def test_count_below_half_is_bernoulli():
    """dNi = 0.3: one ion with probability 0.3 per step (:555-560)"""
    ndot = 0.3 / per_ndot()
    k = np.array([O.ionize_sw(NC, D, DT, 1.0, ndot, VSW, 1.0, SEED, s)[0].size
                  for s in range(4000)])
    assert set(np.unique(k)) <= {0, 1}
    sig = math.sqrt(0.3 * 0.7 / k.size)
    assert abs(k.mean() - 0.3) < 5 * sig


def test_count_between_half_and_one_is_one():
    """dNi = 0.7: rounded up to 1 with probability 0.7, else nint(0.7) = 1"""
    ndot = 0.7 / per_ndot()
    k = [O.ionize_sw(NC, D, DT, 1.0, ndot, VSW, 1.0, SEED, s)[0].size for s in range(300)]
    assert set(k) == {1}


def _many(nions=20000, step=7):
    ndot = nions / per_ndot()
    return O.ionize_sw(NC, D, DT, 1.0, ndot, VSW, 1.0, SEED, step)


def test_ring_velocity():
    """v = (vsw + vsw cos th, vsw sin th, 0) (:570-573): on the ring of radius
    vsw about (vsw, 0, 0), vz = 0, phase uniform on [0, 2 PI) with the
    reference's PI = 3.14159 (so the ring misses its last 5e-6 rad)"""
    _, _, v = _many()
    assert np.allclose(np.hypot(v[0] - VSW, v[1]), VSW, rtol=1e-12)
    assert np.all(v[2] == 0)
    ph = np.mod(np.arctan2(v[1], v[0] - VSW), 2 * math.pi)
    assert ph.max() < 2 * 3.14159
    assert stats.kstest(ph / (2 * 3.14159), "uniform").pvalue > 1e-3
    # the ring's centre is the bulk velocity of the new ions
    n = v.shape[1]
    assert abs(v[0].mean() - VSW) < 5 * VSW / math.sqrt(2 * n)
    assert abs(v[1].mean()) < 5 * VSW / math.sqrt(2 * n)


def test_positions_uniform_in_the_domain():
    """x = q(1) + (1 - u)(q(n-1) - q(1)) per axis (:575-577): cell + fraction
    uniform on (0, nc], the split right-closed (fraction in (0, 1])"""
    cell, frac, _ = _many()
    ijk = [cell & 2047, (cell >> 11) & 1023, cell >> 21]
    for a in range(3):
        assert np.all(ijk[a] < NC[a])
        assert np.all((frac[a] > 0) & (frac[a] <= 1))
        t = ijk[a] + frac[a]
        assert stats.kstest(t / NC[a], "uniform").pvalue > 1e-3
        counts = np.bincount(ijk[a], minlength=NC[a])
        assert stats.chisquare(counts).pvalue > 1e-3


def test_deterministic_per_seed_and_step():
    a = _many(step=3)
    b = _many(step=3)
    c = _many(step=4)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[0], c[0])
