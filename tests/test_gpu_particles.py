"""GPU parity of the particle routines (gutsp.f) against the FP64 oracle.

Every test feeds the same seeded load to both codes: the device draws it,
the oracle receives the device's particles as reference-layout absolute
positions (xp(n,3)) and velocities.
"""
import numpy as np
import pytest

import oracle as O
from helpers import TILE, Pair, rel, sort_perm, tiles, unpack
from swap_amd import api

pytestmark = pytest.mark.gpu


def _face_ambiguous(frac):
    # particles exactly on (or within FP64 rounding of) a cell face can be
    # classified into either neighbour by the FP64 oracle's q(i) < x test
    return np.any((frac > 1 - 1e-9) | (frac < 1e-9), axis=0)


def _face_ambiguous_bound(n, nc):
    # The seeded load puts t = (1 - k 2^-24) nc_a (k an integer) and stores
    # the fraction t - ceil(t) + 1 in FP32: a fraction is either a multiple
    # of nc_a 2^-24 >= 6e-8 or exactly 1 (t an integer, which for nc_a a
    # power of two has probability nc_a 2^-24 per axis), so within 1e-9 of
    # a face means exactly 1.0.  Expected count n sum_a nc_a 2^-24; the
    # bound allows the Poisson tail to about 1e-6.
    lam = n * sum(nc) * 2.0 ** -24
    return max(2, int(np.ceil(lam + 6 * np.sqrt(lam) + 3)))


def test_seeded_load_matches_oracle_sampler():
    pr = Pair(load=False)
    p = pr.p
    pr.s.load_maxwellian(pr.n, p["vsw"], p["vth"])
    cell, frac, v = O.load_maxwellian(pr.n, pr.s.cfg.seed, (p["ncx"], p["ncy"], p["ncz"]),
                                      p["vsw"], p["vth"])
    assert np.array_equal(pr.s.particle_array(api.P_CELL), cell)
    assert np.array_equal(pr.s.particle_array(api.P_FRAC), frac.astype(np.float32))
    vg = pr.s.vp()
    # device and host libm may differ in the last FP64 bit; after rounding to
    # FP32 that leaves at most one ulp
    ulp = np.abs(vg.view(np.int32).astype(np.int64) - v.astype(np.float32).view(np.int32))
    assert ulp.max() <= 1
    assert np.array_equal(pr.s.vp1(), vg)
    # drifting Maxwellian moments (sigma = vth/sqrt(2), part_init.f:204-206)
    assert abs(vg[0].mean() - p["vsw"]) < 1.0
    assert abs(vg[1].std() - p["vth"] / np.sqrt(2)) < 0.5


def test_get_interp_weights_and_deposits():
    pr = Pair()
    s, o, st = pr.s, pr.o, pr.st
    s.get_interp_weights()
    o.get_interp_weights(pr.n, pr.n, st["xp"], st["ijkp"], st["wght"])
    ijk_g = s.particle_array(api.P_IJKP)
    amb = _face_ambiguous(s.particle_array(api.P_FRAC))
    assert amb.sum() <= _face_ambiguous_bound(pr.n, (pr.p["ncx"], pr.p["ncy"], pr.p["ncz"]))
    same = np.all(ijk_g == st["ijkp"], axis=0)
    assert np.all(same | amb), "cell index mismatch away from faces"
    w = s.particle_array(api.P_WGHT)
    assert np.allclose(w[:, same], st["wght"][:, same], atol=2e-6, rtol=0)
    assert np.allclose(w.sum(axis=0), 1.0, atol=1e-6)
    # update_np / update_up (vp) / extrapol_up
    s.update_np()
    o.update_np(pr.n, pr.n, st["ijkp"], st["wght"], None, st["np"])
    assert rel(s.get_field("np"), st["np"]) < 2e-6
    assert abs(st["np"][1:-1, 1:-1, 1:-1].mean() / pr.p["n0"] - 1) < 1e-9
    s.update_up(api.V_CUR)
    o.update_up(pr.n, pr.n, st["ijkp"], st["wght"], None, st["vp"], st["np"], st["up"])
    assert rel(s.get_field("up"), st["up"]) < 2e-6
    s.extrapol_up()
    o.extrapol_up(pr.n, pr.n, st["ijkp"], st["wght"], None, st["vp"], st["vp1"], st["np"],
                  st["up"])
    assert rel(s.get_field("up"), st["up"]) < 2e-6


def _smooth_fields(pr, rng):
    """periodic node/edge fields with O(1) structure for the gather tests"""
    o, st, p = pr.o, pr.st, pr.p
    for nm, scale, off in (("aj", 30.0, 0.0), ("up", 40.0, 400.0), ("btc", 0.004, 0.0)):
        a = (rng.standard_normal(st[nm].shape) * scale).astype(o.dtype)
        a[0] += off
        if nm == "btc":
            a[2] += p["b0z"]
        o.periodic(a)
        st[nm][...] = a
        pr.s.set_field(nm, a)


def test_push_routines_match_oracle():
    pr = Pair()
    s, o, st, n = pr.s, pr.o, pr.st, pr.n
    rng = np.random.default_rng(1)
    _smooth_fields(pr, rng)
    s.get_interp_weights()
    o.get_interp_weights(n, n, st["xp"], st["ijkp"], st["wght"])
    s.update_np()
    o.update_np(n, n, st["ijkp"], st["wght"], None, st["np"])
    keep = np.all(s.particle_array(api.P_IJKP) == st["ijkp"], axis=0)
    vscale = 400.0
    # get_Ep
    s.get_Ep()
    o.get_Ep(n, n, st["ijkp"], st["wght"], None, st["aj"], st["up"], st["btc"], st["Ep"])
    Ep = s.particle_array(api.P_EP)
    assert rel(Ep[:, keep], st["Ep"][:, keep]) < 2e-5
    # get_vplus_vminus on identical Ep (so the rotation alone is compared)
    s.get_vplus_vminus()
    o.get_vplus_vminus(n, n, st["ijkp"], st["wght"], None, np.ascontiguousarray(
        Ep.astype(o.dtype)), st["btc"], st["vp"], st["vplus"], st["vminus"])
    vpl = s.particle_array(api.P_VPLUS)
    vmi = s.particle_array(api.P_VMINUS)
    assert np.max(np.abs(vmi[:, keep] - st["vminus"][:, keep])) < 1e-5 * vscale
    assert np.max(np.abs(vpl[:, keep] - st["vplus"][:, keep])) < 1e-5 * vscale
    # |v+| = |v-| (rotation)
    assert np.allclose(np.linalg.norm(vpl, axis=0), np.linalg.norm(vmi, axis=0), rtol=3e-6)
    # improve_up: vp1 = (v+ + v-)/2 and its deposit
    s.improve_up()
    o.improve_up(n, n, st["ijkp"], st["wght"], None, st["vp1"], np.ascontiguousarray(
        vpl.astype(o.dtype)), np.ascontiguousarray(vmi.astype(o.dtype)), st["np"], st["up"])
    assert np.max(np.abs(s.vp1() - st["vp1"])) < 1e-5 * vscale
    assert rel(s.get_field("up"), st["up"]) < 2e-6
    # get_vp_final
    vp_old = s.vp()
    s.get_vp_final()
    o.get_vp_final(n, n, np.ascontiguousarray(Ep.astype(o.dtype)), st["vp"], st["vp1"],
                   np.ascontiguousarray(vpl.astype(o.dtype)))
    assert np.array_equal(s.vp1(), vp_old)
    assert np.max(np.abs(s.vp() - st["vp"])) < 1e-5 * vscale
    # move_ion_half on identical velocities
    st["vp"][...] = s.vp()
    s.move_ion_half()
    o.move_ion_half(n, n, st["xp"], st["vp"])
    xg = s.xp()
    L = np.array([pr.p["ncx"], pr.p["ncy"], pr.p["ncz"]])[:, None] * pr.p["dx"]
    d = np.abs(xg - st["xp"])
    d = np.minimum(d, L - d)  # periodic image
    assert np.max(d / L) < 1e-6


def test_split_move_is_bit_exact():
    pr = Pair(load=True)
    s, p = pr.s, pr.p
    cell = s.particle_array(api.P_CELL)
    f = s.particle_array(api.P_FRAC).copy()
    v = s.vp()
    sc = np.float32([0.5 * p["dt"] / p["dx"]] * 3)
    s.move_ion_half()
    c2 = cell.copy()
    fx, fy, fz = (np.ascontiguousarray(f[a]) for a in range(3))
    O.split_move(c2, fx, fy, fz, *(np.ascontiguousarray(v[a]) for a in range(3)), sc,
                 (p["ncx"], p["ncy"], p["ncz"]))
    assert np.array_equal(s.particle_array(api.P_CELL), c2)
    fg = s.particle_array(api.P_FRAC)
    assert np.array_equal(fg, np.stack([fx, fy, fz]))
    assert fg.min() > 0 and fg.max() <= 1


def test_sort_permutation_is_stable_and_bit_exact():
    """the sort's permutation is the stable order of (tile, source chunk,
    cell in tile), in the random order of a load and again after a move
    (6 chunks of 4096 here)"""
    pr = Pair(ncx=32, ncy=16, ncz=8, ppc=6)
    s, p = pr.s, pr.p
    nc = (p["ncx"], p["ncy"], p["ncz"])
    cell = s.particle_array(api.P_CELL)
    perm = sort_perm(cell, nc)
    s.sort()
    assert np.array_equal(s.particle_array(api.P_ID), perm)
    assert np.array_equal(s.particle_array(api.P_CELL), cell[perm])
    # sorting the sorted arrays again reorders only within tiles (a tile's
    # runs from several source chunks that now share a chunk merge)
    perm2 = sort_perm(cell[perm], nc)
    s.sort()
    assert np.array_equal(s.particle_array(api.P_ID), perm[perm2])
    # after a half move
    s.move_ion_half()
    ids = s.particle_array(api.P_ID)
    cell = s.particle_array(api.P_CELL)
    perm3 = sort_perm(cell, nc)
    s.sort()
    assert np.array_equal(s.particle_array(api.P_ID), ids[perm3])
    assert np.array_equal(s.particle_array(api.P_CELL), cell[perm3])


def test_fixed_point_deposit_accuracy():
    """200k particles in one cell: the fixed-point deposit (each FP32
    contribution rounded to a 64-bit count of 2^-24, the counts added as
    integers) matches the exact sum of the FP32 weights to a few ulps (plain
    FP32 accumulation would be off by ~1e-5)."""
    from swap_amd import physics
    p = physics.workload(8, 8, 8, 1)
    n = 200_000
    rng = np.random.default_rng(7)
    cell = np.full(n, 3 | (3 << 11) | (3 << 21), np.uint32)
    frac = rng.uniform(1e-3, 1.0, (3, n)).astype(np.float32)
    v = np.zeros((3, n), np.float32)
    s2 = api.Swap(8, 8, 8, p["dx"], p["dt"], p["alpha"], p["beta"], p["nu"], n)
    s2.load_split(cell, frac, v, v)
    s2.update_np()
    npg = s2.get_field("np")
    one = np.float32(1)
    g = [one - frac[a] for a in range(3)]
    f = [frac[a] for a in range(3)]
    w000 = (g[0] * g[1]) * g[2]   # FP32 products as the device forms them
    qn = 1.0 / (p["dx"] ** 3 * p["beta"])
    # cell (3,3,3) has its lower corner at storage node 3 and upper at 4
    # (storage index = Fortran index - 1, ghost layer at 0)
    exact = np.sum(w000.astype(np.float64)) * qn
    assert abs(npg[3, 3, 3] / exact - 1) < 4e-7
    w111 = (f[0] * f[1]) * f[2]
    exact = np.sum(w111.astype(np.float64)) * qn
    assert abs(npg[4, 4, 4] / exact - 1) < 4e-7


def _unwrap(c, centre, nc):
    """the periodic image of cell c nearest to a box centre (k_pass.hip unwrap)"""
    d = c - centre
    return np.where(2 * d > nc, c - nc, np.where(2 * d < -nc, c + nc, c))


@pytest.mark.gpu
def test_step_stores_each_wavefront_in_cell_order():
    """P3 stores the 64 particles of each wavefront in the order of their new
    cells, inside the wavefront's own slots: after a step from a sorted load,
    every 64-particle block of every work item (a tile's run) holds the same
    particles as before, in non-decreasing cell key"""
    pr = Pair(ncx=32, ncy=16, ncz=16, ppc=16, sort_every=8)
    s, p = pr.s, pr.p
    nc = np.array([p["ncx"], p["ncy"], p["ncz"]])
    s.sort()
    cell0 = s.particle_array(api.P_CELL)
    ids0 = s.particle_array(api.P_ID)
    tile, _, _ = tiles(cell0, nc)
    assert np.all(np.diff(tile) >= 0)
    s.step(1)
    assert s.diagnostics()["sorts"] == 1  # the step itself did not sort
    cell1 = s.particle_array(api.P_CELL)
    ids1 = s.particle_array(api.P_ID)
    i1, j1, k1 = unpack(cell1)
    ntx, nty = -(-nc[0] // TILE[0]), -(-nc[1] // TILE[1])
    starts = np.flatnonzero(np.r_[True, tile[1:] != tile[:-1]])
    ends = np.r_[starts[1:], len(tile)]
    nblocks = 0
    for a, b in zip(starts, ends):
        assert b - a <= 32768  # one work item per tile here
        t = tile[a]
        tc = np.array([t % ntx, (t // ntx) % nty, t // (ntx * nty)])
        lo = tc * TILE
        hi = np.minimum(lo + TILE, nc) - 1
        cen = (lo + hi) // 2
        ux = _unwrap(i1[a:b], cen[0], nc[0])
        uy = _unwrap(j1[a:b], cen[1], nc[1])
        uz = _unwrap(k1[a:b], cen[2], nc[2])
        key = ((uz & 255) << 16) | ((uy & 255) << 8) | (ux & 255)
        for q in range(0, b - a, 64):
            blk = slice(q, min(q + 64, b - a))
            assert np.all(np.diff(key[blk]) >= 0), (a, q)
            assert np.array_equal(np.sort(ids1[a:b][blk]), np.sort(ids0[a:b][blk])), (a, q)
            nblocks += 1
    assert nblocks >= pr.n // 64
