"""Inflow/outflow x (SWAP_OUTFLOW_X) on the device against the oracle's open x.

The reference is periodic in every axis; the open x is build-defined and the
oracle's restatement of it is pinned in test_oracle_open_x.py.  Here the
device's fused step with escapes, its end-of-step compaction and its inflow
are checked against it."""
import numpy as np
import pytest

from helpers import Pair, rel, tiles, unpack
from swap_amd import api

pytestmark = pytest.mark.gpu


def open_pair(precision, room=2, **kw):
    """a Pair with an open x; room: particle capacity over the load, on both
    sides (the oracle's arrays grow to it)"""
    npart = kw.get("ncx", 16) * kw.get("ncy", 8) * kw.get("ncz", 8) * kw.get("ppc", 8)
    pr = Pair(precision=precision, flags=api.OUTFLOW_X | api.TRACK_IDS, capacity=room * npart,
              **kw)
    pr.o.set_open_x(True)
    n, cap = pr.n, pr.s.cfg.capacity
    st = pr.st
    for k in ("xp", "vp", "vp1", "vplus", "vminus", "Ep"):
        a = np.zeros((3, cap), pr.o.dtype)
        if k in st:
            a[:, :n] = st[k]
        st[k] = a
    st["wght"] = np.zeros((8, cap), pr.o.dtype)
    st["ijkp"] = np.zeros((3, cap), np.int32)
    st["id"] = np.zeros(cap, np.uint32)
    st["id"][:n] = np.arange(n)
    return pr


def _face_dist(x, p):
    """distance (cells) of absolute x positions to the nearer open face"""
    t = x / p["dx"] - 1.0           # cell coordinate: the faces are 0 and ncx
    return np.minimum(np.abs(t), np.abs(t - p["ncx"]))


@pytest.mark.parametrize("inflow", [False, True])
@pytest.mark.parametrize("precision,xtol,vtol,gtol", [(32, 1e-5, 1e-5, 2e-3),
                                                       (64, 1e-11, 1e-11, 1e-8)])
def test_open_x_step_matches_oracle(precision, xtol, vtol, gtol, inflow):
    # without inflow the plasma is at rest: a drift would empty the cells
    # behind the inflow face, where E ~ 1/np has no precision to compare
    pr = open_pair(precision, vsw=400.0 if inflow else 0.0)
    s, o, st, p = pr.s, pr.o, pr.st, pr.p
    Lyz = np.array([p["ncy"], p["ncz"]])[:, None] * p["dx"]
    Lx = p["ncx"] * p["dx"]
    n0, gone, added = pr.n, 0, 0
    if inflow:
        s.set_inflow(p["ppc"], p["vsw"], p["vth"])
        o.set_inflow(s.cfg.seed, p["ppc"], p["vsw"], p["vth"], next_id=n0)
    for step in range(2):
        s.step(1)
        o.step(st)
        gone += st["escaped"]
        added += st["injected"]
        m = st["Ni_tot"]
        ids = s.particle_array(api.P_ID)
        oid = st["id"][:m]
        xg_all = s.xp()
        if precision == 64:
            assert s.n == m and np.array_equal(np.sort(ids), np.sort(oid)), step
        else:
            # FP32 positions differ from FP64 by ~1e-6 cells: an ion that
            # close to a face may leave in one code and stay in the other
            only_g = ~np.isin(ids, oid)
            only_o = ~np.isin(oid, ids)
            assert (_face_dist(xg_all[0, only_g], p) < 1e-3).all(), step
            assert (_face_dist(st["xp"][0, :m][only_o], p) < 1e-3).all(), step
            if only_g.any() or only_o.any():
                pytest.skip("an ion on a face: the codes differ by one ion (allowed)")
        og = np.argsort(ids)
        oo = np.argsort(oid)
        xg, xo = xg_all[:, og], st["xp"][:, :m][:, oo]
        vg, vo = s.vp()[:, og], st["vp"][:, :m][:, oo]
        assert np.max(np.abs(xg[0] - xo[0])) / Lx < xtol, step
        d = np.abs(xg[1:] - xo[1:])
        assert np.max(np.minimum(d, Lyz - d) / Lyz) < xtol, step
        assert np.max(np.abs(vg - vo)) / 400.0 < vtol, step
        assert rel(s.get_field("np"), st["np"]) < gtol * 0.1, step
        assert rel(s.get_field("up"), st["up"]) < gtol, step
        b0 = p["b0z"]
        db_g = s.get_field("b1")
        db_g[2] -= b0
        db_o = st["b1"].copy()
        db_o[2] -= b0
        assert rel(db_g, db_o) < gtol, step
        assert rel(s.get_field("E"), st["E"]) < gtol, step
        d = s.diagnostics()
        assert d["escaped"] == gone and d["injected"] == added == m - n0 + gone
    assert gone > 0 and (added > 0) == inflow


def test_open_x_run_keeps_its_books_and_its_density():
    """steps with inflow: counts balance, ids stay unique, the arrays stay
    sorted and compact, and the upstream density is carried across the box"""
    pr = open_pair(32, ncx=32, ncy=16, ncz=16, ppc=8)
    s, p = pr.s, pr.p
    n0 = s.n
    s.set_inflow(p["ppc"], p["vsw"], p["vth"])
    s.step(12)
    d = s.diagnostics()
    assert d["escaped"] > 0 and d["injected"] > 0
    assert s.n == d["nparticles"] == n0 - d["escaped"] + d["injected"]
    ids = s.particle_array(api.P_ID)
    assert len(np.unique(ids)) == s.n and ids.max() < n0 + d["injected"]
    cell = s.particle_array(api.P_CELL)
    i, j, k = unpack(cell)
    assert i.max() < p["ncx"] and j.max() < p["ncy"] and k.max() < p["ncz"]
    # the step ends with the compaction sort: the live ions are in tile order
    tile, _, _ = tiles(cell, (p["ncx"], p["ncy"], p["ncz"]))
    assert (np.diff(tile) >= 0).all()
    npf = s.get_field("np")[1:-1, 1:-1, 1:-1]
    assert abs(npf.mean() / p["n0"] - 1) < 0.05
    # the face nodes are divided by their shares: no step at either face
    prof = npf.mean(axis=(0, 1)) / p["n0"]
    assert abs(prof[0] - 1) < 0.1 and abs(prof[-1] - 1) < 0.1


def test_open_x_refusals():
    pr = open_pair(32, ncx=8, ncy=8, ncz=8, ppc=2)
    with pytest.raises(api.SwapError, match="move_ion_half"):
        pr.s.move_ion_half()
    with pytest.raises(api.SwapError, match="periodic x"):
        pr.s.dn_create(1.0, 1.0)
    periodic = Pair(ncx=8, ncy=8, ncz=8, ppc=2)
    with pytest.raises(api.SwapError, match="OUTFLOW_X"):
        periodic.s.set_inflow(2.0, pr.p["vsw"], pr.p["vth"])
    # more ions than the capacity holds: a clear error, not an overrun
    tight = open_pair(32, room=1, ncx=8, ncy=8, ncz=8, ppc=2)
    tight.s.set_inflow(40.0, tight.p["vsw"], tight.p["vth"])
    with pytest.raises(api.SwapError, match="capacity"):
        tight.s.step(1)
