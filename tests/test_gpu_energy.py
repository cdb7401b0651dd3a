"""Energy_diag on the device (swap_energy_diag) against the oracle's
restatement, the input-energy account through sources and an open x, and the
time-loop driver's records."""
import os

import numpy as np
import pytest

from helpers import Pair
from swap_amd import api, driver, io

pytestmark = pytest.mark.gpu

KEYS = ("input_E", "Evp", "Euf", "EB1", "EB1x", "EB1y", "EB1z", "EE", "total_E", "norm_E")


def _close(d, o, rtol):
    for k in KEYS:
        assert d[k] == pytest.approx(o[k], rel=rtol, abs=1e-30), k


@pytest.mark.parametrize("open_x", [False, True])
def test_energy_diag_matches_oracle(open_x):
    from test_gpu_outflow import open_pair
    pr = open_pair(64) if open_x else Pair(precision=64)
    s, o, st, p = pr.s, pr.o, pr.st, pr.p
    b0 = (0.0, 0.0, p["b0z"])
    st["input_k"] = o.ke(st)
    e0 = s.energy_diag(p["mion"], b0)
    _close(e0, o.energy_diag(st, p["mion"], b0), 1e-12)
    if open_x:
        s.set_inflow(p["ppc"], p["vsw"], p["vth"])
        o.set_inflow(s.cfg.seed, p["ppc"], p["vsw"], p["vth"], next_id=pr.n)
    for step in range(2):
        s.step(1)
        o.step(st)
        d = s.energy_diag(p["mion"], b0)
        _close(d, o.energy_diag(st, p["mion"], b0), 1e-10)
    assert d["EB1"] > 0 and d["EE"] > 0 and d["Euf"] > 0
    if open_x:   # ions left and entered: the account moved with them
        assert st["escaped"] > 0 and st["injected"] > 0
        assert abs(d["input_E"] / e0["input_E"] - 1) > 1e-6
    else:        # a periodic box without sources keeps its account
        assert d["input_E"] == e0["input_E"]


def test_input_account_follows_sources():
    pr = Pair(precision=64, flags=api.TRACK_IDS | api.SPECIES, capacity=4 * 16 * 8 * 8 * 8)
    s, p = pr.s, pr.p
    b0 = (0.0, 0.0, p["b0z"])
    e0 = s.energy_diag(p["mion"], b0)
    assert e0["input_E"] == pytest.approx(e0["Evp"], rel=1e-13)     # a load
    n0 = s.n
    per_ndot = p["dt"] * p["beta"] * 0.5 * p["dx"] ** 3 * (p["ncz"] + 2)
    m = s.ionize_sw(1, 500.0 / per_ndot, p["vsw"], 0.5)   # about 500 new ions
    assert m == 500
    # the new ions' energy joins the account: input = Evp again (no step yet)
    e1 = s.energy_diag(p["mion"], b0)
    assert e1["input_E"] == pytest.approx(e1["Evp"], rel=1e-13) and s.n == n0 + m
    s.set_input_energy(p["mion"], 2.5)
    assert s.energy_diag(p["mion"], b0)["input_E"] == pytest.approx(2.5, rel=1e-15)


def test_driver_writes_the_reference_records(tmp_path):
    pr = Pair(ncx=16, ncy=8, ncz=8, ppc=4, flags=0, capacity=16 * 8 * 8 * 4)
    s, p = pr.s, pr.p
    out = str(tmp_path / "run")
    recs = driver.run(s, 3, p["mion"], (0.0, 0.0, p["b0z"]), out_dir=out, nout=2, p=p,
                      mrestart=2, restart_path=str(tmp_path / "restart.npz"))
    assert [r["m"] for r in recs] == [1, 2, 3]
    for f in ("para.dat", "c.coord.dat", "c.energy.dat", "c.np.dat", "c.b1.dat", "c.xp_0.dat"):
        assert os.path.exists(os.path.join(out, f)), f
    rec = io.read_energy(out)
    assert [m for m, _ in rec] == [1, 2, 3]
    for (m, e), r in zip(rec, recs):
        for k in ("input_E", "Evp", "Euf", "EB1", "EE"):
            assert e[k] == pytest.approx(r[k], rel=1e-6), k
        assert e["EeP"] == 0.0
    assert [m for m, _ in io.read_output(out, "np", (s.nz, s.ny, s.nx))] == [2]
    # the checkpoint at m = 2 carries the input account
    r = io.load_checkpoint(str(tmp_path / "restart.npz"))
    assert r.energy_diag(p["mion"], (0.0, 0.0, p["b0z"]))["input_E"] == pytest.approx(
        recs[1]["input_E"], rel=1e-12)
