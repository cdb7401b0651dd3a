"""Output files and checkpoints of a device run: the per-output records hold
the device fields and particles (FP32, reference layouts), and a run resumed
from a checkpoint continues bit for bit like the uninterrupted run."""
import numpy as np
import pytest

from swap_amd import api, io, physics

pytestmark = pytest.mark.gpu


def _run(nc=(12, 8, 6), ppc=8, flags=api.TRACK_IDS | api.SPECIES):
    p = physics.workload(*nc, ppc)
    s = physics.make_swap(p, flags=flags, capacity=p["npart"] + 5000)
    s.load_maxwellian(p["npart"], p["vsw"], p["vth"])
    s.prime()
    return p, s


def test_output_records_hold_the_device_state(tmp_path):
    p, s = _run()
    s.step(2)
    with io.OutputWriter(str(tmp_path), s, p["alpha"]) as w:
        w.write(2)
        s.step(1)
        w.write(3)
    shp3, shp1 = (3, s.nz, s.ny, s.nx), (s.nz, s.ny, s.nx)
    recs = io.read_output(str(tmp_path), "np", shp1)
    assert [m for m, _ in recs] == [2, 3]
    assert np.array_equal(recs[1][1], s.get_field("np").astype(np.float32))
    for name in ("b1", "E", "up"):
        assert np.array_equal(io.read_output(str(tmp_path), name, shp3)[1][1],
                              s.get_field(name).astype(np.float32)), name
    cap = s.cfg.capacity
    xp = io.read_output(str(tmp_path), "xp", (3, cap))[1][1]
    assert np.array_equal(xp[:, :s.n], s.xp().astype(np.float32))
    assert not xp[:, s.n:].any()
    bp = io.read_output(str(tmp_path), "beta_p", (cap,))[1][1]
    assert np.all(bp[:s.n] == 1.0)


def test_checkpoint_resume_is_bit_exact(tmp_path):
    p, a = _run()
    cl = physics.cloud(p, N0_rate=1.0, r0_cells=2.0, H_cells=4.0)
    for k in range(3):
        a.ionize_cloud(k, cl)
        a.step(1)
    ck = str(tmp_path / "ck.npz")
    io.save_checkpoint(ck, a)
    for k in range(3, 5):
        a.ionize_cloud(k, cl)
        a.step(1)
    b = io.load_checkpoint(ck)
    for k in range(3, 5):
        b.ionize_cloud(k, cl)
        b.step(1)
    assert a.n == b.n
    for f in ("b1", "b12", "np", "up", "E", "aj"):
        assert np.array_equal(a.get_field(f), b.get_field(f)), f
    ia, ib = a.particle_array(api.P_ID), b.particle_array(api.P_ID)
    oa, ob = np.argsort(ia), np.argsort(ib)
    assert np.array_equal(ia[oa], ib[ob])
    for which in (api.P_CELL, api.P_FRAC, api.P_VP, api.P_VP1, api.P_BETA):
        xa, xb = a.particle_array(which), b.particle_array(which)
        assert np.array_equal(xa[..., oa], xb[..., ob]), which
