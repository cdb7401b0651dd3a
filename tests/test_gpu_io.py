"""Output files and checkpoints of a device run: the per-output records hold
the device fields and particles (FP32, reference layouts), and a run resumed
from a checkpoint continues bit for bit like the uninterrupted run; the load
file is the identical seeded load for both codes."""
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


@pytest.mark.parametrize("precision", [32, 64])
def test_load_file_is_the_identical_seeded_load(tmp_path, precision):
    """the device's seeded load, saved as a load file, is the oracle's seeded
    load when the oracle's C reader reads it, and a second context loaded
    from the file steps bit for bit like the first"""
    import oracle as O
    from helpers import Pair
    a = Pair(precision=precision, flags=api.TRACK_IDS)
    p = a.p
    path = str(tmp_path / "load.bin")
    io.save_load(path, a.s)
    cell, frac, v, v1, _, _, ids = O.read_load(path)
    oc, of, ov = O.load_maxwellian(a.n, a.s.cfg.seed, (p["ncx"], p["ncy"], p["ncz"]),
                                   p["vsw"], p["vth"])
    rd = np.float32 if precision == 32 else np.float64
    assert np.array_equal(cell, oc)
    assert np.array_equal(frac, of.astype(rd).astype(np.float64))
    assert np.max(np.abs(v - ov.astype(rd))) <= 1e-9 * p["vsw"]
    assert np.array_equal(ids, np.arange(a.n, dtype=np.uint32))
    b = Pair(precision=precision, flags=api.TRACK_IDS, load=False)
    io.load_file(b.s, path)
    for s in (a.s, b.s):
        s.step(2)
    # FP32 moments are exact fixed-point sums (bit for bit); the FP64 build
    # adds doubles with atomics, whose order may differ between runs
    for f in ("np", "up", "E", "b1"):
        x, y = a.s.get_field(f), b.s.get_field(f)
        if precision == 32:
            assert np.array_equal(x, y), f
        else:
            assert np.max(np.abs(x - y)) <= 1e-12 * np.max(np.abs(y)), f
    if precision == 32:
        assert np.array_equal(a.s.particle_array(api.P_CELL), b.s.particle_array(api.P_CELL))
