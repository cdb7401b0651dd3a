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


@pytest.mark.parametrize("open_x", [True, False])
@pytest.mark.parametrize("precision", [32, 64])
def test_checkpoint_mid_sort_interval_resumes_bit_for_bit(tmp_path, precision, open_x):
    """Sorting every 4 steps, periodic or an open x with inflow: a checkpoint
    after step 2 holds the step count (it keys the inflow draws), the sort
    schedule with its work items, boxes and velocity bounds, the id counter
    and the counts, so the resumed run takes the same steps as the
    uninterrupted one: the same books, and in the FP32 build the same arrays
    in the same order and the same fields bit for bit.  The FP64 build sums
    its moments with FP64 atomics, whose order varies from run to run, so two
    FP64 runs -- resumed or not -- agree to rounding (1e-12 here)."""
    p = physics.workload(16, 8, 8, 8)
    n = p["npart"]
    a = physics.make_swap(p, capacity=2 * n, precision=precision, sort_every=4,
                          flags=api.TRACK_IDS | (api.OUTFLOW_X if open_x else 0))
    a.load_maxwellian(n, p["vsw"], p["vth"])
    if open_x:
        a.set_inflow(p["ppc"], p["vsw"], p["vth"])
    a.step(2)
    ck = str(tmp_path / "ck.npz")
    io.save_checkpoint(ck, a)
    a.step(3)
    b = io.load_checkpoint(ck)
    b.step(3)
    da, db = a.diagnostics(), b.diagnostics()
    for k in ("nparticles", "escaped", "injected", "sorts", "window_misses"):
        assert da[k] == db[k], k
    assert da["sorts"] >= 2
    assert (da["injected"] > 0 and da["escaped"] > 0) == open_x
    if precision == 32:
        for f in ("b1", "b12", "np", "up", "E", "aj"):
            assert np.array_equal(a.get_field(f), b.get_field(f)), f
        for which in (api.P_ID, api.P_CELL, api.P_FRAC, api.P_VP, api.P_VP1):
            assert np.array_equal(a.particle_array(which), b.particle_array(which)), which
    else:
        for f in ("b1", "b12", "np", "up", "E", "aj"):
            fa, fb = a.get_field(f), b.get_field(f)
            assert np.max(np.abs(fa - fb)) <= 1e-12 * np.max(np.abs(fa)), f
        ia, ib = a.particle_array(api.P_ID), b.particle_array(api.P_ID)
        oa, ob = np.argsort(ia), np.argsort(ib)
        assert np.array_equal(ia[oa], ib[ob])
        for arr_a, arr_b in ((a.xp(), b.xp()), (a.vp(), b.vp())):
            d = np.abs(arr_a[:, oa] - arr_b[:, ob])
            assert np.max(d) <= 1e-12 * np.max(np.abs(arr_a)), d.max()
    ea, eb = a.energy_diag(p["mion"], (0, 0, p["b0z"])), b.energy_diag(p["mion"], (0, 0, p["b0z"]))
    assert ea["input_E"] == pytest.approx(eb["input_E"], rel=1e-14)


def test_run_state_refuses_another_load():
    p, a = _run()
    st = a.run_state()
    _, b = _run(nc=(12, 8, 6), ppc=4)
    with pytest.raises(api.SwapError, match="particles"):
        b.restore_run_state(st)
    bad = st.copy()
    bad[:8] = 0
    with pytest.raises(api.SwapError, match="not a run state"):
        a.restore_run_state(bad)


@pytest.mark.parametrize("f", [0.0, 1.5, -0.25, float("nan")])
def test_load_split_refuses_fractions_outside_the_cell(f):
    """fractions lie in (0, 1]: anything else would give negative or
    extrapolated weights and silently wrong deposits"""
    p = physics.workload(8, 8, 8, 1)
    s = physics.make_swap(p, capacity=16)
    cell = np.zeros(4, np.uint32)
    frac = np.full((3, 4), 0.5, np.float32)
    v = np.zeros((3, 4), np.float32)
    s.load_split(cell, frac, v, v)
    frac[1, 2] = f
    with pytest.raises(api.SwapError, match="fraction"):
        s.load_split(cell, frac, v, v)


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
