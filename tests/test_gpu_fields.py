"""GPU parity of the field routines (gutsf.f, grid_interp.f, boundary.f)
against the FP64 oracle, on random periodic fields."""
import numpy as np
import pytest

from helpers import Pair, rel
from swap_amd import api, physics

pytestmark = pytest.mark.gpu


def _setup(seed=3):
    pr = Pair(load=False)
    o, st, p, s = pr.o, pr.st, pr.p, pr.s
    rng = np.random.default_rng(seed)
    b = o.grid()
    b[...] = rng.standard_normal(b.shape) * 0.1 * p["b0z"]
    b[2] += p["b0z"]
    o.periodic(b)
    st["b1"][...] = b
    st["bt"][...] = b
    b12 = b + rng.standard_normal(b.shape) * 0.01 * p["b0z"]
    o.periodic(b12)
    st["b12"][...] = b12
    npf = o.grid(1)
    npf[...] = p["n0"] * (1 + 0.3 * rng.uniform(-1, 1, npf.shape))
    o.periodic_scalar(npf)
    st["np"][...] = npf
    up = o.grid()
    up[...] = rng.standard_normal(up.shape) * 30
    up[0] += 400
    o.periodic(up)
    st["up"][...] = up
    for nm in ("b1", "b12", "np", "up"):
        s.set_field(nm, st[nm])
    return pr


def test_periodic_and_interpolations():
    pr = _setup()
    s, o, st = pr.s, pr.o, pr.st
    rng = np.random.default_rng(5)
    a = rng.standard_normal(st["up"].shape)
    s.set_field("tmp0", a)
    s.periodic("tmp0")
    ao = a.copy()
    o.periodic(ao)
    assert np.array_equal(s.get_field("tmp0"), ao.astype(np.float32).astype(np.float64))
    # face_to_center (edge field -> nodes)
    s.face_to_center("up", "tmp1")
    vc = o.grid()
    o.face_to_center(st["up"], vc)
    assert rel(s.get_field("tmp1"), vc) < 2e-7
    # edge_to_center (B faces -> nodes)
    s.edge_to_center("b1", "btc")
    btc = o.grid()
    o.edge_to_center(st["b1"], btc)
    assert rel(s.get_field("btc"), btc) < 2e-7
    # edge_to_face
    s.edge_to_face("b1", "tmp2")
    bmf = o.grid()
    o.edge_to_face(st["b1"], bmf)
    assert rel(s.get_field("tmp2"), bmf) < 2e-7
    # crossf2
    s.crossf2("tmp1", "btc", "tmp0")
    cc = o.grid()
    o.crossf2(vc.copy(), btc.copy(), cc)
    assert rel(s.get_field("tmp0"), cc) < 1e-6


def test_curls_get_E_predict_correct():
    pr = _setup()
    s, o, st = pr.s, pr.o, pr.st
    # curlB(bt, np, aj)
    s.curlB("b1", "aj")
    o.curlB(st["bt"], st["np"], st["aj"])
    aj = s.get_field("aj")
    inner = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
    assert rel(aj[inner], st["aj"][inner]) < 2e-6
    # the reference leaves aj's ghosts stale; the device fills them with the
    # periodic images, which is what every reader sees after its periodic()
    o.periodic(st["aj"])
    assert rel(aj, st["aj"]) < 2e-6
    # curlE
    E = o.grid()
    rng = np.random.default_rng(9)
    E[...] = rng.standard_normal(E.shape) * 5
    o.periodic(E)
    s.set_field("E", E)
    s.curlE("E", "tmp0")
    cE = o.grid()
    o.curlE(E, cE)
    assert rel(s.get_field("tmp0")[inner], cE[inner]) < 2e-6
    # get_E
    s.get_E()
    o.get_E(st["E"], st["bt"], st["aj"], st["up"], st["np"], st["nu"])
    assert rel(s.get_field("E"), st["E"]) < 2e-6
    # predict_B: compare the increment b1p2 - b12
    s.set_field("aj", st["aj"])
    s.predict_B()
    o.predict_B(st["b1"], st["b12"], st["b1p2"], st["bt"], st["E"], st["aj"], st["up"],
                st["np"], st["nu"])
    g = s.get_field("b1p2")
    assert rel(g - st["b12"], st["b1p2"] - st["b12"]) < 1e-5
    assert rel(s.get_field("E"), st["E"]) < 2e-6
    # correct_B (from the same b1p2 on both sides)
    s.set_field("b1p2", st["b1p2"])
    s.correct_B()
    o.correct_B(st["b1"], st["b1p2"], st["E"], st["aj"], st["up"], st["np"], st["nu"])
    g = s.get_field("b1p2")
    o.periodic(st["aj"])  # stale ghosts on the reference side (gutsf.f:241)
    assert rel(s.get_field("aj"), st["aj"]) < 2e-6
    assert rel(s.get_field("E"), st["E"]) < 2e-6
    assert rel(g - st["b1"], st["b1p2"] - st["b1"]) < 1e-5
    # f_update_tlev: bt = b1 = b1p2, b12 = old b1
    b1_old = s.get_field("b1")
    s.f_update_tlev()
    assert np.array_equal(s.get_field("b12"), b1_old)
    assert np.array_equal(s.get_field("b1"), g)
    assert np.array_equal(s.get_field("b1p2"), g)


def test_check_time_step_halving():
    pr = _setup()
    s, o, st, p = pr.s, pr.o, pr.st, pr.p
    assert s.check_time_step() == (10, pytest.approx(p["dt"] / 10, rel=1e-6))
    o.check_time_step(st["bt"], st["np"])
    assert o.dtsub()[1] == 10
    # drive a few nodes far past the whistler bound: n -> n/5000 there
    npf = st["np"].copy()
    npf[3, 4, 5] /= 5000
    npf[6, 2, 9] /= 20000
    npf[1, 1, 1] /= 5000
    s.set_field("np", npf)
    ntf, dtsub = s.check_time_step()
    o.set_dtsub(p["dt"] / 10, 10)
    o.check_time_step(st["bt"], npf)
    d_o, n_o = o.dtsub()
    assert ntf == int(n_o) and ntf > 10
    assert dtsub == pytest.approx(d_o, rel=1e-6)


def _periodic_np(a, open_x=False):
    """ghost layers of a storage array [..., nz, ny, nx] (boundary.f:9-52);
    open_x: zero-gradient copies along x instead"""
    for ax in (-1, -2, -3):
        n = a.shape[ax]
        d = [slice(None)] * a.ndim
        s = [slice(None)] * a.ndim
        d[ax], s[ax] = 0, (1 if open_x and ax == -1 else n - 2)
        a[tuple(d)] = a[tuple(s)]
        d[ax], s[ax] = n - 1, (n - 2 if open_x and ax == -1 else 1)
        a[tuple(d)] = a[tuple(s)]
    return a


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("nc,open_x", [((16, 8, 8), False), ((20, 6, 37), False),
                                       ((2, 3, 70), False), ((40, 9, 33), False),
                                       ((16, 8, 8), True), ((4, 6, 37), True),
                                       ((40, 9, 33), True), ((33, 5, 9), True)])
def test_fused_subcycles_equal_per_routine_kernels(precision, nc, open_x):
    """the LDS-tiled predictor/corrector pair (two launches per subcycle), and
    the same tiles in one cooperative launch with grid barriers, give bit for
    bit what the nine per-routine kernels give: B levels, and the E and aj the
    last corrector leaves.  Shapes off the 32x4 tile and the z chunking, and a
    2-cell axis (every halo wraps more than once); with an open x
    (zero-gradient x ghosts), tiles at either face, a grid narrower than one
    tile and one a cell past it."""
    p = physics.workload(*nc, 2)
    out = []
    ox = api.OUTFLOW_X if open_x else 0
    for flags in (ox, ox | api.UNFUSED_FIELDS, ox | api.PERSISTENT_FIELDS):
        s = physics.make_swap(p, precision=precision, flags=flags, capacity=16)
        rng = np.random.default_rng(5)
        shp = (s.nz, s.ny, s.nx)
        b = rng.standard_normal((3,) + shp) * 0.1 * p["b0z"]
        b[2] += p["b0z"]
        b12 = b + rng.standard_normal((3,) + shp) * 0.01 * p["b0z"]
        npf = p["n0"] * (1 + 0.3 * rng.uniform(-1, 1, shp))
        up = rng.standard_normal((3,) + shp) * 30
        up[0] += 400
        for nm, a in (("b1", b), ("b12", b12), ("np", npf), ("up", up)):
            s.set_field(nm, _periodic_np(a, open_x))
        s.field_subcycles(3, p["dt"] / 10)
        out.append({f: s.get_field(f) for f in ("b1", "b12", "b1p2", "E", "aj")})
        s.close()
    for o, name in zip(out[1:], ("per-routine", "cooperative")):
        for f in out[0]:
            assert np.array_equal(out[0][f], o[f]), (name, f)
    # and the subcycles did move B
    assert not np.array_equal(out[0]["b1"], out[0]["b12"])
