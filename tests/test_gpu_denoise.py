"""The learned moment denoiser on the device: its forward and backward kernels
against the CPU restatement (oracle/denoise_ref.py, same BF16 roundings), and
what training buys: on held-out pairs the trained net's moment error is below
that of the plain low-ppc deposit; an untrained net keeps the gate closed."""
import numpy as np
import pytest

import denoise_ref as D
from swap_amd import api, physics

pytestmark = pytest.mark.gpu


def _ctx(nc, ppc=2, flags=api.TRACK_IDS, **kw):
    p = physics.workload(*nc, ppc)
    s = physics.make_swap(p, flags=flags, **kw)
    return p, s


def _grid_to_v4(a):
    """[4][nz][ny][nx] -> [V][4]"""
    return np.ascontiguousarray(a.reshape(4, -1).T)


def test_forward_and_gradients_match_cpu():
    nc = (6, 5, 4)
    p, s = _ctx(nc)
    s.dn_create(p["n0"], p["vsw"], max_pairs=1, stride=1)
    rng = np.random.default_rng(11)
    params = (rng.standard_normal(D.n_params()) * 0.5 / np.sqrt(27 * 8)).astype(np.float32)
    s.dn_set_params(params)
    shape = (nc[2], nc[1], nc[0])
    x = (rng.standard_normal((4,) + shape) * 0.5).astype(np.float32)
    t = (x + rng.standard_normal((4,) + shape) * 0.1).astype(np.float32)
    y = s.dn_infer(x)
    loss, g = s.dn_grad(x, t)
    lo, out, go = D.forward_backward(params, _grid_to_v4(x), _grid_to_v4(t), nc, bf=True)
    assert np.max(np.abs(_grid_to_v4(y) - out)) < 2e-3 * np.max(np.abs(out))
    assert abs(loss - lo) < 1e-3 * lo
    o = 0
    for l in range(D.NL):
        nw = 27 * D.CH[l] * D.CH[l + 1]
        for part in (slice(o, o + nw), slice(o + nw, o + nw + D.CH[l + 1])):
            a, b = g[part].astype(np.float64), go[part]
            err = np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30)
            cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-300)
            assert err < 2e-2 and cos > 0.9999, (l, part, err, cos)
        o += nw + D.CH[l + 1]


def test_disjoint_target_and_pair_transfer():
    """the disjoint target deposits exactly the particles the input leaves
    out: S * np(sub) + (S - 1) * np(rest) = S * np(all) at every node.  The
    edge velocity up averages two nodal ratios (momentum / np), which no
    such identity combines; its mean over the grid is the drift either way.
    Pairs copied out and back in are unchanged."""
    nc, S = (8, 6, 5), 4
    p, s = _ctx(nc, ppc=16, capacity=8 * 6 * 5 * 16)
    s.load_maxwellian(p["npart"], p["vsw"], p["vth"])
    s.dn_create(p["n0"], p["vsw"], max_pairs=4, stride=S)
    s.prime()
    s.step(1)
    s.dn_add_pair()               # target: all particles
    s.dn_add_pair(api.DN_DISJOINT)  # target: id % S != 0
    x0, t_all = s.dn_get_pair(0)
    x1, t_rest = s.dn_get_pair(1)
    assert np.array_equal(x0, x1)
    n_sub, n_rest, n_all = x0[0] + 1.0, t_rest[0] + 1.0, t_all[0] + 1.0  # np / n0
    assert np.allclose(n_sub + (S - 1) * n_rest, S * n_all, rtol=2e-5, atol=2e-5 * S)
    # the drift (up_x = vsw, normalised 1) with thermal noise 30/400/sqrt(N)
    assert abs(t_rest[1].mean() - t_all[1].mean()) < 5e-3
    assert abs(t_rest[1].mean() - 1.0) < 2e-2
    # a disjoint target differs from the full one: the subsample is not in it
    assert not np.allclose(t_all, t_rest)
    s.dn_put_pair(x0, t_rest)
    x2, t2 = s.dn_get_pair(2)
    assert np.array_equal(x2, x0) and np.array_equal(t2, t_rest)
    raw_a, _ = s.dn_eval(1, 2)
    raw_b, _ = s.dn_eval(2, 3)
    assert raw_a == pytest.approx(raw_b, rel=1e-12)  # FP64 atomic sums, any order


def test_training_beats_plain_low_ppc_and_gate():
    """pairs from the code's own particles: a 4-ppc subsample (stride 16) of
    a 64-ppc plasma with a pickup cloud; train on 16 pairs, evaluate on 8"""
    nc = (16, 16, 16)
    p, s = _ctx(nc, ppc=64, flags=api.TRACK_IDS | api.SPECIES,
                capacity=16 * 16 * 16 * 64 + 400000)
    s.load_maxwellian(p["npart"], p["vsw"], p["vth"])
    cl = physics.cloud(p, N0_rate=2.0, r0_cells=2.0, H_cells=4.0)
    s.dn_create(p["n0"], p["vsw"], max_pairs=24, stride=16, seed=3)
    s.prime()
    for k in range(24):
        s.ionize_cloud(k, cl)
        s.step(1)
        s.dn_add_pair()
    assert s.dn_num_pairs() == 24
    raw0, net0 = s.dn_eval(16, 24)
    # a net that provably loses: all weights zero and an output bias of 1/2
    # make the residual net f(x) = x + 1/2, whose error exceeds the raw one
    # by 1/4 per value; the gate closes and the step keeps plain deposition
    init = s.dn_params()
    bad = np.zeros_like(init)
    bad[-4:] = 0.5
    s.dn_set_params(bad)
    raw_b, net_b = s.dn_eval(16, 24)
    assert raw_b == pytest.approx(raw0) and net_b > raw_b + 0.2
    with pytest.raises(api.SwapError, match="gate"):
        s.dn_enable(True)
    s.dn_set_params(init)
    s.dn_train(400, 2e-3, 16)
    raw, net = s.dn_eval(16, 24)
    assert raw == pytest.approx(raw0)
    assert net < 0.5 * raw, (raw, net)
    s.dn_enable(True)
    np_before = s.get_field("np").copy()
    s.ionize_cloud(99, cl)
    s.step(1)  # the step now runs on denoised moments
    assert np.isfinite(s.get_field("np")).all() and not np.array_equal(np_before,
                                                                       s.get_field("np"))
    s.dn_enable(False)


def test_disjoint_target_from_the_step_sums():
    """SWAP_DN_DISJOINT_STEP: the input is the subsample deposited where the
    step's final deposit took place (x^{n+1/2}, half a move back), checked
    against the FP64 oracle's update_np of those positions; the target is
    the step's sums minus the subsample's, so S np(sub) + (S-1) np(rest) =
    S np of the step; the mode is refused before a step and once particles
    or moment sums changed after it"""
    import oracle as O
    from helpers import rel
    nc, S = (12, 8, 6), 4
    p = physics.workload(*nc, 16)
    s = physics.make_swap(p, flags=api.TRACK_IDS | api.SPECIES, capacity=p["npart"] + 20000)
    s.load_maxwellian(p["npart"], p["vsw"], p["vth"])
    s.dn_create(p["n0"], p["vsw"], max_pairs=4, stride=S)
    with pytest.raises(api.SwapError):
        s.dn_add_pair(api.DN_DISJOINT_STEP)
    s.prime()
    s.step(1)
    np_step = s.get_field("np")
    s.dn_add_pair(api.DN_DISJOINT_STEP)
    s.dn_add_pair(api.DN_DISJOINT_STEP)  # nothing changed in between
    x0, t0 = s.dn_get_pair(0)
    x1, t1 = s.dn_get_pair(1)
    assert np.array_equal(x0, x1) and np.array_equal(t0, t1)
    assert np.array_equal(s.get_field("np"), np_step)  # the context keeps its np
    inner = (slice(1, -1),) * 3
    n_sub, n_rest = x0[0] + 1.0, t0[0] + 1.0
    assert np.allclose(n_sub + (S - 1) * n_rest, S * np_step[inner] / p["n0"], rtol=2e-5,
                       atol=2e-5 * S)
    # the subsample at x^{n+1/2}, deposited by the oracle with weights x S
    sel = s.particle_array(api.P_ID) % S == 0
    xp = np.ascontiguousarray(s.xp()[:, sel])
    back = np.ascontiguousarray(-s.vp()[:, sel].astype(np.float64))
    n = xp.shape[1]
    o = O.Oracle(s.nx, s.ny, s.nz, p["dx"], p["dx"], p["dx"], p["dt"], p["ntsub"], p["alpha"],
                 p["beta"] / S)
    o.move_ion_half(n, n, xp, back)
    ijkp, wght, npo = np.zeros((3, n), np.int32), o.parts(n, 8), o.grid(1)
    o.get_interp_weights(n, n, xp, ijkp, wght)
    o.update_np(n, n, ijkp, wght, None, npo)
    assert rel(n_sub * p["n0"], npo[inner]) < 1e-5
    # refused once the sums or the particles changed
    s.update_np()
    with pytest.raises(api.SwapError):
        s.dn_add_pair(api.DN_DISJOINT_STEP)
    s.step(1)
    s.dn_add_pair(api.DN_DISJOINT_STEP)
    s.ionize_cloud(3, physics.cloud(p, N0_rate=1.0, r0_cells=2.0, H_cells=4.0))
    with pytest.raises(api.SwapError):
        s.dn_add_pair(api.DN_DISJOINT_STEP)
