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


def test_training_beats_plain_low_ppc_and_gate():
    """pairs from the code's own particles: a 4-ppc subsample (stride 16) of
    a 64-ppc plasma with a pickup cloud; train on 16 pairs, judge on 8"""
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
    # untrained: the residual net starts near the identity, no better than raw
    if not net0 < raw0:
        with pytest.raises(api.SwapError):
            s.dn_enable(True)
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
