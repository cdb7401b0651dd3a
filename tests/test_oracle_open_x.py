"""The oracle's open x (build-defined inflow/outflow; the reference is periodic
in every axis, gutsp.f:490-510, boundary.f:9-52), pinned by what the rules
fix without restating them:

* a quiet lattice (a uniform plasma) has uniform density and drift up to both
  faces, which pins the fold target and the two face share factors;
* the charge of any load equals the particle count once each x node is
  weighted by its share of the particle domain;
* with no fields an ion keeps its velocity, so the ions that leave are
  exactly those whose straight path crosses a face;
* the inflow's count and mean speed follow the kinetic flux of a drifting
  Maxwellian through a plane (closed forms with the normal CDF).
"""
import math

import numpy as np
import pytest

import oracle as O

PAR = O.initparameters()
NC = (8, 6, 4)


def make(nc=NC, npart=None, width=64, open_x=True):
    n_p = npart or 8 * nc[0] * nc[1] * nc[2]
    vol = nc[0] * nc[1] * nc[2] * PAR["dx"] ** 3
    beta = n_p / (vol * 0.1e15)
    o = O.Oracle(nc[0] + 2, nc[1] + 2, nc[2] + 2, PAR["dx"], PAR["dx"], PAR["dx"], PAR["dt"],
                 10, PAR["alpha"], beta, width=width)
    o.set_open_x(open_x)
    return o, beta


def lattice(o, k):
    """k^3 ions per cell at the centres of a k-fold sub-lattice"""
    q = [o.coord(a) for a in ("qx", "qy", "qz")]
    axes = [q[a][0] + (np.arange(k * (len(q[a]) - 2)) + 0.5) / k * PAR["dx"] for a in range(3)]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    return np.ascontiguousarray(np.stack([X.ravel(), Y.ravel(), Z.ravel()]))


def deposit(o, xp, vp=None):
    n = xp.shape[1]
    ijkp = np.zeros((3, n), np.int32)
    w = o.parts(n, 8)
    o.get_interp_weights(n, n, xp, ijkp, w)
    npg = o.grid(1)
    bp = np.ones(n)
    o.update_np(n, n, ijkp, w, bp, npg)
    up = None
    if vp is not None:
        up = o.grid()
        o.update_up(n, n, ijkp, w, bp, np.ascontiguousarray(vp), npg, up)
    return npg, up


def test_open_x_ghosts_are_zero_gradient_and_yz_stay_periodic():
    o, _ = make()
    g = np.random.default_rng(1).standard_normal(o.grid().shape)
    o.periodic(g)
    assert np.array_equal(g[..., 0], g[..., 1])
    assert np.array_equal(g[..., -1], g[..., -2])
    assert np.array_equal(g[:, :, 0, :], g[:, :, -2, :])   # y still periodic
    assert np.array_equal(g[:, 0, :, :], g[:, -2, :, :])   # z still periodic
    s = np.random.default_rng(2).standard_normal(o.grid(1).shape)
    o.periodic_scalar(s)
    assert np.array_equal(s[..., 0], s[..., 1]) and np.array_equal(s[..., -1], s[..., -2])


@pytest.mark.parametrize("k", [2, 3])
def test_quiet_lattice_is_uniform_up_to_both_faces(k):
    o, beta = make()
    xp = lattice(o, k)
    n = xp.shape[1]
    V = np.array([400.0, -12.0, 3.5])
    vp = np.repeat(V[:, None], n, axis=1)
    npg, up = deposit(o, xp, vp)
    n0 = k ** 3 / (PAR["dx"] ** 3 * beta)
    assert np.allclose(npg[1:-1, 1:-1, 1:-1], n0, rtol=1e-12)
    # without the face factors the first and last x nodes would be 1.5 n0 and n0 / 2
    for c in range(3):
        assert np.allclose(up[c, 1:-1, 1:-1, 1:-1], V[c], rtol=1e-12, atol=1e-12)


def test_charge_with_face_shares_equals_particle_count():
    o, beta = make(npart=3000)
    rng = np.random.default_rng(3)
    q = [o.coord(a) for a in ("qx", "qy", "qz")]
    xp = np.stack([q[a][0] + (q[a][-2] - q[a][0]) * (1.0 - rng.random(3000))
                   for a in range(3)])
    npg, _ = deposit(o, np.ascontiguousarray(xp))
    share = np.ones(NC[0])
    share[0], share[-1] = 1.5, 0.5
    total = (npg[1:-1, 1:-1, 1:-1] * share).sum() * PAR["dx"] ** 3 * beta
    assert total == pytest.approx(3000, rel=1e-12)


def zero_field_state(o, n, rng):
    q = [o.coord(a) for a in ("qx", "qy", "qz")]
    xp = np.stack([q[a][0] + (q[a][-2] - q[a][0]) * (1.0 - rng.random(n)) for a in range(3)])
    vp = rng.standard_normal((3, n)) * 150.0 + np.array([[40.0], [0.0], [0.0]])
    st = {k: o.grid() for k in ("b1", "b12", "b1p2", "bt", "btc", "up", "aj", "E")}
    st["np"] = o.grid(1)
    st["nu"] = np.full(o.grid(1).shape, PAR["nu_init"])
    st.update(xp=np.ascontiguousarray(xp), vp=np.ascontiguousarray(vp), vp1=vp.copy(),
              vplus=o.parts(n), vminus=o.parts(n), Ep=o.parts(n), wght=o.parts(n, 8),
              ijkp=np.zeros((3, n), np.int32), Ni_tot=n, id=np.arange(n, dtype=np.uint32))
    return st


def test_field_free_ions_leave_exactly_where_their_path_crosses_a_face():
    o, _ = make(npart=2000)
    rng = np.random.default_rng(5)
    n = 2000
    st = zero_field_state(o, n, rng)
    x0, v0 = st["xp"].copy(), st["vp"].copy()
    o.step(st)
    qx = o.coord("qx")
    dth = PAR["dt"] / 2
    x1 = x0[0] + dth * v0[0]
    x2 = x1 + dth * v0[0]
    inside = lambda x: (x > qx[0]) & (x <= qx[-2])  # noqa: E731
    keep = inside(x1) & inside(x2)
    m = st["Ni_tot"]
    assert st["escaped"] == n - m == int((~keep).sum()) > 0
    assert np.array_equal(st["id"][:m], np.nonzero(keep)[0].astype(np.uint32))
    assert np.array_equal(st["vp"][:, :m], v0[:, keep])      # no field: v unchanged
    assert np.array_equal(st["xp"][0, :m], x2[keep])
    # both faces lose ions: the spread of speeds reaches back against the drift
    assert (x2[~keep] > qx[-2]).any() and (x2[~keep] <= qx[0]).any()


def _phi(a):
    return math.exp(-0.5 * a * a) / math.sqrt(2 * math.pi)


def _Phi(a):
    return 0.5 * (1 + math.erf(a / math.sqrt(2)))


@pytest.mark.parametrize("vsw,vth", [(400.0, 30.0), (50.0, 80.0)])
def test_inflow_follows_the_kinetic_flux(vsw, vth):
    nc = (16, 8, 8)
    d = (PAR["dx"],) * 3
    dt, ppc = PAR["dt"], 6.0
    s = dt / d[0]
    counts, vx = [], []
    for step in range(40):
        th, te, v, flag = O.inflow(7, step, nc, dt, d, ppc, vsw, vth)
        keep = (flag & 2) > 0
        dep = (flag & 1) > 0
        counts.append(int(keep.sum()))
        vx.append(v[0, keep])
        # kept: entered during the step; deposits: entered in its first half
        assert (te[0, keep] > 0).all() and (te[0, keep] <= v[0, keep] * s + 1e-12).all()
        assert (th[0, dep] > 0).all() and (th[0, dep] <= v[0, dep] * s / 2 + 1e-12).all()
        assert np.allclose(te[0] - th[0], v[0] * s / 2, rtol=0, atol=1e-12)
        for a in (1, 2):
            assert (te[a] > 0).all() and (te[a] <= nc[a]).all()
    # the sampler draws v = vsw + vth sqrt(-ln u) cos(pi u'): a normal of sd vth / sqrt 2
    sig = vth / math.sqrt(2)
    a = vsw / sig
    flux = vsw * _Phi(a) + sig * _phi(a)             # E[v; v > 0]
    mean = ppc * nc[1] * nc[2] * s * flux
    assert abs(np.mean(counts) - mean) < 4 * math.sqrt(mean / len(counts))
    m2 = (vsw ** 2 + sig ** 2) * _Phi(a) + vsw * sig * _phi(a)  # E[v^2; v > 0]
    allv = np.concatenate(vx)
    sd = math.sqrt(np.var(allv) / len(allv))
    assert abs(allv.mean() - m2 / flux) < 5 * sd
    # a fixed (seed, step) repeats; another step draws anew
    a1 = O.inflow(7, 3, nc, dt, d, ppc, vsw, vth)
    a2 = O.inflow(7, 3, nc, dt, d, ppc, vsw, vth)
    assert all(np.array_equal(x, y) for x, y in zip(a1, a2))
    a3 = O.inflow(7, 4, nc, dt, d, ppc, vsw, vth)
    assert a3[0].shape != a1[0].shape or not np.array_equal(a1[0], a3[0])


def test_field_free_inflow_keeps_the_half_step_density_uniform():
    """a uniform drifting plasma with matching inflow stays uniform up to the
    inflow face at x^{n+1/2}, where the step's moments are taken: the ions
    entering in the first half of the step are in them"""
    nc = (12, 4, 4)
    ppc = 64
    n = ppc * nc[0] * nc[1] * nc[2]
    o, beta = make(nc=nc, npart=n)
    rng = np.random.default_rng(9)
    vsw, vth = 300.0, 40.0
    st = zero_field_state(o, n, rng)
    # the loader's distribution: a normal of sd vth / sqrt 2 about (vsw, 0, 0)
    st["vp"][:] = vth / math.sqrt(2) * rng.standard_normal((3, n))
    st["vp"][0] += vsw
    st["vp1"] = st["vp"].copy()
    cap = 3 * n
    for k in ("xp", "vp", "vp1", "vplus", "vminus", "Ep"):
        a = np.zeros((3, cap))
        a[:, :n] = st[k]
        st[k] = a
    st["wght"] = np.zeros((8, cap))
    st["ijkp"] = np.zeros((3, cap), np.int32)
    ids = np.zeros(cap, np.uint32)
    ids[:n] = np.arange(n)
    st["id"] = ids
    o.set_inflow(11, ppc, vsw, vth, next_id=n)
    n0 = ppc / (PAR["dx"] ** 3 * beta)
    profs, added = [], 0
    for step in range(6):
        o.step(st)
        assert st["injected"] > 0 and st["escaped"] > 0
        added += st["injected"]
        profs.append(st["np"][1:-1, 1:-1, 1:-1].mean(axis=(0, 1)) / n0)
    prof = np.mean(profs[2:], axis=0)
    # noise of a face node average: ~1 / sqrt(ppc ny nz share) ~ 2-4 %
    assert np.abs(prof - 1).max() < 0.12, prof
    m = st["Ni_tot"]
    live = st["id"][:m]
    assert len(np.unique(live)) == m and live.max() < n + added


def test_cloud_ions_mirror_at_an_open_x_face():
    nc = (6, 4, 4)
    cnt = np.zeros(nc[0] * nc[1] * nc[2], np.uint32)
    cnt[nc[0] - 1::nc[0]] = 50      # the last x node of every (y, z) column
    cp, fp = O.cloud_ions(nc, cnt, 5, 3)
    co, fo = O.cloud_ions(nc, cnt, 5, 3, open_x=True)
    tp = (cp & 2047) + fp[0]
    to = (co & 2047) + fo[0]
    wrapped = tp < 1.0                  # periodic: past the face, into cell 0
    assert wrapped.any()
    assert np.allclose(to[wrapped], 2 * nc[0] - (tp[wrapped] + nc[0]), rtol=0, atol=1e-12)
    assert np.array_equal(to[~wrapped], tp[~wrapped])
    assert (to > nc[0] - 0.5).all() and (to <= nc[0]).all()
