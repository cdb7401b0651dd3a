"""The oracle's Energy_diag (part_init.f:14-139) pinned by closed forms:
uniform fields and uniform ions give each energy term in one line of
arithmetic, and with no fields an open box with inflow keeps every ion's
energy, so the input account equals the ions' energy at every step
(normalised energy 1)."""
import math

import numpy as np
import pytest

import oracle as O

PAR = O.initparameters()
Q, MU0, EPS = 1.6e-19, 3.14159 * 4e-7, 8.85e-12
NC = (6, 4, 5)


def make(npart, open_x=False, nc=NC):
    vol = nc[0] * nc[1] * nc[2] * PAR["dx"] ** 3
    beta = npart / (vol * 0.1e15)
    o = O.Oracle(nc[0] + 2, nc[1] + 2, nc[2] + 2, PAR["dx"], PAR["dx"], PAR["dx"], PAR["dt"],
                 10, PAR["alpha"], beta, width=64)
    o.set_open_x(open_x)
    return o, beta


def state(o, n, rng=None, v=(0.0, 0.0, 0.0)):
    st = {k: o.grid() for k in ("b1", "b12", "b1p2", "bt", "btc", "up", "aj", "E")}
    st["np"] = o.grid(1)
    st["nu"] = np.full(o.grid(1).shape, PAR["nu_init"])
    q = [o.coord(a) for a in ("qx", "qy", "qz")]
    rng = rng or np.random.default_rng(1)
    xp = np.stack([q[a][0] + (q[a][-2] - q[a][0]) * (1.0 - rng.random(n)) for a in range(3)])
    vp = np.repeat(np.asarray(v, np.float64)[:, None], n, axis=1)
    st.update(xp=np.ascontiguousarray(xp), vp=np.ascontiguousarray(vp), vp1=vp.copy(),
              vplus=o.parts(n), vminus=o.parts(n), Ep=o.parts(n), wght=o.parts(n, 8),
              ijkp=np.zeros((3, n), np.int32), Ni_tot=n)
    return st


def test_energy_terms_of_uniform_fields_and_ions():
    n = 500
    o, beta = make(n)
    mion = PAR["mion"]
    st = state(o, n, v=(300.0, -40.0, 12.0))
    b = np.array([0.003, -0.001, 0.0192])
    b0 = np.array([0.0, 0.0, 0.0192])
    E = np.array([0.5, 0.2, -0.1])
    U = np.array([410.0, 3.0, -7.0])
    n0 = 0.1e15
    for c in range(3):
        st["b1"][c] = b[c]
        st["E"][c] = E[c]
        st["up"][c] = U[c]
    st["np"][:] = n0
    st["beta_p"] = np.full(n, 0.5)
    st["mrat"] = np.full(n, 0.25)
    d = o.energy_diag(st, mion, b0)
    nodes = NC[0] * NC[1] * NC[2]
    vol = PAR["dx"] ** 3 * 1e9
    mq = mion / Q
    for c, k in enumerate(("EB1x", "EB1y", "EB1z")):
        assert d[k] == pytest.approx(nodes * vol / (2 * MU0) * (mq * b[c]) ** 2, rel=1e-12)
    assert d["EB1"] == pytest.approx(nodes * vol / (2 * MU0) * mq ** 2 *
                                     np.sum((b - b0) ** 2), rel=1e-12)
    assert d["EE"] == pytest.approx(nodes * EPS * vol / 2 * mq ** 2 * 1e6 * np.sum(E ** 2),
                                    rel=1e-12)
    assert d["Euf"] == pytest.approx(nodes * 0.5 * mion * n0 / 1e9 * vol * 1e6 * np.sum(U ** 2),
                                     rel=1e-12)
    # each ion: mass mion / mrat, weight 1 / (beta beta_p)
    V2 = 300.0 ** 2 + 40.0 ** 2 + 12.0 ** 2
    assert d["Evp"] == pytest.approx(n * 0.5 * (mion / 0.25) * V2 * 1e6 / (beta * 0.5),
                                     rel=1e-12)
    assert d["total_E"] == pytest.approx(d["Evp"] + d["EE"] + d["EB1"], rel=1e-14)


def test_field_free_open_box_with_inflow_keeps_the_energy_account():
    nc = (10, 4, 4)
    ppc = 32
    n = ppc * nc[0] * nc[1] * nc[2]
    o, beta = make(n, open_x=True, nc=nc)
    rng = np.random.default_rng(4)
    vsw, vth = 250.0, 60.0
    st = state(o, n, rng)
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
    st["id"] = np.zeros(cap, np.uint32)
    st["input_k"] = o.ke(st)
    o.set_inflow(3, ppc, vsw, vth, next_id=n)
    mion = PAR["mion"]
    for step in range(5):
        o.step(st)
        assert st["escaped"] > 0 and st["injected"] > 0
        d = o.energy_diag(st, mion, (0.0, 0.0, 0.0))
        assert d["EB1"] == 0 and d["EE"] == 0
        assert d["norm_E"] == pytest.approx(1.0, rel=1e-12), step
