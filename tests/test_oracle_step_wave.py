"""Closed-form pins of the oracle's composite step (or_step, maind.f:385-479).

1. A small circularly polarised wave along B0 in a cold ion plasma.  The
   step's equations -- ions dv/dt = mrat (E + v x B), E = (aj - up) x B
   (massless electrons, nu = 0), aj = curl B / (alpha n), dB/dt = -curl E --
   linearise, for B_+ = Bx + i By ~ exp(i(kz - omega t)) about B0 z, to
   w^2 + kappa^2 w - kappa^2 = 0 with w = omega / Omega_h, Omega_h = mrat B0
   and kappa^2 = k^2 / (alpha n mrat): the ion-cyclotron branch (0 < w < 1)
   and the whistler branch (w < -1/2 here).  The ions' velocity of the mode is
   v_+ = -(k mrat B0 / (omega alpha n)) B_+.  A quiet start (ions on a lattice
   of 2x2x2 per cell, velocities of the mode) turns through about one radian;
   the measured frequency must be within 2.5 % of the closed form and the
   amplitude within 5 % (measured: within 1.4 % and 1.3 %).  Heavy ions
   (mrat = 1/2) carried at beta_p = 1/2 (each ion counts double) are
   included.
   Planted in scratch copies of the oracle, each of these fails it by 5 % or
   more: the Hall term's sign, the mrat factor of get_Ep or of the rotation
   in get_vplus_vminus, 1/beta_p in update_np, a full instead of a half dt Ep
   in v- (gutsp.f:302-305), no half dt Ep in get_vp_final (gutsp.f:433-449),
   and the field subcycles moved before the particle push.  The second
   predictor pass and the time level of the moments move the frequency by
   less than 1 % at Omega dt = 0.05 and are not resolved by this pin.
2. A uniform cold drift is an exact equilibrium of the step: E = -V x B0
   cancels the force, so every ion moves by exactly V dt per step (two half
   moves of V dt / 2, maind.f:413,479) with v, np, up and B unchanged.
"""
import math

import numpy as np
import pytest

import oracle as O

PAR = O.initparameters()
N0 = 0.1e15


def _wave_run(mrat, bp, branch, vz0=0.0, ncz=32, amp=1e-3):
    d, dt = PAR["dx"], PAR["dt"]
    nc = (2, 2, ncz)
    lat = (2, 2, 2)
    n = int(np.prod(nc)) * int(np.prod(lat))
    beta = n / (np.prod(nc) * d ** 3 * N0)  # ions of weight 1/(beta bp): density N0/bp
    o = O.Oracle(nc[0] + 2, nc[1] + 2, nc[2] + 2, d, d, d, dt, 10, PAR["alpha"], beta)
    B0 = PAR["omega_p"]  # B carried as q B / m_p
    dens = N0 / bp
    k = 2 * math.pi / (ncz * d)
    kap2 = k * k / (PAR["alpha"] * dens * mrat)
    w = (-kap2 + branch * math.sqrt(kap2 * kap2 + 4 * kap2)) / 2
    om = w * mrat * B0
    om_lab = om + k * vz0
    # quiet start: a 2x2x2 lattice per cell (fractions 1/4, 3/4)
    cells = np.stack(np.meshgrid(*[np.arange(c) for c in nc], indexing="ij"), -1).reshape(-1, 3)
    sub = np.stack(np.meshgrid(*[(np.arange(m) + 0.5) / m for m in lat], indexing="ij"),
                   -1).reshape(-1, 3)
    pos = (cells[:, None, :] + sub[None, :, :]).reshape(-1, 3)
    xp = np.ascontiguousarray(((pos + 1.0) * d).T)  # q(i) = i d
    z = xp[2]
    C = -k * mrat * B0 / (om * PAR["alpha"] * dens)
    A = amp * B0
    vp, vp1 = np.zeros((3, n)), np.zeros((3, n))
    vp[0], vp[1], vp[2] = C * A * np.cos(k * z), C * A * np.sin(k * z), vz0
    ph1 = k * (z - vz0 * dt) + om_lab * dt  # the same ions one step earlier
    vp1[0], vp1[1], vp1[2] = C * A * np.cos(ph1), C * A * np.sin(ph1), vz0
    st = {nm: o.grid() for nm in ("b1", "b12", "b1p2", "bt", "btc", "up", "aj", "E")}
    st["np"] = o.grid(1)
    st["nu"] = np.zeros((o.nz, o.ny, o.nx))
    zz = (np.arange(1, o.nz + 1) * d)[:, None, None] * np.ones((o.nz, o.ny, o.nx))
    h = d / 2  # Bx, By sit at z + d/2 (face-centred B, grid_interp.f:35-51)
    for nm, t in (("b1", 0.0), ("bt", 0.0), ("b1p2", 0.0), ("b12", -dt / 10)):
        b = st[nm]
        b[2] = B0
        b[0] = A * np.cos(k * (zz + h) - om_lab * t)
        b[1] = A * np.sin(k * (zz + h) - om_lab * t)
    st.update(xp=xp, vp=vp, vp1=vp1, vplus=o.parts(n), vminus=o.parts(n), Ep=o.parts(n),
              wght=o.parts(n, 8), ijkp=np.zeros((3, n), np.int32), Ni_tot=n,
              beta_p=np.full(n, float(bp)), mrat=np.full(n, float(mrat)))
    zf = (np.arange(1, ncz + 1) + 1) * d + h

    def mode():
        bpl = st["b1"][0, 1:-1, 1, 1] + 1j * st["b1"][1, 1:-1, 1, 1]
        return np.sum(bpl * np.exp(-1j * k * zf)) / ncz
    a = [mode()]
    nsteps = int(math.ceil(1.0 / (abs(om) * dt)))  # about one radian in the plasma frame
    for _ in range(nsteps):
        o.step(st)
        a.append(mode())
    a = np.array(a)
    ph = np.unwrap(np.angle(a))
    t = np.arange(a.size) * dt
    return -np.polyfit(t, ph, 1)[0], om_lab, abs(a[-1]) / abs(a[0]), abs(ph[-1] - ph[0])


@pytest.mark.parametrize("mrat,bp", [(1.0, 1.0), (0.5, 0.5)])
@pytest.mark.parametrize("branch", [1, -1])
def test_parallel_wave_follows_the_hybrid_dispersion(mrat, bp, branch):
    om_meas, om_want, amp, turned = _wave_run(mrat, bp, branch)
    assert turned > 0.9  # the mode turned far enough to measure its rate
    assert om_meas == pytest.approx(om_want, rel=0.025)
    assert amp == pytest.approx(1.0, abs=0.05)


def test_uniform_cold_drift_moves_each_ion_by_v_dt():
    d, dt = PAR["dx"], PAR["dt"]
    nc = (8, 6, 4)
    n = 8 * int(np.prod(nc))
    beta = n / (np.prod(nc) * d ** 3 * N0)
    o = O.Oracle(nc[0] + 2, nc[1] + 2, nc[2] + 2, d, d, d, dt, 10, PAR["alpha"], beta)
    # a lattice of 2x2x2 ions per cell: uniform np, so up = V exactly
    cells = np.stack(np.meshgrid(*[np.arange(c) for c in nc], indexing="ij"), -1).reshape(-1, 3)
    sub = np.stack(np.meshgrid(*[(np.arange(2) + 0.5) / 2] * 3, indexing="ij"), -1).reshape(-1, 3)
    pos = (cells[:, None, :] + sub[None, :, :]).reshape(-1, 3)
    xp = np.ascontiguousarray(((pos + 1.0) * d).T)
    V = np.array([400.0, 30.0, -20.0])
    vp = np.ascontiguousarray(np.repeat(V[:, None], n, 1))
    st = {nm: o.grid() for nm in ("b1", "b12", "b1p2", "bt", "btc", "up", "aj", "E")}
    for b in ("b1", "b12", "b1p2", "bt"):
        st[b][2] = PAR["omega_p"]
    st.update(np=o.grid(1), nu=np.full((o.nz, o.ny, o.nx), PAR["nu_init"]), xp=xp.copy(),
              vp=vp.copy(), vp1=vp.copy(), vplus=o.parts(n), vminus=o.parts(n),
              Ep=o.parts(n), wght=o.parts(n, 8), ijkp=np.zeros((3, n), np.int32), Ni_tot=n)
    steps = 3
    for _ in range(steps):
        o.step(st)
    L = np.array(nc)[:, None] * d
    want = (xp - d + steps * V[:, None] * dt) % L + d
    dist = np.abs(st["xp"] - want)
    assert np.max(np.minimum(dist, L - dist)) < 1e-9 * d
    assert np.allclose(st["vp"], vp, rtol=0, atol=1e-9)
    for c in range(3):
        assert np.allclose(st["up"][c, 1:-1, 1:-1, 1:-1], V[c], rtol=1e-12, atol=1e-9)
    assert np.allclose(st["b1"][2], PAR["omega_p"], rtol=1e-12)
    assert np.allclose(st["b1"][:2], 0, atol=1e-15)
    assert np.allclose(st["np"][1:-1, 1:-1, 1:-1], N0, rtol=1e-12)
