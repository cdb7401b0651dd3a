"""The Fortran boundary: maind.f's call sequence written against the reference
routine names (swap_amd/fortran/swap_f.f90, strict mode over the C-ABI) runs
one step on the device, optionally opened by Ionize_sw_mp as maind.f:379-383
does; the result must match the FP64 oracle's step on the same seeded load
(and the oracle's restatement of Ionize_sw_mp), to the FP32 tolerances of the
fused-step test."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from helpers import rel
from swap_amd import physics

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "swap_amd", "fortran", "_build", "swap_step_strict")
SEED = 20261203


@pytest.mark.parametrize("ionize", [False, True])
def test_fortran_strict_step_matches_oracle(tmp_path, ionize):
    if not os.path.exists(EXE):
        pytest.fail("swap_step_strict not built: run __graft_entry__.build()")
    nc = (16, 8, 8)
    ppc = 4
    p = physics.workload(*nc, ppc)
    n = p["npart"]
    pf = tmp_path / "params.txt"
    # Ionize_sw_mp's rate ndot = 1e-6 nf_init omega_p (chem_rates.f:551): an
    # nf_init that gives about 300 ring protons of weight beta_pu = 1/2
    beta_pu, mstep = 0.5, 3
    per_nf = 1e-6 * p["omega_p"] * p["dt"] * p["beta"] * beta_pu * p["dx"] ** 3 * (nc[2] + 2)
    nf_init = 300.0 / per_nf if ionize else 0.0
    pf.write_text("%d %d %d %d %.17g %.17g %d %.17g %.17g %.17g %.17g %.17g %d %.17g %.17g %d\n"
                  % (nc[0], nc[1], nc[2], ppc, p["dx"], p["dt"], p["ntsub"], p["alpha"],
                     p["beta"], p["nu"], p["vsw"], p["vth"], SEED, nf_init, beta_pu, mstep))
    of = tmp_path / "out.bin"
    r = subprocess.run([EXE, str(pf), str(of)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    nx, ny, nz = nc[0] + 2, nc[1] + 2, nc[2] + 2
    m = int(np.fromfile(of, np.int32, count=1)[0])
    raw = np.fromfile(of, np.float32, offset=4)
    sizes = [nx * ny * nz, 3 * nx * ny * nz, 3 * nx * ny * nz, 3 * m, 3 * m, m, m]
    assert raw.size == sum(sizes)
    parts = np.split(raw, np.cumsum(sizes)[:-1])
    np_f = parts[0].reshape(nz, ny, nx)
    up_f = parts[1].reshape(3, nz, ny, nx)
    b1_f = parts[2].reshape(3, nz, ny, nx)
    xp_f = parts[3].reshape(3, m)
    vp_f = parts[4].reshape(3, m)
    bp_f, mr_f = parts[5], parts[6]

    # the oracle on the same seeded load (the device sampler is pinned to it
    # by the loader test: identical cells and fractions, v within one ulp)
    cell, frac, v = O.load_maxwellian(n, SEED, nc, p["vsw"], p["vth"])
    i, j, k = cell & 2047, (cell >> 11) & 1023, cell >> 21
    xp = np.stack([(i + 1 + frac[0]) * p["dx"], (j + 1 + frac[1]) * p["dx"],
                   (k + 1 + frac[2]) * p["dx"]])
    bp = np.ones(n)
    if ionize:
        ndot = 1e-6 * nf_init * p["omega_p"]
        c1, f1, v1 = O.ionize_sw(nc, (p["dx"],) * 3, p["dt"], p["beta"], ndot, p["vsw"],
                                 beta_pu, SEED, mstep)
        i1, j1, k1 = c1 & 2047, (c1 >> 11) & 1023, c1 >> 21
        xp = np.concatenate([xp, np.stack([(i1 + 1 + f1[0]) * p["dx"],
                                           (j1 + 1 + f1[1]) * p["dx"],
                                           (k1 + 1 + f1[2]) * p["dx"]])], axis=1)
        v = np.concatenate([v, v1], axis=1)
        bp = np.concatenate([bp, np.full(c1.size, beta_pu)])
        assert c1.size > 200
    assert m == xp.shape[1]
    assert np.all(bp_f == bp.astype(np.float32)) and np.all(mr_f == 1.0)
    o = O.Oracle(nx, ny, nz, p["dx"], p["dx"], p["dx"], p["dt"], p["ntsub"], p["alpha"],
                 p["beta"])
    st = {nm: o.grid() for nm in ("b1", "b12", "b1p2", "bt", "btc", "up", "aj", "E")}
    for b in ("b1", "b12", "b1p2", "bt"):
        st[b][2] = p["b0z"]
    st.update(np=o.grid(1), nu=np.full((nz, ny, nx), p["nu"]), xp=np.ascontiguousarray(xp),
              vp=np.ascontiguousarray(v), vp1=v.copy(), vplus=o.parts(m), vminus=o.parts(m),
              Ep=o.parts(m), wght=o.parts(m, 8), ijkp=np.zeros((3, m), np.int32), Ni_tot=m,
              beta_p=bp, mrat=np.ones(m))
    o.step(st)

    L = np.array(nc)[:, None] * p["dx"]
    d = np.abs(xp_f - st["xp"])
    d = np.minimum(d, L - d)
    assert np.max(d / L) < 1e-5
    assert np.max(np.abs(vp_f - st["vp"])) / 400.0 < 1e-5
    assert rel(np_f, st["np"]) < 2e-4
    assert rel(up_f, st["up"]) < 2e-3
    # b1 comes back as REAL*4, as the reference holds it: a few FP32 ulps of |B|
    assert rel(b1_f, st["b1"]) < 1e-6
