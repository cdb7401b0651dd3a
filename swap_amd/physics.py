"""Physical set-up of a run: inputs.f's derived parameters and the workloads.

initparameters (inputs.f:125-172) in FP64 with the reference's constants
(q = 1.6e-19 C, m_p = 1.67e-27 kg, PI = 3.14159, mu0 = PI 4e-7,
eps0 = 8.85e-12, c = 3e8).  B is carried as q B / m_p (s^-1) and positions
in km (initial.f:36,63).
"""
import math

Q = 1.6e-19
MP = 1.67e-27
PI_REF = 3.14159
MU0 = PI_REF * 4.0e-7


def initparameters(b0_init=0.2e-9, ion_amu=1, nf_init=0.1e15, dt_frac=0.05, dx_frac=0.5,
                   nu_init_frac=0.002, ntsub=10):
    mion = ion_amu * MP
    omega_p = Q * b0_init / mion
    lambda_i = (3e8 / math.sqrt((nf_init / 1e9) * Q * Q / (8.85e-12 * mion))) / 1e3
    dx = lambda_i * dx_frac
    dt = dt_frac * mion / (Q * b0_init)
    return {
        "mion": mion, "omega_p": omega_p, "lambda_i": lambda_i, "dx": dx, "dt": dt,
        "dtsub": dt / ntsub, "nu": nu_init_frac * omega_p,
        "alpha": (MU0 / 1e3) * Q * (Q / mion), "b0z": b0_init * Q / mion,
        "n0": nf_init, "ntsub": ntsub,
    }


def workload(ncx, ncy, ncz, ppc, vsw=400.0, vth=30.0, **kw):
    """Uniform drifting-Maxwellian proton run, e.g. 64x32x32 cells x 8 ppc
    or 256x128x128 cells x 64 ppc.  beta = N_p / (V n0) makes
    the deposited density n0 (the explicit form of get_beta, misc.f:462-475)."""
    p = initparameters(**kw)
    npart = ncx * ncy * ncz * ppc
    vol = ncx * ncy * ncz * p["dx"] ** 3
    p.update(ncx=ncx, ncy=ncy, ncz=ncz, ppc=ppc, npart=npart, vsw=vsw, vth=vth,
             beta=npart / (vol * p["n0"]))
    return p


def cloud(p, N0_rate=50.0, r0_cells=4.0, H_cells=16.0, tau_photo=1.0e6, m_pu=28.0,
          max_per_cell=8, beta_min=0.05, rmin_cells=0.5):
    """The pickup-ion source of the cloud workload: a spherical neutral cloud
    n_n(r) = N0 (r0/r)^2 exp(-r/H) at the domain centre, photoionization time
    tau_photo, charge exchange sigma = 8e-26 km^2 (chem_rates.f:817), 28-amu
    ions (mrat = 1/28, cf. chem_rates.f:379-380).  N0 is set so that a node at
    r = r0 produces N0_rate new ions per step; at most max_per_cell (heavier,
    beta_p < 1) ions are born per node and step, so ppc near the core grows
    to hundreds while the solar-wind background stays at its own ppc."""
    d = p["dx"]
    rate = d ** 3 * p["beta"] * p["dt"] / tau_photo  # new ions per unit density
    r0 = r0_cells * d
    H = H_cells * d
    N0 = N0_rate / (rate * math.exp(-r0 / H))
    return {"cx": (p["ncx"] / 2 + 1) * d, "cy": (p["ncy"] / 2 + 1) * d,
            "cz": (p["ncz"] / 2 + 1) * d, "N0": N0, "r0": r0, "H": H,
            "tau_photo": tau_photo, "sigma_chex": 8e-26, "mrat_pu": 1.0 / m_pu,
            "beta_min": beta_min, "max_per_cell": max_per_cell, "rmin": rmin_cells * d}


def make_swap(p, capacity=None, **kw):
    """A Swap context with B = B0 zhat everywhere (grd6_setup + f_update_tlev)."""
    from .api import Swap, B1, B12
    import numpy as np
    s = Swap(p["ncx"], p["ncy"], p["ncz"], p["dx"], p["dt"], p["alpha"], p["beta"], p["nu"],
             capacity or p["npart"], ntsub=p["ntsub"], **kw)
    b = np.zeros((3, s.nz, s.ny, s.nx))
    b[2] = p["b0z"]
    s.set_field(B1, b)
    s.set_field(B12, b)
    return s
