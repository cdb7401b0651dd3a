"""The time loop of maind.f:372-613 over the device library.

Per step m (maind.f:372-613): the sources (Ionize_sw_mp, maind.f:379-382, or
the neutral cloud), the step (maind.f:385-479 as swap_step), the energy
record (Energy_diag, maind.f:390-392, written to c.energy.dat by
maind.f:503-504), the field and particle records every nout steps
(maind.f:506-580) and the checkpoint at mrestart (maind.f:588-607).  The
energy record is taken at the end of the step, from the step's moments at
x^{n+1/2} and its B and E (the reference takes it after the x^n moments).

    python -m swap_amd.driver --cells 64 32 32 --ppc 8 --steps 20 --out run1

runs the uniform drifting-Maxwellian workload (the reference's set-up of
maind.f:120-230 with the constants of inputs.f) and writes the reference's
output files into --out.
"""
import argparse
import json
import os
import sys

from . import api, io, physics


def run(s, nt, mion, b0, *, mstart=0, out_dir=None, nout=0, p=None, sw=None, cloud=None,
        mrestart=0, restart_path=None, log=None):
    """Advance s from step mstart to nt.

    s: a loaded context; mion: the ion mass (kg); b0: the background field
    (b1 units) for EB1; out_dir: where the reference's records go (None: no
    files); nout: field/particle records every nout steps (0: none); p: the
    workload dict (for para.dat); sw: dict(ndot, vsw, beta_pu) for
    Ionize_sw_mp; cloud: a neutral-cloud dict (physics.cloud); mrestart: save
    a checkpoint to restart_path when m == mrestart.  Returns the list of
    energy records."""
    records = []
    ew = ow = None
    if out_dir:
        os.makedirs(out_dir, exist_ok=True)
        if p is not None:
            io.write_para(out_dir, s, p, nt, nout)
        io.write_coord(out_dir, s)
        ew = io.EnergyWriter(out_dir)
        if nout:
            ow = io.OutputWriter(out_dir, s, s.cfg.alpha)
    try:
        s.prime()  # the moments at x^n of the first step
        for m in range(mstart + 1, nt + 1):
            if sw is not None and s.n < s.cfg.capacity:  # maind.f:379-382
                s.ionize_sw(m, sw["ndot"], sw["vsw"], sw.get("beta_pu", 1.0))
            if cloud is not None:
                s.ionize_cloud(m, cloud)
            s.step(1)
            e = s.energy_diag(mion, b0)
            e["m"] = m
            records.append(e)
            if ew is not None:
                ew.write(m, e)
            if ow is not None and m % nout == 0:
                ow.write(m)
            if mrestart and m == mrestart and restart_path:
                io.save_checkpoint(restart_path, s)
            if log:
                log("step %d: %d ions, normalised energy %.6f" % (m, s.n, e["norm_E"]))
    finally:
        if ew is not None:
            ew.close()
        if ow is not None:
            ow.close()
    return records


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--cells", type=int, nargs=3, default=[64, 32, 32])
    ap.add_argument("--ppc", type=int, default=8)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--nout", type=int, default=10)
    ap.add_argument("--out", default=None)
    ap.add_argument("--precision", type=int, default=32)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--open-x", action="store_true",
                    help="inflow/outflow x with the upstream plasma flowing in at x = 0")
    a = ap.parse_args(argv)
    p = physics.workload(*a.cells, a.ppc)
    flags = api.OUTFLOW_X if a.open_x else 0
    cap = p["npart"] * (2 if a.open_x else 1)
    s = physics.make_swap(p, capacity=cap, precision=a.precision, flags=flags, seed=a.seed)
    s.load_maxwellian(p["npart"], p["vsw"], p["vth"])
    if a.open_x:
        s.set_inflow(a.ppc, p["vsw"], p["vth"])
    recs = run(s, a.steps, p["mion"], (0.0, 0.0, p["b0z"]), out_dir=a.out, nout=a.nout, p=p,
               log=lambda t: print("driver: " + t, file=sys.stderr, flush=True))
    print(json.dumps(recs[-1]))


if __name__ == "__main__":
    main()
