"""On-disk formats of the reference's output, and a checkpoint of the device
state.

Reference formats (Fortran unformatted sequential, as written by maind.f
with -i4 -real-size 32: each WRITE is one record framed by 4-byte length
markers; INTEGER and REAL are 4 bytes; arrays in Fortran order, which is the
memory order of this package's numpy arrays, e.g. a grid (nx,ny,nz,3) is
(3,nz,ny,nx) here):

  para.dat        maind.f:237-255   (nx,ny,nz,dx,dy,delz) (nt,dtsub_init,ntsub,dt,nout)
                                    (out_dir, character(50)) (vtop,vbottom) (Ni_max)
                                    (mion,m_pu,m_heavy) (np_top,np_bottom)
                                    (b0_top,b0_bottom) (vth_top,vth_bottom) (alpha,beta)
  c.coord.dat     initial.f:285-297 nx; ny; nz; qx; qy; qz; dz_grid; dz_cell
  c.np.dat        maind.f:529-530   m; np(nx,ny,nz)
  c.b1.dat        maind.f:537-539   m; bt(nx,ny,nz,3)
  c.aj.dat        maind.f:540-541   m; aj*alpha*np3 (the curl of B, edge current)
  c.E.dat         maind.f:542-543   m; E(nx,ny,nz,3)
  c.up.dat        maind.f:545-546   m; up(nx,ny,nz,3)
  c.xp_0.dat      maind.f:549-550   m; xp(Ni_max,3)
  c.vp_0.dat      maind.f:551-552   m; vp(Ni_max,3)
  c.mrat_0.dat    maind.f:553-554   m; mrat(Ni_max)
  c.beta_p_0.dat  maind.f:555-556   m; beta_p(Ni_max)
  c.energy.dat    maind.f:503-504   m; input_E, input_EeP, Evp, Euf, EB1, EB1x, EB1y, EB1z,
                                    EE, EeP, input_chex, input_bill (12 REAL per step)

The energy record's electron-pressure and chemistry terms (input_EeP, EeP,
input_chex, input_bill) are not tracked by the time-step path and are
written as 0.

Particle records hold Ni_max (the context capacity) entries, the first n
live, the rest zero, as the reference writes its whole arrays.  The other
files of maind.f:268-333 hold quantities outside the time-step path
(temperatures, separated populations, energy budgets) and are not written.

Checkpoint (replaces restart.vars, maind.f:588-607): an .npz of the
configuration, the B levels b1 and b12, the split particle arrays (cell,
fraction, vp, vp1, beta_p, mrat, id) in the device's order, and the
library's run state (swap_save_run_state: step count, sort schedule and work
items, id counter, injected/escaped counts, inflow parameters, input-energy
account).  The moments, E and the node sources are recomputed by the next
step, so a run resumed from a checkpoint continues as the uninterrupted run
does, also in the middle of a sort interval and with an open x and inflow:
bit for bit in the FP32 build (exact fixed-point moment sums), to rounding in
the FP64 build (its FP64 atomics add in a varying order, so no two FP64 runs
are bit-identical).  The learned
denoiser's weights are not part of it (swap_dn_get_params / set_params).
"""
import os
import struct

import numpy as np

from . import api

_NAMES = {"np": "c.np.dat", "b1": "c.b1.dat", "aj": "c.aj.dat", "E": "c.E.dat",
          "up": "c.up.dat", "xp": "c.xp_0.dat", "vp": "c.vp_0.dat", "mrat": "c.mrat_0.dat",
          "beta_p": "c.beta_p_0.dat"}


def write_record(f, *items):
    """one Fortran unformatted record: 4-byte length, payload, 4-byte length"""
    payload = b"".join(_pack(x) for x in items)
    n = struct.pack("<i", len(payload))
    f.write(n + payload + n)


def _pack(x):
    if isinstance(x, (bytes, bytearray)):
        return bytes(x)
    if isinstance(x, str):
        return x.encode()
    if isinstance(x, (bool, int, np.integer)):
        return struct.pack("<i", int(x))
    if isinstance(x, (float, np.floating)):
        return struct.pack("<f", float(x))
    a = np.asarray(x)
    if a.dtype.kind in "iu":
        return np.ascontiguousarray(a, "<i4").tobytes()
    return np.ascontiguousarray(a, "<f4").tobytes()


def read_records(path):
    """all records of a Fortran unformatted sequential file, as bytes"""
    out = []
    with open(path, "rb") as f:
        while True:
            h = f.read(4)
            if not h:
                return out
            (n,) = struct.unpack("<i", h)
            out.append(f.read(n))
            (t,) = struct.unpack("<i", f.read(4))
            if t != n:
                raise ValueError("%s: record markers %d != %d" % (path, n, t))


def write_para(out_dir, s, p, nt, nout, m_pu=28.0, m_heavy=1.0):
    """para.dat (maind.f:237-255); p is a workload dict (physics.workload)"""
    os.makedirs(out_dir, exist_ok=True)
    name = (out_dir.rstrip("/") + "/")[:50].ljust(50)
    with open(os.path.join(out_dir, "para.dat"), "wb") as f:
        write_record(f, s.nx, s.ny, s.nz, p["dx"], p["dx"], p["dx"])
        write_record(f, nt, p["dt"] / p["ntsub"], p["ntsub"], p["dt"], nout)
        write_record(f, name)
        write_record(f, p["vsw"], -p["vsw"])
        write_record(f, s.cfg.capacity)
        write_record(f, p["mion"], m_pu, m_heavy)
        write_record(f, p["n0"], p["n0"] / m_heavy)
        write_record(f, p["b0z"], p["b0z"])
        write_record(f, p["vth"], p["vth"])
        write_record(f, p["alpha"], p["beta"])


def write_coord(out_dir, s):
    """c.coord.dat (initial.f:285-297): uniform grid, dz_grid = dz_cell = dz"""
    os.makedirs(out_dir, exist_ok=True)
    qx, qy, qz = (s.coords(a) for a in range(3))
    dz = np.full(s.nz, qz[2] - qz[1])
    with open(os.path.join(out_dir, "c.coord.dat"), "wb") as f:
        for x in (s.nx, s.ny, s.nz, qx, qy, qz, dz, dz):
            write_record(f, x)


class OutputWriter:
    """the per-output records of maind.f:527-556 for one context"""

    def __init__(self, out_dir, s, alpha):
        os.makedirs(out_dir, exist_ok=True)
        self.s, self.alpha = s, alpha
        self.f = {k: open(os.path.join(out_dir, v), "wb") for k, v in _NAMES.items()}

    def close(self):
        for f in self.f.values():
            f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _particles(self, arr, ncomp):
        cap = self.s.cfg.capacity
        out = np.zeros((ncomp, cap) if ncomp > 1 else (cap,), np.float32)
        out[..., :arr.shape[-1]] = arr
        return out

    def write(self, m):
        s = self.s
        npf = s.get_field("np")
        fields = {"np": npf, "b1": s.get_field("b1"), "E": s.get_field("E"),
                  "up": s.get_field("up")}
        # aj * alpha * np3, np3 the edge density of curlB (gutsf.f:196-207)
        aj = s.get_field("aj")
        np3 = np.empty_like(aj)
        for ax, m_ in ((-1, 0), (-2, 1), (-3, 2)):
            np3[m_] = 0.5 * (npf + np.roll(npf, -1, axis=ax))
        fields["aj"] = aj * self.alpha * np3
        n = s.n
        parts = {"xp": self._particles(s.xp(), 3), "vp": self._particles(s.vp(), 3)}
        flags = s.cfg.flags
        ones = np.ones(n, np.float32)
        parts["mrat"] = self._particles(s.particle_array(api.P_MRAT) if flags & api.SPECIES
                                        else ones, 1)
        parts["beta_p"] = self._particles(s.particle_array(api.P_BETA) if flags & api.SPECIES
                                          else ones, 1)
        for k, a in list(fields.items()) + list(parts.items()):
            write_record(self.f[k], m)
            write_record(self.f[k], a)


ENERGY_RECORD = ("input_E", "input_EeP", "Evp", "Euf", "EB1", "EB1x", "EB1y", "EB1z", "EE",
                 "EeP", "input_chex", "input_bill")


class EnergyWriter:
    """c.energy.dat (maind.f:287, 503-504): per step, m then 12 REAL"""

    def __init__(self, out_dir):
        os.makedirs(out_dir, exist_ok=True)
        self.f = open(os.path.join(out_dir, "c.energy.dat"), "wb")

    def write(self, m, e):
        """e: a swap_energy_diag dict"""
        write_record(self.f, m)
        write_record(self.f, np.array([e.get(k, 0.0) for k in ENERGY_RECORD], np.float32))

    def close(self):
        self.f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_energy(out_dir):
    """(m, dict) per record of c.energy.dat"""
    recs = read_records(os.path.join(out_dir, "c.energy.dat"))
    out = []
    for a, b in zip(recs[::2], recs[1::2]):
        (m,) = struct.unpack("<i", a)
        out.append((m, dict(zip(ENERGY_RECORD, np.frombuffer(b, "<f4").tolist()))))
    return out


def read_output(out_dir, name, shape):
    """(m, array) pairs of one c.*.dat file; shape of one record's array"""
    recs = read_records(os.path.join(out_dir, _NAMES.get(name, name)))
    out = []
    for a, b in zip(recs[::2], recs[1::2]):
        (m,) = struct.unpack("<i", a)
        out.append((m, np.frombuffer(b, "<f4").reshape(shape)))
    return out


# ---------------- load file ----------------
#
# The seeded initial load in a form any code can read with plain binary I/O
# (the oracle reads it in C, oracle/swap_oracle.c or_read_load), so that
# both codes start from one file.  Little-endian:
#   header (64 bytes): char magic[8] = "SWAPLD01"; int64 n; int32 ncx, ncy,
#     ncz; int32 flags (1: beta_p and mrat follow, 2: ids follow); uint64
#     seed; float64 dx, dy, dz
#   uint32 cell[n] (i | j<<11 | k<<21, 0-based cells)
#   float64 frac[3][n] (in-cell offsets in (0,1], right-closed cells)
#   float64 vp[3][n], vp1[3][n] (km/s)
#   [float32 beta_p[n], mrat[n]] [uint32 id[n]]
# The absolute position is x = (i + 1 + frac) dx (q(i) = i dx, grd7
# initial.f:193-196 on a uniform grid).

LOAD_MAGIC = b"SWAPLD01"


def write_load(path, cell, frac, vp, vp1, nc, d, seed=0, beta_p=None, mrat=None, ids=None):
    cell = np.ascontiguousarray(cell, "<u4")
    n = cell.size
    flags = (1 if beta_p is not None else 0) | (2 if ids is not None else 0)
    with open(path, "wb") as f:
        f.write(LOAD_MAGIC)
        f.write(struct.pack("<q3ii", n, nc[0], nc[1], nc[2], flags))
        f.write(struct.pack("<Q3d", seed, d[0], d[1], d[2]))
        f.write(cell.tobytes())
        for a in (frac, vp, vp1):
            f.write(np.ascontiguousarray(np.asarray(a, np.float64).reshape(3, n), "<f8").tobytes())
        if beta_p is not None:
            f.write(np.ascontiguousarray(beta_p, "<f4").tobytes())
            f.write(np.ascontiguousarray(np.ones(n) if mrat is None else mrat, "<f4").tobytes())
        if ids is not None:
            f.write(np.ascontiguousarray(ids, "<u4").tobytes())


def read_load(path):
    """dict of the load file's header fields and arrays"""
    with open(path, "rb") as f:
        if f.read(8) != LOAD_MAGIC:
            raise ValueError("%s: not a SWAPLD01 load file" % path)
        n, ncx, ncy, ncz, flags = struct.unpack("<q3ii", f.read(24))
        seed, dx, dy, dz = struct.unpack("<Q3d", f.read(32))
        out = {"n": n, "nc": (ncx, ncy, ncz), "seed": seed, "d": (dx, dy, dz)}
        out["cell"] = np.frombuffer(f.read(4 * n), "<u4").copy()
        for k in ("frac", "vp", "vp1"):
            out[k] = np.frombuffer(f.read(24 * n), "<f8").reshape(3, n).copy()
        if flags & 1:
            out["beta_p"] = np.frombuffer(f.read(4 * n), "<f4").copy()
            out["mrat"] = np.frombuffer(f.read(4 * n), "<f4").copy()
        if flags & 2:
            out["id"] = np.frombuffer(f.read(4 * n), "<u4").copy()
    return out


def save_load(path, s):
    """the context's particles as a load file"""
    sp = s.cfg.flags & api.SPECIES
    write_load(path, s.particle_array(api.P_CELL), s.particle_array(api.P_FRAC),
               s.particle_array(api.P_VP), s.particle_array(api.P_VP1),
               (s.cfg.ncx, s.cfg.ncy, s.cfg.ncz), (s.cfg.dx, s.cfg.dy, s.cfg.dz), s.cfg.seed,
               s.particle_array(api.P_BETA) if sp else None,
               s.particle_array(api.P_MRAT) if sp else None,
               s.particle_array(api.P_ID) if s.cfg.flags & api.TRACK_IDS else None)


def load_file(s, path):
    """load a load file into the context (same grid)"""
    z = read_load(path)
    if tuple(z["nc"]) != (s.cfg.ncx, s.cfg.ncy, s.cfg.ncz):
        raise ValueError("%s: grid %s, context %s" % (path, z["nc"],
                                                     (s.cfg.ncx, s.cfg.ncy, s.cfg.ncz)))
    s.load_split(z["cell"], z["frac"], z["vp"], z["vp1"], z.get("beta_p"), z.get("mrat"))
    if "id" in z and s.cfg.flags & api.TRACK_IDS:
        s.set_particle_array(api.P_ID, z["id"])


# ---------------- checkpoint ----------------

_CFG = ("ncx", "ncy", "ncz", "dx", "dy", "dz", "dt", "ntsub", "alpha", "beta", "nu",
        "capacity", "precision", "flags", "sort_every", "seed")


def save_checkpoint(path, s):
    cfg = {k: getattr(s.cfg, k) for k in _CFG}
    arrays = {"b1": s.get_field("b1"), "b12": s.get_field("b12"),
              "cell": s.particle_array(api.P_CELL), "frac": s.particle_array(api.P_FRAC),
              "vp": s.particle_array(api.P_VP), "vp1": s.particle_array(api.P_VP1)}
    if s.cfg.flags & api.SPECIES:
        arrays["beta_p"] = s.particle_array(api.P_BETA)
        arrays["mrat"] = s.particle_array(api.P_MRAT)
    if s.cfg.flags & api.TRACK_IDS:
        arrays["id"] = s.particle_array(api.P_ID)
    arrays["run_state"] = s.run_state()
    np.savez(path, cfg=np.array([cfg[k] for k in _CFG], np.float64),
             cfg_int=np.array([int(cfg[k]) for k in ("ncx", "ncy", "ncz", "ntsub", "capacity",
                                                    "precision", "flags", "sort_every")],
                              np.int64),
             seed=np.array([s.cfg.seed], np.uint64), **arrays)


def load_checkpoint(path):
    """a new context in the checkpointed state"""
    z = np.load(path, allow_pickle=False)
    c = dict(zip(_CFG, z["cfg"]))
    ncx, ncy, ncz, ntsub, cap, prec, flags, sort_every = (int(v) for v in z["cfg_int"])
    s = api.Swap(ncx, ncy, ncz, c["dx"], c["dt"], c["alpha"], c["beta"], c["nu"], cap,
                 dy=c["dy"], dz=c["dz"], ntsub=ntsub, precision=prec, flags=flags,
                 sort_every=sort_every, seed=int(z["seed"][0]))
    s.set_field("b1", z["b1"])
    s.set_field("b12", z["b12"])
    s.load_split(z["cell"], z["frac"], z["vp"], z["vp1"],
                 z["beta_p"] if "beta_p" in z.files else None,
                 z["mrat"] if "mrat" in z.files else None)
    if "id" in z.files:
        s.set_particle_array(api.P_ID, z["id"])
    s.restore_run_state(z["run_state"])
    return s
