"""ctypes binding of the CPU oracle (oracle/swap_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package swap_amd.

Arrays follow the reference's Fortran layouts, held as C-ordered numpy arrays
with the axes reversed: a vector grid (nx,ny,nz,3) is np.empty((3,nz,ny,nx)),
a particle array (Ni_max,3) is np.empty((3,Ni_max)), ijkp is int32 (3,Ni_max)
with 1-based values, wght is (8,Ni_max).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")


def build():
    """Compile both oracle widths (gcc; a few seconds)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load(width):
    path = os.path.join(_BUILD, "liboracle_f%d.so" % width)
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    real = C.c_double if width == 64 else C.c_float
    vp = C.c_void_p
    i32, i64, dbl = C.c_int, C.c_int64, C.c_double
    sig = {
        "or_create": ([i32, i32, i32, dbl, dbl, dbl, dbl, i32, dbl, dbl], vp),
        "or_destroy": ([vp], None),
        "or_real_size": ([], i32),
        "or_get_coord": ([vp, i32, vp], None),
        "or_set_dtsub": ([vp, dbl, dbl], None),
        "or_get_dtsub": ([vp, vp, vp], None),
        "or_initparameters": ([dbl, i32, dbl, dbl, dbl, dbl, i32, vp], None),
        "or_grd6_b0": ([vp, dbl, dbl, vp], None),
        "or_periodic": ([vp, vp], None),
        "or_periodic_scalar": ([vp, vp], None),
        "or_edge_to_center": ([vp, vp, vp], None),
        "or_edge_to_face": ([vp, vp, vp], None),
        "or_face_to_center": ([vp, vp, vp], None),
        "or_crossf2": ([vp, vp, vp, vp], None),
        "or_curlB": ([vp, vp, vp, vp], None),
        "or_curlE": ([vp, vp, vp], None),
        "or_get_E": ([vp, vp, vp, vp, vp, vp, vp], None),
        "or_predict_B": ([vp] + [vp] * 9, None),
        "or_get_Ep1": ([vp] + [vp] * 7, None),
        "or_correct_B": ([vp] + [vp] * 7, None),
        "or_f_update_tlev": ([vp, vp, vp, vp, vp], None),
        "or_check_time_step": ([vp, vp, vp], None),
        "or_get_interp_weights": ([vp, i32, i32, vp, vp, vp], None),
        "or_update_np": ([vp, i32, i32, vp, vp, vp, vp], None),
        "or_update_up": ([vp, i32, i32, vp, vp, vp, vp, vp, vp], None),
        "or_extrapol_up": ([vp, i32, i32, vp, vp, vp, vp, vp, vp, vp], None),
        "or_get_Ep": ([vp, i32, i32, vp, vp, vp, vp, vp, vp, vp], None),
        "or_get_vplus_vminus": ([vp, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp], None),
        "or_improve_up": ([vp, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp], None),
        "or_get_vp_final": ([vp, i32, i32, vp, vp, vp, vp], None),
        "or_move_ion_half": ([vp, i32, i32, vp, vp], None),
        "or_step": ([vp, vp], None),
        "or_set_open_x": ([vp, i32], None),
        "or_set_threads": ([vp, i32], None),
        "or_read_load": ([C.c_char_p, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp], i32),
        "or_energy_diag": ([vp, vp, dbl, vp, vp], None),
        "or_ke": ([vp, vp, i32, i32], dbl),
        "or_inflow": ([C.c_uint64, C.c_uint32, vp, dbl, vp, dbl, dbl, dbl, vp, vp, vp, vp,
                       i64], i64),
        "or_set_inflow": ([vp, C.c_uint64, dbl, dbl, dbl, dbl, C.c_uint32], None),
        "or_philox4x32_10": ([vp, vp, vp], None),
        "or_split_move": ([i64, vp, vp, vp, vp, vp, vp, vp, vp, vp], None),
        "or_stable_sort": ([i64, vp, C.c_uint32, vp], None),
        "or_load_maxwellian": ([i64, C.c_uint64, C.c_uint32, vp, dbl, dbl, vp, vp, vp], None),
        "or_ionize_sw": ([vp, dbl, dbl, dbl, dbl, dbl, dbl, dbl, dbl, C.c_uint64, C.c_uint32,
                          vp, vp, vp, i64], i64),
        "or_cloud_counts": ([vp, dbl, dbl, dbl, dbl, dbl, vp, i32, C.c_uint64, C.c_uint32, vp,
                             vp], None),
        "or_cloud_ions": ([vp, vp, C.c_uint64, C.c_uint32, i32, vp, vp, i64], None),
        "or_chex": ([i64, vp, vp, vp, vp, vp, vp, vp, vp, dbl, dbl, dbl, dbl, vp, C.c_uint64,
                     C.c_uint32, vp], None),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    return lib, real


def _p(a):
    return None if a is None else a.ctypes.data


ENERGY_KEYS = ("input_E", "Evp", "Euf", "EB1", "EB1x", "EB1y", "EB1z", "EE", "total_E",
               "norm_E")


class OrState(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "b1", "b12", "b1p2", "bt", "btc", "np", "up", "aj", "E", "nu",
        "xp", "vp", "vp1", "vplus", "vminus", "Ep", "wght", "beta_p", "mrat",
        "ijkp", "id")] + [("Ni_tot", C.c_int), ("Ni_max", C.c_int), ("escaped", C.c_int),
                          ("injected", C.c_int), ("input_k", C.c_double)]


class Oracle:
    """One oracle grid context.  `width` is 64 (the FP64 reference arithmetic) or 32."""

    COORDS = ("qx", "qy", "qz", "dx_grid", "dy_grid", "dz_grid",
              "dx_cell", "dy_cell", "dz_cell", "xrat", "yrat", "zrat")

    def __init__(self, nx, ny, nz, dx, dy, dz, dt, ntsub, alpha, beta, width=64):
        self.lib, self._real = _load(width)
        self.dtype = np.float64 if width == 64 else np.float32
        self.nx, self.ny, self.nz = nx, ny, nz
        self.h = self.lib.or_create(nx, ny, nz, dx, dy, dz, dt, ntsub, alpha, beta)
        self.dt = dt

    def __del__(self):
        try:
            self.lib.or_destroy(self.h)
        except Exception:
            pass

    # -- allocation helpers --
    def grid(self, ncomp=3):
        shape = (self.nz, self.ny, self.nx) if ncomp == 1 else (ncomp, self.nz, self.ny, self.nx)
        return np.zeros(shape, self.dtype)

    def parts(self, n, ncomp=3):
        return np.zeros((ncomp, n), self.dtype)

    def coord(self, name):
        axis = name[0] if name.endswith("rat") else name[1]  # qx, dx_grid, dx_cell, xrat
        n = {"x": self.nx, "y": self.ny, "z": self.nz}[axis]
        out = np.zeros(n, self.dtype)
        self.lib.or_get_coord(self.h, self.COORDS.index(name), _p(out))
        return out

    def _chk(self, *arrs):
        for a in arrs:
            if a is not None:
                assert a.flags.c_contiguous, "oracle arrays must be C-contiguous"
                assert a.dtype in (self.dtype, np.int32, np.uint32), a.dtype

    def dtsub(self):
        d, n = C.c_double(), C.c_double()
        self.lib.or_get_dtsub(self.h, C.byref(d), C.byref(n))
        return d.value, n.value

    def set_dtsub(self, dtsub, ntf):
        self.lib.or_set_dtsub(self.h, dtsub, ntf)

    # -- reference routines --
    def __getattr__(self, name):
        # or_<name> with numpy arrays converted to pointers
        if name.startswith("_") or name in ("lib", "h"):
            raise AttributeError(name)
        fn = getattr(self.lib, "or_" + name)

        def call(*args):
            conv = []
            for a in args:
                if isinstance(a, np.ndarray):
                    self._chk(a)
                    conv.append(a.ctypes.data)
                else:
                    conv.append(a)
            return fn(self.h, *conv)
        return call

    def step(self, st):
        """st: dict of numpy arrays (see OrState); advances one step in place."""
        s = self._state(st)
        self.lib.or_step(self.h, C.byref(s))
        # an open x removes escaped ions: the live ones are the first Ni_tot
        st["Ni_tot"] = s.Ni_tot
        st["escaped"] = s.escaped
        st["injected"] = s.injected
        st["input_k"] = s.input_k

    def _state(self, st):
        s = OrState()
        for n, _ in OrState._fields_[:-5]:
            a = st.get(n)
            if a is not None:
                want = np.int32 if n == "ijkp" else (np.uint32 if n == "id" else self.dtype)
                if a.dtype != want or not a.flags.c_contiguous:
                    raise TypeError("oracle state %r: need a C-contiguous %s array, got %s"
                                    % (n, np.dtype(want).name, a.dtype))
            setattr(s, n, _p(a))
        s.Ni_tot = st["Ni_tot"]
        s.Ni_max = st["xp"].shape[1]
        s.input_k = st.get("input_k", 0.0)
        return s

    def ke(self, st, l0=1, l1=None):
        """kinetic energy / ion mass of ions l0 <= l < l1 (1-based)"""
        s = self._state(st)
        return self.lib.or_ke(self.h, C.byref(s), l0, st["Ni_tot"] + 1 if l1 is None else l1)

    def energy_diag(self, st, mion, b0):
        """Energy_diag's record as a dict (J)"""
        s = self._state(st)
        out = np.zeros(10)
        b0 = np.asarray(b0, np.float64)
        self.lib.or_energy_diag(self.h, C.byref(s), mion, b0.ctypes.data, out.ctypes.data)
        return dict(zip(ENERGY_KEYS, out.tolist()))

    def set_threads(self, n):
        """OpenMP threads of the oracle's loops (1: serial)"""
        self.lib.or_set_threads(self.h, int(n))

    def set_open_x(self, on=True):
        """build-defined inflow/outflow x instead of periodic"""
        self.lib.or_set_open_x(self.h, 1 if on else 0)

    def set_inflow(self, seed, ppc, vsw, vth, beta_p=1.0, next_id=0):
        """inflow through the open x face in every later step"""
        self.lib.or_set_inflow(self.h, seed, ppc, vsw, vth, beta_p, next_id)


def initparameters(b0_init=0.2e-9, ion_amu=1, nf_init=0.1e15, dt_frac=0.05,
                   dx_frac=0.5, nu_init_frac=0.002, ntsub=10, width=64):
    lib, _ = _load(width)
    out = np.zeros(8)
    lib.or_initparameters(b0_init, ion_amu, nf_init, dt_frac, dx_frac, nu_init_frac,
                          ntsub, out.ctypes.data)
    keys = ("mion", "omega_p", "lambda_i", "dx", "dt", "dtsub_init", "nu_init", "alpha")
    return dict(zip(keys, out.tolist()))


def philox4x32_10(ctr, key):
    lib, _ = _load(64)
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib.or_philox4x32_10(c.ctypes.data, k.ctypes.data, o.ctypes.data)
    return o


def split_move(cell, fx, fy, fz, vx, vy, vz, s, nc):
    lib, _ = _load(64)
    s = np.asarray(s, np.float32)
    nc = np.asarray(nc, np.int32)
    lib.or_split_move(cell.size, cell.ctypes.data, fx.ctypes.data, fy.ctypes.data,
                      fz.ctypes.data, vx.ctypes.data, vy.ctypes.data, vz.ctypes.data,
                      s.ctypes.data, nc.ctypes.data)


def stable_sort(key, nkeys):
    lib, _ = _load(64)
    key = np.ascontiguousarray(key, np.uint32)
    perm = np.zeros(key.size, np.uint32)
    lib.or_stable_sort(key.size, key.ctypes.data, nkeys, perm.ctypes.data)
    return perm


CLOUD_KEYS = ("cx", "cy", "cz", "N0", "r0", "H", "tau_photo", "rmin", "sigma_chex", "mrat_pu",
              "beta_min")


def cloud_counts(nc, d, beta, dt, cloud, seed, step):
    """(count uint32[nin], beta_p float32[nin]) per interior node, i fastest"""
    lib, _ = _load(64)
    nc = np.asarray(nc, np.int32)
    cl = np.array([cloud[k] for k in CLOUD_KEYS], np.float64)
    nin = int(np.prod(nc))
    cnt = np.zeros(nin, np.uint32)
    bp = np.zeros(nin, np.float32)
    lib.or_cloud_counts(nc.ctypes.data, d[0], d[1], d[2], beta, dt, cl.ctypes.data,
                        int(cloud["max_per_cell"]), seed, step, cnt.ctypes.data, bp.ctypes.data)
    return cnt, bp


def ionize_sw(nc, d, dt, beta, ndot, vsw, beta_pu, seed, step):
    """Ionize_sw_mp's new ions: (cell uint32[m], frac float64[3,m],
    v float64[3,m])"""
    lib, _ = _load(64)
    nc = np.asarray(nc, np.int32)
    args = (nc.ctypes.data, d[0], d[1], d[2], dt, beta, ndot, vsw, beta_pu, seed, step)
    m = lib.or_ionize_sw(*args, None, None, None, 0)
    cell = np.zeros(max(m, 1), np.uint32)
    frac, v = np.zeros((3, max(m, 1))), np.zeros((3, max(m, 1)))
    lib.or_ionize_sw(*args, cell.ctypes.data, frac.ctypes.data, v.ctypes.data, max(m, 1))
    return cell[:m].copy(), frac[:, :m].copy(), v[:, :m].copy()


def cloud_ions(nc, cnt, seed, step, open_x=False):
    """(cell uint32[m], frac float64[3,m]) of the new ions in node order"""
    lib, _ = _load(64)
    nc = np.asarray(nc, np.int32)
    cnt = np.ascontiguousarray(cnt, np.uint32)
    m = int(cnt.sum())
    cell = np.zeros(m, np.uint32)
    frac = np.zeros((3, m))
    lib.or_cloud_ions(nc.ctypes.data, cnt.ctypes.data, seed, step, 1 if open_x else 0,
                      cell.ctypes.data, frac.ctypes.data, m)
    return cell, frac


def inflow(seed, step, nc, dt, d, ppc, vsw, vth, cap=1 << 22):
    """the inflow candidates of one step that deposit or are appended, in
    slot order: (th float64[3,m], te float64[3,m], v float64[3,m],
    flag uint8[m]); th, te are cell coordinates after the first and second
    half move, flag bit 0 = deposits at x^{n+1/2}, bit 1 = appended"""
    lib, _ = _load(64)
    nc = np.asarray(nc, np.int32)
    d = np.asarray(d, np.float64)
    th, te, v = np.zeros((3, cap)), np.zeros((3, cap)), np.zeros((3, cap))
    flag = np.zeros(cap, np.uint8)
    m = lib.or_inflow(seed, step, nc.ctypes.data, dt, d.ctypes.data, ppc, vsw, vth,
                      th.ctypes.data, te.ctypes.data, v.ctypes.data, flag.ctypes.data, cap)
    if m < 0 or m > cap:
        raise ValueError("or_inflow: bad parameters or more than %d candidates" % cap)
    return th[:, :m].copy(), te[:, :m].copy(), v[:, :m].copy(), flag[:m].copy()


def chex(cell, f, v, d, dt, cloud, seed, step, ids=None):
    """charge-exchange decision (uint8) per ion; f, v float32 [3, n]; the draw
    of an ion is keyed by its id when ids are given, else by its index"""
    lib, _ = _load(64)
    cell = np.ascontiguousarray(cell, np.uint32)
    f = [np.ascontiguousarray(a, np.float32) for a in f]
    v = [np.ascontiguousarray(a, np.float32) for a in v]
    idp = None if ids is None else np.ascontiguousarray(ids, np.uint32)
    cl = np.array([cloud[k] for k in CLOUD_KEYS], np.float64)
    hit = np.zeros(cell.size, np.uint8)
    lib.or_chex(cell.size, cell.ctypes.data, f[0].ctypes.data, f[1].ctypes.data,
                f[2].ctypes.data, v[0].ctypes.data, v[1].ctypes.data, v[2].ctypes.data,
                None if idp is None else idp.ctypes.data, d[0], d[1], d[2], dt, cl.ctypes.data,
                seed, step, hit.ctypes.data)
    return hit


def load_maxwellian(n, seed, nc, vsw, vth, stream=1):
    """(cell uint32[n], frac float64[3,n], v float64[3,n]) of the seeded load."""
    lib, _ = _load(64)
    nc = np.asarray(nc, np.int32)
    cell = np.zeros(n, np.uint32)
    frac = np.zeros((3, n))
    v = np.zeros((3, n))
    lib.or_load_maxwellian(n, seed, stream, nc.ctypes.data, vsw, vth, cell.ctypes.data,
                           frac.ctypes.data, v.ctypes.data)
    return cell, frac, v


def read_load(path):
    """the load file read by the oracle's C reader: (cell, frac[3,n], vp[3,n],
    vp1[3,n], beta_p, mrat, id); beta_p, mrat and id are zeros where the
    file holds none"""
    lib, _ = _load(64)
    n = C.c_int64()
    nc = np.zeros(3, np.int32)
    p = path.encode()
    if lib.or_read_load(p, 0, C.byref(n), nc.ctypes.data, *([None] * 7)) != 0:
        raise ValueError("%s: not a load file" % path)
    m = n.value
    cell = np.zeros(m, np.uint32)
    frac, vp, vp1 = np.zeros((3, m)), np.zeros((3, m)), np.zeros((3, m))
    bp, mr, ids = np.zeros(m, np.float32), np.zeros(m, np.float32), np.zeros(m, np.uint32)
    if lib.or_read_load(p, m, C.byref(n), nc.ctypes.data, cell.ctypes.data, frac.ctypes.data,
                        vp.ctypes.data, vp1.ctypes.data, bp.ctypes.data, mr.ctypes.data,
                        ids.ctypes.data) != 0:
        raise ValueError("%s: short load file" % path)
    return cell, frac, vp, vp1, bp, mr, ids
