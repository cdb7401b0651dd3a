"""Host-side mirror of SWAP's time-step routines over the gfx950 C-ABI.

The method names and argument meanings follow the reference's Fortran
routines (gutsp.f, gutsf.f, grid_interp.f, boundary.f, chem_rates.f); the
state they act on lives on the GPU inside one `Swap` context, as the
reference's module state lives in global.f.  Arrays handed back use the
reference's Fortran layouts seen from C order: a vector grid (nx,ny,nz,3) is
a numpy array (3,nz,ny,nx), a scalar grid (nz,ny,nx), particle arrays (3,n).

There is no CPU fallback: constructing a context without the built library
or without a GPU raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libswap_hip.so")

# field ids (include/swap_hip.h)
B1, B12, B1P2, BTC, NP, UP, AJ, E, NU, AJC, UPC, TMP0, TMP1, TMP2 = range(14)
FIELD_NAMES = {"b1": B1, "bt": B1, "b12": B12, "b1p2": B1P2, "btc": BTC, "np": NP, "up": UP,
               "aj": AJ, "E": E, "nu": NU, "ajc": AJC, "upc": UPC, "tmp0": TMP0, "tmp1": TMP1,
               "tmp2": TMP2}
P_XP, P_VP, P_VP1, P_EP, P_VPLUS, P_VMINUS, P_IJKP, P_WGHT, P_CELL, P_FRAC, P_BETA, P_MRAT, \
    P_ID = range(13)
V_CUR, V_EXTRAP, V_VP1 = 0, 1, 2
FP32, FP64 = 32, 64
SPECIES, OUTFLOW_X, TRACK_IDS, UNFUSED_FIELDS, PERSISTENT_FIELDS = 1, 2, 4, 8, 16
DN_ALL, DN_DISJOINT, DN_DISJOINT_STEP = 0, 1, 2


class SwapError(RuntimeError):
    pass


class Config(C.Structure):
    _fields_ = [("ncx", C.c_int), ("ncy", C.c_int), ("ncz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("dt", C.c_double), ("ntsub", C.c_int), ("alpha", C.c_double),
                ("beta", C.c_double), ("nu", C.c_double), ("capacity", C.c_int64),
                ("precision", C.c_int), ("flags", C.c_int), ("sort_every", C.c_int),
                ("seed", C.c_uint64)]


ENERGY_KEYS = ("input_E", "Evp", "Euf", "EB1", "EB1x", "EB1y", "EB1z", "EE", "total_E",
               "norm_E")


class Energy(C.Structure):
    _fields_ = [(k, C.c_double) for k in ENERGY_KEYS]


class Diag(C.Structure):
    _fields_ = [("charge", C.c_double), ("momentum", C.c_double * 3), ("kinetic", C.c_double),
                ("divb_l2", C.c_double), ("b_energy", C.c_double), ("nparticles", C.c_int64),
                ("escaped", C.c_int64), ("injected", C.c_int64),
                ("window_misses", C.c_int64), ("shrunk_windows", C.c_int64),
                ("sorts", C.c_int64)]


class Cloud(C.Structure):
    _fields_ = [("cx", C.c_double), ("cy", C.c_double), ("cz", C.c_double),
                ("N0", C.c_double), ("r0", C.c_double), ("H", C.c_double),
                ("tau_photo", C.c_double), ("sigma_chex", C.c_double),
                ("mrat_pu", C.c_double), ("beta_min", C.c_double),
                ("max_per_cell", C.c_int), ("rmin", C.c_double)]


class DnConfig(C.Structure):
    _fields_ = [("n_scale", C.c_double), ("v_scale", C.c_double), ("max_pairs", C.c_int),
                ("stride", C.c_int), ("seed", C.c_uint64)]


_lib = None


def load_library():
    """dlopen the in-tree library; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SwapError("libswap_hip.so not built: run python -c 'import __graft_entry__ as g; "
                        "g.build()' (or swap_amd/build.py)")
    lib = C.CDLL(LIB_PATH)
    vp, i32, i64, dbl = C.c_void_p, C.c_int, C.c_int64, C.c_double
    sig = {
        "swap_create": ([C.POINTER(Config), C.POINTER(vp)], i32),
        "swap_destroy": ([vp], i32),
        "swap_last_error": ([vp], C.c_char_p),
        "swap_sync": ([vp], i32),
        "swap_num_particles": ([vp], i64),
        "swap_get_coords": ([vp, i32, vp], i32),
        "swap_load_particles": ([vp, i64, vp, vp, vp, vp, vp], i32),
        "swap_load_particles_split": ([vp, i64, vp, vp, vp, vp, vp, vp], i32),
        "swap_load_maxwellian": ([vp, i64, dbl, dbl, dbl, dbl], i32),
        "swap_get_particle_array": ([vp, i32, vp], i32),
        "swap_set_particle_array": ([vp, i32, vp], i32),
        "swap_set_field": ([vp, i32, vp], i32),
        "swap_get_field": ([vp, i32, vp], i32),
        "swap_set_field_f32": ([vp, i32, vp], i32),
        "swap_get_field_f32": ([vp, i32, vp], i32),
        "swap_periodic": ([vp, i32], i32),
        "swap_get_interp_weights": ([vp], i32),
        "swap_update_np": ([vp], i32),
        "swap_update_up": ([vp, i32], i32),
        "swap_extrapol_up": ([vp], i32),
        "swap_get_Ep": ([vp], i32),
        "swap_get_vplus_vminus": ([vp], i32),
        "swap_improve_up": ([vp], i32),
        "swap_get_vp_final": ([vp], i32),
        "swap_move_ion_half": ([vp], i32),
        "swap_curlB": ([vp, i32, i32], i32),
        "swap_curlE": ([vp, i32, i32], i32),
        "swap_crossf2": ([vp, i32, i32, i32], i32),
        "swap_get_E": ([vp], i32),
        "swap_predict_B": ([vp], i32),
        "swap_get_Ep1": ([vp], i32),
        "swap_correct_B": ([vp], i32),
        "swap_f_update_tlev": ([vp], i32),
        "swap_check_time_step": ([vp, C.POINTER(i32), C.POINTER(dbl)], i32),
        "swap_edge_to_center": ([vp, i32, i32], i32),
        "swap_edge_to_face": ([vp, i32, i32], i32),
        "swap_face_to_center": ([vp, i32, i32], i32),
        "swap_field_subcycles": ([vp, i32, dbl], i32),
        "swap_sort": ([vp], i32),
        "swap_prime": ([vp], i32),
        "swap_step": ([vp, i32], i32),
        "swap_ionize_sw": ([vp, i32, dbl, dbl, dbl, C.POINTER(i64)], i32),
        "swap_set_inflow": ([vp, dbl, dbl, dbl, dbl], i32),
        "swap_ionize_cloud": ([vp, i32, C.POINTER(Cloud), C.POINTER(i64), C.POINTER(i64)], i32),
        "swap_diagnostics": ([vp, C.POINTER(Diag)], i32),
        "swap_energy_diag": ([vp, dbl, vp, C.POINTER(Energy)], i32),
        "swap_set_input_energy": ([vp, dbl, dbl], i32),
        "swap_get_particle_array_noalloc": ([vp, i32, vp], i32),
        "swap_run_state_size": ([vp, C.POINTER(i64)], i32),
        "swap_save_run_state": ([vp, vp], i32),
        "swap_restore_run_state": ([vp, vp], i32),
        "swap_profile": ([vp, i32], i32),
        "swap_profile_read": ([vp, vp, vp, i32], i32),
        "swap_set_seed": ([vp, C.c_uint64], i32),
        "swap_dn_num_params": ([], i64),
        "swap_dn_create": ([vp, C.POINTER(DnConfig)], i32),
        "swap_dn_add_pair": ([vp, i32], i32),
        "swap_dn_num_pairs": ([vp], i32),
        "swap_dn_get_pair": ([vp, i32, vp, vp], i32),
        "swap_dn_put_pair": ([vp, vp, vp], i32),
        "swap_dn_train": ([vp, i32, dbl, i32, C.POINTER(dbl)], i32),
        "swap_dn_eval": ([vp, i32, i32, C.POINTER(dbl), C.POINTER(dbl)], i32),
        "swap_dn_enable": ([vp, i32], i32),
        "swap_dn_grad": ([vp, vp, vp, C.POINTER(dbl), vp], i32),
        "swap_dn_infer": ([vp, vp, vp], i32),
        "swap_dn_get_params": ([vp, vp], i32),
        "swap_dn_set_params": ([vp, vp], i32),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def _fid(f):
    return FIELD_NAMES[f] if isinstance(f, str) else int(f)


class Swap:
    """One device-resident hybrid-PIC state (the reference's global.f + maind arrays)."""

    def __init__(self, ncx, ncy, ncz, dx, dt, alpha, beta, nu, capacity, dy=None, dz=None,
                 ntsub=10, precision=FP32, flags=0, sort_every=2, seed=20261203):
        self.lib = load_library()
        cfg = Config(ncx, ncy, ncz, dx, dy or dx, dz or dx, dt, ntsub, alpha, beta, nu,
                     capacity, precision, flags, sort_every, seed)
        self.cfg = cfg
        h = C.c_void_p()
        rc = self.lib.swap_create(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise SwapError("swap_create failed (%d)" % rc)
        self.h = h
        self.nx, self.ny, self.nz = ncx + 2, ncy + 2, ncz + 2
        self.precision = precision
        self.rdtype = np.float32 if precision == FP32 else np.float64
        self.dt = dt

    def close(self):
        if getattr(self, "h", None):
            self.lib.swap_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc, what):
        if rc != 0:
            msg = self.lib.swap_last_error(self.h)
            raise SwapError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    # ---- state transfer ----
    @property
    def n(self):
        return self.lib.swap_num_particles(self.h)

    def load_particles(self, xp, vp, vp1, beta_p=None, mrat=None):
        xp = np.ascontiguousarray(xp, np.float64)
        vp = np.ascontiguousarray(vp, np.float64)
        vp1 = np.ascontiguousarray(vp1, np.float64)
        n = xp.shape[1]
        bp = None if beta_p is None else np.ascontiguousarray(beta_p, np.float64)
        mr = None if mrat is None else np.ascontiguousarray(mrat, np.float64)
        self._ck(self.lib.swap_load_particles(self.h, n, xp.ctypes.data, vp.ctypes.data,
                                              vp1.ctypes.data,
                                              None if bp is None else bp.ctypes.data,
                                              None if mr is None else mr.ctypes.data),
                 "load_particles")

    def load_split(self, cell, frac, vp, vp1, beta_p=None, mrat=None):
        cell = np.ascontiguousarray(cell, np.uint32)
        frac = np.ascontiguousarray(frac, self.rdtype)
        vp = np.ascontiguousarray(vp, self.rdtype)
        vp1 = np.ascontiguousarray(vp1, self.rdtype)
        bp = None if beta_p is None else np.ascontiguousarray(beta_p, np.float32)
        mr = None if mrat is None else np.ascontiguousarray(mrat, np.float32)
        self._ck(self.lib.swap_load_particles_split(
            self.h, cell.size, cell.ctypes.data, frac.ctypes.data, vp.ctypes.data,
            vp1.ctypes.data, None if bp is None else bp.ctypes.data,
            None if mr is None else mr.ctypes.data), "load_split")

    def load_maxwellian(self, n, vsw, vth, beta_p=1.0, mrat=1.0):
        self._ck(self.lib.swap_load_maxwellian(self.h, n, vsw, vth, beta_p, mrat),
                 "load_maxwellian")

    def _parray_spec(self, which):
        ncomp = {P_XP: 3, P_VP: 3, P_VP1: 3, P_EP: 3, P_VPLUS: 3, P_VMINUS: 3, P_IJKP: 3,
                 P_WGHT: 8, P_CELL: 1, P_FRAC: 3, P_BETA: 1, P_MRAT: 1, P_ID: 1}[which]
        if which == P_XP:
            dt = np.float64
        elif which in (P_IJKP,):
            dt = np.int32
        elif which in (P_CELL, P_ID):
            dt = np.uint32
        elif which in (P_BETA, P_MRAT):
            dt = np.float32
        else:
            dt = self.rdtype
        return ((ncomp, self.n) if ncomp > 1 else (self.n,)), dt

    def particle_array(self, which):
        shape, dt = self._parray_spec(which)
        out = np.zeros(shape, dt)
        self._ck(self.lib.swap_get_particle_array(self.h, which, out.ctypes.data),
                 "get_particle_array")
        return out

    def set_particle_array(self, which, arr):
        shape, dt = self._parray_spec(which)
        a = np.ascontiguousarray(np.broadcast_to(arr, shape), dt)
        self._ck(self.lib.swap_set_particle_array(self.h, which, a.ctypes.data),
                 "set_particle_array")

    def xp(self):
        return self.particle_array(P_XP)

    def vp(self):
        return self.particle_array(P_VP)

    def vp1(self):
        return self.particle_array(P_VP1)

    def _fshape(self, f):
        return (self.nz, self.ny, self.nx) if f in (NP, NU) else (3, self.nz, self.ny, self.nx)

    def get_field(self, f):
        f = _fid(f)
        out = np.zeros(self._fshape(f), np.float64)
        self._ck(self.lib.swap_get_field(self.h, f, out.ctypes.data), "get_field")
        return out

    def set_field(self, f, arr):
        f = _fid(f)
        a = np.ascontiguousarray(np.broadcast_to(arr, self._fshape(f)), np.float64)
        self._ck(self.lib.swap_set_field(self.h, f, a.ctypes.data), "set_field")

    def coords(self, axis):
        n = (self.nx, self.ny, self.nz)[axis]
        q = np.zeros(n)
        self._ck(self.lib.swap_get_coords(self.h, axis, q.ctypes.data), "get_coords")
        return q

    def set_seed(self, seed):
        self._ck(self.lib.swap_set_seed(self.h, seed), "set_seed")
        self.cfg.seed = seed

    def sync(self):
        self._ck(self.lib.swap_sync(self.h), "sync")

    # ---- boundary.f / grid_interp.f ----
    def periodic(self, f):
        self._ck(self.lib.swap_periodic(self.h, _fid(f)), "periodic")

    def edge_to_center(self, src="bt", dst="btc"):
        self._ck(self.lib.swap_edge_to_center(self.h, _fid(src), _fid(dst)), "edge_to_center")

    def edge_to_face(self, src, dst):
        self._ck(self.lib.swap_edge_to_face(self.h, _fid(src), _fid(dst)), "edge_to_face")

    def face_to_center(self, src, dst):
        self._ck(self.lib.swap_face_to_center(self.h, _fid(src), _fid(dst)), "face_to_center")

    # ---- gutsp.f ----
    def get_interp_weights(self):
        self._ck(self.lib.swap_get_interp_weights(self.h), "get_interp_weights")

    def update_np(self):
        self._ck(self.lib.swap_update_np(self.h), "update_np")

    def update_up(self, vsrc=V_CUR):
        self._ck(self.lib.swap_update_up(self.h, vsrc), "update_up")

    def extrapol_up(self):
        self._ck(self.lib.swap_extrapol_up(self.h), "extrapol_up")

    def get_Ep(self):
        self._ck(self.lib.swap_get_Ep(self.h), "get_Ep")

    def get_vplus_vminus(self):
        self._ck(self.lib.swap_get_vplus_vminus(self.h), "get_vplus_vminus")

    def improve_up(self):
        self._ck(self.lib.swap_improve_up(self.h), "improve_up")

    def get_vp_final(self):
        self._ck(self.lib.swap_get_vp_final(self.h), "get_vp_final")

    def move_ion_half(self):
        self._ck(self.lib.swap_move_ion_half(self.h), "move_ion_half")

    # ---- gutsf.f ----
    def curlB(self, b="bt", aj="aj"):
        self._ck(self.lib.swap_curlB(self.h, _fid(b), _fid(aj)), "curlB")

    def curlE(self, e="E", out="tmp0"):
        self._ck(self.lib.swap_curlE(self.h, _fid(e), _fid(out)), "curlE")

    def crossf2(self, aa, btc, out):
        self._ck(self.lib.swap_crossf2(self.h, _fid(aa), _fid(btc), _fid(out)), "crossf2")

    def get_E(self):
        self._ck(self.lib.swap_get_E(self.h), "get_E")

    def predict_B(self):
        self._ck(self.lib.swap_predict_B(self.h), "predict_B")

    def get_Ep1(self):
        self._ck(self.lib.swap_get_Ep1(self.h), "get_Ep1")

    def correct_B(self):
        self._ck(self.lib.swap_correct_B(self.h), "correct_B")

    def f_update_tlev(self):
        self._ck(self.lib.swap_f_update_tlev(self.h), "f_update_tlev")

    def check_time_step(self):
        ntf, dtsub = C.c_int(), C.c_double()
        self._ck(self.lib.swap_check_time_step(self.h, C.byref(ntf), C.byref(dtsub)),
                 "check_time_step")
        return ntf.value, dtsub.value

    # ---- fused path ----
    def field_subcycles(self, ntf, dtsub):
        self._ck(self.lib.swap_field_subcycles(self.h, ntf, dtsub), "field_subcycles")

    def sort(self):
        self._ck(self.lib.swap_sort(self.h), "sort")

    def prime(self):
        self._ck(self.lib.swap_prime(self.h), "prime")

    def step(self, nsteps=1):
        self._ck(self.lib.swap_step(self.h, nsteps), "step")

    # ---- chem_rates.f ----
    def ionize_sw(self, step, ndot, vsw, beta_pu):
        n = C.c_int64()
        self._ck(self.lib.swap_ionize_sw(self.h, step, ndot, vsw, beta_pu, C.byref(n)),
                 "ionize_sw")
        return n.value

    def set_inflow(self, ppc, vsw, vth, beta_p=1.0):
        """every later step of an OUTFLOW_X context adds the ions of an
        upstream drifting Maxwellian (ppc per cell) that cross x = 0 during
        the step; ppc = 0 switches it off"""
        self._ck(self.lib.swap_set_inflow(self.h, ppc, vsw, vth, beta_p), "set_inflow")

    def ionize_cloud(self, step, cloud):
        """photoionization + charge exchange of a neutral cloud; cloud is a
        Cloud or a dict of its fields; returns (injected, charge exchanges)"""
        if isinstance(cloud, dict):
            cloud = Cloud(**cloud)
        n, x = C.c_int64(), C.c_int64()
        self._ck(self.lib.swap_ionize_cloud(self.h, step, C.byref(cloud), C.byref(n),
                                            C.byref(x)), "ionize_cloud")
        return n.value, x.value

    # ---- learned moment denoiser ----
    def dn_create(self, n_scale, v_scale, max_pairs=64, stride=32, seed=7):
        cfg = DnConfig(n_scale, v_scale, max_pairs, stride, seed)
        self._ck(self.lib.swap_dn_create(self.h, C.byref(cfg)), "dn_create")

    def dn_add_pair(self, mode=DN_ALL):
        self._ck(self.lib.swap_dn_add_pair(self.h, int(mode)), "dn_add_pair")

    def dn_num_pairs(self):
        return self.lib.swap_dn_num_pairs(self.h)

    def dn_get_pair(self, k):
        x = np.zeros((4, self.nz - 2, self.ny - 2, self.nx - 2), np.float32)
        t = np.zeros_like(x)
        self._ck(self.lib.swap_dn_get_pair(self.h, k, x.ctypes.data, t.ctypes.data),
                 "dn_get_pair")
        return x, t

    def dn_put_pair(self, x, target):
        x, t = self._grid4(x), self._grid4(target)
        self._ck(self.lib.swap_dn_put_pair(self.h, x.ctypes.data, t.ctypes.data), "dn_put_pair")

    def dn_train(self, iters, lr, ntrain):
        loss = C.c_double()
        self._ck(self.lib.swap_dn_train(self.h, iters, lr, ntrain, C.byref(loss)), "dn_train")
        return loss.value

    def dn_eval(self, p0, p1):
        a, b = C.c_double(), C.c_double()
        self._ck(self.lib.swap_dn_eval(self.h, p0, p1, C.byref(a), C.byref(b)), "dn_eval")
        return a.value, b.value

    def dn_enable(self, on=True):
        self._ck(self.lib.swap_dn_enable(self.h, 1 if on else 0), "dn_enable")

    def _grid4(self, a):
        return np.ascontiguousarray(np.broadcast_to(a, (4, self.nz - 2, self.ny - 2, self.nx - 2)),
                                    np.float32)

    def dn_grad(self, x, target):
        x, t = self._grid4(x), self._grid4(target)
        g = np.zeros(self.lib.swap_dn_num_params(), np.float32)
        loss = C.c_double()
        self._ck(self.lib.swap_dn_grad(self.h, x.ctypes.data, t.ctypes.data, C.byref(loss),
                                       g.ctypes.data), "dn_grad")
        return loss.value, g

    def dn_infer(self, x):
        x = self._grid4(x)
        y = np.zeros_like(x)
        self._ck(self.lib.swap_dn_infer(self.h, x.ctypes.data, y.ctypes.data), "dn_infer")
        return y

    def dn_params(self):
        p = np.zeros(self.lib.swap_dn_num_params(), np.float32)
        self._ck(self.lib.swap_dn_get_params(self.h, p.ctypes.data), "dn_get_params")
        return p

    def dn_set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.lib.swap_dn_num_params()
        self._ck(self.lib.swap_dn_set_params(self.h, p.ctypes.data), "dn_set_params")

    STAGES = ("P1", "P2", "P3", "sort", "field", "glue")

    def profile(self, enable=True):
        self._ck(self.lib.swap_profile(self.h, 1 if enable else 0), "profile")

    def profile_read(self):
        ms = np.zeros(len(self.STAGES))
        cnt = np.zeros(len(self.STAGES), np.int64)
        self._ck(self.lib.swap_profile_read(self.h, ms.ctypes.data, cnt.ctypes.data, len(ms)),
                 "profile_read")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(self.STAGES)}

    def energy_diag(self, mion, b0):
        """Energy_diag's record (part_init.f:14-139) as a dict, J; b0: the
        uniform background field (b1 units)"""
        e = Energy()
        b = np.ascontiguousarray(b0, np.float64)
        self._ck(self.lib.swap_energy_diag(self.h, mion, b.ctypes.data, C.byref(e)),
                 "energy_diag")
        return {k: getattr(e, k) for k in ENERGY_KEYS}

    def set_input_energy(self, mion, input_E):
        self._ck(self.lib.swap_set_input_energy(self.h, mion, input_E), "set_input_energy")

    def run_state(self):
        """the run state a checkpoint needs beside particles and B (opaque uint8)"""
        nb = C.c_int64()
        self._ck(self.lib.swap_run_state_size(self.h, C.byref(nb)), "run_state_size")
        buf = np.zeros(nb.value, np.uint8)
        self._ck(self.lib.swap_save_run_state(self.h, buf.ctypes.data), "save_run_state")
        return buf

    def restore_run_state(self, buf):
        buf = np.ascontiguousarray(buf, np.uint8)
        self._ck(self.lib.swap_restore_run_state(self.h, buf.ctypes.data), "restore_run_state")

    def diagnostics(self):
        d = Diag()
        self._ck(self.lib.swap_diagnostics(self.h, C.byref(d)), "diagnostics")
        return {"charge": d.charge, "momentum": list(d.momentum), "kinetic": d.kinetic,
                "divb_l2": d.divb_l2, "b_energy": d.b_energy, "nparticles": d.nparticles,
                "escaped": d.escaped, "injected": d.injected,
                "window_misses": d.window_misses, "shrunk_windows": d.shrunk_windows,
                "sorts": d.sorts}
