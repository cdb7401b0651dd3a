"""The C-ABI boundary without a GPU: the library loads, exports every entry
point include/swap_hip.h declares, and the ctypes mirrors of the ABI structs
have the C layout (checked by compiling the header with gcc)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "swap_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(swap_\w+)\s*\(", text,
                                 flags=re.M)))


def test_header_declares_the_reference_routines():
    names = declared_functions()
    # one entry point per hot-path routine maind.f calls (gutsp.f, gutsf.f,
    # grid_interp.f, boundary.f, chem_rates.f)
    for r in ("get_interp_weights", "update_np", "update_up", "extrapol_up", "get_Ep",
              "get_vplus_vminus", "improve_up", "get_vp_final", "move_ion_half", "curlB",
              "curlE", "crossf2", "get_E", "predict_B", "get_Ep1", "correct_B",
              "f_update_tlev", "check_time_step", "edge_to_center", "edge_to_face",
              "face_to_center", "periodic", "ionize_sw", "ionize_cloud"):
        assert "swap_" + r in names, r
    assert len(names) >= 40


@pytest.fixture(scope="module")
def lib():
    from swap_amd import api
    if not os.path.exists(api.LIB_PATH):
        from swap_amd import build
        build.build()
    return C.CDLL(api.LIB_PATH)


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from swap_amd import api
    lib = api.load_library()
    assert lib is not None
    src = open(os.path.join(ROOT, "swap_amd", "api.py")).read()
    unbound = [n for n in declared_functions() if '"%s"' % n not in src]
    assert not unbound, unbound


def test_fortran_shim_binds_declared_symbols_and_reference_names():
    """swap_amd/fortran/swap_f.f90 binds only symbols the header declares and
    gives every hot-path routine of the reference under its own name"""
    src = open(os.path.join(ROOT, "swap_amd", "fortran", "swap_f.f90")).read().lower()
    bound = set(re.findall(r'bind\(c,\s*name="(swap_\w+)"\)', src))
    declared = {n.lower() for n in declared_functions()}
    assert bound and bound <= declared, bound - declared
    subs = set(re.findall(r"^\s*subroutine\s+(\w+)", src, flags=re.M))
    for r in ("get_interp_weights", "update_np", "update_up", "extrapol_up", "get_ep",
              "get_vplus_vminus", "improve_up", "get_vp_final", "move_ion_half", "curlb",
              "curle", "predict_b", "correct_b", "f_update_tlev", "check_time_step",
              "edge_to_center", "edge_to_face", "face_to_center", "periodic",
              "periodic_scalar"):
        assert r in subs, r


def c_layout(struct, fields):
    body = "\n".join('  printf("%%s %%zu\\n", "%s", offsetof(%s, %s));' % (f, struct, f)
                     for f in fields)
    prog = ('#include <stdio.h>\n#include <stddef.h>\n#include "swap_hip.h"\nint main(void){\n'
            '  printf("sizeof %%zu\\n", sizeof(%s));\n%s\n  return 0;\n}\n' % (struct, body))
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "l.c"), os.path.join(d, "l")
        open(src, "w").write(prog)
        subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(HEADER), src, "-o", exe],
                       check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return {k: int(v) for k, v in (ln.split() for ln in out.splitlines())}


@pytest.mark.parametrize("cname, pyname", [("swap_config", "Config"), ("swap_diag", "Diag"),
                                           ("swap_cloud", "Cloud"),
                                           ("swap_dn_config", "DnConfig")])
def test_struct_layouts_match(cname, pyname):
    from swap_amd import api
    cls = getattr(api, pyname)
    fields = [f[0] for f in cls._fields_]
    lay = c_layout(cname, fields)
    assert lay["sizeof"] == C.sizeof(cls)
    for f in fields:
        assert lay[f] == getattr(cls, f).offset, f


def c_constants(names):
    prog = ('#include <stdio.h>\n#include "swap_hip.h"\nint main(void){\n%s\n  return 0;\n}\n'
            % "\n".join('  printf("%%s %%d\\n", "%s", (int)(%s));' % (n, n) for n in names))
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "k.c"), os.path.join(d, "k")
        open(src, "w").write(prog)
        subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(HEADER), src, "-o", exe],
                       check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return {k: int(v) for k, v in (ln.split() for ln in out.splitlines())}


def test_python_constants_match_the_header():
    """the ctypes mirror's config flags, field ids, particle-array ids,
    velocity selectors and denoiser pair modes carry the header's values"""
    from swap_amd import api
    pairs = {
        "SWAP_SPECIES": api.SPECIES, "SWAP_OUTFLOW_X": api.OUTFLOW_X,
        "SWAP_TRACK_IDS": api.TRACK_IDS, "SWAP_UNFUSED_FIELDS": api.UNFUSED_FIELDS,
        "SWAP_PERSISTENT_FIELDS": api.PERSISTENT_FIELDS,
        "SWAP_FP32": api.FP32, "SWAP_FP64": api.FP64,
        "SWAP_DN_ALL": api.DN_ALL, "SWAP_DN_DISJOINT": api.DN_DISJOINT,
        "SWAP_DN_DISJOINT_STEP": api.DN_DISJOINT_STEP,
        "SWAP_V_CUR": api.V_CUR, "SWAP_V_EXTRAP": api.V_EXTRAP, "SWAP_V_VP1": api.V_VP1,
    }
    for f in ("B1", "B12", "B1P2", "BTC", "NP", "UP", "AJ", "E", "NU", "AJC", "UPC", "TMP0",
              "TMP1", "TMP2"):
        pairs["SWAP_F_" + f] = getattr(api, f)
    for p in ("XP", "VP", "VP1", "EP", "VPLUS", "VMINUS", "IJKP", "WGHT", "CELL", "FRAC",
              "BETA", "MRAT", "ID"):
        pairs["SWAP_P_" + p] = getattr(api, "P_" + p)
    got = c_constants(sorted(pairs))
    bad = {k: (got[k], v) for k, v in pairs.items() if got[k] != v}
    assert not bad, bad
