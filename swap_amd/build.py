"""Build the gfx950 HIP library libswap_hip.so in-tree (hipcc, no JIT).

Every translation unit compiles for gfx950 only, with packed-FP32 VALU forms
disabled and hardware FP atomics (-munsafe-fp-atomics).  The FP32 build's
moment sums are exact fixed-point integer atomics (swap_common.h) and need
no FP atomics; the flag serves the FP64 build's deposit (FP64 atomic adds
into the LDS windows and HBM accumulators) and the FP64 diagnostic
reductions, which would otherwise compile to compare-and-swap loops.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libswap_hip.so")
BUILD = os.path.join(HERE, "_build")
SOURCES = ["swap_api.hip", "k_fields.hip", "k_particles.hip", "k_pass.hip", "k_sort.hip", "k_sources.hip",
           "k_denoise.hip", "swap_noalloc.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-munsafe-fp-atomics",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-ffp-contract=fast", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
         "-Wno-unused-command-line-argument"]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s%s" % (src, r.stdout, r.stderr))
    return obj, r.stderr


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(6, len(srcs))) as ex:
        res = list(ex.map(_compile, srcs))
    if verbose:
        for _, err in res:
            if err.strip():
                sys.stderr.write(err)
    objs = [o for o, _ in res]
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s%s" % (r.stdout, r.stderr))
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
