"""Write tests/golden/b0_fixture.json from the reference's b0.dat.

b0.dat (shipped with the reference) is the initial field written by grd6_setup
(initial.f:10-142) as two Fortran unformatted sequential records: the int32
grid length nz, then b0(3,3,nz,3) in FP32 (component slowest).  It is read
here strictly as data (record markers, int32, float32); the fixture keeps the
header and, per component, the distinct values with their bit patterns and
counts.  Run in a checkout where the reference tree is present:

    python tests/golden/make_b0_fixture.py /root/reference/b0.dat
"""
import json
import os
import struct
import sys

import numpy as np


def read_records(path):
    raw = open(path, "rb").read()
    recs, off = [], 0
    while off < len(raw):
        (n,) = struct.unpack_from("<i", raw, off)
        body = raw[off + 4: off + 4 + n]
        (m,) = struct.unpack_from("<i", raw, off + 4 + n)
        assert m == n, "record markers disagree"
        recs.append(body)
        off += 8 + n
    return recs


def main(path):
    recs = read_records(path)
    (nz,) = struct.unpack("<i", recs[0])
    b0 = np.frombuffer(recs[1], "<f4").reshape(3, nz, 3, 3)  # (comp, k, j, i)
    comps = []
    for c in range(3):
        vals, counts = np.unique(b0[c], return_counts=True)
        comps.append([{"value": float(v), "bits": "0x%08x" % np.float32(v).view(np.uint32),
                       "count": int(n)} for v, n in zip(vals, counts)])
    out = {"source": "b0.dat", "nz": nz, "shape_fortran": [3, 3, nz, 3],
           "record_bytes": [len(r) for r in recs], "components": comps}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "b0_fixture.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/b0.dat")
