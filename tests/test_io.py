"""The reference's on-disk formats without a GPU: Fortran unformatted
records (4-byte length markers around each WRITE, -i4 -real-size 32) and the
record layout of para.dat (maind.f:237-255) and c.coord.dat
(initial.f:285-297)."""
import struct
import types

import numpy as np

from swap_amd import io, physics


def test_record_framing_round_trip(tmp_path):
    f = tmp_path / "r.dat"
    a = np.arange(24, dtype=np.float64).reshape(2, 3, 4)
    with open(f, "wb") as h:
        io.write_record(h, 7)
        io.write_record(h, a)
        io.write_record(h, 3, 2.5, "ab")
    raw = f.read_bytes()
    # first record: marker 4, int 7, marker 4
    assert raw[:12] == struct.pack("<iii", 4, 7, 4)
    recs = io.read_records(f)
    assert [len(r) for r in recs] == [4, 96, 10]
    assert np.array_equal(np.frombuffer(recs[1], "<f4").reshape(2, 3, 4), a.astype(np.float32))
    assert struct.unpack("<if", recs[2][:8]) == (3, 2.5) and recs[2][8:] == b"ab"


def _fake_ctx(nc, cap):
    p = physics.workload(*nc, 2)
    s = types.SimpleNamespace(nx=nc[0] + 2, ny=nc[1] + 2, nz=nc[2] + 2,
                              cfg=types.SimpleNamespace(capacity=cap))
    s.coords = lambda a: np.arange((s.nx, s.ny, s.nz)[a]) * p["dx"]
    return p, s


def test_para_and_coord_layout(tmp_path):
    p, s = _fake_ctx((8, 6, 4), 1000)
    io.write_para(str(tmp_path), s, p, nt=100, nout=10)
    r = io.read_records(tmp_path / "para.dat")
    assert len(r) == 10
    nx, ny, nz, dx, dy, dz = struct.unpack("<iiifff", r[0])
    assert (nx, ny, nz) == (10, 8, 6) and abs(dx - p["dx"]) < 1e-3 * p["dx"]
    nt, dts, ntsub, dt, nout = struct.unpack("<ififi", r[1])
    assert (nt, ntsub, nout) == (100, p["ntsub"], 10)
    assert abs(dt - p["dt"]) < 1e-6 * p["dt"] and abs(dts * ntsub - dt) < 1e-6 * dt
    assert len(r[2]) == 50  # character(50) out_dir (inputs.f:30)
    assert struct.unpack("<i", r[4]) == (1000,)  # Ni_max
    assert np.allclose(struct.unpack("<ff", r[9]), (p["alpha"], p["beta"]), rtol=1e-6)
    io.write_coord(str(tmp_path), s)
    c = io.read_records(tmp_path / "c.coord.dat")
    assert [struct.unpack("<i", c[k])[0] for k in range(3)] == [10, 8, 6]
    qx = np.frombuffer(c[3], "<f4")
    assert qx.size == 10 and np.allclose(np.diff(qx), p["dx"], rtol=1e-6)
    assert [len(x) // 4 for x in c[3:]] == [10, 8, 6, 6, 6]
