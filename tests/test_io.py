"""The reference's on-disk formats without a GPU: Fortran unformatted
records (4-byte length markers around each WRITE, -i4 -real-size 32) and the
record layout of para.dat (maind.f:237-255) and c.coord.dat
(initial.f:285-297); the load file read by both codes."""
import struct
import types

import numpy as np
import pytest

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


def test_load_file_round_trip_through_the_oracle_reader(tmp_path):
    """the seeded load written by io.write_load reads back bit for bit in the
    oracle's C reader and in io.read_load"""
    import oracle as O
    nc = (12, 6, 4)
    n = 3000
    cell, frac, v = O.load_maxwellian(n, 77, nc, 400.0, 30.0)
    rng = np.random.default_rng(2)
    bp = rng.random(n).astype(np.float32)
    ids = np.arange(n, dtype=np.uint32)[::-1].copy()
    path = str(tmp_path / "load.bin")
    io.write_load(path, cell, frac, v, v * 0.5, nc, (360.0, 360.0, 360.0), seed=77,
                  beta_p=bp, mrat=np.full(n, 0.25), ids=ids)
    c2, f2, v2, v12, bp2, mr2, id2 = O.read_load(path)
    assert np.array_equal(c2, cell) and np.array_equal(f2, frac)
    assert np.array_equal(v2, v) and np.array_equal(v12, v * 0.5)
    assert np.array_equal(bp2, bp) and np.all(mr2 == np.float32(0.25))
    assert np.array_equal(id2, ids)
    z = io.read_load(path)
    assert z["n"] == n and tuple(z["nc"]) == nc and z["seed"] == 77
    assert np.array_equal(z["cell"], cell) and np.array_equal(z["frac"], frac)
    # a short file is refused by both readers
    with open(path, "rb") as f:
        head = f.read(64 + 4 * n + 100)
    short = str(tmp_path / "short.bin")
    with open(short, "wb") as f:
        f.write(head)
    with pytest.raises(ValueError):
        O.read_load(short)
