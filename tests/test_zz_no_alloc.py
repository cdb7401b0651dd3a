"""swap_get_particle_array_noalloc (no counterpart in the reference; for
callers that cannot allocate device memory) against swap_get_particle_array:
the absolute positions staged through the library's fixed buffer are the
same bits as those staged through a per-call allocation, on two calls in a
row (the second with fewer particles, so stale contents of the shared buffer
would show), and a load larger than the buffer is refused."""
import numpy as np
import pytest

from swap_amd import api, physics

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 1 << 20  # SWAP_NOALLOC_MAX of include/swap_hip.h
SWAP_EINVAL = -1


def _noalloc(s, which, shape, dtype):
    out = np.full(shape, np.nan if dtype == np.float64 else 0, dtype)
    rc = s.lib.swap_get_particle_array_noalloc(s.h, which, out.ctypes.data)
    return rc, out


def test_noalloc_export_matches_the_allocating_one_on_two_calls():
    p = physics.workload(32, 16, 16, 8)
    n1 = p["npart"]
    s = physics.make_swap(p, capacity=n1, flags=api.TRACK_IDS)
    s.load_maxwellian(n1, p["vsw"], p["vth"])
    s.step(2)
    rc, xa = _noalloc(s, api.P_XP, (3, s.n), np.float64)
    assert rc == 0, s.lib.swap_last_error(s.h).decode()
    assert np.array_equal(xa, s.xp())
    # second call, fewer particles (another load into the same context)
    n2 = n1 // 3
    s.load_maxwellian(n2, p["vsw"], p["vth"])
    s.step(1)
    rc, xb = _noalloc(s, api.P_XP, (3, n2), np.float64)
    assert rc == 0
    assert np.array_equal(xb, s.xp())
    # arrays without staging go through as swap_get_particle_array copies them
    rc, vp = _noalloc(s, api.P_VP, (3, n2), np.float32)
    assert rc == 0 and np.array_equal(vp, s.vp())
    rc, ids = _noalloc(s, api.P_ID, (n2,), np.uint32)
    assert rc == 0 and np.array_equal(ids, s.particle_array(api.P_ID))


def test_noalloc_export_refuses_more_than_its_buffer():
    p = physics.workload(16, 16, 16, 1)
    n = NOALLOC_MAX + 1
    s = physics.make_swap(p, capacity=n)
    s.load_maxwellian(n, p["vsw"], p["vth"])
    rc, _ = _noalloc(s, api.P_XP, (3, n), np.float64)
    assert rc == SWAP_EINVAL
    msg = s.lib.swap_last_error(s.h).decode()
    assert "at most %d" % NOALLOC_MAX in msg
    # the allocating entry point still takes it
    assert s.xp().shape == (3, n)
