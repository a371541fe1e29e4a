"""ising_reduce_ising_noalloc: the allocation-free variant of
ising_reduce_ising (reduce_gpu.rs:23-66), whose partial sums live in a fixed
per-device array, against the allocating entry point on two calls in a row,
the second on a smaller lattice (the array keeps the first call's partials)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ising_hip = pytest.importorskip("ising_hip")
hipmem = pytest.importorskip("hipmem")


def _noalloc():
    return ising_hip.lib.ising_reduce_ising_noalloc


def test_noalloc_equals_allocating_reduce_on_two_calls():
    f = _noalloc()
    out = hipmem.DeviceArray(nbytes=16)
    for n, seed, j in ((96, 1, 1.0), (20, 2, 0.5)):
        rng = np.random.default_rng(seed)
        sp = np.where(rng.random(n ** 3) < 0.6, 1, -1).astype(np.int8)
        d = hipmem.DeviceArray(sp)
        rc = f(d.address, n, j, out.address, None)
        assert rc == 0, ising_hip.lib.ising_last_error()
        got = out.download(np.float64, 2)
        want = ising_hip.reduce_ising(d.address, n, j)
        assert got[0] == want[0] and got[1] == want[1], (n, got, want)
        assert got[1] == float(sp.astype(np.int64).sum())


def test_noalloc_rejects_lattices_beyond_its_buffer():
    f = _noalloc()
    out = hipmem.DeviceArray(nbytes=16)
    d = hipmem.DeviceArray(np.ones(8, np.int8))
    assert f(d.address, 646, 1.0, out.address, None) == -1
    assert b"645" in ising_hip.lib.ising_last_error()
    assert f(None, 2, 1.0, out.address, None) == -1
    assert f(d.address, 2, 1.0, out.address, None) == 0
    assert np.array_equal(out.download(np.float64, 2), [-24.0, 8.0])
