"""reduce_gpu::* (src/cuda/reduce_gpu.rs:8-178) on caller-held device lattices.

Ising totals are exact (int64) and must equal the oracle's bond and spin sums
bit for bit; continuous totals are FP64 sums of the same FP32 spins and must
equal the oracle's FP64 sums to 1e-12 relative (only the summation order
differs), and be bitwise repeatable from call to call.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu
ising_hip = pytest.importorskip("ising_hip")
hipmem = pytest.importorskip("hipmem")


def _oracle_ising(n, sp):
    b, m = C.c_int64(), C.c_int64()
    o.lib.o_ising_measure(n, np.ascontiguousarray(sp, np.int8), C.byref(b), C.byref(m))
    return b.value, m.value


def _oracle_cont(n, nc, sp, J):
    e = C.c_double()
    m = np.zeros(3)
    o.lib.o_cont_measure(n, nc, np.ascontiguousarray(sp, np.float64), J, C.byref(e), m)
    return e.value, m


def _unit_vectors(rng, count, nc):
    v = rng.normal(size=(count, nc))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32).reshape(-1)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 16, 33, 64, 130])
def test_reduce_ising_exact_against_oracle(n):
    rng = np.random.default_rng(n)
    sp = np.where(rng.random(n ** 3) < 0.55, 1, -1).astype(np.int8)
    d = hipmem.DeviceArray(sp)
    ising_hip.load_reduce_kernels()
    bond, mag = _oracle_ising(n, sp)
    for j in (1.0, 0.5, -1.0):
        e, m = ising_hip.reduce_ising(d.address, n, j)
        assert e == -j * bond and m == mag
    e1, m1 = ising_hip.measure_ising_gpu(d.address, n, 1.0)
    assert e1 == -bond / n ** 3 and m1 == abs(mag) / n ** 3
    # ground states: E/N = -3, |M|/N = 1; Neel state: E/N = +3 (even n)
    for val in (1, -1):
        d.upload(np.full(n ** 3, val, np.int8))
        e1, m1 = ising_hip.measure_ising_gpu(d.address, n, 1.0)
        assert e1 == -3.0 and m1 == 1.0
    if n % 2 == 0:
        z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
        d.upload(np.where((x + y + z) % 2 == 0, 1, -1).astype(np.int8).reshape(-1))
        e1, m1 = ising_hip.measure_ising_gpu(d.address, n, 1.0)
        assert e1 == 3.0 and m1 == 0.0


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("n", [1, 2, 7, 16, 40])
def test_reduce_continuous_against_oracle(nc, n):
    rng = np.random.default_rng(100 * nc + n)
    sp = _unit_vectors(rng, n ** 3, nc)
    d = hipmem.DeviceArray(sp)
    for j in (1.0, 0.75):
        e_ref, m_ref = _oracle_cont(n, nc, sp.astype(np.float64), j)
        got = ising_hip.measure_continuous_gpu(d.address, n, nc, j)
        assert abs(got[0] - e_ref) <= 1e-12 * max(1.0, abs(e_ref))
        for k in range(nc):
            assert abs(got[1 + k] - m_ref[k]) <= 1e-12 * n ** 3
        if nc == 2:
            assert got[3] == 0.0
        assert ising_hip.measure_continuous_gpu(d.address, n, nc, j) == got  # same bits every call
        assert ising_hip.reduce_continuous_energy(d.address, n, nc, j) == got[0]
    # aligned state: E = -3 j N exactly
    al = np.zeros((n ** 3, nc), np.float32)
    al[:, 0] = 1.0
    d.upload(al.reshape(-1))
    assert ising_hip.reduce_continuous_energy(d.address, n, nc, 2.0) == -6.0 * n ** 3


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("n_sites", [1, 5, 1000, 4096, 4097, 70001])
def test_reduce_continuous_mag_any_site_count(nc, n_sites):
    rng = np.random.default_rng(n_sites)
    sp = _unit_vectors(rng, n_sites, nc)
    d = hipmem.DeviceArray(sp)
    mx, my, mz = ising_hip.reduce_continuous_mag(d.address, n_sites, nc)
    want = sp.reshape(-1, nc).astype(np.float64).sum(axis=0)
    assert abs(mx - want[0]) <= 1e-12 * n_sites and abs(my - want[1]) <= 1e-12 * n_sites
    assert mz == (0.0 if nc == 2 else mz) and (nc == 2 or abs(mz - want[2]) <= 1e-12 * n_sites)
    assert ising_hip.reduce_continuous_mag(d.address, 0, nc) == (0.0, 0.0, 0.0)


def test_reduce_argument_errors():
    d = hipmem.DeviceArray(np.ones(8, np.int8))
    with pytest.raises(ising_hip.IsingError):
        ising_hip.reduce_ising(d.address, 0)
    with pytest.raises(ising_hip.IsingError):
        ising_hip.reduce_ising(d.address, 1025)
    with pytest.raises(ising_hip.IsingError):
        ising_hip.reduce_ising(0, 2)
    with pytest.raises(ising_hip.IsingError):
        ising_hip.measure_continuous_gpu(d.address, 1, 4)
