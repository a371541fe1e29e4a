"""GPU parity of the multispin-coded Ising path against its own CPU replay (bit-exact)."""
import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu
ising_hip = pytest.importorskip("ising_hip")


@pytest.mark.parametrize("L,R", [(2, 3), (4, 32), (6, 5), (8, 33), (16, 32), (12, 7)])
def test_msc_init_matches_int8_start(L, R):
    temps = np.linspace(4.4, 4.62, R)
    pt = ising_hip.PTEnsemble("ising_msc", [L], R, temps, 42)
    for r in range(R):
        assert np.array_equal(pt.get_spins(0, r), o.ising_init(L, 42, r)), r


@pytest.mark.parametrize("sizes,R,warmup,samples,ee,me", [([4, 8], 32, 5, 40, 10, 5), ([6, 16], 5, 10, 30, 3, 1),
                                                          ([8], 40, 0, 25, 5, 5), ([32], 32, 3, 12, 4, 2)])
def test_msc_pt_records_are_bit_exact(sizes, R, warmup, samples, ee, me):
    temps = o.linspace(4.40, 4.62, R)
    pt = ising_hip.PTEnsemble("ising_msc", sizes, R, temps, 7, exchange_every=ee, measure_every=me)
    pt.warmup(warmup)
    pt.sample(samples)
    tidx, bond, mag = pt.read_ising()
    for gi, L in enumerate(sizes):
        rt, rb, rm, final = o.msc_pt_run(L, R, temps, 7, warmup, samples, ee, me)
        assert np.array_equal(bond[:, gi, :].reshape(-1), rb), L
        assert np.array_equal(mag[:, gi, :].reshape(-1), rm), L
        assert np.array_equal(tidx[:, gi, :].reshape(-1), rt), L
        for r in range(R):
            assert np.array_equal(pt.get_spins(gi, r), final[r]), (L, r)


def test_msc_set_spins_touches_one_bit_lane():
    L, R = 8, 32
    pt = ising_hip.PTEnsemble("ising_msc", [L], R, np.linspace(4.4, 4.6, R), 3)
    before = [pt.get_spins(0, r) for r in range(R)]
    pt.set_spins(0, 5, np.ones(L ** 3, np.int8))
    for r in range(R):
        got = pt.get_spins(0, r)
        assert np.array_equal(got, np.ones(L ** 3, np.int8) if r == 5 else before[r])


def test_msc_rejects_field():
    with pytest.raises(ising_hip.IsingError):
        ising_hip.PTEnsemble("ising_msc", [8], 4, np.linspace(4.4, 4.6, 4), 1, h=0.1)
