"""2D square-lattice Ising on the GPU: bit-exact against the oracle, exact
enumeration of 4x4 (the reference's validation fixture), and the Binder
crossing at the exact Onsager point T_c = 2 / ln(1 + sqrt 2)."""
import csv
import math
import os

import numpy as np
import pytest

import oracle_lib as o
from test_gpu_equilibrium import binder, mean_err, reference_crossing

pytestmark = pytest.mark.gpu
ising_hip = pytest.importorskip("ising_hip")
HERE = os.path.dirname(os.path.abspath(__file__))
T_ONSAGER = 2.0 / math.log(1.0 + math.sqrt(2.0))


@pytest.mark.parametrize("sizes,R,h", [([4, 8], 4, 0.0), ([6, 16], 3, 0.2), ([32], 5, 0.0)])
def test_2d_pt_is_bit_exact(sizes, R, h):
    temps = o.linspace(2.1, 2.4, R)
    pt = ising_hip.PTEnsemble("ising2d", sizes, R, temps, 13, exchange_every=4, measure_every=2, h=h)
    pt.warmup(6)
    pt.sample(24)
    tidx, bond, mag = pt.read_ising()
    for gi, L in enumerate(sizes):
        rt, rb, rm, final = o.ising2d_pt_run(L, R, temps, 13, 6, 24, 4, 2, h=h)
        assert np.array_equal(tidx[:, gi, :].reshape(-1), rt)
        assert np.array_equal(bond[:, gi, :].reshape(-1), rb)
        assert np.array_equal(mag[:, gi, :].reshape(-1), rm)
        for r in range(R):
            assert np.array_equal(pt.get_spins(gi, r), final[r])


def test_2d_4x4_matches_exact_enumeration_fixture():
    rows = [r for r in csv.DictReader(open(os.path.join(HERE, "golden", "exact_enumeration_2d_4x4.csv")))
            if float(r["T"]) in (1.5, 2.0, 2.5, 3.0)]
    temps = np.array([float(r["T"]) for r in rows])
    R = len(temps)
    pt = ising_hip.PTEnsemble("ising2d", [4], R, temps, 99, exchange_every=10 ** 9, measure_every=1,
                              record_capacity=200002)
    pt.warmup(1000)
    pt.sample(200000)
    tidx, bond, mag = pt.read_ising()
    for k, r in enumerate(rows):
        e = -bond[:, 0, k] / 16.0
        m = np.abs(mag[:, 0, k]) / 16.0
        for ser, col in ((e, "E_exact"), (m, "M_exact"), (m * m, "M2_exact")):
            g, ge = mean_err(ser)
            assert abs(g - float(r[col])) < 4.5 * ge + 1e-9, (r["T"], col, g, float(r[col]), ge)


def test_2d_binder_crossing_at_onsager_point():
    R = 24
    temps = o.linspace(2.20, 2.34, R)
    U = {}
    for L, n in ((16, 200000), (32, 200000)):
        pt = ising_hip.PTEnsemble("ising2d", [L], R, temps, 7, exchange_every=5, measure_every=5,
                                  record_capacity=n // 5 + 2)
        pt.warmup(10000)
        pt.sample(n)
        tidx, bond, mag = pt.read_ising()
        t = tidx[:, 0, :].reshape(-1)
        m = np.abs(mag[:, 0, :].reshape(-1)) / L ** 2
        U[L] = np.array([binder(m[t == k])[0] for k in range(R)])
    tc = reference_crossing(np.array(temps), U[16], U[32])
    assert tc is not None and abs(tc - T_ONSAGER) < 0.01, (tc, T_ONSAGER)
    assert ising_hip.binder_crossings({16: (temps, U[16]), 32: (temps, U[32])})[0][2] == pytest.approx(tc, abs=1e-10)
