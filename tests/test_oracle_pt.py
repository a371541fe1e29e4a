"""Pins the oracle's replica exchange, linspace and jackknife with invariants and closed forms."""
import math

import numpy as np
import pytest

import oracle_lib as o


def _exchange(R, temps, energies, r2t, t2r, rnd, seed=42, stream=0):
    betas = 1.0 / np.asarray(temps, np.float64)
    return o.lib.o_exchange(R, betas, np.asarray(energies, np.float64), r2t, t2r, rnd, seed, stream, 0)


def test_linspace_matches_reference_definition():
    t = o.linspace(4.0, 5.0, 5)
    assert t[0] == 4.0 and t[4] == 5.0 and abs(t[2] - 4.5) < 1e-12  # parallel_tempering.rs:70-76
    assert list(o.linspace(4.4, 4.6, 1)) == [4.4]


def test_exchange_keeps_a_bijection():
    R = 17
    rng = np.random.default_rng(0)
    r2t = np.arange(R, dtype=np.int32)
    t2r = np.arange(R, dtype=np.int32)
    temps = np.linspace(4.4, 4.62, R)
    for rnd in range(2000):
        e = rng.normal(-1000, 30, R)
        _exchange(R, temps, e, r2t, t2r, rnd)
        assert np.array_equal(t2r[r2t], np.arange(R))
        assert sorted(r2t) == list(range(R))


def test_exchange_pairs_follow_round_parity():
    R = 6
    temps = np.linspace(1.0, 2.0, R)
    e = np.zeros(R)  # delta = 0 -> always accepted
    for rnd, pairs in ((0, [(0, 1), (2, 3), (4, 5)]), (1, [(1, 2), (3, 4)])):
        r2t = np.arange(R, dtype=np.int32)
        t2r = np.arange(R, dtype=np.int32)
        acc = _exchange(R, temps, e, r2t, t2r, rnd)
        assert acc == len(pairs)
        want = list(range(R))
        for i, j in pairs:
            want[i], want[j] = want[j], want[i]
        assert list(t2r) == want


def test_exchange_acceptance_frequency_is_exp_minus_delta():
    # two replicas, colder one holding the higher energy: delta > 0 every even round
    temps = np.array([1.0, 2.0])
    beta = 1.0 / temps
    e = np.array([-10.0, -12.0])  # replica 0 at T0 has E=-10, replica 1 at T1 has E=-12
    delta = (beta[0] - beta[1]) * (e[1] - e[0])
    assert delta < 0  # this arrangement always swaps
    e_hot_first = np.array([-12.0, -10.0])
    delta2 = (beta[0] - beta[1]) * (e_hot_first[1] - e_hot_first[0])
    assert delta2 > 0
    n, acc = 40000, 0
    for rnd in range(0, 2 * n, 2):
        r2t = np.arange(2, dtype=np.int32)
        t2r = np.arange(2, dtype=np.int32)
        acc += _exchange(2, temps, e_hot_first, r2t, t2r, rnd, seed=11)
    p = math.exp(-delta2)
    assert abs(acc / n - p) < 5 * math.sqrt(p * (1 - p) / n)


def test_jackknife_constant_series_has_zero_error():
    e = np.full(400, -1.25)
    m = np.full(400, 0.5)
    out = o.jackknife(e, m, beta=0.2, n3=512.0, n_blocks=20)
    assert out[0] == -1.25 and out[2] == 0.5 and out[4] == 0.25 and out[6] == 0.0625
    assert np.all(out[1::2] == 0.0)
    assert abs(out[8]) < 1e-12 and abs(out[10]) < 1e-12


def test_jackknife_two_blocks_closed_form():
    a, b, k = 1.0, 3.0, 50
    e = np.array([a] * k + [b] * k)
    m = np.array([0.25] * (2 * k))
    out = o.jackknife(e, m, beta=1.0, n3=1.0, n_blocks=2)
    assert out[0] == 2.0 and abs(out[1] - abs(a - b) / 2) < 1e-15


def test_jackknife_linear_series_closed_form():
    nb, bs = 20, 37
    n = nb * bs + 5  # trailing samples beyond n_used are dropped
    e = np.arange(n, dtype=np.float64)
    out = o.jackknife(e, np.zeros(n), beta=1.0, n3=1.0, n_blocks=nb)
    assert abs(out[0] - (nb * bs - 1) / 2) < 1e-9
    assert abs(out[1] - bs * math.sqrt((nb + 1) / 12.0)) < 1e-9


def test_jackknife_susceptibility_closed_form():
    # each block holds the same alternating pattern -> all jackknife estimates equal the full one
    x, y, beta, n3 = 0.9, 0.3, 0.25, 64.0
    m = np.array([x, y] * 200)
    out = o.jackknife(np.zeros_like(m), m, beta=beta, n3=n3, n_blocks=20)
    assert abs(out[10] - beta * n3 * (x - y) ** 2 / 4) < 1e-12 and out[11] < 1e-12


def test_jackknife_too_few_samples_returns_zeros():
    out = o.jackknife(np.ones(5), np.ones(5), 1.0, 1.0, 20)
    assert np.all(out == 0.0)
    out = o.jackknife(np.ones(5), np.ones(5), 1.0, 1.0, 1)
    assert np.all(out == 0.0)


def test_pt_run_records_layout():
    temps = o.linspace(4.4, 4.6, 4)
    tidx, bond, mag, final = o.ising_pt_run(4, 4, temps, 42, warmup=10, samples=40, exchange_every=10,
                                            measure_every=5)
    assert len(tidx) == 8 * 4
    for k in range(8):
        assert sorted(tidx[4 * k: 4 * k + 4]) == [0, 1, 2, 3]
    assert np.all(np.abs(mag) <= 64) and np.all(np.abs(bond) <= 192)
    for r in range(4):
        assert o.ising_measure(4, final[r].copy())[0] % 2 == 0


@pytest.mark.parametrize("rho", [0.0, 0.5, 0.9])
def test_tau_int_of_ar1_matches_closed_form(rho):
    # AR(1): x_t = rho x_{t-1} + e_t has rho(t) = rho^t and tau_int = (1 + rho) / (2 (1 - rho))
    rng = np.random.default_rng(1)
    n = 200000
    e = rng.normal(size=n)
    x = np.empty(n)
    x[0] = e[0] / math.sqrt(1 - rho * rho)
    for i in range(1, n):
        x[i] = rho * x[i - 1] + e[i]
    tau, w = o.tau_int(x)
    exact = (1 + rho) / (2 * (1 - rho))
    # Sokal: relative error ~ sqrt(2 (2W + 1) / n)
    assert abs(tau - exact) < 4 * exact * math.sqrt(2 * (2 * w + 1) / n) + 0.02
    assert w >= 6 * tau - 1


def test_tau_int_library_matches_oracle():
    import ising_hip
    rng = np.random.default_rng(2)
    x = np.cumsum(rng.normal(size=5000)) * 0.01 + rng.normal(size=5000)
    assert ising_hip.tau_int(x) == o.tau_int(x)
    assert ising_hip.tau_int(np.ones(10)) == (0.5, 0)
