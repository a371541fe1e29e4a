"""Pins the oracle's Ising checkerboard replay.

Checks: exact enumeration (reference fixture, 2D 4x4), brute-force Boltzmann
distribution on the 2x2x2 cube, T->0 and T->inf limits (metropolis.rs:70-100),
ground-state energies (observables.rs:189-243), determinism and thresholds.
"""
import csv
import itertools
import math
import os

import numpy as np
import pytest

import oracle_lib as o

HERE = os.path.dirname(os.path.abspath(__file__))


def jk_mean_err(x, nb=20):
    x = np.asarray(x, np.float64)
    bs = len(x) // nb
    x = x[: bs * nb]
    full = x.mean()
    jk = np.array([(x.sum() - x[b * bs:(b + 1) * bs].sum()) / (len(x) - bs) for b in range(nb)])
    return full, math.sqrt((nb - 1) / nb * ((jk - full) ** 2).sum())


def enumerate_2d(L):
    n = L * L
    states = ((np.arange(2 ** n)[:, None] >> np.arange(n)[None, :]) & 1).astype(np.int64) * 2 - 1
    s = states.reshape(-1, L, L)
    B = (s * np.roll(s, -1, axis=2) + s * np.roll(s, -1, axis=1)).sum(axis=(1, 2))
    M = s.sum(axis=(1, 2))
    return B, M


def exact_observables(B, M, n, T):
    beta = 1.0 / T
    w = np.exp(beta * (B - B.max()))
    w /= w.sum()
    e = -B / n
    m = np.abs(M) / n
    E = (w * e).sum()
    Mabs = (w * m).sum()
    M2 = (w * m * m).sum()
    M4 = (w * m ** 4).sum()
    Cv = beta ** 2 * n * ((w * e * e).sum() - E * E)
    chi = beta * n * (M2 - Mabs ** 2)
    return E, Mabs, M2, M4, Cv, chi


def test_exact_enumeration_matches_reference_fixture():
    B, M = enumerate_2d(4)
    rows = list(csv.DictReader(open(os.path.join(HERE, "golden", "exact_enumeration_2d_4x4.csv"))))
    assert len(rows) >= 20
    for r in rows:
        got = exact_observables(B, M, 16, float(r["T"]))
        want = [float(r[k]) for k in ("E_exact", "M_exact", "M2_exact", "M4_exact", "Cv_exact", "chi_conn_exact")]
        assert np.allclose(got, want, rtol=1e-9, atol=1e-12), (r["T"], got, want)


@pytest.mark.parametrize("T", [1.8, 2.3, 3.0])
def test_2d_checkerboard_chain_matches_exact_enumeration(T):
    B, M = enumerate_2d(4)
    E, Mabs, M2, _, _, _ = exact_observables(B, M, 16, T)
    b, m = o.ising_chain(2, 4, 1234 + int(T * 10), 1.0 / T, 1000, 200000)
    for series, want in ((-b / 16.0, E), (np.abs(m) / 16.0, Mabs), ((m / 16.0) ** 2, M2)):
        got, err = jk_mean_err(series)
        assert abs(got - want) < 4.5 * err + 1e-12, (T, got, want, err)


def test_3d_2x2x2_chain_matches_brute_force_boltzmann():
    L, n, T = 2, 8, 3.0
    sp = ((np.arange(256)[:, None] >> np.arange(8)[None, :]) & 1).astype(np.int8) * 2 - 1
    Bs = np.array([o.ising_measure(L, s.copy())[0] for s in sp])
    beta = 1.0 / T
    w = np.exp(beta * (Bs - Bs.max()))
    w /= w.sum()
    b, m = o.ising_chain(3, L, 99, beta, 500, 300000)
    # probability of every energy level, with block-jackknife errors
    for level in np.unique(Bs):
        p_exact = w[Bs == level].sum()
        got, err = jk_mean_err((b == level).astype(np.float64))
        assert abs(got - p_exact) < 4.5 * err + 2e-4, (level, got, p_exact, err)
    got, err = jk_mean_err((m / 8.0) ** 2)
    M2 = (w * (sp.sum(axis=1) / 8.0) ** 2).sum()
    assert abs(got - M2) < 4.5 * err


@pytest.mark.parametrize("kind", ["metropolis", "wolff"])
def test_reference_cpu_chains_match_brute_force_boltzmann(kind):
    """The restatements of the reference's CPU algorithms (random-site
    Metropolis, metropolis.rs:8-27, which the CPU baseline times; Wolff,
    wolff.rs:27-65) sample the Boltzmann distribution of the 2x2x2 cube."""
    L, T = 2, 3.0
    sp = ((np.arange(256)[:, None] >> np.arange(8)[None, :]) & 1).astype(np.int8) * 2 - 1
    Bs = np.array([o.ising_measure(L, s.copy())[0] for s in sp])
    beta = 1.0 / T
    w = np.exp(beta * (Bs - Bs.max()))
    w /= w.sum()
    b, m = o.ref_ising_chain(kind, L, 7, beta, 500, 300000)
    for level in np.unique(Bs):
        p_exact = w[Bs == level].sum()
        got, err = jk_mean_err((b == level).astype(np.float64))
        assert abs(got - p_exact) < 4.5 * err + 2e-4, (kind, level, got, p_exact, err)
    for series, want in (((m / 8.0) ** 2, (w * (sp.sum(axis=1) / 8.0) ** 2).sum()),
                         (np.abs(m) / 8.0, (w * np.abs(sp.sum(axis=1)) / 8.0).sum())):
        got, err = jk_mean_err(series)
        assert abs(got - want) < 4.5 * err, (kind, got, want, err)


def test_ground_state_is_fixed_at_low_temperature():
    L = 4
    sp = np.ones(L ** 3, np.int8)
    for t in range(100):
        o.ising_sweep(L, sp, t, 42, beta=100.0)
    assert (sp == 1).all()
    B, M = o.ising_measure(L, sp)
    assert -B / L ** 3 == -3.0 and M == L ** 3  # E_ground = -3 J per spin


def test_antiferromagnet_energy_and_high_temperature_disorder():
    L = 4
    x, y, z = np.meshgrid(np.arange(L), np.arange(L), np.arange(L), indexing="ij")
    neel = np.where(((x + y + z) & 1) == 0, 1, -1).astype(np.int8).transpose(2, 1, 0).reshape(-1)
    B, M = o.ising_measure(L, neel)
    assert -B / L ** 3 == 3.0 and M == 0  # observables.rs:232-243
    sp = np.ones(8 ** 3, np.int8)
    for t in range(200):
        o.ising_sweep(8, sp, t, 5, beta=0.01)
    assert abs(sp.astype(np.int64).sum()) / 8 ** 3 < 0.3  # metropolis.rs:89-100
    assert set(np.unique(sp)) <= {-1, 1}


def test_thresholds_encode_metropolis_probabilities():
    beta = 1.0 / 4.5115
    hi, lo = o.thresholds(beta)
    assert np.all(hi <= 128) and np.all(lo < 2 ** 25)
    thr = (hi.astype(np.uint64) << 25) + lo.astype(np.uint64)
    for b in (0, 1):
        for n_up in range(7):
            s = 1 if b else -1
            dE = 2.0 * s * (2 * n_up - 6)
            p = 1.0 if dE <= 0 else math.exp(-beta * dE)
            k = b * 7 + n_up
            if p >= 1.0:
                assert thr[k] == 1 << 32
            else:
                assert thr[k] == math.floor(p * 2 ** 32)
    # with a field the table is no longer symmetric
    hi2, _ = o.thresholds(beta, 1.0, 0.5)
    assert not np.array_equal(hi2[:7], hi2[7:][::-1])


def test_init_is_iid_pm1_and_deterministic():
    a = o.ising_init(16, 42, 0)
    b = o.ising_init(16, 42, 0)
    c = o.ising_init(16, 42, 1)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert set(np.unique(a)) == {-1, 1}
    assert abs(a.astype(np.float64).mean()) < 4.0 / math.sqrt(a.size)


def test_sweep_is_deterministic_and_changes_state():
    L = 6
    s1 = o.ising_init(L, 3)
    s2 = s1.copy()
    start = s1.copy()
    for t in range(10):
        o.ising_sweep(L, s1, t, 3, 0.5)
        o.ising_sweep(L, s2, t, 3, 0.5)
    assert np.array_equal(s1, s2) and not np.array_equal(s1, start)


def test_split_acceptance_is_the_32bit_comparison():
    """All-up lattice, colour 0: every site has dE = 12 J and the same threshold
    T = hi * 2^25 + lo with lo / 2^25 = 1/2, so the tie path (r7 == hi, drawn
    with probability 1/128) carries 1/256 of the acceptance.  The flip
    frequency must be exp(-12 beta) to within 4 sigma; dropping the tie draw
    (ties never or always accepted) moves it by about 11 sigma."""
    L, reps = 16, 1000
    p = 70.5 / 128
    beta = -math.log(p) / 12.0
    hi, lo = o.thresholds(beta)
    assert hi[13] == 70 and abs(lo[13] / 2 ** 25 - 0.5) < 1e-6
    st = o.lib.o_stream(L, 0)
    flips = 0
    for t in range(reps):
        sp = np.ones(L ** 3, np.int8)
        o.lib.o_ising_halfsweep(L, sp, 0, t, 77, st, hi, lo)
        flips += int((sp < 0).sum())
    n = reps * L ** 3 // 2
    sigma = math.sqrt(p * (1 - p) / n)
    assert abs(flips / n - p) < 4 * sigma, (flips / n, p, sigma)
