"""Annealing schedules and the other LatticeGpu callers (Kibble-Zurek quench,
coarsening, annealed FSS) as batched drivers: bit-exact against the oracle's
replay of the same sweeps (same Philox counters, per-sweep acceptance tables)."""
import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu
ising_hip = pytest.importorskip("ising_hip")


def f32_beta(T):
    return float(np.float32(1.0 / T))


def replay(L, R, seed, betas, J=1.0, h=0.0, t0=0, states=None):
    """Oracle: replica r's lattice after sweeps t0.. with betas[k]; per-sweep (B, M)."""
    sp = states if states is not None else [o.ising_init(L, seed, r) for r in range(R)]
    B = np.zeros((len(betas), R), np.int64)
    M = np.zeros_like(B)
    for k, b in enumerate(betas):
        for r in range(R):
            o.ising_sweep(L, sp[r], t0 + k, seed, b, J, h, r)
            B[k, r], M[k, r] = o.ising_measure(L, sp[r])
    return sp, B, M


def test_anneal_schedule_is_bit_exact():
    sizes, R, seed = [16, 64], 3, 11
    pt = ising_hip.PTEnsemble("ising", sizes, R, [5.0] * R, seed)
    betas = [f32_beta(T) for T in np.linspace(6.0, 3.5, 9)]
    bond, mag = pt.anneal(betas, measure=True)
    for gi, L in enumerate(sizes):
        sp, B, M = replay(L, R, seed, betas)
        assert np.array_equal(bond[:, gi * R:(gi + 1) * R], B) and np.array_equal(mag[:, gi * R:(gi + 1) * R], M)
        sp, _, _ = replay(L, R, seed, [0.3, 0.31], t0=9, states=sp)
        if gi == 0:
            pt.anneal([0.3, 0.31])  # no measurement
        for r in range(R):
            assert np.array_equal(pt.get_spins(gi, r), sp[r]), (L, r)
    assert pt.sweeps() == 11


def test_kz_quench_densities_match_the_replay():
    n, tau, trials, warm, freeze = 8, 6, 3, 4, 2
    rho = ising_hip.kz_run(n, 1.0, 6.0, 1.0, tau, trials, warm, freeze, 0.5, seed=17)
    ramp = [6.0 + (1.0 - 6.0) * (k + 1) / tau for k in range(tau)]
    betas = [f32_beta(6.0)] * warm + [f32_beta(max(T, 0.01)) for T in ramp] + [f32_beta(0.5)] * freeze
    _, B, _ = replay(n, trials, 17, betas)
    bonds = 3.0 * n ** 3
    assert np.array_equal(rho, (bonds - B[-1]) / (2 * bonds))
    pts = ising_hip.run_kz_sweep(n, 1.0, 6.0, 1.0, [tau, 3], trials, warm, freeze, 0.5, base_seed=11)
    r6 = ising_hip.kz_run(n, 1.0, 6.0, 1.0, tau, trials, warm, freeze, 0.5, seed=11 + tau)
    assert pts[0][0] == tau and pts[0][1] == pytest.approx(r6.mean(), abs=1e-15)
    assert pts[0][2] == pytest.approx(r6.std(ddof=1) / np.sqrt(trials), abs=1e-15)


def test_coarsening_samples_match_the_replay():
    n, trials = 8, 2
    steps, rho = ising_hip.run_coarsening(n, 1.0, 10.0, 0.5, warmup_sweeps=3, total_steps=10, sample_every=3,
                                          seed=5, n_trials=trials)
    assert list(steps) == [0, 3, 6, 9]
    betas = [f32_beta(10.0)] * 3 + [f32_beta(0.5)] * 10
    _, B, _ = replay(n, trials, 5, betas)
    bonds = 3.0 * n ** 3
    want = (bonds - B[3 + np.array([0, 3, 6, 9])]) / (2 * bonds)
    assert np.array_equal(rho, want)


def test_annealed_fss_observables_match_the_replay():
    sizes, t_steps, warm, samp, seed, J, h = [4, 8], 3, 2, 5, 23, 1.0, 0.1
    res = ising_hip.run_fss(sizes, J, h, 4.0, 5.0, t_steps, warm, samp, seed)
    Jf, hf = float(np.float32(J)), float(np.float32(h))
    for gi, (L, obs) in enumerate(res):
        assert L == sizes[gi]
        sp = [o.ising_init(L, seed, 0)]
        t0 = 0
        want = {}
        for step in range(t_steps - 1, -1, -1):
            T = 4.0 + (5.0 - 4.0) * step / (t_steps - 1)
            b = f32_beta(T)
            sp, _, _ = replay(L, 1, seed, [b] * warm, Jf, hf, t0, sp)
            t0 += warm
            sp, B, M = replay(L, 1, seed, [b] * samp, Jf, hf, t0, sp)
            t0 += samp
            n3 = float(L ** 3)
            e = (-J * B[:, 0] - h * M[:, 0]) / n3
            ms = M[:, 0] / n3
            ma = np.abs(ms)
            bt = 1.0 / T
            want[step] = (T, e.mean(), ma.mean(), bt * bt * ((e * e).mean() - e.mean() ** 2) * n3,
                          bt * ((ms * ms).mean() - ms.mean() ** 2) * n3, (ma * ma).mean(), (ma ** 4).mean())
        for step in range(t_steps):
            got = obs[step]
            exp = want[step]
            vals = (got.temperature, got.energy, got.magnetisation, got.heat_capacity, got.susceptibility, got.m2,
                    got.m4)
            assert np.allclose(vals, exp, rtol=1e-12, atol=1e-12), (L, step, vals, exp)
