"""Pins the FP64 restatement of reweighting.py:45-108 (oracle/reweight.py)
on cases with closed-form answers."""
import math
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import reweight as ref  # noqa: E402


def test_two_level_system_closed_form():
    betas = np.array([0.2, 0.7, 1.3, 2.0])
    lists = []
    for b in betas:
        p1 = math.exp(-b) / (1 + math.exp(-b))
        n = 100000
        k1 = int(round(p1 * n))
        lists.append(np.array([1.0] * k1 + [0.0] * (n - k1)))
    targets = np.linspace(0.1, 2.5, 13)
    f, it, out = ref.wham(lists, betas, [lists], targets, n_iter=1000, tol=1e-12)
    exact = np.exp(-targets) / (1 + np.exp(-targets))
    assert np.max(np.abs(out[:, 0] - exact)) < 2e-5
    z = 1 + np.exp(-betas)
    assert np.max(np.abs(f - (-np.log(z / z[0])))) < 1e-4
    assert f[0] == 0.0


def test_single_run_is_single_histogram_reweighting():
    rng = np.random.default_rng(3)
    e = rng.normal(-300.0, 5.0, 20000)
    b0, bt = 0.22, 0.225
    f, it, out = ref.wham([e], [b0], [[e]], [bt])
    w = np.exp(-(bt - b0) * e - np.max(-(bt - b0) * e))
    w /= w.sum()
    assert abs(out[0, 0] - (w * e).sum()) < 1e-9 * abs(out[0, 0])


def test_gaussian_energies_reweight_to_the_exact_mean():
    # energy density of states g(E) ~ exp(-E^2 / (2 s^2)) gives, at beta,
    # Gaussian energies with mean -beta s^2 and variance s^2
    rng = np.random.default_rng(5)
    s = 3.0
    betas = np.array([0.5, 0.8, 1.1])
    lists = [rng.normal(-b * s * s, s, 200000) for b in betas]
    targets = np.array([0.6, 0.9, 1.0])
    f, it, out = ref.wham(lists, betas, [lists], targets)
    assert np.max(np.abs(out[:, 0] - (-targets * s * s))) < 0.03
    # f_k = log(Z(b_0)/Z(b_k)) with Z(b) ~ exp(b^2 s^2 / 2)
    exact_f = -(betas ** 2 - betas[0] ** 2) * s * s / 2
    assert np.max(np.abs(f - exact_f)) < 0.02
