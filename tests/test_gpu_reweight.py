"""On-device FP64 WHAM reweighting against the FP64 restatement of
reweighting.py:45-108 and against the closed-form two-level system."""
import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import reweight as ref  # noqa: E402

pytestmark = pytest.mark.gpu
ising_hip = pytest.importorskip("ising_hip")


def test_two_level_system_closed_form():
    # E in {0, 1}, each with one state: <E>(beta) = e^-b / (1 + e^-b); exact counts per run
    betas = np.array([0.2, 0.7, 1.3, 2.0])
    lists = []
    for b in betas:
        p1 = math.exp(-b) / (1 + math.exp(-b))
        n = 100000
        k1 = int(round(p1 * n))
        lists.append(np.array([1.0] * k1 + [0.0] * (n - k1)))
    targets = np.linspace(0.1, 2.5, 13)
    f, it, out = ising_hip.wham(lists, betas, [lists], targets, max_iter=1000, tol=1e-12)
    exact = np.exp(-targets) / (1 + np.exp(-targets))
    assert np.max(np.abs(out[:, 0] - exact)) < 2e-5
    # f_k - f_0 = -log(Z(b_k) / Z(b_0)), Z(b) = 1 + e^-b (up to count rounding)
    z = 1 + np.exp(-betas)
    assert np.max(np.abs(f - (-np.log(z / z[0])))) < 1e-4


def test_matches_fp64_restatement_on_pt_like_data():
    rng = np.random.default_rng(0)
    betas = 1.0 / np.linspace(4.4, 4.62, 8)
    N = 512
    lists, mags = [], []
    for b in betas:
        mu = -N * (1.0 + 0.8 * (b - betas.mean()) / (betas.max() - betas.min()))
        e = rng.normal(mu, 0.03 * N, 3000)
        lists.append(e)
        mags.append(np.abs(rng.normal(0.4, 0.1, 3000)))
    targets = 1.0 / np.linspace(4.40, 4.62, 40)
    f_g, it_g, out_g = ising_hip.wham(lists, betas, [lists, mags], targets)
    f_r, it_r, out_r = ref.wham(lists, betas, [lists, mags], targets)
    assert it_g == it_r
    assert np.max(np.abs(f_g - f_r)) < 1e-9 * max(1.0, np.max(np.abs(f_r)))
    assert np.max(np.abs(out_g - out_r) / np.maximum(np.abs(out_r), 1e-300)) < 1e-10


def test_single_run_is_single_histogram_reweighting():
    rng = np.random.default_rng(3)
    e = rng.normal(-300.0, 5.0, 20000)
    b0, bt = 0.22, 0.225
    f, it, out = ising_hip.wham([e], [b0], [[e]], [bt])
    w = np.exp(-(bt - b0) * e - np.max(-(bt - b0) * e))  # reweighting.py:15-38
    w /= w.sum()
    assert abs(out[0, 0] - (w * e).sum()) < 1e-9 * abs(out[0, 0])
