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


def test_fine_grid_table_wham_equals_single_histogram_for_one_run():
    rng = np.random.default_rng(8)
    e = rng.normal(-900.0, 12.0, 20000)
    m = np.abs(rng.normal(0.35, 0.06, 20000))
    temps = np.linspace(4.45, 4.55, 7)
    a = ising_hip.reweight_fine([e], [m], [1.0 / 4.5], temps, n=8, method="wham", tol=1e-12)
    b = ising_hip.reweight_fine([e], [m], [1.0 / 4.5], temps, n=8, method="single")
    for key in ("M", "M2", "M4", "chi_conn", "Cv", "U", "E", "chi_err", "M_err", "U_err"):
        assert np.allclose(a[key], b[key], rtol=1e-9, atol=1e-12), key


def test_fine_grid_binder_crossing_from_a_pt_run():
    """Ensemble records -> per-temperature series -> WHAM on a fine grid ->
    Binder crossing of L = 8 and 16 (the reference's analysis chain) lands
    near the 3D Ising T_c (finite-size shifts at these sizes are ~0.02)."""
    R = 16
    temps = np.linspace(4.40, 4.62, R)
    pt = ising_hip.PTEnsemble("ising", [8, 16], R, temps, 17, exchange_every=5, measure_every=2,
                              record_capacity=60000)
    pt.warmup(5000)
    pt.sample(100000)
    tidx, bond, mag = pt.read_ising()
    fine = np.linspace(4.42, 4.60, 200)
    curves = {}
    for g, L in enumerate((8, 16)):
        e_l, m_l = ising_hip.series_by_temperature(tidx[:, g, :], bond[:, g, :], mag[:, g, :], L)
        out = ising_hip.reweight_fine(e_l, m_l, 1.0 / temps, fine, n=L, method="wham")
        assert np.all(np.isfinite(out["U"])) and np.all(out["chi_conn"] > 0)
        curves[L] = (fine, out["U"])
    (l1, l2, tc), = ising_hip.binder_crossings(curves)
    assert (l1, l2) == (8, 16) and abs(tc - 4.5115) < 0.05, tc
