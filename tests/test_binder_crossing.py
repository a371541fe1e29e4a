"""Binder-cumulant crossing finder (the rule of analyze_gpu_fss.py:256-277):
closed forms that the not-a-knot spline reproduces exactly (lines, cubics),
scipy's interp1d(kind="cubic") on noisy and unevenly sampled curves, and the
edge cases (no crossing, too few temperatures, unsorted input).  Host code:
no GPU needed."""
import numpy as np
import pytest
from scipy.interpolate import interp1d

import ising_hip


def scipy_crossing(T1, U1, T2, U2):
    lo, hi = max(np.min(T1), np.min(T2)), min(np.max(T1), np.max(T2))
    Tc = np.linspace(lo, hi, 2000)
    d = interp1d(T1, U1, kind="cubic")(Tc) - interp1d(T2, U2, kind="cubic")(Tc)
    for j in range(len(d) - 1):
        if d[j] * d[j + 1] < 0:
            return Tc[j] - d[j] * (Tc[j + 1] - Tc[j]) / (d[j + 1] - d[j])
    return None


def test_lines_cross_at_the_closed_form():
    T = np.linspace(4.40, 4.62, 32)
    tc = ising_hip.binder_crossing(T, 0.50 - 0.3 * (T - 4.5115), T, 0.50 - 0.9 * (T - 4.5115))
    assert tc == pytest.approx(4.5115, abs=1e-12)


def test_cubics_are_reproduced_exactly():
    # a not-a-knot spline interpolates a cubic exactly, so the only error is the
    # linear root between grid points (spacing 1.1e-4): below 1e-7 here
    T = np.linspace(2.0, 3.0, 9)
    p1 = np.poly1d([0.4, -0.2, -0.5, 0.61])
    p2 = p1 - np.poly1d([1.0, -2.37]) * np.poly1d([0.5, 0.0, 1.0])  # U1 - U2 = (T - 2.37)(T^2 / 2 + 1)
    tc = ising_hip.binder_crossing(T, p1(T), T, p2(T))
    assert tc == pytest.approx(2.37, abs=1e-7)


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_noisy_curves_match_scipy(seed):
    rng = np.random.default_rng(seed)
    T1 = np.sort(rng.uniform(4.40, 4.62, 24))
    T2 = np.linspace(4.41, 4.63, 17)
    U1 = 0.6 - 0.8 * (T1 - 4.51) + 0.01 * rng.normal(size=T1.size)
    U2 = 0.6 - 1.9 * (T2 - 4.51) + 0.01 * rng.normal(size=T2.size)
    want = scipy_crossing(T1, U1, T2, U2)
    perm = rng.permutation(T1.size)  # input order does not matter
    got = ising_hip.binder_crossing(T1[perm], U1[perm], T2, U2)
    assert want is not None and got == pytest.approx(want, rel=1e-12, abs=1e-12)


def test_edge_cases():
    T = np.linspace(1.0, 2.0, 8)
    assert ising_hip.binder_crossing(T, T, T, T + 1.0) is None  # parallel curves
    assert ising_hip.binder_crossing(T, T, T + 5.0, T) is None  # no common range
    with pytest.raises(ising_hip.IsingError):
        ising_hip.binder_crossing(T[:3], T[:3], T, T)
    with pytest.raises(ising_hip.IsingError):
        ising_hip.binder_crossing(np.array([1.0, 1.0, 2.0, 3.0]), np.zeros(4), T, T)  # repeated temperature
    with pytest.raises(ValueError):
        ising_hip.binder_crossing(T, T[:5], T, T)


def test_adjacent_size_pairs():
    T = np.linspace(4.40, 4.62, 32)
    curves = {L: (T, 0.6 - 0.1 * L * (T - 4.5115)) for L in (32, 8, 16)}
    out = ising_hip.binder_crossings(curves)
    assert [(a, b) for a, b, _ in out] == [(8, 16), (16, 32)]
    assert all(tc == pytest.approx(4.5115, abs=1e-12) for _, _, tc in out)
