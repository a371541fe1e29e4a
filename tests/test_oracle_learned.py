"""Pins the FP64 restatement of the learned proposal (oracle/learned.py):
component densities normalise, their mean cosines match numerical integrals,
the Metropolis-Hastings ratio satisfies detailed balance, and the finite-
difference gradient agrees with directional derivatives (on samples whose
hidden pre-activations all stay clear of the ReLU kinks, where the loss is
smooth)."""
import math
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import learned as ref  # noqa: E402

KAPPAS = np.array([0.01, 0.5, 3.0, 30.0])


def _xy_density(kap, theta):
    c = np.cos(theta)
    return np.exp(ref.log_component(2, np.array([[kap, kap]] * len(c)), c)[:, 0])


def _heis_density(kap, c):
    return np.exp(ref.log_component(3, np.array([[kap, kap]] * len(c)), c)[:, 0])


def test_component_densities_normalise_and_mean_cosines():
    th = np.linspace(0.0, 2 * math.pi, 200001)
    c = np.linspace(-1.0, 1.0, 200001)
    for kap in KAPPAS:
        fx = _xy_density(kap, th)           # density in theta on [0, 2 pi)
        assert abs(np.trapezoid(fx, th) - 1.0) < 1e-7, kap
        mx = np.trapezoid(fx * np.cos(th), th)
        assert abs(mx - ref.mean_cos(2, np.array([kap]))[0]) < 1e-7, kap
        fh = _heis_density(kap, c)          # density in solid angle: dOmega = 2 pi dc
        assert abs(2 * math.pi * np.trapezoid(fh, c) - 1.0) < 1e-6, kap
        mh = 2 * math.pi * np.trapezoid(fh * c, c)
        assert abs(mh - ref.mean_cos(3, np.array([kap]))[0]) < 1e-6, kap


def test_mh_ratio_satisfies_detailed_balance():
    rng = np.random.default_rng(2)
    n = 4000
    for ncomp in (2, 3):
        o = rng.normal(size=(n, 3)) * 2.0
        c0 = rng.uniform(-1, 1, n)
        c1 = rng.uniform(-1, 1, n)
        beta, J, hn = 0.7, 1.0, rng.uniform(0.1, 6.0, n)
        log_pi = lambda cc: beta * J * hn * cc  # target, same base measure as q
        lr01 = ref.mh_log_ratio(ncomp, beta, J, hn, o, c0, c1)
        lr10 = ref.mh_log_ratio(ncomp, beta, J, hn, o, c1, c0)
        # pi(c0) q(c1) min(1, r01) == pi(c1) q(c0) min(1, r10)
        lhs = log_pi(c0) + ref.log_q(ncomp, o, c1) + np.minimum(0.0, lr01)
        rhs = log_pi(c1) + ref.log_q(ncomp, o, c0) + np.minimum(0.0, lr10)
        assert np.max(np.abs(lhs - rhs)) < 1e-10
        assert np.max(np.abs(lr01 + lr10)) < 1e-10


def test_finite_difference_gradient_matches_directional_derivatives():
    rng = np.random.default_rng(4)
    H = 8
    for ncomp in (2, 3):
        p = ref.init_params(H, 3)
        nb = rng.normal(size=(64, 6, ncomp))
        nb /= np.linalg.norm(nb, axis=2, keepdims=True)
        x, hhat, _ = ref.features(0.5, nb)
        s = hhat + 0.5 * rng.normal(size=hhat.shape)
        s /= np.linalg.norm(s, axis=1, keepdims=True)
        c = (s * hhat).sum(axis=1)
        keep = ref.kink_margin(p, x) > 1e-3
        assert keep.sum() > 48
        x, c = x[keep], c[keep]
        g = ref.pack_params(ref.grad_fd(p, ncomp, x, c), H)
        flat = ref.pack_params(p, H)
        for _ in range(3):
            d = rng.normal(size=flat.shape)
            eps = 1e-5
            lp = ref.loss(ref.unpack_params(flat + eps * d, H), ncomp, x, c)
            lm = ref.loss(ref.unpack_params(flat - eps * d, H), ncomp, x, c)
            dd = (lp - lm) / (2 * eps)
            assert abs(dd - g @ d) < 1e-6 * max(1.0, abs(dd)), (dd, g @ d)
        # pack/unpack round trip
        assert np.array_equal(ref.pack_params(ref.unpack_params(flat, H), H), flat)
