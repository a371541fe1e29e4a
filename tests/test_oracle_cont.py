"""Pins the oracle's continuous-spin (XY / Heisenberg) restatement.

* overrelaxation conserves energy and is an involution (heisenberg/overrelax.rs:74-120 analogue);
* Metropolis keeps unit norm (heisenberg/metropolis.rs:125-136);
* with the other colour frozen, each site's chain samples the exact Boltzmann
  conditional: <cos theta> = coth(k) - 1/k (Heisenberg), I1(k)/I0(k) (XY),
  k = beta J |h|.  A non-symmetric proposal or a wrong acceptance rule fails this;
* the random start is uniform (mean 0, <z^2> = 1/3, heisenberg/mod.rs:139-161).
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.special import i0, i1

import oracle_lib as o


def init(L, nc, seed=7, replica=0):
    sp = np.zeros(L ** 3 * nc)
    o.lib.o_cont_init(L, nc, seed, o.lib.o_stream(L, replica), sp)
    return sp


def energy(L, nc, sp, J=1.0):
    e = C.c_double()
    m = np.zeros(3)
    o.lib.o_cont_measure(L, nc, sp, J, C.byref(e), m)
    return e.value, m


def field(L, nc, sp, s):
    z, rem = divmod(s, L * L)
    y, x = divmod(rem, L)
    S = sp.reshape(-1, nc)
    nb = [((z * L + y) * L + (x + 1) % L), ((z * L + y) * L + (x - 1) % L), ((z * L + (y + 1) % L) * L + x),
          ((z * L + (y - 1) % L) * L + x), ((((z + 1) % L) * L + y) * L + x), ((((z - 1) % L) * L + y) * L + x)]
    return S[nb].sum(axis=0)


@pytest.mark.parametrize("nc", [2, 3])
def test_init_is_unit_and_uniform(nc):
    sp = init(16, nc).reshape(-1, nc)
    assert np.allclose((sp ** 2).sum(axis=1), 1.0, atol=1e-12)
    n = len(sp)
    assert np.all(np.abs(sp.mean(axis=0)) < 5 / math.sqrt(n))
    if nc == 3:
        assert abs((sp[:, 2] ** 2).mean() - 1 / 3) < 5 * math.sqrt(4 / 45 / n)


@pytest.mark.parametrize("nc", [2, 3])
def test_overrelaxation_conserves_energy_and_is_an_involution(nc):
    L = 6
    sp = init(L, nc, 3)
    e0, _ = energy(L, nc, sp)
    for c in (0, 1, 0, 1):
        o.lib.o_cont_overrelax_half(L, nc, sp, c, 1.0)
        e1, _ = energy(L, nc, sp)
        assert abs(e1 - e0) < 1e-10 * abs(e0)
    before = sp.copy()
    o.lib.o_cont_overrelax_half(L, nc, sp, 1, 1.0)
    o.lib.o_cont_overrelax_half(L, nc, sp, 1, 1.0)
    assert np.allclose(sp, before, atol=1e-12)
    # and it is not the identity
    o.lib.o_cont_overrelax_half(L, nc, sp, 0, 1.0)
    assert not np.allclose(sp, before, atol=1e-6)


@pytest.mark.parametrize("nc", [2, 3])
def test_metropolis_keeps_unit_norm_and_accepts_downhill(nc):
    L = 6
    sp = init(L, nc, 11)
    u = np.zeros(L ** 3 // 2)
    p = np.zeros(L ** 3 // 2)
    for t in range(5):
        for c in (0, 1):
            o.lib.o_cont_metropolis_half(L, nc, sp, c, t, 11, o.lib.o_stream(L, 0), 0.7, 1.0, 0.8, 0,
                                         u.ctypes.data, p.ctypes.data)
    assert np.allclose((sp.reshape(-1, nc) ** 2).sum(axis=1), 1.0, atol=1e-12)
    assert np.all((p > 0) & (p <= 1)) and np.all((u >= 0) & (u < 1))


@pytest.mark.parametrize("nc,beta,delta", [(3, 0.35, 1.2), (3, 0.15, 3.0), (2, 0.35, 1.5), (2, 0.2, 3.1)])
def test_single_site_chain_samples_the_boltzmann_conditional(nc, beta, delta):
    """Colour 1 frozen: every colour-0 site runs an independent chain in a fixed field."""
    L = 4
    sp = init(L, nc, 5)
    sites = [s for s in range(L ** 3) if (s % L + (s // L) % L + s // (L * L)) % 2 == 0]
    hs = [field(L, nc, sp, s) for s in sites]
    n_iter = 6000
    cos = np.zeros((n_iter, len(sites)))
    for t in range(n_iter):
        o.lib.o_cont_metropolis_half(L, nc, sp, 0, t, 5, o.lib.o_stream(L, 0), beta, 1.0, delta, 0, None, None)
        S = sp.reshape(-1, nc)
        for k, s in enumerate(sites):
            cos[t, k] = S[s] @ hs[k] / np.linalg.norm(hs[k])
    cos = cos[200:]
    kap = np.array([beta * np.linalg.norm(h) for h in hs])
    exact = 1.0 / np.tanh(kap) - 1.0 / kap if nc == 3 else i1(kap) / i0(kap)
    got = cos.mean(axis=0)
    # block errors per site (20 blocks)
    nb = 20
    bs = len(cos) // nb
    bm = cos[: nb * bs].reshape(nb, bs, -1).mean(axis=1)
    err = bm.std(axis=0, ddof=1) / math.sqrt(nb)
    z = (got - exact) / err
    assert np.all(np.abs(z) < 5.0), (z, got, exact)
    assert abs(z.mean()) * math.sqrt(len(z)) < 5.0
