"""Pins the oracle's cluster updates (Swendsen-Wang / Wolff, Ising and the
continuous-spin embedding).

Checks, none of which restates the oracle's own formulas:
* the bond threshold against its closed form (p = 1/2 at beta J = ln 2 / 2,
  beta = 0, the beta -> infinity clamp, J -> -J symmetry);
* clusters = connected components of the returned bonds found by a
  breadth-first search here; bonds only on satisfied pairs; the flip rule
  (Swendsen-Wang: the component's smallest site's coin; Wolff: the component
  of the Philox-drawn seed site);
* the frequency of active bonds among satisfied pairs equals floor(2^32 p) / 2^32,
  and among the bonds decided by the 16-bit tie draw equals the threshold's
  low half / 2^16;
* the Boltzmann distribution is stationary: the joint (B, M) histogram of long
  chains on the 2x2x2 cube and the 4x4 square matches exact enumeration, also
  for an antiferromagnet (src/wolff.rs:27-77 samples the same distribution);
* Wolff's improved estimator <|C|> = <M^2> / N;
* continuous spins: unit norm, bond energies inside / outside clusters kept,
  and equilibrium against the reference's XY Wolff (src/xy/wolff.rs:16-73) and
  Heisenberg Metropolis + overrelaxation restatements.
"""
import collections
import math

import numpy as np
import pytest

import oracle_lib as o

SEED = 0x5EED_C1A5


def jk_mean_err(x, nb=20):
    x = np.asarray(x, np.float64)
    bs = len(x) // nb
    x = x[: bs * nb]
    full = x.mean()
    jk = np.array([(x.sum() - x[b * bs:(b + 1) * bs].sum()) / (len(x) - bs) for b in range(nb)])
    return full, math.sqrt((nb - 1) / nb * ((jk - full) ** 2).sum())


def forward(dim, L, i, d):
    x, y, z = i % L, (i // L) % L, i // (L * L)
    if d == 0:
        x = (x + 1) % L
    elif d == 1:
        y = (y + 1) % L
    else:
        z = (z + 1) % L
    return (z * L + y) * L + x


def components(dim, L, bonds):
    n = L ** dim
    adj = [[] for _ in range(n)]
    for i in range(n):
        for d in range(dim):
            if (bonds[i] >> d) & 1:
                j = forward(dim, L, i, d)
                adj[i].append(j)
                adj[j].append(i)
    comp = -np.ones(n, np.int64)
    for s in range(n):
        if comp[s] >= 0:
            continue
        comp[s] = s  # s is the smallest site of its component (visited in order)
        q = collections.deque([s])
        while q:
            i = q.popleft()
            for j in adj[i]:
                if comp[j] < 0:
                    comp[j] = s
                    q.append(j)
    return comp


def wolff_seed_site(u, stream, model, n):
    w = o.philox([0, u, stream, (10 << 8) | model], [SEED & 0xFFFFFFFF, SEED >> 32])
    v = (int(w[0]) << 32) | int(w[1])
    return (v * n) >> 64


# --- bond threshold ----------------------------------------------------------

def test_bond_threshold_closed_form():
    assert o.bond_threshold(0.0) == 0
    assert o.bond_threshold(1e3) == 2 ** 32 - 1
    half = math.log(2.0) / 2.0  # 1 - exp(-2 beta J) = 1/2
    assert abs(o.bond_threshold(half) - 2 ** 31) <= 1
    for beta in (0.1, 0.2216544, 0.35, 0.9):
        t = o.bond_threshold(beta)
        assert t == o.bond_threshold(beta, -1.0)
        p = 1.0 - math.exp(-2.0 * beta)
        assert -1e-15 <= p - t / 2 ** 32 < 2.0 ** -32 + 1e-15  # floor(2^32 p) / 2^32
    ts = [o.bond_threshold(b) for b in np.linspace(0.01, 2.0, 50)]
    assert all(a < b for a, b in zip(ts, ts[1:]))


# --- cluster structure -------------------------------------------------------

@pytest.mark.parametrize("dim,L,J,kind", [(3, 6, 1.0, 0), (3, 6, 1.0, 1), (2, 10, 1.0, 0), (2, 10, -1.0, 1),
                                          (3, 2, 1.0, 0), (3, 6, -1.0, 0)])
def test_clusters_are_components_and_flip_rule(dim, L, J, kind):
    rng = np.random.default_rng(dim * 100 + L)
    n = L ** dim
    stream = (L << 16) | 1
    for u in range(6):
        before = (rng.integers(0, 2, n) * 2 - 1).astype(np.int8)
        sp = before.copy()
        flipped, bonds = o.ising_cluster(dim, L, sp, kind, u, SEED, stream, beta=0.35, J=J)
        comp = components(dim, L, bonds)
        for i in range(n):
            for d in range(dim):
                if (bonds[i] >> d) & 1:
                    j = forward(dim, L, i, d)
                    assert (before[i] == before[j]) == (J > 0)
        changed = sp != before
        assert np.all(sp[changed] == -before[changed])
        assert changed.sum() == flipped
        if kind == 0:
            want = np.array([(bonds[comp[i]] >> 3) & 1 for i in range(n)], bool)
        else:
            seed = wolff_seed_site(u, stream, 0, n)
            want = comp == comp[seed]
        assert np.array_equal(changed, want)


def test_bond_frequency_and_coins():
    L, dim = 16, 3
    beta = 0.2216544
    thr = o.bond_threshold(beta)
    p = thr / 2 ** 32
    rng = np.random.default_rng(7)
    roots = coins = 0
    per_dir = np.zeros((2, 3))
    for u in range(700):
        sp = (rng.integers(0, 2, L ** 3) * 2 - 1).astype(np.int8)
        before = sp.copy().reshape(L, L, L)
        _, bonds = o.ising_cluster(dim, L, sp, 0, u, SEED, (L << 16), beta=beta)
        b = bonds.reshape(L, L, L)
        for d, ax in enumerate((2, 1, 0)):
            s = before == np.roll(before, -1, axis=ax)
            a = ((b >> d) & 1).astype(bool)
            assert not np.any(a & ~s)
            per_dir[0, d] += s.sum()
            per_dir[1, d] += (a & s).sum()
        coins += int(((bonds >> 3) & 1).sum())
        roots += L ** 3
    sat, act = per_dir[0].sum(), per_dir[1].sum()
    sigma = math.sqrt(p * (1 - p) / sat)
    assert abs(act / sat - p) < 4 * sigma, (act / sat, p, sigma)
    for d in range(3):
        sd = math.sqrt(p * (1 - p) / per_dir[0, d])
        assert abs(per_dir[1, d] / per_dir[0, d] - p) < 4.5 * sd
    assert abs(coins / roots - 0.5) < 4 * math.sqrt(0.25 / roots)


def test_tie_draw_decides_with_the_low_threshold_bits():
    # a 16-bit draw equal to the threshold's high half (1 bond in 65536) is
    # decided by 16 more bits against the low half: the active fraction among
    # those bonds is lo / 65536 (0.81 here; 0 or 1 if the tie path were dropped
    # or forced, 0.19 if its comparison were reversed)
    L, dim, beta = 64, 3, 0.211
    thr = o.bond_threshold(beta)
    q = (thr & 0xFFFF) / 65536
    assert 0.75 < q < 0.85
    rng = np.random.default_rng(3)
    ties = active = 0
    u = 0
    while ties < 800:
        sp = (rng.integers(0, 2, L ** 3) * 2 - 1).astype(np.int8)
        _, bonds = o.ising_cluster(dim, L, sp, 0, u, SEED, (L << 16) | 5, beta=beta)
        for d in range(3):
            t = ((bonds >> (5 + d)) & 1).astype(bool)
            ties += int(t.sum())
            active += int((((bonds >> d) & 1).astype(bool) & t).sum())
        u += 1
    f = active / ties
    assert abs(f - q) < 4 * math.sqrt(q * (1 - q) / ties), (f, q, ties)


# --- stationarity against exact enumeration ---------------------------------

def exact_joint(dim, L, beta, J):
    n = L ** dim
    states = ((np.arange(2 ** n)[:, None] >> np.arange(n)[None, :]) & 1).astype(np.int64) * 2 - 1
    s = states.reshape((-1,) + (L,) * dim)
    B = np.zeros(len(s), np.int64)
    for ax in range(1, dim + 1):
        B += (s * np.roll(s, -1, axis=ax)).sum(axis=tuple(range(1, dim + 1)))
    M = s.sum(axis=tuple(range(1, dim + 1)))
    logw = beta * J * B
    w = np.exp(logw - logw.max())
    joint = collections.Counter()
    for b, m, ww in zip(B, M, w):
        joint[(int(b), int(m))] += ww
    z = sum(joint.values())
    return {k: v / z for k, v in joint.items()}


@pytest.mark.parametrize("dim,L,beta,J,kind", [(3, 2, 0.3, 1.0, 0), (3, 2, 0.3, 1.0, 1), (2, 4, 0.44, 1.0, 0),
                                               (2, 4, 0.44, 1.0, 1), (3, 2, 0.25, -1.0, 0)])
def test_boltzmann_distribution_is_stationary(dim, L, beta, J, kind):
    n = 200_000
    b, m, _ = o.ising_cluster_chain(dim, L, SEED + kind, kind, beta, 100, n, J=J)
    exact = exact_joint(dim, L, beta, J)
    seen = collections.Counter(zip(b.tolist(), m.tolist()))
    assert set(seen) <= set(exact)
    chi2, dof = 0.0, 0
    for k, p in exact.items():
        e = p * n
        if e < 20:
            continue
        chi2 += (seen.get(k, 0) - e) ** 2 / e
        dof += 1
    # successive updates are correlated, so allow a generous margin over dof
    assert chi2 < 3.0 * dof + 30, (chi2, dof)
    # first moments against exact values, within 5 blocked errors
    eb = sum(p * k[0] for k, p in exact.items())
    mb, err = jk_mean_err(b.astype(np.float64))
    assert abs(mb - eb) < 5 * err + 1e-9, (mb, eb, err)


def test_wolff_mean_cluster_size_equals_m2_over_n():
    L, beta = 8, 0.2216544
    b, m, f = o.ising_cluster_chain(3, L, SEED, 1, beta, 2000, 40000)
    n = L ** 3
    c, ce = jk_mean_err(f.astype(np.float64))
    m2, me = jk_mean_err(m.astype(np.float64) ** 2 / n)
    # the magnetisation after an update is that of an equilibrium state too
    assert abs(c - m2) < 4 * math.hypot(ce, me), (c, ce, m2, me)


def test_sw_equilibrium_matches_reference_wolff():
    L, beta = 8, 1.0 / 4.5115
    b, m, _ = o.ising_cluster_chain(3, L, SEED, 0, beta, 1000, 30000)
    rb, rm = o.ref_ising_chain("wolff", L, SEED + 1, beta, 2000, 60000)
    for x, y in ((b, rb), (np.abs(m), np.abs(rm))):
        a, ae = jk_mean_err(x.astype(np.float64))
        r, re_ = jk_mean_err(y.astype(np.float64))
        assert abs(a - r) < 4.5 * math.hypot(ae, re_), (a, ae, r, re_)


# --- continuous spins --------------------------------------------------------

def random_unit(rng, n, nc):
    v = rng.normal(size=(n, nc))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).reshape(-1)


@pytest.mark.parametrize("nc,model,kind", [(2, o.MODEL_XY, 0), (2, o.MODEL_XY, 1), (3, o.MODEL_HEIS, 0),
                                           (3, o.MODEL_HEIS, 1)])
def test_cont_cluster_structure(nc, model, kind):
    L = 5 if nc == 2 else 4
    n = L ** 3
    rng = np.random.default_rng(nc + kind)
    stream = (L << 16) | 2
    for u in range(4):
        before = random_unit(rng, n, nc)
        sp = before.copy()
        flipped, bonds, _ = o.cont_cluster(L, nc, sp, kind, u, SEED, stream, model, beta=0.7)
        S0, S1 = before.reshape(n, nc), sp.reshape(n, nc)
        assert np.allclose(np.linalg.norm(S1, axis=1), 1.0, atol=1e-12)
        changed = np.any(np.abs(S1 - S0) > 1e-12, axis=1)
        comp = components(3, L, bonds)
        if kind == 0:
            want = np.array([(bonds[comp[i]] >> 3) & 1 for i in range(n)], bool)
        else:
            want = comp == comp[wolff_seed_site(u, stream, model, n)]
        # a site on the reflection plane (a_i = 0) does not move even when its cluster flips
        assert np.all(~changed | want)
        assert flipped == want.sum()
        # a reflection is an isometry: bond energies within the flipped set and within
        # the rest are unchanged
        for i in range(n):
            for d in range(3):
                j = forward(3, L, i, d)
                if want[i] == want[j]:
                    assert abs(S0[i] @ S0[j] - S1[i] @ S1[j]) < 1e-12
                if (bonds[i] >> d) & 1:
                    assert want[i] == want[j]


def cont_cluster_chain(L, nc, model, beta, kind, warmup, n, seed):
    rng = np.random.default_rng(seed)
    sp = random_unit(rng, L ** 3, nc)
    stream = (L << 16)
    e = np.zeros(n)
    for k in range(warmup + n):
        o.cont_cluster(L, nc, sp, kind, k, seed, stream, model, beta)
        if k >= warmup:
            ee = o.C.c_double()
            m = np.zeros(3)
            o.lib.o_cont_measure(L, nc, sp, 1.0, o.C.byref(ee), m)
            e[k - warmup] = ee.value
    return e



def test_xy_sw_equilibrium_matches_reference_xy_wolff():
    L, beta = 4, 0.45
    e = cont_cluster_chain(L, 2, o.MODEL_XY, beta, 0, 500, 20000, 11)
    re_, _ = o.ref_xy_chain(L, 12, beta, 1000, 20000)
    a, ae = jk_mean_err(e)
    r, rerr = jk_mean_err(re_)
    assert abs(a - r) < 4.5 * math.hypot(ae, rerr), (a, ae, r, rerr)


def test_heisenberg_wolff_equilibrium_matches_reference_metropolis():
    L, beta = 4, 0.7
    e = cont_cluster_chain(L, 3, o.MODEL_HEIS, beta, 1, 500, 30000, 13)
    re_, _ = o.ref_heis_chain(L, 14, beta, 1.0, 2, 1000, 20000)
    a, ae = jk_mean_err(e)
    r, rerr = jk_mean_err(re_)
    assert abs(a - r) < 4.5 * math.hypot(ae, rerr), (a, ae, r, rerr)
