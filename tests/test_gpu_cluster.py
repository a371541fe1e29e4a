"""Cluster updates on the device (Swendsen-Wang and Wolff): Ising lattices,
flip counts and measurements bit-exact against the oracle's replay with the
same Philox draws; continuous spins against the FP64 replay (bonds whose
uniform lies within 1e-5 of their probability are excluded by construction);
equilibrium and Wolff's improved estimator near T_c."""
import math

import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu
ising_hip = pytest.importorskip("ising_hip")

KINDS = {"sw": 0, "wolff": 1}


def stream(L, r):
    return int(o.lib.o_stream(L, r))


def jk_mean_err(x, nb=20):
    x = np.asarray(x, np.float64)
    bs = len(x) // nb
    x = x[: bs * nb]
    full = x.mean()
    jk = np.array([(x.sum() - x[b * bs:(b + 1) * bs].sum()) / (len(x) - bs) for b in range(nb)])
    return full, math.sqrt((nb - 1) / nb * ((jk - full) ** 2).sum())


def measure(dim, L, sp):
    return o.ising_measure(L, sp) if dim == 3 else o.ising2d_measure(L, sp)


@pytest.mark.parametrize("model,dim,sizes", [("ising", 3, [2, 6, 8, 12, 34, 40]), ("ising2d", 2, [4, 10, 66, 130])])
@pytest.mark.parametrize("kind", ["sw", "wolff"])
def test_ensemble_cluster_updates_are_bit_exact(model, dim, sizes, kind):
    R, seed, n = 3, 21, 3
    temps = [3.5, 4.5, 6.0] if dim == 3 else [1.8, 2.27, 3.0]
    pt = ising_hip.PTEnsemble(model, sizes, R, temps, seed)
    flipped, bond, mag = pt.cluster(kind, n, measure=True)
    assert pt.cluster_count() == n
    for gi, L in enumerate(sizes):
        for r in range(R):
            sp = o.ising_init(L, seed, r)[: L ** dim].copy()  # the square lattice's start is the cube's first L^2
            for k in range(n):
                f, _ = o.ising_cluster(dim, L, sp, KINDS[kind], k, seed, stream(L, r), beta=1.0 / temps[r])
                gr = gi * R + r
                assert flipped[k, gr] == f, (L, r, k)
                assert (bond[k, gr], mag[k, gr]) == measure(dim, L, sp), (L, r, k)
            assert np.array_equal(pt.get_spins(gi, r), sp), (L, r)


def test_large_lattices_bit_exact():
    sizes, R, seed = [64, 128], 2, 5
    temps = [4.5115, 4.6]
    pt = ising_hip.PTEnsemble("ising", sizes, R, temps, seed)
    f_sw = pt.cluster("sw", 1)
    f_w = pt.cluster("wolff", 1)
    for gi, L in enumerate(sizes):
        for r in range(R):
            sp = o.ising_init(L, seed, r)
            a, _ = o.ising_cluster(3, L, sp, 0, 0, seed, stream(L, r), beta=1.0 / temps[r])
            b, _ = o.ising_cluster(3, L, sp, 1, 1, seed, stream(L, r), beta=1.0 / temps[r])
            assert (f_sw[0, gi * R + r], f_w[0, gi * R + r]) == (a, b)
            assert np.array_equal(pt.get_spins(gi, r), sp), (L, r)


def test_antiferromagnet_bit_exact():
    sizes, R, seed = [6, 8], 2, 9
    temps = [2.0, 3.0]
    pt = ising_hip.PTEnsemble("ising", sizes, R, temps, seed, j=-1.0)
    flipped = pt.cluster("sw", 2)
    for gi, L in enumerate(sizes):
        for r in range(R):
            sp = o.ising_init(L, seed, r)
            for k in range(2):
                f, _ = o.ising_cluster(3, L, sp, 0, k, seed, stream(L, r), beta=1.0 / temps[r], J=-1.0)
                assert flipped[k, gi * R + r] == f
            assert np.array_equal(pt.get_spins(gi, r), sp)


def test_cluster_updates_interleave_with_sweeps_and_follow_the_permutation():
    L, R, seed = 8, 4, 33
    temps = list(np.linspace(4.0, 5.0, R))
    pt = ising_hip.PTEnsemble("ising", [L], R, temps, seed, exchange_every=1, measure_every=1)
    pt.warmup(3)
    pt.sample(12)
    pt.read_ising()
    r2t, _ = pt.get_perm(0)
    assert sorted(r2t.tolist()) == list(range(R))
    states = [pt.get_spins(0, r) for r in range(R)]
    flipped = pt.cluster("sw", 2)
    for r in range(R):
        sp = states[r].copy()
        for k in range(2):
            f, _ = o.ising_cluster(3, L, sp, 0, k, seed, stream(L, r), beta=1.0 / temps[r2t[r]])
            assert flipped[k, r] == f
        assert np.array_equal(pt.get_spins(0, r), sp), r
    # sweeps continue from the clustered states with their own counters
    pt.sample(1)
    sweeps = pt.sweeps()
    for r in range(R):
        sp = states[r].copy()
        for k in range(2):
            o.ising_cluster(3, L, sp, 0, k, seed, stream(L, r), beta=1.0 / temps[r2t[r]])
        o.ising_sweep(L, sp, sweeps - 1, seed, 1.0 / temps[r2t[r]], replica=r)
        assert np.array_equal(pt.get_spins(0, r), sp), r


def test_single_lattice_steps_match_the_replay():
    L, seed = 10, 77
    beta = float(np.float32(1.0 / 4.5115))
    lat = ising_hip.LatticeGpu(L, seed)
    w = ising_hip.WolffGpuLattice(L, seed + 1)
    sp = o.ising_init(L, seed, 0)
    spw = o.ising_init(L, seed + 1, 0)
    for k in range(4):
        kind = k % 2
        f = lat.cluster_step(kind, beta)
        fo, _ = o.ising_cluster(3, L, sp, kind, k, seed, stream(L, 0), beta=beta)
        assert f == fo
        assert w.step(beta) == o.ising_cluster(3, L, spw, 1, k, seed + 1, stream(L, 0), beta=beta)[0]
    assert np.array_equal(lat.get_spins(), sp)
    assert np.array_equal(w.get_spins(), spw)
    e, m = w.measure_gpu(1.0)
    b, mm = o.ising_measure(L, spw)
    assert e == pytest.approx(-b / L ** 3, abs=1e-15) and m == pytest.approx(abs(mm) / L ** 3, abs=1e-15)


def cont_replay_case(L, nc, kind, beta, gpu_state, u, seed, r, model):
    sp = gpu_state.astype(np.float64)
    f, _, marg = o.cont_cluster(L, nc, sp, kind, u, seed, stream(L, r), model, beta)
    return f, sp, marg


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("kind", ["sw", "wolff"])
def test_continuous_single_lattice_matches_fp64_replay(nc, kind):
    L, seed = 6, 101
    model = o.MODEL_XY if nc == 2 else o.MODEL_HEIS
    beta = float(np.float32(0.45 if nc == 2 else 0.7))
    lat = ising_hip.ContinuousGpuLattice(L, nc, seed)
    rng = np.random.default_rng(nc * 10 + KINDS[kind])
    compared = 0
    for u in range(6):
        v = rng.normal(size=(L ** 3, nc))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        lat.set_spins(v.astype(np.float32).reshape(-1))
        start = lat.get_spins()
        f = lat.cluster_step(kind, beta)
        fo, want, marg = cont_replay_case(L, nc, KINDS[kind], beta, start, u, seed, 0, model)
        if marg:
            continue
        compared += 1
        assert f == fo, u
        got = lat.get_spins().astype(np.float64)
        assert np.max(np.abs(got - want)) < 1e-5, u
    assert compared >= 4


@pytest.mark.parametrize("model,nc", [("xy", 2), ("heisenberg", 3)])
def test_continuous_ensemble_matches_fp64_replay(model, nc):
    sizes, R, seed = [4, 8], 2, 55
    temps = [1.0, 2.5] if nc == 2 else [0.8, 1.5]
    om = o.MODEL_XY if nc == 2 else o.MODEL_HEIS
    pt = ising_hip.PTEnsemble(model, sizes, R, temps, seed)
    start = {(gi, r): pt.get_spins(gi, r) for gi in range(len(sizes)) for r in range(R)}
    flipped, meas = pt.cluster("sw", 1, measure=True)
    for gi, L in enumerate(sizes):
        for r in range(R):
            f, want, marg = cont_replay_case(L, nc, 0, float(1.0 / temps[r]), start[(gi, r)], 0, seed, r, om)
            if marg:
                continue
            gr = gi * R + r
            assert flipped[0, gr] == f
            got = pt.get_spins(gi, r).astype(np.float64)
            assert np.max(np.abs(got - want)) < 1e-5
            e = o.C.c_double()
            m = np.zeros(3)
            o.lib.o_cont_measure(L, nc, got, 1.0, o.C.byref(e), m)
            assert meas[0, gr, 0] == pytest.approx(e.value, rel=1e-9, abs=1e-9)
            assert np.allclose(meas[0, gr, 1:1 + nc], m[:nc], rtol=1e-9, atol=1e-9)


def test_sw_equilibrium_and_autocorrelation_near_tc():
    L, R, seed = 16, 4, 2024
    T = 4.5115
    pt = ising_hip.PTEnsemble("ising", [L], R, [T] * R, seed)
    pt.cluster("sw", 200)
    _, bond, mag = pt.cluster("sw", 3000, measure=True)
    n = L ** 3
    e = (-bond / n).reshape(-1)
    am = (np.abs(mag) / n)
    rb, rm = o.ref_ising_chain("wolff", L, seed, 1.0 / T, 2000, 40000)
    a, ae = jk_mean_err(e)
    r, re_ = jk_mean_err(-rb / n)
    assert abs(a - r) < 4.5 * math.hypot(ae, re_), (a, ae, r, re_)
    a, ae = jk_mean_err(am.reshape(-1))
    r, re_ = jk_mean_err(np.abs(rm) / n)
    assert abs(a - r) < 4.5 * math.hypot(ae, re_), (a, ae, r, re_)
    # Swendsen-Wang decorrelates |M| within a few updates at T_c
    tau, _ = o.tau_int(am[:, 0])
    assert tau < 8.0, tau


def test_wolff_improved_estimator_on_the_device():
    L, R, seed = 16, 4, 99
    pt = ising_hip.PTEnsemble("ising", [L], R, [4.5115] * R, seed)
    pt.cluster("wolff", 500)
    flipped, _, mag = pt.cluster("wolff", 6000, measure=True)
    n = L ** 3
    c, ce = jk_mean_err(flipped.mean(axis=1))
    m2, me = jk_mean_err((mag.astype(np.float64) ** 2 / n).mean(axis=1))
    assert abs(c - m2) < 4.5 * math.hypot(ce, me), (c, ce, m2, me)


def test_invalid_cluster_calls_are_rejected():
    pt = ising_hip.PTEnsemble("ising_msc", [8], 32, [4.5] * 32, 1)
    with pytest.raises(ising_hip.IsingError):
        pt.cluster("sw", 1)
    pt = ising_hip.PTEnsemble("ising", [8], 2, [4.5] * 2, 1, h=0.1)
    with pytest.raises(ising_hip.IsingError):
        pt.cluster("sw", 1)
    pt = ising_hip.PTEnsemble("ising", [8], 2, [4.5] * 2, 1)
    with pytest.raises(KeyError):
        pt.cluster("bogus", 1)
    with pytest.raises(ising_hip.IsingError):
        ising_hip._check(ising_hip.lib.ising_pt_cluster(pt._h, 7, 1, None, None, None))
