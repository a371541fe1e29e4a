"""GPU parity of the Ising path against the oracle (bit-exact), through the C ABI.

Small sizes are replayed sweep for sweep by the oracle; the full size (L = 8 .. 128 x 32 replicas)
(L=128 x 32 replicas) is checked through size-independent properties: fused
records equal a standalone reduction of the downloaded lattice, spins stay
+-1, the temperature permutation stays a bijection.
"""
import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu

ising_hip = pytest.importorskip("ising_hip")


def _lat(L, seed):
    return ising_hip.LatticeGpu(L, seed)


@pytest.mark.parametrize("L", [2, 4, 6, 8, 12, 16, 32, 48, 64])
def test_init_matches_oracle(L):
    lat = _lat(L, 42)
    assert np.array_equal(lat.get_spins(), o.ising_init(L, 42, 0))


@pytest.mark.parametrize("L,beta,h,sweeps", [(4, 0.2216, 0.0, 20), (6, 0.30, 0.0, 12), (8, 0.2216, 0.0, 10),
                                             (12, 0.5, 0.3, 8), (16, 0.2216, 0.0, 8), (32, 0.2216, 0.0, 4),
                                             (48, 0.18, -0.2, 3), (64, 0.2216, 0.0, 6)])
def test_sweeps_are_bit_exact(L, beta, h, sweeps):
    seed = 1000 + L
    lat = _lat(L, seed)
    ref = o.ising_init(L, seed, 0)
    for t in range(sweeps):
        lat.step(beta, 1.0, h)
        o.ising_sweep(L, ref, t, seed, float(np.float32(beta)), 1.0, float(np.float32(h)))
    got = lat.get_spins()
    assert np.array_equal(got, ref), (L, int((got != ref).sum()))
    assert lat.measure_raw() == o.ising_measure(L, ref)


def test_measure_helpers_match_oracle():
    L = 16
    lat = _lat(L, 5)
    for t in range(5):
        lat.step(0.25, 1.0, 0.0)
    sp = lat.get_spins()
    B, M = o.ising_measure(L, sp)
    e, m = lat.measure_gpu(1.0)
    assert e == -B / L ** 3 and m == abs(M) / L ** 3
    assert lat.energy_gpu(1.0) == -float(B)
    walls = 0
    s3 = sp.reshape(L, L, L)
    for ax in range(3):
        walls += int((s3 != np.roll(s3, -1, axis=ax)).sum())
    assert abs(lat.domain_wall_density() - walls / (3 * L ** 3)) < 1e-15
    lat.set_spins(np.ones(L ** 3, np.int8))
    assert lat.measure_raw() == (3 * L ** 3, L ** 3)


def test_ground_state_preserved_and_measure_observables():
    L = 8
    lat = _lat(L, 1)
    lat.set_spins(np.ones(L ** 3, np.int8))
    obs = lat.measure(0.01, 1.0, 0.0, 50)  # beta = 100
    assert obs.energy == -3.0 and obs.magnetisation == 1.0 and obs.m4 == 1.0


@pytest.mark.parametrize("sizes,R,warmup,samples,ee,me", [([4, 8], 4, 10, 60, 10, 5), ([8, 16, 6], 8, 20, 100, 10, 1),
                                                          ([32], 6, 5, 40, 3, 2), ([24, 40], 3, 4, 12, 3, 2),
                                                          # more replicas per size than a block has threads:
                                                          # the exchange closing reads the permutation per block
                                                          ([4, 6], 260, 2, 12, 1, 2)])
def test_pt_ensemble_records_are_bit_exact(sizes, R, warmup, samples, ee, me):
    seed = 42
    temps = o.linspace(4.40, 4.62, R)
    pt = ising_hip.PTEnsemble("ising", sizes, R, temps, seed, exchange_every=ee, measure_every=me)
    pt.warmup(warmup)
    pt.sample(samples)
    tidx, bond, mag = pt.read_ising()
    assert tidx.shape == (samples // me, len(sizes), R)
    for gi, L in enumerate(sizes):
        rt, rb, rm, final = o.ising_pt_run(L, R, temps, seed, warmup, samples, ee, me)
        assert np.array_equal(tidx[:, gi, :].reshape(-1), rt), L
        assert np.array_equal(bond[:, gi, :].reshape(-1), rb), L
        assert np.array_equal(mag[:, gi, :].reshape(-1), rm), L
        for r in range(R):
            assert np.array_equal(pt.get_spins(gi, r), final[r]), (L, r)
        r2t, t2r = pt.get_perm(gi)
        assert np.array_equal(t2r[r2t], np.arange(R))
    tries, acc = pt.exchange_stats()
    assert tries > 0 and 0 <= acc <= tries


def test_batched_equals_unbatched():
    temps = o.linspace(4.4, 4.62, 4)
    both = ising_hip.PTEnsemble("ising", [8, 16], 4, temps, 9, exchange_every=5, measure_every=5)
    one = ising_hip.PTEnsemble("ising", [16], 4, temps, 9, exchange_every=5, measure_every=5)
    for p in (both, one):
        p.warmup(7)
        p.sample(30)
    a = both.read_ising()
    b = one.read_ising()
    for x, y in zip(a, b):
        assert np.array_equal(x[:, 1, :], y[:, 0, :])


def test_record_buffer_drain_and_overflow_guard():
    temps = o.linspace(4.4, 4.6, 2)
    pt = ising_hip.PTEnsemble("ising", [4], 2, temps, 3, measure_every=1, record_capacity=8)
    pt.sample(6)
    with pytest.raises(ising_hip.IsingError):
        pt.sample(5)  # 11 > 8
    t1, b1, m1 = pt.read_ising(4)
    t2, b2, m2 = pt.read_ising()
    assert t1.shape[0] == 4 and t2.shape[0] == 2
    pt.sample(8)
    assert pt.pending() == 8
    rt, rb, rm, _ = o.ising_pt_run(4, 2, temps, 3, 0, 14, 10, 1)
    got_b = np.concatenate([b1.reshape(-1), b2.reshape(-1), pt.read_ising()[1].reshape(-1)])
    assert np.array_equal(got_b, rb)


def test_large_lattice_sweeps_are_bit_exact():
    """L=128 (16-site chunks in one row) and L=80 (two 8-site runs per chunk):
    whole-lattice bytes after a few sweeps equal the oracle's, ties included
    (about 1 site in 128 takes the tie path every half-sweep)."""
    for L, sweeps in ((128, 3), (80, 4)):
        seed = 7 + L
        lat = _lat(L, seed)
        ref = o.ising_init(L, seed, 0)
        for t in range(sweeps):
            lat.step(0.2216, 1.0, 0.0)
            o.ising_sweep(L, ref, t, seed, float(np.float32(0.2216)))
        got = lat.get_spins()
        assert np.array_equal(got, ref), (L, int((got != ref).sum()))
        assert lat.measure_raw() == o.ising_measure(L, ref)


def test_thread_step_layouts_do_not_change_results():
    temps = o.linspace(4.40, 4.62, 4)
    out = []
    for width, runs, spt in ((0, 0, 0), (16, 1, 1), (16, 2, 1), (32, 1, 1), (32, 1, 2), (16, 2, 4), (16, 1, 3)):
        pt = ising_hip.PTEnsemble("ising", [8, 16, 24, 32, 64, 96, 128], 4, temps, 5, exchange_every=3,
                                  measure_every=2, ising_spt=spt, ising_width=width, ising_runs=runs)
        pt.warmup(5)
        pt.sample(12)
        out.append(pt.read_ising() + tuple(pt.get_spins(g, r) for g in (0, 2, 6) for r in range(4)))
    for other in out[1:]:
        for x, y in zip(out[0], other):
            assert np.array_equal(x, y)


def test_autotune_keeps_the_chain():
    """The autotuner's timed sweeps are ordinary warm-up sweeps: the tuned
    ensemble continues exactly like an untuned one warmed up as long."""
    temps = o.linspace(4.40, 4.62, 4)
    tuned = ising_hip.PTEnsemble("ising", [32, 64], 4, temps, 9, exchange_every=3, measure_every=2)
    width, runs, spt = tuned.autotune(3)
    assert (width, runs) in ((16, 1), (16, 2), (32, 1)) and spt in (1, 2, 4)
    assert tuned.get_layout(1) == (width, runs, spt)  # the largest size's
    n = tuned.sweeps()
    assert n >= 9 * (3 + 1)  # nine layouts for both sizes, then size by size; one untimed sweep each
    plain = ising_hip.PTEnsemble("ising", [32, 64], 4, temps, 9, exchange_every=3, measure_every=2)
    plain.warmup(n)
    for p in (tuned, plain):
        p.sample(10)
    for x, y in zip(tuned.read_ising(), plain.read_ising()):
        assert np.array_equal(x, y)
    with pytest.raises(ising_hip.IsingError):
        tuned.autotune(1)  # sampling has started


def test_autotune_per_size_and_cache(tmp_path, monkeypatch):
    """Per-size layouts keep the chain; the tuning cache applies a stored choice
    without timing anything (no sweeps run) for the same model, replicas,
    sizes and device."""
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("ISING_HIP_TUNE_CACHE", str(cache))
    sizes, R = [8, 12, 16, 32, 64], 4
    temps = o.linspace(4.40, 4.62, R)
    kw = dict(exchange_every=3, measure_every=2)
    tuned = ising_hip.PTEnsemble("ising", sizes, R, temps, 5, **kw)
    tuned.autotune(2)
    layouts = [tuned.get_layout(g) for g in range(len(sizes))]
    for w, r, s_ in layouts:
        assert (w, r) in ((16, 1), (16, 2), (32, 1)) and s_ in (1, 2, 4)
    lines = cache.read_text().splitlines()
    assert len(lines) == 1 and "sizes=8,12,16,32,64" in lines[0] and "R=4" in lines[0]
    n = tuned.sweeps()
    again = ising_hip.PTEnsemble("ising", sizes, R, temps, 5, **kw)
    assert again.autotune(2) == layouts[-1]
    assert again.sweeps() == 0  # nothing timed
    assert [again.get_layout(g) for g in range(len(sizes))] == layouts
    plain = ising_hip.PTEnsemble("ising", sizes, R, temps, 5, **kw)
    plain.warmup(n)
    again.warmup(n)
    for p in (tuned, plain, again):
        p.sample(12)
    ref = plain.read_ising()
    for p in (tuned, again):
        for x, y in zip(p.read_ising(), ref):
            assert np.array_equal(x, y)
    other = ising_hip.PTEnsemble("ising", [8, 16], R, temps, 5, **kw)
    other.autotune(1)
    assert other.sweeps() > 0 and len(cache.read_text().splitlines()) == 2  # another key


@pytest.mark.parametrize("model", ["xy", "heisenberg"])
def test_continuous_autotune_keeps_the_chain(model):
    """The overrelaxation pass's sites per thread do not change any result."""
    temps = np.linspace(2.15, 2.25, 4) if model == "xy" else np.linspace(1.40, 1.50, 4)
    kw = dict(exchange_every=3, measure_every=2, n_overrelax=2, delta=1.0)
    tuned = ising_hip.PTEnsemble(model, [8, 16], 4, temps, 6, **kw)
    spt, _, _ = tuned.autotune(3)
    assert spt in (1, 2) and tuned.get_layout(0)[0] == spt
    plain = ising_hip.PTEnsemble(model, [8, 16], 4, temps, 6, **kw)
    plain.warmup(tuned.sweeps())
    for p in (tuned, plain):
        p.sample(10)
    for x, y in zip(tuned.read_cont(), plain.read_cont()):
        assert np.array_equal(x, y)
    for g in range(2):
        assert np.array_equal(tuned.get_spins(g, 1), plain.get_spins(g, 1))


def test_full_size_config_properties():
    """L=128 x 32 replicas (the largest size of the benchmark ensemble): fused records vs standalone reduction."""
    R, L = 32, 128
    temps = o.linspace(4.40, 4.62, R)
    pt = ising_hip.PTEnsemble("ising", [L], R, temps, 42, exchange_every=2, measure_every=1)
    pt.warmup(3)
    pt.sample(4)
    tidx, bond, mag = pt.read_ising()
    assert tidx.shape == (4, 1, R)
    for k in range(4):
        assert sorted(tidx[k, 0]) == list(range(R))
    for r in (0, 17, 31):
        sp = pt.get_spins(0, r)
        assert set(np.unique(sp)) <= {-1, 1}
        s3 = sp.reshape(L, L, L).astype(np.int64)
        B = int((s3 * (np.roll(s3, -1, 0) + np.roll(s3, -1, 1) + np.roll(s3, -1, 2))).sum())
        assert (B, int(s3.sum())) == (int(bond[-1, 0, r]), int(mag[-1, 0, r]))
    r2t, t2r = pt.get_perm(0)
    assert np.array_equal(t2r[r2t], np.arange(R))


def test_graph_replays_match_direct_launches():
    """Captured blocks of sweeps, replayed several times with eager sweeps and
    host synchronisations in between, give the same records as direct launches."""
    temps = o.linspace(4.4, 4.62, 6)
    kw = dict(exchange_every=3, measure_every=2)
    g = ising_hip.PTEnsemble("ising", [8, 16], 6, temps, 21, graph_sweeps=4, **kw)
    d = ising_hip.PTEnsemble("ising", [8, 16], 6, temps, 21, graph_sweeps=0, **kw)
    out_g, out_d = [], []
    for p, out in ((g, out_g), (d, out_d)):
        p.warmup(9)   # 2 replays + 1 direct
        for n in (8, 3, 12, 1, 4):
            p.sample(n)
            p.sync()
            out.append(p.read_ising())
    for a, b in zip(out_g, out_d):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    rt, rb, rm, _ = o.ising_pt_run(16, 6, temps, 21, 9, 28, 3, 2)
    assert np.array_equal(np.concatenate([x[1][:, 1, :].reshape(-1) for x in out_g]), rb)
