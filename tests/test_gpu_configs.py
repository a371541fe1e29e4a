"""The batched and continuous workloads at their full sizes on the GPU.

Batched uneven PT: Ising + XY + Heisenberg, L = 8, 12, 16, 24, 32, 48, 64, 96,
128 x 32 replicas each (1.85 GB of spins), advanced by one batch graph, either
of each member's own launches or of work-list launches over the members'
blocks.  At
this size the oracle cannot replay the continuous members, so they are checked
through size-independent properties: every record equals a standalone FP64
reduction of the downloaded lattice, spins stay unit vectors, permutations stay
bijections and every exchange, replayed by the oracle from the recorded
energies, gives the next record's temperatures.  The Ising members up to L=48
are replayed sweep for sweep by the oracle (bit-exact records and lattices).

Heisenberg PT at L=64 x 32 (the largest size of its ladder) with step
adaptation during warm-up and renormalisation: the same properties.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu
ising_hip = pytest.importorskip("ising_hip")

SIZES = [8, 12, 16, 24, 32, 48, 64, 96, 128]
R = 32


def o_cont_measure(L, nc, sp):
    e = C.c_double()
    m = np.zeros(3)
    o.lib.o_cont_measure(L, nc, np.ascontiguousarray(sp, np.float64), 1.0, C.byref(e), m)
    return e.value, m


def check_exchanges(tidx, e, temps, seed, L, model, q):
    """Replays the exchanges from the recorded energies: with exchange_every =
    q * measure_every, record k was taken at an exchange sweep when
    (k + 1) % q == 0, of round (k + 1) // q - 1, and the next record's
    temperatures are the permutation after it."""
    betas = 1.0 / np.asarray(temps, np.float64)
    r2t = np.ascontiguousarray(tidx[0], np.int32)
    t2r = np.zeros(len(r2t), np.int32)
    t2r[r2t] = np.arange(len(r2t), dtype=np.int32)
    acc = 0
    for k in range(tidx.shape[0] - 1):
        if (k + 1) % q == 0:
            acc += o.lib.o_exchange(len(r2t), betas, np.ascontiguousarray(e[k], np.float64), r2t, t2r,
                                    (k + 1) // q - 1, seed, o.lib.o_stream(L, 0), model)
        assert np.array_equal(r2t, tidx[k + 1]), (L, k)
    return acc


def check_cont_member(pt, nc, temps, seed, model, sizes, replicas, exchange_every, measure_every):
    tidx, e, mx, my, mz = pt.read_cont()
    n = tidx.shape[0]
    for g, L in enumerate(sizes):
        for k in range(n):
            assert sorted(tidx[k, g]) == list(range(R))
        check_exchanges(tidx[:, g], e[:, g], temps, seed, L, model, exchange_every // measure_every)
        for r in replicas(L):
            sp = pt.get_spins(g, r).astype(np.float64)
            ee, mm = o_cont_measure(L, nc, sp)
            assert abs(e[-1, g, r] - ee) <= 1e-10 * abs(ee), (L, r)
            assert abs(mx[-1, g, r] - mm[0]) <= 1e-12 * L ** 3 and abs(my[-1, g, r] - mm[1]) <= 1e-12 * L ** 3
            if nc == 3:
                assert abs(mz[-1, g, r] - mm[2]) <= 1e-12 * L ** 3
            norms = (sp.reshape(-1, nc) ** 2).sum(axis=1)
            assert np.max(np.abs(norms - 1.0)) < 1e-5, (L, r)
        r2t, t2r = pt.get_perm(g)
        assert np.array_equal(t2r[r2t], np.arange(R))


@pytest.mark.parametrize("mode", [ising_hip.Batch.BRANCHES, ising_hip.Batch.WORKLIST])
def test_batched_uneven_three_models_full_size(mode):
    seed = 2024
    t_ising = o.linspace(4.40, 4.62, R)
    t_xy = np.linspace(2.15, 2.25, R)
    t_heis = np.linspace(1.40, 1.50, R)
    kw = dict(exchange_every=10, measure_every=5, graph_sweeps=0)
    ckw = dict(kw, n_overrelax=2, renorm_every=100)
    members = [ising_hip.PTEnsemble("ising", SIZES, R, t_ising, seed, **kw),
               ising_hip.PTEnsemble("xy", SIZES, R, t_xy, seed, delta=1.5, **ckw),
               ising_hip.PTEnsemble("heisenberg", SIZES, R, t_heis, seed, delta=1.0, **ckw)]
    assert sum(p.device_bytes() for p in members) > 1.8e9
    batch = ising_hip.Batch(members, graph_sweeps=5, mode=mode)
    warm, samp = 5, 20
    batch.warmup(warm)
    batch.sample(samp)
    batch.sync()
    for p in members:
        assert p.sweeps() == warm + samp
    # Ising: the sizes the oracle replays in seconds, bit for bit
    tidx, bond, mag = members[0].read_ising()
    assert tidx.shape == (samp // 5, len(SIZES), R)
    for g, L in enumerate(SIZES):
        if L > 48:
            for k in range(tidx.shape[0]):
                assert sorted(tidx[k, g]) == list(range(R))
            for r in (0, R - 1):
                sp = members[0].get_spins(g, r)
                assert set(np.unique(sp)) <= {-1, 1}
                assert (bond[-1, g, r], mag[-1, g, r]) == o.ising_measure(L, sp)
            continue
        rt, rb, rm, final = o.ising_pt_run(L, R, t_ising, seed, warm, samp, 10, 5)
        assert np.array_equal(tidx[:, g, :].reshape(-1), rt), L
        assert np.array_equal(bond[:, g, :].reshape(-1), rb), L
        assert np.array_equal(mag[:, g, :].reshape(-1), rm), L
        for r in (0, 7, R - 1):
            assert np.array_equal(members[0].get_spins(g, r), final[r]), (L, r)
    # XY and Heisenberg: properties at every size
    for p, nc, temps, model in ((members[1], 2, t_xy, o.MODEL_XY), (members[2], 3, t_heis, o.MODEL_HEIS)):
        check_cont_member(p, nc, temps, seed, model, SIZES, lambda L: (0, R - 1) if L >= 96 else (0, 13, R - 1),
                          10, 5)
    batch.close()


def test_heisenberg_pt_L64_x32_full_ladder():
    seed = 9
    temps = np.linspace(1.40, 1.50, R)
    pt = ising_hip.PTEnsemble("heisenberg", [64], R, temps, seed, exchange_every=5, measure_every=5, n_overrelax=2,
                              delta=1.2, adapt_every=10, adapt_target=0.5, renorm_every=100)
    d0 = pt.get_deltas(0).copy()
    pt.warmup(60)
    d1 = pt.get_deltas(0).copy()
    assert not np.array_equal(d0, d1) and np.all(d1 > 1e-3) and np.all(d1 < np.pi)
    pt.sample(60)  # renormalised at t + 1 = 100
    assert np.array_equal(pt.get_deltas(0), d1)  # frozen while sampling
    pt.sync()
    check_cont_member(pt, 3, temps, seed, o.MODEL_HEIS, [64], lambda L: (0, 5, 17, R - 1), 5, 5)
    tries, acc = pt.exchange_stats()
    assert tries == 12 * 16 - 6 and acc > 0  # alternating parity: 16 pairs, then 15
