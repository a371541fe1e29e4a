"""One launch per int8 Ising sweep (z planes of several replicas streamed
through LDS, both colours updated, colour 0 of the planes beyond a block's ends
recomputed; small lattices whole): records, permutations and lattices
bit-exact against the oracle and against the two-launch sweeps, with graphs
and annealing."""
import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu
ising_hip = pytest.importorskip("ising_hip")


@pytest.mark.parametrize("planes,sizes,R", [(8, [8, 16, 32, 64, 128], 3), (16, [16, 64, 128], 3), (8, [24, 96], 3),
                                           (32, [8, 32, 96], 3), (16, [32, 64], 21)])
def test_fused_pt_records_are_bit_exact(planes, sizes, R):
    seed, warmup, samples, ee, me = 11, 4, 12, 3, 2
    temps = o.linspace(4.40, 4.62, R)
    pt = ising_hip.PTEnsemble("ising", sizes, R, temps, seed, exchange_every=ee, measure_every=me, ising_fused=planes)
    pt.warmup(warmup)
    pt.sample(samples)
    tidx, bond, mag = pt.read_ising()
    for gi, L in enumerate(sizes):
        rt, rb, rm, final = o.ising_pt_run(L, R, temps, seed, warmup, samples, ee, me)
        assert np.array_equal(tidx[:, gi, :].reshape(-1), rt), L
        assert np.array_equal(bond[:, gi, :].reshape(-1), rb), L
        assert np.array_equal(mag[:, gi, :].reshape(-1), rm), L
        for r in range(R):
            assert np.array_equal(pt.get_spins(gi, r), final[r]), (L, r)


def test_fused_equals_two_launches_with_graphs_and_host_calls():
    temps = o.linspace(4.4, 4.62, 6)
    kw = dict(exchange_every=3, measure_every=2)
    outs = []
    for fused, graph in ((0, 0), (8, 0), (8, 4), (16, 4), (32, 4)):
        p = ising_hip.PTEnsemble("ising", [16, 32, 64], 6, temps, 21, graph_sweeps=graph, ising_fused=fused, **kw)
        p.warmup(9)
        p.sample(10)
        p.sample(7)
        outs.append(p.read_ising() + tuple(p.get_spins(2, r) for r in range(6)) + tuple(p.get_perm(2)))
    for other in outs[1:]:
        for x, y in zip(outs[0], other):
            assert np.array_equal(x, y)


def test_fused_annealing_matches_two_launches():
    betas = [float(np.float32(1.0 / T)) for T in np.linspace(6.0, 3.8, 7)]
    res = []
    for fused in (0, 8):
        p = ising_hip.PTEnsemble("ising", [32, 64], 2, [5.0, 5.0], 3, ising_fused=fused)
        b, m = p.anneal(betas, measure=True)
        res.append((b, m, p.get_spins(1, 1)))
    for x, y in zip(*res):
        assert np.array_equal(x, y)


def test_ineligible_sizes_fall_back_to_two_launches():
    temps = o.linspace(4.4, 4.6, 2)
    a = ising_hip.PTEnsemble("ising", [6, 40], 2, temps, 5, ising_fused=8)
    b = ising_hip.PTEnsemble("ising", [6, 40], 2, temps, 5)
    for p in (a, b):
        p.warmup(3)
        p.sample(4)
    for x, y in zip(a.read_ising(), b.read_ising()):
        assert np.array_equal(x, y)
    with pytest.raises(ising_hip.IsingError):
        ising_hip.PTEnsemble("ising", [8], 2, temps, 5, ising_fused=12)
