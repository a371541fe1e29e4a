"""Several ensembles of different models advanced together by one batch graph
give exactly the records each ensemble gives when run alone."""
import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu
ising_hip = pytest.importorskip("ising_hip")


def make_members(seed):
    t_ising = o.linspace(4.4, 4.62, 4)
    t_xy = np.linspace(2.1, 2.3, 4)
    t_heis = np.linspace(1.40, 1.46, 4)
    kw = dict(exchange_every=3, measure_every=2, graph_sweeps=0)
    ckw = dict(kw, delta=1.0, n_overrelax=2)
    return [
        ising_hip.PTEnsemble("ising", [8, 16], 4, t_ising, seed, **kw),
        ising_hip.PTEnsemble("xy", [8], 4, t_xy, seed + 1, **ckw),
        ising_hip.PTEnsemble("heisenberg", [8], 4, t_heis, seed + 2, **ckw),
        ising_hip.PTEnsemble("ising_msc", [8], 32, o.linspace(4.4, 4.62, 32), seed + 3, **kw),
    ]


def read(p, model):
    return p.read_ising() if model.startswith("ising") else p.read_cont()


def test_batch_equals_members_run_alone():
    models = ["ising", "xy", "heisenberg", "ising_msc"]
    alone = make_members(5)
    batched = make_members(5)
    batch = ising_hip.Batch(batched, graph_sweeps=4)
    out_a, out_b = [], []
    for p in alone:
        p.warmup(9)
    batch.warmup(9)      # two replays of the 4-sweep graph and one direct sweep
    for n in (8, 3, 13):
        for p in alone:
            p.sample(n)
        batch.sample(n)
        batch.sync()
        out_a.append([read(p, m) for p, m in zip(alone, models)])
        out_b.append([read(p, m) for p, m in zip(batched, models)])
    for ra, rb in zip(out_a, out_b):
        for xa, xb in zip(ra, rb):
            for u, v in zip(xa, xb):
                assert np.array_equal(u, v)
    for p, q in zip(alone, batched):
        assert np.array_equal(p.get_perm(0), q.get_perm(0))
        assert p.sweeps() == q.sweeps() == 9 + 8 + 3 + 13


def test_batch_orders_member_eager_work():
    """Eager work on a member between batch calls (set_spins, a member's own
    sample) is ordered with the batch graph."""
    temps = o.linspace(4.4, 4.62, 2)
    a = ising_hip.PTEnsemble("ising", [8], 2, temps, 1, graph_sweeps=0)
    b = ising_hip.PTEnsemble("ising", [8], 2, temps, 1, graph_sweeps=0)
    batch = ising_hip.Batch([b], graph_sweeps=2)
    up = np.ones(8 ** 3, np.int8)
    for p in (a, b):
        p.set_spins(0, 1, up)
    a.warmup(6)
    batch.warmup(6)
    a.sample(5)
    b.sample(5)        # the member's own stream, after the batch graph
    batch.sample(4)
    a.sample(4)
    batch.sync()
    for x, y in zip(a.read_ising(), b.read_ising()):
        assert np.array_equal(x, y)
    assert np.array_equal(a.get_spins(0, 0), b.get_spins(0, 0))


def test_batch_rejects_shared_members():
    temps = o.linspace(4.4, 4.62, 2)
    a = ising_hip.PTEnsemble("ising", [8], 2, temps, 1)
    batch = ising_hip.Batch([a])
    with pytest.raises(ising_hip.IsingError):
        ising_hip.Batch([a])
    batch.close()
    ising_hip.Batch([a]).close()  # free again once the first batch is gone
