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


def make_worklist_members(seed):
    kw = dict(exchange_every=4, measure_every=2, graph_sweeps=0)
    ckw = dict(kw, n_overrelax=2, renorm_every=7)
    return [
        ising_hip.PTEnsemble("ising", [8, 16, 32], 6, o.linspace(4.4, 4.62, 6), seed, **kw),
        ising_hip.PTEnsemble("xy", [8, 12], 6, np.linspace(2.1, 2.3, 6), seed + 1, delta=1.5, adapt_every=3, **ckw),
        ising_hip.PTEnsemble("heisenberg", [8, 6], 6, np.linspace(1.40, 1.46, 6), seed + 2, delta=1.0, adapt_every=3,
                             **ckw),
    ]


@pytest.mark.parametrize("mode", [ising_hip.Batch.WORKLIST, ising_hip.Batch.PERSISTENT])
def test_worklist_batch_equals_branch_batch(mode):
    """One launch per pass over a work list of every member's blocks (and the
    same list on a persistent grid) gives the branch batch's lattices,
    permutations, step sizes and records bit for bit (every block runs its
    member's own body, XY's with two sites per thread in half blocks)."""
    ref, wl = make_worklist_members(31), make_worklist_members(31)
    b_ref = ising_hip.Batch(ref, graph_sweeps=3)
    b_wl = ising_hip.Batch(wl, graph_sweeps=3, mode=mode)
    for b in (b_ref, b_wl):
        b.warmup(10)   # step-size adaptation in warm-up
    outs = []
    for b, ms in ((b_ref, ref), (b_wl, wl)):
        recs = []
        for n in (6, 1, 8):  # graph replays and direct launches
            b.sample(n)
            b.sync()
            recs.append([ms[0].read_ising(), ms[1].read_cont(), ms[2].read_cont()])
        outs.append(recs)
    for ra, rb in zip(*outs):
        for xa, xb in zip(ra, rb):
            for u, v in zip(xa, xb):
                assert np.array_equal(u, v)
    for p, q in zip(ref, wl):
        assert p.sweeps() == q.sweeps() == 10 + 15
        for g in range(len(p.sizes)):
            assert np.array_equal(p.get_perm(g)[0], q.get_perm(g)[0])
            for r in range(6):
                assert np.array_equal(p.get_spins(g, r), q.get_spins(g, r))
    for i in (1, 2):
        assert np.array_equal(ref[i].get_deltas(0), wl[i].get_deltas(0))


def test_worklist_batch_rejects_other_members():
    temps = o.linspace(4.4, 4.62, 32)
    msc = ising_hip.PTEnsemble("ising_msc", [8], 32, temps, 1)
    b = ising_hip.Batch([msc])
    with pytest.raises(ising_hip.IsingError):
        b.set_mode(ising_hip.Batch.WORKLIST)
    b.close()
    xy = ising_hip.PTEnsemble("xy", [8], 2, np.linspace(2.1, 2.3, 2), 1, n_overrelax=2)
    heis = ising_hip.PTEnsemble("heisenberg", [8], 2, np.linspace(1.4, 1.5, 2), 1, n_overrelax=1)
    b = ising_hip.Batch([xy, heis])
    with pytest.raises(ising_hip.IsingError):
        b.set_mode(ising_hip.Batch.WORKLIST)  # unequal overrelaxation counts
    b.close()
    b = ising_hip.Batch([xy], mode=ising_hip.Batch.WORKLIST)
    b.sample(2)
    xy.heat_bath(True)  # a member switched to passes the work list does not run
    with pytest.raises(ising_hip.IsingError):
        b.sample(2)
    b.close()


@pytest.mark.parametrize("mode", [ising_hip.Batch.WORKLIST, ising_hip.Batch.PERSISTENT])
@pytest.mark.parametrize("which", ["ising", "xy", "heisenberg_xy"])
def test_worklist_batch_of_member_subsets(mode, which):
    """Work lists with the Ising member alone (its colour passes and closing
    kernel only), one continuous member, or both continuous members in the
    other order, against the branch batch."""
    def members(seed):
        kw = dict(exchange_every=4, measure_every=2, graph_sweeps=0)
        ckw = dict(kw, n_overrelax=1, renorm_every=7)
        out = []
        if which == "ising":
            out.append(ising_hip.PTEnsemble("ising", [8, 16], 6, o.linspace(4.4, 4.62, 6), seed, **kw))
        if which.startswith("heisenberg"):
            out.append(ising_hip.PTEnsemble("heisenberg", [8], 6, np.linspace(1.40, 1.46, 6), seed + 2, delta=1.0,
                                            **ckw))
        if which.endswith("xy"):
            out.append(ising_hip.PTEnsemble("xy", [8, 12], 6, np.linspace(2.1, 2.3, 6), seed + 1, delta=1.5, **ckw))
        return out
    ref, wl = members(5), members(5)
    b_ref = ising_hip.Batch(ref, graph_sweeps=3)
    b_wl = ising_hip.Batch(wl, graph_sweeps=3, mode=mode)
    for b in (b_ref, b_wl):
        b.warmup(5)
        b.sample(7)
        b.sync()
    for p, q in zip(ref, wl):
        ra = p.read_ising() if which == "ising" else p.read_cont()
        rb = q.read_ising() if which == "ising" else q.read_cont()
        for u, v in zip(ra, rb):
            assert np.array_equal(u, v)
        assert p.sweeps() == q.sweeps() == 12
        assert np.array_equal(p.get_spins(0, 3), q.get_spins(0, 3))
