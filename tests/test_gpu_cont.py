"""GPU parity of the XY / Heisenberg path against the FP64 oracle replay.

FP32 on the device vs FP64 on the host: one pass from an identical input state
is compared per site (local energy to 1e-5 relative); sites whose Metropolis
uniform lies within 1e-4 of the acceptance probability are masked as
near-threshold and counted.  Reductions are FP64 and must match the oracle's
FP64 sums over the same FP32 state to 1e-12.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu
ising_hip = pytest.importorskip("ising_hip")

MODEL = {2: "xy", 3: "heisenberg"}


def o_init(L, nc, seed, replica=0):
    sp = np.zeros(L ** 3 * nc)
    o.lib.o_cont_init(L, nc, seed, o.lib.o_stream(L, replica), sp)
    return sp


def o_measure(L, nc, sp, J=1.0):
    e = C.c_double()
    m = np.zeros(3)
    o.lib.o_cont_measure(L, nc, np.ascontiguousarray(sp, np.float64), J, C.byref(e), m)
    return e.value, m


def site_energy(L, nc, sp):
    out = np.zeros(L ** 3)
    o.lib.o_cont_site_energy(L, nc, np.ascontiguousarray(sp, np.float64), 1.0, out)
    return out


def colour_sites(L, c):
    s = np.arange(L ** 3)
    x, y, z = s % L, (s // L) % L, s // (L * L)
    return s[((x + y + z) & 1) == c]


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("L", [4, 6, 16, 32])
def test_init_matches_oracle(nc, L):
    lat = ising_hip.ContinuousGpuLattice(L, nc, 42)
    got = lat.get_spins().astype(np.float64)
    assert np.max(np.abs(got - o_init(L, nc, 42))) < 2e-6


@pytest.mark.parametrize("nc,beta,delta", [(3, 1.0 / 1.44, 0.9), (3, 0.3, 2.5), (2, 1.0 / 2.2, 1.2), (2, 1.5, 0.4)])
@pytest.mark.parametrize("L", [8, 12, 16, 24, 48])
def test_one_metropolis_pass_matches_fp64_replay(nc, beta, delta, L):
    seed = 77
    lat = ising_hip.ContinuousGpuLattice(L, nc, seed)
    # move away from the random start so the acceptance is non-trivial
    for _ in range(3):
        lat.sweep(beta, 1.0, delta, 1)
    t = lat.sweeps()
    for c in (0, 1):
        before = lat.get_spins().astype(np.float64)
        lat.metropolis_pass(c, beta, 1.0, delta)
        got = lat.get_spins().astype(np.float64)
        ref = before.copy()
        u = np.zeros(L ** 3 // 2)
        p = np.zeros(L ** 3 // 2)
        o.lib.o_cont_metropolis_half(L, nc, ref, c, t, seed, o.lib.o_stream(L, 0), float(np.float32(beta)), 1.0,
                                     float(np.float32(delta)), 0, u.ctypes.data, p.ctypes.data)
        sites = colour_sites(L, c)
        near = np.abs(u - p) < 1e-4
        ok = sites[~near]
        e_got = site_energy(L, nc, got)[ok]
        e_ref = site_energy(L, nc, ref)[ok]
        assert np.all(np.abs(e_got - e_ref) <= 1e-5 * np.maximum(np.abs(e_ref), 1.0)), \
            np.max(np.abs(e_got - e_ref))
        # the untouched colour is unchanged, and the moved colour agrees component-wise
        other = colour_sites(L, 1 - c)
        g3, r3 = got.reshape(-1, nc), ref.reshape(-1, nc)
        assert np.array_equal(g3[other], before.reshape(-1, nc)[other])
        assert np.max(np.abs(g3[ok] - r3[ok])) < 1e-5
        assert near.sum() <= max(2, 0.01 * len(sites))
        # lattice carried forward from the GPU state
        lat.set_spins(got.astype(np.float32))


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("L", [12, 16, 24, 48])
def test_one_overrelaxation_pass_matches_fp64_replay(nc, L):
    lat = ising_hip.ContinuousGpuLattice(L, nc, 5)
    for c in (0, 1):
        before = lat.get_spins().astype(np.float64)
        lat.overrelax_pass(c, 1.0)
        got = lat.get_spins().astype(np.float64)
        ref = before.copy()
        o.lib.o_cont_overrelax_half(L, nc, ref, c, 1.0)
        # energies of the reflected sites to 1e-5 relative (the reflection keeps
        # them: S'.h = S.h; the other colour's sites see the moved neighbours,
        # whose conditioning is bounded below)
        mine = colour_sites(L, c)
        e_got, e_ref = site_energy(L, nc, got)[mine], site_energy(L, nc, ref)[mine]
        assert np.all(np.abs(e_got - e_ref) <= 1e-5 * np.maximum(np.abs(e_ref), 1.0))
        # components: S' = 2 (S.h) h / |h|^2 - S turns an absolute FP32 rounding
        # error d of the field sum (|d| <~ 2e-6 for 6 unit vectors) into ~2 d / |h|,
        # so the bound is 1e-5 where |h| >= 1 and grows as 1 / |h| below
        s4 = before.reshape(L, L, L, nc)
        hf = sum(np.roll(s4, sh, axis=ax) for ax in range(3) for sh in (1, -1)).reshape(-1, nc)
        hn = np.linalg.norm(hf, axis=1)
        tol = 1e-5 * np.maximum(1.0, 1.0 / np.maximum(hn, 1e-3))
        assert np.all(np.abs(got - ref).reshape(-1, nc).max(axis=1) <= tol)
        e0, _ = o_measure(L, nc, before)
        e1, _ = o_measure(L, nc, got)
        assert abs(e1 - e0) < 1e-5 * abs(e0)


@pytest.mark.parametrize("nc", [2, 3])
def test_measure_gpu_matches_fp64_sums(nc):
    L = 16
    lat = ising_hip.ContinuousGpuLattice(L, nc, 9)
    for _ in range(4):
        lat.sweep(0.6, 1.0, 0.8, 2)
    sp = lat.get_spins().astype(np.float64)
    e, m = o_measure(L, nc, sp, 1.0)
    got = lat.measure_gpu(1.0)
    assert abs(got[0] - e) < 1e-12 * abs(e)
    for k in range(nc):
        assert abs(got[1 + k] - m[k]) < 1e-9
    eh = lat.measure_gpu(0.5)
    assert abs(eh[0] - 0.5 * e) < 1e-12 * abs(e)
    assert lat.measure_raw() == got


@pytest.mark.parametrize("nc", [2, 3])
def test_pt_ensemble_records_and_renormalisation(nc):
    sizes, R = [8, 16], 8
    temps = np.linspace(1.40, 1.50, R) if nc == 3 else np.linspace(2.15, 2.25, R)
    pt = ising_hip.PTEnsemble(MODEL[nc], sizes, R, temps, 42, exchange_every=2, measure_every=1, n_overrelax=2,
                              renorm_every=3, delta=0.8)
    pt.warmup(5)
    pt.sample(9)  # last sample sweep is t=13; (13+1) % 7 ... renormalised at t+1 = 3,6,9,12
    tidx, e, mx, my, mz = pt.read_cont()
    assert tidx.shape == (9, 2, R)
    for k in range(9):
        for g in range(2):
            assert sorted(tidx[k, g]) == list(range(R))
    for g, L in enumerate(sizes):
        for r in (0, R - 1):
            sp = pt.get_spins(g, r).astype(np.float64)
            ee, mm = o_measure(L, nc, sp)
            assert abs(e[-1, g, r] - ee) < 1e-10 * abs(ee)
            assert abs(mx[-1, g, r] - mm[0]) < 1e-9 and abs(my[-1, g, r] - mm[1]) < 1e-9
            if nc == 3:
                assert abs(mz[-1, g, r] - mm[2]) < 1e-9
            norms = (sp.reshape(-1, nc) ** 2).sum(axis=1)
            assert np.max(np.abs(norms - 1.0)) < 1e-5
        r2t, t2r = pt.get_perm(g)
        assert np.array_equal(t2r[r2t], np.arange(R))
    tries, acc = pt.exchange_stats()
    assert tries > 0


def test_delta_adaptation_runs_in_warmup_only():
    R = 4
    temps = np.linspace(1.40, 1.50, R)
    pt = ising_hip.PTEnsemble("heisenberg", [16], R, temps, 3, delta=3.0, adapt_every=10, adapt_target=0.5,
                              n_overrelax=0, exchange_every=1000)
    assert np.all(pt.get_deltas(0) == np.float32(3.0))
    pt.warmup(200)
    d = pt.get_deltas(0)
    # near T_c a 3-radian cap accepts far less than half the moves, so the cap shrinks
    assert np.all(d < 2.5) and np.all(d > 0.05)
    pt.sample(50)
    assert np.array_equal(pt.get_deltas(0), d)  # frozen while sampling


def _helicity_numpy(sp, L):
    s = sp.reshape(L, L, L, 2).astype(np.float64)  # [z, y, x, comp]
    out = []
    for ax in (2, 1, 0):  # +x, +y, +z
        t = np.roll(s, -1, axis=ax)
        out.append((s * t).sum())
    for ax in (2, 1, 0):
        t = np.roll(s, -1, axis=ax)
        out.append((s[..., 1] * t[..., 0] - s[..., 0] * t[..., 1]).sum())
    return np.array(out)


def test_xy_helicity_sums_closed_forms_and_random_state():
    L, R = 8, 3
    pt = ising_hip.PTEnsemble("xy", [L], R, [2.2] * R, 4)
    n = L ** 3
    z, y, x = np.meshgrid(np.arange(L), np.arange(L), np.arange(L), indexing="ij")
    aligned = np.stack([np.ones(n), np.zeros(n)], axis=1).astype(np.float32).reshape(-1)
    theta = 2 * np.pi * 1 * x.reshape(-1) / L  # one full twist along x
    twisted = np.stack([np.cos(theta), np.sin(theta)], axis=1).astype(np.float32).reshape(-1)
    pt.set_spins(0, 0, aligned)
    pt.set_spins(0, 1, twisted)
    pt.sweeps()
    hel = pt.helicity()
    assert np.allclose(hel[0], [n, n, n, 0, 0, 0], atol=1e-9)
    c, s_ = np.cos(2 * np.pi / L), np.sin(-2 * np.pi / L)
    assert np.allclose(hel[1], [n * c, n, n, n * s_, 0, 0], rtol=1e-6, atol=1e-6)
    rnd = pt.get_spins(0, 2)  # the random start
    want = _helicity_numpy(rnd, L)
    assert np.allclose(hel[2], want, rtol=1e-10, atol=1e-9)
    assert np.allclose(hel[1], _helicity_numpy(pt.get_spins(0, 1), L), rtol=1e-10, atol=1e-9)


@pytest.mark.parametrize("nc", [2, 3])
def test_pt_multiblock_close_between_unmeasured_sweeps(nc):
    """48 replicas (12 closing blocks), measuring and exchanging every 5th sweep.

    Sweeps that neither measure nor exchange are closed by the closing launch's
    last block too; the device's sweep counter must equal the host's, every
    record must be there, and each exchange replayed by the oracle from the
    recorded energies (same key, round and pair parity) must give the next
    record's temperature indices bit for bit.
    """
    sizes, R, seed = [8, 12, 16], 16, 11
    temps = np.linspace(1.40, 1.50, R) if nc == 3 else np.linspace(2.15, 2.25, R)
    pt = ising_hip.PTEnsemble(MODEL[nc], sizes, R, temps, seed, exchange_every=5, measure_every=5, n_overrelax=1,
                              renorm_every=7, delta=0.8, adapt_every=5)
    pt.warmup(13)
    pt.sample(100)
    pt.sync()  # fails unless the device's sweep counter equals the sweeps enqueued
    assert pt.sweeps() == 113
    tidx, e, mx, my, mz = pt.read_cont()
    assert tidx.shape == (20, len(sizes), R)
    betas = 1.0 / np.asarray(temps, np.float64)
    model = o.MODEL_XY if nc == 2 else o.MODEL_HEIS
    accepted = 0
    for g, L in enumerate(sizes):
        r2t = np.ascontiguousarray(tidx[0, g], np.int32)
        t2r = np.zeros(R, np.int32)
        t2r[r2t] = np.arange(R, dtype=np.int32)
        for k in range(19):
            accepted += o.lib.o_exchange(R, betas, np.ascontiguousarray(e[k, g], np.float64), r2t, t2r, k, seed,
                                         o.lib.o_stream(L, 0), model)
            assert np.array_equal(r2t, tidx[k + 1, g]), (L, k)
        for r in (0, R // 2, R - 1):
            sp = pt.get_spins(g, r).astype(np.float64)
            ee, mm = o_measure(L, nc, sp)
            assert abs(e[-1, g, r] - ee) < 1e-10 * abs(ee)
            assert abs(mx[-1, g, r] - mm[0]) < 1e-9 and abs(my[-1, g, r] - mm[1]) < 1e-9
    assert accepted > 0
    tries, acc = pt.exchange_stats()
    assert acc >= accepted and tries >= 19 * len(sizes) * ((R - 1) // 2)


def _ensembles_with_offsets(nc, wide):
    """A continuous ensemble whose passes use 32-bit (default at these sizes)
    or 64-bit offsets (ISING_HIP_OFFSETS64 at creation: the path of lattices
    whose half-lattices reach 2^30 sites)."""
    import os
    if wide:
        os.environ["ISING_HIP_OFFSETS64"] = "1"
    try:
        return ising_hip.PTEnsemble(MODEL[nc], [8, 12, 16], 6, np.linspace(1.4, 2.3, 6), 21, exchange_every=3,
                                    measure_every=2, n_overrelax=2, delta=1.0, adapt_every=4, renorm_every=5,
                                    graph_sweeps=0)
    finally:
        os.environ.pop("ISING_HIP_OFFSETS64", None)


@pytest.mark.parametrize("nc", [2, 3])
def test_64bit_offset_passes_equal_32bit_ones(nc):
    """The passes with 64-bit addressing (kept for half-lattices of 2^30 sites
    or more) give the 32-bit-offset passes' lattices and records bit for bit,
    per ensemble and in a work-list batch."""
    a, b = _ensembles_with_offsets(nc, False), _ensembles_with_offsets(nc, True)
    for p in (a, b):
        p.warmup(9)
        p.sample(11)
    for x, y in zip(a.read_cont(), b.read_cont()):
        assert np.array_equal(x, y)
    for g in range(3):
        for r in range(6):
            assert np.array_equal(a.get_spins(g, r), b.get_spins(g, r))
    ba = ising_hip.Batch([a], graph_sweeps=2, mode=ising_hip.Batch.WORKLIST)
    bb = ising_hip.Batch([b], graph_sweeps=2, mode=ising_hip.Batch.WORKLIST)
    for bt in (ba, bb):
        bt.sample(5)
        bt.sync()
    for x, y in zip(a.read_cont(), b.read_cont()):
        assert np.array_equal(x, y)
    assert np.array_equal(a.get_spins(2, 5), b.get_spins(2, 5))
    ba.close()
    bb.close()
    for p in (a, b):  # heat-bath passes, then the (untrained) learned proposal
        p.heat_bath(True)
        p.sample(4)
        p.heat_bath(False)
        p.learn_init(16, seed=3)
        p.learn_enable(True)
        p.sample(4)
    for x, y in zip(a.read_cont(), b.read_cont()):
        assert np.array_equal(x, y)
    for g in range(3):
        assert np.array_equal(a.get_spins(g, 0), b.get_spins(g, 0))
