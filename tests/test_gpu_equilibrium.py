"""Equilibrium agreement of the GPU PT path with the reference's CPU algorithms.

* 3D Ising: GPU PT (checkerboard Metropolis + exchange) against the oracle's
  restatement of the reference Wolff algorithm (wolff.rs:27-65), within
  combined jackknife errors; and against the reference's published GPU
  summary for L=8 (analysis/data/gpu_windows_pipeline/publication).
* Binder crossing of L=16 and L=32 with the reference's crossing rule
  (analysis/scripts/analyze_gpu_fss.py:256-277) -> T_c ~ 4.5115.
* Heisenberg / XY: GPU cap/rotation Metropolis + overrelaxation against the
  restated CPU cap Metropolis + OR (heisenberg/) and XY Wolff (xy/wolff.rs).
* E, |M|, chi and U4 per temperature (Ising, XY, Heisenberg at L = 8 and 16)
  within 2 combined jackknife sigma of the CPU chains: at least 90% of the
  z-scores within 2, none beyond 4; the reference's published L=16 summary.
"""
import csv
import math
import os

import numpy as np
import pytest
from scipy.interpolate import interp1d

import oracle_lib as o

pytestmark = pytest.mark.gpu
ising_hip = pytest.importorskip("ising_hip")
HERE = os.path.dirname(os.path.abspath(__file__))


def blocks(x, nb=20):
    x = np.asarray(x, np.float64)
    bs = len(x) // nb
    return x[: nb * bs].reshape(nb, bs)


def mean_err(x, nb=20):
    b = blocks(x, nb).mean(axis=1)
    return b.mean(), b.std(ddof=1) / math.sqrt(nb)


def binder(m, nb=20):
    """U4 = 1 - <m^4>/(3 <m^2>^2) with a block-jackknife error."""
    m = np.asarray(m, np.float64)
    full = 1 - (m ** 4).mean() / (3 * (m ** 2).mean() ** 2)
    bm = blocks(m, nb)
    jk = []
    for i in range(nb):
        keep = np.delete(bm, i, axis=0).reshape(-1)
        jk.append(1 - (keep ** 4).mean() / (3 * (keep ** 2).mean() ** 2))
    jk = np.array(jk)
    return full, math.sqrt((nb - 1) / nb * ((jk - full) ** 2).sum())


def ising_pt_series(L, R, temps, seed, warmup, samples, me=1, ee=10):
    pt = ising_hip.PTEnsemble("ising", [L], R, temps, seed, exchange_every=ee, measure_every=me,
                              record_capacity=min(samples // me + 2, 20000))
    pt.warmup(warmup)
    e = [[] for _ in range(R)]
    m = [[] for _ in range(R)]
    done = 0
    chunk = 20000 * me
    while done < samples:
        n = min(chunk, samples - done)
        pt.sample(n)
        tidx, bond, mag = pt.read_ising()
        t = tidx[:, 0, :].reshape(-1)
        eb = (-bond[:, 0, :].reshape(-1)) / L ** 3
        mb = np.abs(mag[:, 0, :].reshape(-1)) / L ** 3
        for k in range(R):
            sel = t == k
            e[k].append(eb[sel])
            m[k].append(mb[sel])
        done += n
    return [np.concatenate(x) for x in e], [np.concatenate(x) for x in m]


def test_ising_pt_matches_cpu_wolff():
    L, R = 8, 8
    temps = o.linspace(4.40, 4.62, R)
    e, m = ising_pt_series(L, R, temps, 11, 2000, 60000)
    for k in (0, 3, 7):
        b, mg = o.ref_ising_chain("wolff", L, 100 + k, 1.0 / temps[k], 2000, 60000)
        ew, mw = -b / L ** 3, np.abs(mg) / L ** 3
        for gpu, cpu in ((e[k], ew), (m[k], mw), (m[k] ** 2, mw ** 2)):
            g, ge = mean_err(gpu)
            c, ce = mean_err(cpu)
            assert abs(g - c) < 4.5 * math.hypot(ge, ce), (temps[k], g, c, ge, ce)


def test_ising_pt_matches_published_reference_summary_L8():
    rows = list(csv.DictReader(open(os.path.join(HERE, "golden", "published_gpu_fss_ising_N8_summary.csv"))))
    temps = np.array([float(r["T"]) for r in rows])
    R = len(temps)
    assert R == 32 and abs(temps[0] - 4.40) < 1e-6 and abs(temps[-1] - 4.62) < 1e-6
    e, m = ising_pt_series(8, R, o.linspace(4.40, 4.62, R), 5, 5000, 50000)
    z = []
    for k, r in enumerate(rows):
        for col, ser in (("E", e[k]), ("M", m[k]), ("M2", m[k] ** 2)):
            g, ge = mean_err(ser)
            z.append((g - float(r[col])) / math.hypot(ge, float(r[col + "_err"])))
    z = np.array(z)
    # the published 20-block errors are themselves estimates: require agreement in distribution
    assert np.mean(z ** 2) < 3.0 and np.max(np.abs(z)) < 6.0, (np.mean(z ** 2), np.max(np.abs(z)))
    assert abs(np.mean(z)) < 4.0 / math.sqrt(len(z)) + 0.3


def observables_z(gpu_e, gpu_m, cpu_e, cpu_m, beta, n3):
    """z-scores of E, |M|, chi and U4 between two series (E and |M| per spin),
    errors by the reference's 20-block jackknife (gpu_fss.rs:327-412) and, for
    U4 = 1 - <m^4> / (3 <m^2>^2), the same blocking."""
    jg = o.jackknife(gpu_e, gpu_m, beta, n3, 20)
    jc = o.jackknife(cpu_e, cpu_m, beta, n3, 20)
    z = [(jg[i] - jc[i]) / math.hypot(jg[i + 1], jc[i + 1]) for i in (0, 2, 10)]  # E, |M|, chi
    (ug, uge), (uc, uce) = binder(gpu_m), binder(cpu_m)
    z.append((ug - uc) / math.hypot(uge, uce))
    return z


def check_two_sigma(z, what):
    """At least 90% of the z-scores within 2 and none beyond 4."""
    a = np.abs(np.asarray(z))
    assert np.all(np.isfinite(a)), (what, z)
    assert np.mean(a < 2.0) >= 0.9 and a.max() < 4.0, (what, np.round(z, 2).tolist())


def _parallel(fn, args):
    from concurrent.futures import ThreadPoolExecutor  # the oracle's C loops release the GIL
    with ThreadPoolExecutor(max_workers=16) as ex:
        return list(ex.map(lambda a: fn(*a), args))


def ising_z():
    """GPU PT (L = 8, 16; 8 temperatures in [4.40, 4.62]) against the restated
    reference Wolff (wolff.rs:27-65) at each temperature."""
    R = 8
    temps = o.linspace(4.40, 4.62, R)
    z = []
    for L in (8, 16):
        e, m = ising_pt_series(L, R, temps, 21 + L, 5000, 100000)
        cpu = _parallel(lambda k: o.ref_ising_chain("wolff", L, 300 + 10 * L + k, 1.0 / temps[k], 2000, 60000),
                        [(k,) for k in range(R)])
        for k in range(R):
            b, mg = cpu[k]
            z += observables_z(e[k], m[k], -b / L ** 3, np.abs(mg) / L ** 3, 1.0 / temps[k], float(L ** 3))
    return z


def continuous_z(model):
    """GPU PT (L = 8, 16; 4 temperatures) against the restated reference CPU
    chains: XY Wolff (xy/wolff.rs:16-73), Heisenberg cap Metropolis + 2
    overrelaxation sweeps (heisenberg/metropolis.rs:10-92, overrelax.rs:15-48;
    4 independent chains per temperature, concatenated)."""
    R = 4
    temps = o.linspace(2.15, 2.25, R) if model == "xy" else o.linspace(1.40, 1.50, R)
    kw = dict(delta=1.5, n_overrelax=2, exchange_every=10) if model == "xy" else \
        dict(delta=1.0, n_overrelax=2, exchange_every=10, adapt_every=0)
    z = []
    for L in (8, 16):
        e, m = cont_pt_series(model, L, R, temps, 31 + L, 3000, 120000, **kw)
        if model == "xy":
            cpu = _parallel(lambda k: o.ref_xy_chain(L, 500 + 10 * L + k, 1.0 / temps[k], 2000, 60000),
                            [(k,) for k in range(R)])
        else:
            n = 40000 if L == 8 else 10000  # ~13 s of CPU per chain at L = 16
            runs = _parallel(lambda k, c: o.ref_heis_chain(L, 700 + 40 * L + 4 * k + c, 1.0 / temps[k], 1.0, 2, 2000,
                                                           n), [(k, c) for k in range(R) for c in range(4)])
            cpu = [tuple(np.concatenate([runs[4 * k + c][q] for c in range(4)]) for q in (0, 1)) for k in range(R)]
        for k in range(R):
            ec, mc = cpu[k]
            z += observables_z(e[k], m[k], ec / L ** 3, mc / L ** 3, 1.0 / temps[k], float(L ** 3))
    return z


def test_e_m_chi_u4_within_two_sigma_of_cpu_reference_chains():
    """E, |M|, chi and U4 per temperature for Ising, XY and Heisenberg at L = 8
    and 16 (128 z-scores): at least 90% within 2 over all of them, none beyond
    4 for any model.  The 90% is pooled: E and |M| (and chi, U4) of one
    temperature are strongly correlated, so a per-model fraction of 32 would
    fail by chance several times in a hundred runs."""
    zs = {"ising": ising_z(), "xy": continuous_z("xy"), "heisenberg": continuous_z("heisenberg")}
    for model, z in zs.items():
        a = np.abs(np.asarray(z))
        assert np.all(np.isfinite(a)) and a.max() < 4.0, (model, np.round(z, 2).tolist())
    check_two_sigma(sum(zs.values(), []), "pooled")
    print({m: round(float(np.mean(np.abs(z) < 2.0)), 3) for m, z in zs.items()})


def test_ising_pt_matches_published_reference_summary_L16():
    """The reference's published L=16 summary (32 temperatures in [4.40, 4.62],
    20-block jackknife errors) against GPU PT on the same ladder: E, |M|, M^2
    and chi, in distribution (the published errors are estimates too)."""
    rows = list(csv.DictReader(open(os.path.join(HERE, "golden", "published_gpu_fss_ising_N16_summary.csv"))))
    temps = np.array([float(r["T"]) for r in rows])
    R = len(temps)
    assert R == 32 and abs(temps[0] - 4.40) < 1e-6 and abs(temps[-1] - 4.62) < 1e-6
    L = 16
    lt = o.linspace(4.40, 4.62, R)
    e, m = ising_pt_series(L, R, lt, 6, 5000, 100000)
    z = []
    for k, r in enumerate(rows):
        j = o.jackknife(e[k], m[k], 1.0 / lt[k], float(L ** 3), 20)
        for col, i in (("E", 0), ("M", 2), ("M2", 4), ("chi", 10)):
            z.append((j[i] - float(r[col])) / math.hypot(j[i + 1], float(r[col + "_err"])))
    z = np.array(z)
    assert np.mean(z ** 2) < 3.0 and np.max(np.abs(z)) < 6.0, (np.mean(z ** 2), np.max(np.abs(z)))
    assert abs(np.mean(z)) < 4.0 / math.sqrt(len(z)) + 0.3


def reference_crossing(T, U1, U2):
    """analyze_gpu_fss.py:256-277: cubic interpolation on a 2000-point common grid,
    first sign change of U1 - U2, linear root between the bracketing points."""
    Tc = np.linspace(T.min(), T.max(), 2000)
    d = interp1d(T, U1, kind="cubic")(Tc) - interp1d(T, U2, kind="cubic")(Tc)
    for j in range(len(d) - 1):
        if d[j] * d[j + 1] < 0:
            return Tc[j] - d[j] * (Tc[j + 1] - Tc[j]) / (d[j + 1] - d[j])
    return None


def test_binder_crossing_reproduces_3d_ising_tc():
    R = 32
    temps = o.linspace(4.40, 4.62, R)
    U = {}
    for L, samples in ((16, 400000), (32, 400000)):
        _, m = ising_pt_series(L, R, temps, 42, 20000, samples, me=5)
        U[L] = np.array([binder(m[k])[0] for k in range(R)])
    tc = reference_crossing(np.array(temps), U[16], U[32])
    assert tc is not None
    assert abs(tc - 4.5115) < 0.02, tc
    # the library's crossing finder gives the same temperature
    assert ising_hip.binder_crossing(temps, U[16], temps, U[32]) == pytest.approx(tc, abs=1e-10)


def cont_pt_series(model, L, R, temps, seed, warmup, samples, **kw):
    pt = ising_hip.PTEnsemble(model, [L], R, temps, seed, measure_every=1, record_capacity=samples + 2, **kw)
    pt.warmup(warmup)
    pt.sample(samples)
    tidx, e, mx, my, mz = pt.read_cont()
    t = tidx[:, 0, :].reshape(-1)
    eb = e[:, 0, :].reshape(-1) / L ** 3
    mb = np.sqrt(mx ** 2 + my ** 2 + mz ** 2)[:, 0, :].reshape(-1) / L ** 3
    return [eb[t == k] for k in range(R)], [mb[t == k] for k in range(R)]


def test_heisenberg_pt_matches_cpu_cap_metropolis_overrelaxation():
    L, R = 8, 4
    temps = o.linspace(1.40, 1.50, R)
    e, m = cont_pt_series("heisenberg", L, R, temps, 3, 2000, 30000, delta=1.0, n_overrelax=2, exchange_every=10,
                          adapt_every=0)
    for k in (0, R - 1):
        ec, mc = o.ref_heis_chain(L, 50 + k, 1.0 / temps[k], 1.0, 2, 2000, 30000)
        for gpu, cpu in ((e[k], ec / L ** 3), (m[k], mc / L ** 3)):
            g, ge = mean_err(gpu)
            c, ce = mean_err(cpu)
            assert abs(g - c) < 4.5 * math.hypot(ge, ce), (temps[k], g, c, ge, ce)


def test_xy_pt_matches_cpu_wolff():
    L, R = 8, 4
    temps = o.linspace(2.15, 2.25, R)
    e, m = cont_pt_series("xy", L, R, temps, 4, 2000, 30000, delta=1.5, n_overrelax=2, exchange_every=10)
    for k in (0, R - 1):
        ec, mc = o.ref_xy_chain(L, 70 + k, 1.0 / temps[k], 2000, 60000)
        for gpu, cpu in ((e[k], ec / L ** 3), (m[k], mc / L ** 3)):
            g, ge = mean_err(gpu)
            c, ce = mean_err(cpu)
            assert abs(g - c) < 4.5 * math.hypot(ge, ce), (temps[k], g, c, ge, ce)
