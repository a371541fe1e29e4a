"""Throughput of the 3D Ising parallel-tempering path on one MI355X.

Workload (the reference's 3D Ising parallel-tempering run at full size): L in
{8,16,32,64,128}, 32 replicas per L on T = linspace(4.40, 4.62, 32), int8
spins from a random +-1 start, Philox per (seed, replica, sweep, site),
exchange every 10 sweeps, measure every 5.  One step = one exchange interval,
the period at which that schedule repeats: 10 sweeps of every lattice
(766,935,040 spin updates), the 5th and 10th measured, the 10th closed by a
replica-exchange round.

Prints one JSON line (metric, value, configuration) with a `roofline` object for the
dominant kernel (ising_half_kernel) and a `cpu_baseline` from the oracle's
restatement of the reference's random-site Metropolis (metropolis.rs:8-27)
timed on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ising-rs_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
SIZES = [8, 16, 32, 64, 128]
R = 32


def linspace(a, b, n):
    return [a + (b - a) * i / (n - 1) for i in range(n)]


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_leg(model, seed, seconds):
    """One single-threaded chain of the reference's CPU algorithm (oracle
    restatement) run for about `seconds`: (updates, wall seconds, detail)."""
    import ctypes as C

    import oracle_lib as o

    if model.startswith("ising"):
        # config 1 exactly, repeated with seeds seed, seed + 1, ... until `seconds`:
        # L=16, T=4.5115, hot start, 2000 warm-up + 2000 sample sweeps, E, |M|, M^2, M^4 every sweep
        L, n = 16, 2000
        b = np.zeros(n, np.int64)
        m = np.zeros(n, np.int64)
        runs, upd, obs = 0, 0, None
        t0 = time.perf_counter()
        while True:
            o.lib.o_ref_ising_chain(0, L, seed + runs, 1.0 / 4.5115, 2000, n, b, m)
            runs += 1
            upd += 2 * n * L ** 3
            if obs is None:
                e, am = -b / L ** 3, np.abs(m) / L ** 3
                obs = {"E": round(float(e.mean()), 5), "absM": round(float(am.mean()), 5),
                       "M2": round(float((am ** 2).mean()), 6), "M4": round(float((am ** 4).mean()), 7)}
            if time.perf_counter() - t0 >= seconds:
                break
        return upd, time.perf_counter() - t0, {"runs": runs, "first_run_means": obs}
    wall = C.c_double()
    if model == "xy":
        steps = C.c_int64()
        flipped = o.lib.o_bench_ref_xy_wolff(32, 1.0 / 2.2, seed, seconds, C.byref(wall), C.byref(steps))
        return flipped, wall.value, {"wolff_steps": steps.value}
    upd = o.lib.o_bench_ref_heis(32, 1.0 / 1.443, 1.0, 2, seed, seconds, C.byref(wall)) // 3
    return upd, wall.value, {}


def cpu_baseline(seconds, model="ising"):
    """The reference's single-threaded CPU algorithm (oracle restatement) on one
    core, and as independent chains on every core of this process's CPU share
    (capped at 16), each for about `seconds` of wall time.  The C loops run with
    the GIL released, one Python thread per chain."""
    import threading

    upd, wall, detail = _cpu_leg(model, 42, seconds)
    ncores = max(1, min(16, len(os.sched_getaffinity(0))))
    many = [None] * ncores

    def run(k):
        many[k] = _cpu_leg(model, 1000 * (k + 1), seconds)

    threads = [threading.Thread(target=run, args=(i,)) for i in range(ncores)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    all_upd = sum(u for u, _, _ in many)
    all_wall = max(w for _, w, _ in many)
    if model.startswith("ising"):
        what = ("config 1 on the reference's CPU path: oracle restatement of metropolis.rs:8-27 (random-site "
                "Metropolis), L=16, T=4.5115, hot start, 2000 warm-up + 2000 sample sweeps measuring E, |M|, M^2, "
                "M^4 every sweep (16.4M updates per run), %d runs (seeds 42..%d): %d updates in %.1f s on 1 core"
                % (detail["runs"], 41 + detail["runs"], upd, wall))
        unit = "updates/ns"
    elif model == "xy":
        what = ("oracle restatement of xy/wolff.rs:16-73 (Wolff, the reference's only XY CPU path), L=32, T=2.2, "
                "%d steps flipping %d spins in %.1f s on 1 core" % (detail["wolff_steps"], upd, wall))
        unit = "flipped spins/ns"
    else:
        what = ("oracle restatement of heisenberg/metropolis.rs:10-36 + overrelax.rs:15-48 (combined sweep with 2 "
                "OR passes), L=32, T=1.443, %d Metropolis updates in %.1f s on 1 core" % (upd, wall))
        unit = "updates/ns"
    out = {"value": upd / (wall * 1e9), "unit": unit, "cores": 1, "kind": "port", "sample": what,
           "cpu_model": cpu_model(),
           "all_cores": {"value": all_upd / (all_wall * 1e9), "cores": ncores,
                         "sample": "%d independent chains of the same kind, one per core, %d updates in %.1f s"
                                   % (ncores, all_upd, all_wall)}}
    if detail.get("first_run_means"):
        out["first_run_means"] = detail["first_run_means"]
    return out


def graph_length(sweeps, cap=200):
    """Largest divisor of `sweeps` not above `cap`: the timed sweeps are whole graph replays."""
    return max(g for g in range(1, min(cap, sweeps) + 1) if sweeps % g == 0)


# Parallel-tempering workloads time whole exchange intervals: one step is
# `exchange_every` sweeps (10 at every configuration here).
SWEEPS_PER_STEP = 10


def warmup_sweeps(warmup_steps, G):
    """Warm-up sweeps: the requested steps rounded up to whole replays of the
    timed graph of G sweeps, at least one replay."""
    return max(1, -(-warmup_steps * SWEEPS_PER_STEP // G)) * G


def beyond_mall(args, ising_hip, opts, layout):
    """int8 Ising PT with a working set above the 256 MB Infinity Cache: L=256 x 32
    replicas (512 MiB of spins), same ladder, cadence and thread-step layout;
    the main run's sweeps, between 200 and 600."""
    L, steps = 256, min(max(200, args.steps * SWEEPS_PER_STEP), 600)
    kw = dict(opts)
    if layout is not None:
        kw.update(ising_width=layout[0], ising_runs=layout[1], ising_spt=layout[2])
    pt = ising_hip.PTEnsemble("ising", [L], R, linspace(4.40, 4.62, R), args.seed,
                              record_capacity=steps // kw["measure_every"] + 8, **kw)
    G = graph_length(steps)
    pt.prepare_graph(G)
    pt.warmup(warmup_sweeps(args.warmup, G))
    pt.sample(0)
    pt.sync()
    t0 = time.perf_counter()
    pt.sample(steps)
    ev_ms = pt.sync()
    wall = time.perf_counter() - t0
    pool = pt.device_bytes()
    pt.close()
    sites = L ** 3 * R
    launch_s = ev_ms * 1e-3 / (2 * steps)
    achieved = 3 * sites / 2 / launch_s / 1e9
    return {"workload": "3D Ising PT, L=256 x 32 replicas, T=linspace(4.40,4.62,32), exchange every 10, measure "
                        "every 5, int8 spins", "pool_bytes": pool, "graph_sweeps": G,
            "sweeps": steps, "value": round(sites * steps / (wall * 1e9), 3), "unit": "updates/ns",
            "device_updates_per_ns": round(sites * steps / (ev_ms * 1e6), 3),
            "roofline": {"achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4), "alg_bytes_per_launch": 3 * sites // 2,
                         "avg_launch_us": round(launch_s * 1e6, 3)}}


def load_traffic():
    p = os.path.join(ROOT, "profiles", "ising_half_kernel_pt_L8-128_R32_traffic.json")
    if not os.path.exists(p):
        return None
    try:
        return json.load(open(p)).get("hbm_bytes_per_launch")
    except Exception:
        return None


def copy_calibration(nbytes):
    """Runs the HIP runtime's device-to-device copy kernel on `nbytes` (a plain
    copy of known size) so PMC byte counters can be scaled: FETCH_SIZE should
    read nbytes and WRITE_SIZE nbytes for one such dispatch."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_calibration_bytes": nbytes, "dispatches": 3}), flush=True)


def stream_calibration():
    """Achievable HBM/Infinity-Cache rate of the int8 colour pass's access
    pattern without its arithmetic: an in-place byte add (two streams read,
    one written, torch's elementwise kernel) and a device copy, at one colour
    of the config-2 ensemble (38,346,752 bytes per stream), of the L = 256 x 32
    ensemble (268,435,456) and twice that, timed with events over 200 passes."""
    import torch
    out = []
    for n in (38346752, 268435456, 536870912):
        a = torch.randint(0, 2, (n,), dtype=torch.uint8, device="cuda:0")
        b = torch.randint(0, 2, (n,), dtype=torch.uint8, device="cuda:0")
        c = torch.empty_like(a)
        row = {"bytes_per_stream": n}
        for name, fn, streams in (("add_read2_write1", lambda: a.add_(b), 3), ("copy", lambda: c.copy_(a), 2)):
            for _ in range(20):
                fn()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(200):
                fn()
            e1.record()
            torch.cuda.synchronize()
            us = e0.elapsed_time(e1) * 1e3 / 200
            row[name] = {"us": round(us, 2), "TB_s": round(streams * n / (us * 1e6), 3)}
        out.append(row)
        del a, b, c
    print(json.dumps({"stream_calibration": out}), flush=True)


WORKLOADS = {
    # name: (model, sizes, temps, bytes/update, dtype, options, description)
    "ising": ("ising", SIZES, (4.40, 4.62), 3, "u8", dict(exchange_every=10, measure_every=5),
              "3D Ising PT: L=8,16,32,64,128 x 32 replicas, T=linspace(4.40,4.62,32), exchange every 10, "
              "measure every 5, int8 spins, Philox4x32-10"),
    "ising_msc": ("ising_msc", SIZES, (4.40, 4.62), 0.375, "u32 (1 bit/spin)", dict(exchange_every=10, measure_every=5),
                  "3D Ising PT L=8..128 x 32 replicas, multispin coded (32 replicas per 32-bit word), exchange every 10, "
                  "measure every 5"),
    "heisenberg": ("heisenberg", [16, 32, 64], (1.40, 1.50), 36, "f32",
                   dict(exchange_every=10, measure_every=5, n_overrelax=2, delta=1.0, renorm_every=100),
                   "3D Heisenberg PT: L=16,32,64 x 32 replicas, T=linspace(1.40,1.50,32), cap Metropolis "
                   "+ 2 overrelaxation passes per sweep, FP64 sums"),
    "xy": ("xy", [16, 32, 64], (2.15, 2.25), 24, "f32",
           dict(exchange_every=10, measure_every=5, n_overrelax=2, delta=1.5, renorm_every=100),
           "3D XY PT: L=16,32,64 x 32 replicas, T=linspace(2.15,2.25,32), rotation Metropolis + 2 overrelaxation "
           "passes per sweep, FP64 sums"),
}


BATCH_SIZES = [8, 12, 16, 24, 32, 48, 64, 96, 128]
BATCH_MODELS = [
    # model, temperatures, bytes/update, options
    ("ising", (4.40, 4.62), 3, dict(exchange_every=10, measure_every=5)),
    ("xy", (2.15, 2.25), 24, dict(exchange_every=10, measure_every=5, n_overrelax=2, delta=1.5, renorm_every=100)),
    ("heisenberg", (1.40, 1.50), 36,
     dict(exchange_every=10, measure_every=5, n_overrelax=2, delta=1.0, renorm_every=100)),
]


def run_batched(args, ising_hip):
    """Ising + XY + Heisenberg, L = 8..128 (9 sizes) x 32 replicas each, advanced
    together by one captured graph per block of sweeps (a branch per model).
    One step = one exchange interval: 10 PT sweeps of all 3 x 109,041,664
    sites (one Metropolis pass per site and sweep; the continuous models add 2
    overrelaxation passes)."""
    sites = sum(L ** 3 for L in BATCH_SIZES) * R
    sweeps = args.steps * SWEEPS_PER_STEP
    members = []
    for model, (tlo, thi), _, opts in BATCH_MODELS:
        cap = sweeps // opts["measure_every"] + 8
        members.append(ising_hip.PTEnsemble(model, BATCH_SIZES, R, linspace(tlo, thi, R), args.seed,
                                            record_capacity=cap, graph_sweeps=0, **opts))
    G = graph_length(sweeps, cap=50)
    mode = {"branches": ising_hip.Batch.BRANCHES, "worklist": ising_hip.Batch.WORKLIST,
            "persistent": ising_hip.Batch.PERSISTENT}[args.batch_mode]
    batch = ising_hip.Batch(members, graph_sweeps=G, mode=mode)
    warm = max(args.warmup * SWEEPS_PER_STEP, G)  # the batch graph is captured by the first call of >= G sweeps
    batch.warmup(warm)
    batch.sample(0)
    batch.sync()
    t0 = time.perf_counter()
    batch.sample(sweeps)
    ev_ms = batch.sync()
    wall = time.perf_counter() - t0
    for p, (model, _, _, opts) in zip(members, BATCH_MODELS):
        recs = p.read_ising() if model == "ising" else p.read_cont()
        assert recs[0].shape[0] == sweeps // opts["measure_every"]
    updates = 3 * sites * sweeps
    passes = sum(1 + o.get("n_overrelax", 0) for _, _, _, o in BATCH_MODELS)
    alg_bytes_per_sweep = sum(bpu * sites * (1 + o.get("n_overrelax", 0)) for _, _, bpu, o in BATCH_MODELS)
    sweep_s = ev_ms * 1e-3 / sweeps
    achieved = alg_bytes_per_sweep / sweep_s / 1e9
    line = {
        "metric": "spin updates/ns",
        "value": round(updates / (wall * 1e9), 3),
        "unit": "updates/ns",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8 + f32",
        "data": "synthetic",
        "config": {"workload": "batched PT: Ising int8 T=4.40-4.62 + XY T=2.15-2.25 + Heisenberg "
                               "T=1.40-1.50, L=8,12,16,24,32,48,64,96,128 x 32 replicas each, exchange every 10, "
                               "measure every 5, continuous models 1 Metropolis + 2 overrelaxation passes",
                   "sizes": BATCH_SIZES, "replicas": R, "models": [m for m, _, _, _ in BATCH_MODELS],
                   "batch_mode": args.batch_mode,
                   "sweeps_per_step": SWEEPS_PER_STEP, "sweeps": sweeps, "warmup_sweeps": warm,
                   "updates_per_step": 3 * sites * SWEEPS_PER_STEP,
                   "site_passes_per_sweep": passes * sites, "graph_sweeps": G,
                   "combined_updates_per_ns": round(passes * sites * sweeps / (wall * 1e9), 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": {"branches": "whole sweep graph (each model's kernels, concurrent branches)",
                                "worklist": "whole sweep: batch_or_kernel x 4 + batch_pass_kernel x 4 (Ising "
                                            "colour passes; XY and Heisenberg Metropolis passes together) + "
                                            "batch_close_kernel, over work lists of the models' blocks",
                                "persistent": "whole sweep: batch_persistent_kernel x 8 + batch_close_kernel "
                                              "(the same work lists on persistent grids)"}[
                                    args.batch_mode],
                     "alg_bytes_per_sweep": int(alg_bytes_per_sweep), "avg_sweep_us": round(sweep_s * 1e6, 3),
                     "device_updates_per_ns": round(updates / (ev_ms * 1e6), 3)},
    }
    if args.cpu_seconds > 0:
        line["cpu_baseline"] = cpu_baseline(args.cpu_seconds, "ising")
    print(json.dumps(line), flush=True)


def temperature_series(tidx, x):
    """Per (size, temperature) time series from per-replica records: at each
    measured sweep, the value of the replica sitting at that temperature."""
    n, ng, R = tidx.shape
    order = np.argsort(tidx, axis=2)
    return np.take_along_axis(x, order, axis=2)  # [sweep, size, temperature]


def binder_jk(m2, nb=20):
    """U4 = 1 - <m^4> / (3 <m^2>^2) from a series of m^2, with a 20-block jackknife error."""
    m2 = np.asarray(m2, np.float64)
    bs = len(m2) // nb
    x = m2[: bs * nb].reshape(nb, bs)
    full = 1 - (x ** 2).mean() / (3 * x.mean() ** 2)
    jk = np.array([1 - (np.delete(x, i, 0) ** 2).mean() / (3 * np.delete(x, i, 0).mean() ** 2) for i in range(nb)])
    return full, float(np.sqrt((nb - 1) / nb * ((jk - full) ** 2).sum()))


def chain_stats(pt, model, sizes, Rr, temps, steps, ising_hip):
    """Samples `steps` sweeps (in chunks; XY helicity sums between chunks, outside
    the timed sampling) and returns per-(size, temperature) statistics."""
    chunk, wall, dev_ms, done = 10, 0.0, 0.0, 0
    hc = np.zeros((len(sizes), Rr))
    hj2 = np.zeros((len(sizes), Rr))
    nh = 0
    while done < steps:
        m = min(chunk, steps - done)
        t0 = time.perf_counter()
        pt.sample(m)
        dev_ms += pt.sync()
        wall += time.perf_counter() - t0
        done += m
        if model == "xy":
            hel = pt.helicity()
            for g in range(len(sizes)):
                r2t, _ = pt.get_perm(g)
                hg = hel[g * Rr:(g + 1) * Rr]
                hc[g, r2t] += hg[:, :3].mean(axis=1)
                hj2[g, r2t] += (hg[:, 3:] ** 2).mean(axis=1)
            nh += 1
    tidx, e, mx, my, mz = pt.read_cont()
    es = temperature_series(tidx, e)
    m2s = temperature_series(tidx, mx * mx + my * my + mz * mz)
    st = {"tau_E": [], "tau_M2": [], "E": [], "E_err": [], "U4": [], "U4_err": []}
    for g, L in enumerate(sizes):
        n3 = float(L ** 3)
        for k in range(Rr):
            ek, m2k = es[:, g, k] / n3, m2s[:, g, k] / n3 ** 2
            st["tau_E"].append(ising_hip.tau_int(ek)[0])
            st["tau_M2"].append(ising_hip.tau_int(m2k)[0])
            jk = ising_hip.jackknife_observables(ek, np.sqrt(m2k), 1.0 / temps[k], n3, 20)
            st["E"].append(jk[0])
            st["E_err"].append(jk[1])
            u, ue = binder_jk(m2k)
            st["U4"].append(u)
            st["U4_err"].append(ue)
    st = {k: np.array(v) for k, v in st.items()}
    st["wall"] = wall
    st["device_us_per_sweep"] = dev_ms * 1e3 / steps
    if model == "xy":
        st["helicity"] = [float(((hc[g] / nh - np.array([1.0 / t for t in temps]) * hj2[g] / nh) / L ** 3).mean())
                          for g, L in enumerate(sizes)]
    return st


def summarize(st, steps, base=None):
    """Independent samples/s of E and M^2 (steps / (2 tau) per (size, temperature),
    summed, over the wall time) and, against the plain chain, the cost ratio at
    equal error: (err_plain^2 t_plain) / (err^2 t), geometric mean over
    (size, temperature), for E and U4."""
    out = {"sweeps_per_s": round(steps / st["wall"], 2),
           "device_us_per_sweep": round(st["device_us_per_sweep"], 2),
           "tau_int_E_mean": round(float(st["tau_E"].mean()), 3),
           "tau_int_M2_mean": round(float(st["tau_M2"].mean()), 3),
           "independent_samples_per_s_E": round(float(np.sum(steps / (2 * np.maximum(st["tau_E"], 0.5))) / st["wall"]), 1),
           "independent_samples_per_s_M2": round(float(np.sum(steps / (2 * np.maximum(st["tau_M2"], 0.5))) / st["wall"]), 1),
           "E_mean": round(float(st["E"].mean()), 6), "E_err_mean": float("%.3g" % st["E_err"].mean()),
           "U4_mean": round(float(st["U4"].mean()), 5), "U4_err_mean": float("%.3g" % st["U4_err"].mean())}
    if "helicity" in st:
        out["helicity_modulus_mean"] = [round(v, 5) for v in st["helicity"]]
    if base is not None:
        for q in ("E", "U4"):
            r = (base[q + "_err"] ** 2 * base["wall"]) / (st[q + "_err"] ** 2 * st["wall"])
            out["equal_error_speedup_" + q] = round(float(np.exp(np.mean(np.log(r)))), 4)
        # agreement with the plain chain, per (size, temperature), in combined errors
        for q in ("E", "U4"):
            z = (st[q] - base[q]) / np.hypot(st[q + "_err"], base[q + "_err"])
            out["z_vs_plain_" + q] = {"max_abs": round(float(np.abs(z).max()), 2),
                                      "frac_within_2": round(float(np.mean(np.abs(z) < 2)), 3)}
    return out


LEARNED_MODELS = [
    # model, sizes, temperatures, Metropolis step (radians)
    ("xy", [32, 64], (2.18, 2.22), 1.0),
    ("heisenberg", [32], (1.43, 1.46), 1.0),
]


def learned_compare(steps, warmup, seed, hidden, train_steps, learned_or, cluster=True):
    """Learned proposal against plain Metropolis + overrelaxation, and the
    heat-bath ceiling: XY L = 32, 64 and Heisenberg L = 32, 32 replicas each.
    Chains: plain (Metropolis + 2 overrelaxation passes per sweep), learned
    (the trained network's Metropolis-Hastings pass + `learned_or`
    overrelaxation passes), heat bath (the exact single-site conditional + 2
    overrelaxation passes: the best any proposal that depends on the
    neighbours only can do per pass), and with `cluster` for XY embedded
    Swendsen-Wang.  Same ensembles, temperatures and seed; measured every
    sweep; `warmup` sweeps, then `steps` sampled sweeps per chain."""
    import ising_hip
    res = {}
    for model, sizes, (tlo, thi), dlt in LEARNED_MODELS:
        temps = linspace(tlo, thi, R)
        kw = dict(exchange_every=10, measure_every=1, n_overrelax=2, delta=dlt, adapt_every=10, adapt_target=0.5,
                  renorm_every=100, record_capacity=steps + 8)
        out, stats = {}, {}
        for mode in ("plain", "learned", "heat_bath") + (("cluster",) if model == "xy" and cluster else ()):
            nor = learned_or if mode == "learned" else 2
            pt = ising_hip.PTEnsemble(model, sizes, R, temps, seed, **dict(kw, n_overrelax=nor))
            pt.warmup(warmup)
            if mode == "cluster":
                out[mode] = xy_cluster_leg(pt, sizes, R, temps, steps, ising_hip)
                pt.close()
                continue
            train = None
            if mode == "learned":
                pt.learn_init(hidden, seed=7)
                t0 = time.perf_counter()
                loss = pt.learn_train(train_steps, batch=8192, lr=3e-3)
                pt.sync()
                train = {"steps": train_steps, "batch": 8192, "seconds": round(time.perf_counter() - t0, 3),
                         "final_loss": loss}
                pt.learn_enable(True)
                pt.warmup(50)
            elif mode == "heat_bath":
                pt.heat_bath(True)
                pt.warmup(50)
            pt.sync()
            pass_us = None
            if mode in ("learned", "heat_bath"):
                # one pass over one colour (no measurement), timed back to back
                t_now = pt.sweeps()
                pt.sync()
                t0 = time.perf_counter()
                for k in range(200):
                    pt.metropolis_pass(k & 1, t_now + k // 2)
                pt.sync()
                pass_us = (time.perf_counter() - t0) * 1e6 / 200
            stats[mode] = chain_stats(pt, model, sizes, R, temps, steps, ising_hip)
            out[mode] = summarize(stats[mode], steps, None if mode == "plain" else stats["plain"])
            if pass_us is not None:
                out[mode]["pass_us"] = round(pass_us, 2)
            if mode == "learned":
                tries, acc = pt.mh_stats()
                out[mode]["learned_acceptance"] = round(acc / tries, 4) if tries else None
                out[mode]["training"] = train
            pt.close()
        for mode in ("learned", "heat_bath"):
            for q in ("E", "M2"):
                out[mode + "_over_plain_samples_" + q] = round(out[mode]["independent_samples_per_s_" + q]
                                                              / out["plain"]["independent_samples_per_s_" + q], 4)
        out["sizes"], out["temperatures"] = sizes, [tlo, thi]
        res[model] = out
    return res


def learned_digest(res, steps, warmup, hidden, train_steps):
    """The learned-vs-plain comparison in brief, for the headline line: per
    model, the ratios of independent samples/s (E, M^2) and the equal-error
    speed-ups (E, U4) of the learned and heat-bath chains over the plain one."""
    out = {"chains": "plain Metropolis + 2 overrelaxation passes vs learned MLP proposal (exact Metropolis-"
                     "Hastings) vs exact heat bath, each + 2 overrelaxation passes; 32 replicas, measured every "
                     "sweep, exchange every 10",
           "sweeps": steps, "warmup_sweeps": warmup, "hidden": hidden, "train_steps": train_steps}
    for model, r in res.items():
        d = {"sizes": r["sizes"], "temperatures": r["temperatures"]}
        for mode in ("learned", "heat_bath"):
            for q in ("E", "M2"):
                d[mode + "_over_plain_samples_" + q] = r[mode + "_over_plain_samples_" + q]
            for q in ("E", "U4"):
                d[mode + "_equal_error_speedup_" + q] = r[mode]["equal_error_speedup_" + q]
                d[mode + "_z_vs_plain_max_" + q] = r[mode]["z_vs_plain_" + q]["max_abs"]
        d["U4_err_mean"] = {m: r[m]["U4_err_mean"] for m in ("plain", "learned", "heat_bath")}
        d["learned_acceptance"] = r["learned"]["learned_acceptance"]
        out[model] = d
    return out


def run_learned(args, ising_hip):
    """The learned-proposal workload (see learned_compare) as a bench line whose
    value is the learned chains' independent samples/s of E."""
    steps = args.steps
    res = learned_compare(steps, args.warmup, args.seed, args.hidden, args.train_steps, args.learned_or)
    value = sum(res[m]["learned"]["independent_samples_per_s_E"] for m in res)
    # the learned pass (cont_learned_kernel, one colour, no measurement) against
    # the HBM roofline at the plain pass's bytes per update, and its network's
    # matrix-core work: 6 -> H -> H -> H -> 3, 9H + 2H^2 multiply-adds per site
    bpu = {"xy": 24, "heisenberg": 36}
    per_model = {}
    for m, sizes, _, _ in LEARNED_MODELS:
        sites = sum(L ** 3 for L in sizes) * R // 2
        us = res[m]["learned"]["pass_us"]
        per_model[m] = {"sites_per_pass": sites, "pass_us": us,
                        "achieved_GBs": round(bpu[m] * sites / (us * 1e3), 1),
                        "frac": round(bpu[m] * sites / (us * 1e3) / HBM_PEAK_GBS, 4),
                        "mlp_tflops": round(2 * (9 * args.hidden + 2 * args.hidden ** 2) * sites / (us * 1e6), 2)}
    line = {
        "metric": "independent samples/s (learned proposal)",
        "value": round(value, 1),
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 / res["xy"]["learned"]["sweeps_per_s"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (MLP bf16 weights, f32 accumulate)",
        "data": "synthetic",
        "config": {"workload": "learned proposal vs plain Metropolis + 2 overrelaxation passes and the heat-bath "
                               "ceiling: XY L=32,64 x 32 replicas T=linspace(2.18,2.22,32); Heisenberg L=32 x 32 "
                               "T=linspace(1.43,1.46,32); measured every sweep, exchange every 10; value = learned "
                               "independent samples/s of E summed over both models",
                   "hidden": args.hidden, "network_parameters": 6 * args.hidden + args.hidden
                   + 2 * (args.hidden * args.hidden + args.hidden) + 3 * args.hidden + 3,
                   "learned_overrelaxation_passes": args.learned_or, **res},
        "roofline": {"bound": "hbm", "achieved": per_model["xy"]["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": per_model["xy"]["frac"], "traffic": None,
                     "kernel": "cont_learned_kernel (XY learned pass, one colour, 24 B per update)",
                     "per_model": per_model, "bf16_mfma_peak_tflops": 2500.0},
    }
    print(json.dumps(line), flush=True)


def xy_cluster_leg(pt, sizes, Rr, temps, steps, ising_hip):
    """Embedded Swendsen-Wang updates (Wolff's reflection along a random axis,
    every cluster flipped with probability 1/2) of the same XY ensemble: one
    update per step, E and M after each; the statistical-efficiency comparator
    of the learned proposal."""
    pt.cluster("sw", 100)
    chunk, wall, done = 10, 0.0, 0
    E = np.zeros((steps, len(sizes), Rr))
    M2 = np.zeros_like(E)
    hc = np.zeros((len(sizes), Rr))
    hj2 = np.zeros((len(sizes), Rr))
    nh = 0
    while done < steps:
        m = min(chunk, steps - done)
        t0 = time.perf_counter()
        _, meas = pt.cluster("sw", m, measure=True)
        wall += time.perf_counter() - t0
        hel = pt.helicity()
        for g in range(len(sizes)):
            r2t, _ = pt.get_perm(g)
            sl = slice(g * Rr, (g + 1) * Rr)
            E[done:done + m, g, r2t] = meas[:, sl, 0]
            M2[done:done + m, g, r2t] = meas[:, sl, 1] ** 2 + meas[:, sl, 2] ** 2
            hg = hel[sl]
            hc[g, r2t] += hg[:, :3].mean(axis=1)
            hj2[g, r2t] += (hg[:, 3:] ** 2).mean(axis=1)
        nh += 1
        done += m
    ups = [(hc[g] / nh - np.array([1.0 / t for t in temps]) * hj2[g] / nh) / L ** 3 for g, L in enumerate(sizes)]
    taus_e = np.array([ising_hip.tau_int(E[:, g, k])[0] for g in range(len(sizes)) for k in range(Rr)])
    taus_m = np.array([ising_hip.tau_int(M2[:, g, k])[0] for g in range(len(sizes)) for k in range(Rr)])
    isps = float(np.sum(steps / (2.0 * np.maximum(taus_e, 0.5))) / wall)
    u4 = [float(np.mean(1 - (M2[:, g, :] ** 2).mean(axis=0) / (3 * M2[:, g, :].mean(axis=0) ** 2)))
          for g in range(len(sizes))]
    return {"updates_per_s": round(steps / wall, 2), "tau_int_E_mean": round(float(taus_e.mean()), 3),
            "tau_int_M2_mean": round(float(taus_m.mean()), 3), "independent_samples_per_s": round(isps, 1),
            "helicity_modulus_mean": [round(float(u.mean()), 5) for u in ups],
            "binder_u4_mean": [round(v, 5) for v in u4],
            "note": "one embedded Swendsen-Wang update per step, measured after each (no Metropolis / overrelaxation)"}


def run_cluster(args, ising_hip):
    """Cluster updates (Swendsen-Wang, Wolff) against checkerboard Metropolis
    near T_c, 3D Ising.  (a) Throughput: Swendsen-Wang updates of the L=8..128 x 32
    ensemble (L = 8..128 x 32 replicas, T = linspace(4.40, 4.62)), in sites
    per ns.  (b) Efficiency at T_c = 4.5115, L = 32 and 64 x 32 independent
    replicas: independent samples/s of |M| = steps/s / (2 tau_int(|M|)) summed
    over replicas, tau_int by Sokal's window (c = 6)."""
    temps = linspace(4.40, 4.62, R)
    pt = ising_hip.PTEnsemble("ising", SIZES, R, temps, args.seed)
    pt.cluster("sw", 3)
    n_thr = 200
    t0 = time.perf_counter()
    pt.cluster("sw", n_thr)
    wall_thr = time.perf_counter() - t0
    sites = sum(L ** 3 for L in SIZES) * R
    pt.close()
    thr = {"sizes": SIZES, "replicas": R, "sw_updates": n_thr, "sites_per_ns": round(sites * n_thr / (wall_thr * 1e9), 3),
           "ms_per_update": round(wall_thr * 1e3 / n_thr, 4)}
    if args.cluster_part == "throughput":
        print(json.dumps({"metric": "sites/ns (Swendsen-Wang updates)", "value": thr["sites_per_ns"],
                          "unit": "sites/ns", "n_gpus": 1, "config": thr}), flush=True)
        return
    counts = {"metropolis": 40000, "sw": 4000, "wolff": 20000}
    eff = {}
    for L in (32, 64):
        res = {}
        for mode, n in counts.items():
            pt = ising_hip.PTEnsemble("ising", [L], R, [4.5115] * R, args.seed, exchange_every=1 << 30,
                                      measure_every=1, record_capacity=n + 8)
            if mode == "metropolis":
                pt.warmup(2000)
                pt.sync()
                t0 = time.perf_counter()
                pt.sample(n)
                pt.sync()
                wall = time.perf_counter() - t0
                _, _, mag = pt.read_ising()
                series = np.abs(mag[:, 0, :]) / L ** 3
                flipped = None
            else:
                pt.cluster(mode, 2000 if mode == "wolff" else 300)
                t0 = time.perf_counter()
                flipped, _, mag = pt.cluster(mode, n, measure=True)
                wall = time.perf_counter() - t0
                series = np.abs(mag) / L ** 3
            taus = np.array([ising_hip.tau_int(series[:, r])[0] for r in range(R)])
            isps = float(np.sum(n / (2.0 * np.maximum(taus, 0.5))) / wall)
            res[mode] = {"steps": n, "steps_per_s": round(n / wall, 1), "tau_int_absM_mean": round(float(taus.mean()), 3),
                         "independent_samples_per_s": round(isps, 1), "absM_mean": round(float(series.mean()), 5)}
            if flipped is not None:
                res[mode]["mean_sites_flipped"] = round(float(flipped.mean()), 1)
            pt.close()
        for mode in ("sw", "wolff"):
            res[mode + "_over_metropolis"] = round(res[mode]["independent_samples_per_s"]
                                                   / res["metropolis"]["independent_samples_per_s"], 3)
        eff["L%d" % L] = res
    value = eff["L64"]["sw"]["independent_samples_per_s"]
    line = {
        "metric": "independent samples/s (Swendsen-Wang, |M|, 3D Ising L=64 x 32 replicas at T_c)",
        "value": value, "unit": "samples/s", "n_gpus": 1, "steps": counts["sw"], "warmup": 300,
        "ms_per_step": round(1e3 / eff["L64"]["sw"]["steps_per_s"], 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": "cluster updates vs checkerboard Metropolis near T_c: 3D Ising, 32 independent "
                               "replicas at T=4.5115, L=32 and 64; throughput on the L=8..128 x 32 Ising ensemble",
                   "throughput": thr,
                   "efficiency": eff},
    }
    print(json.dumps(line), flush=True)


def run_driver(args, ising_hip):
    """The other LatticeGpu callers as batched drivers (int8 Ising, one launch
    per colour for every trial / size): Kibble-Zurek sweep, coarsening and
    annealed FSS at production-like sizes.  Metric: spin updates/ns over the
    whole driver call (host-side schedule uploads and reductions included)."""
    w = args.workload
    t0 = time.perf_counter()
    if w == "kz":
        n, trials, taus, warm = 64, 64, [250, 500, 1000], 200
        pts = ising_hip.run_kz_sweep(n, 1.0, 6.0, 1.0, taus, trials, warm, 0, 0.01, base_seed=args.seed)
        updates = sum(n ** 3 * trials * (warm + t) for t in taus)
        desc = "Kibble-Zurek sweep: L=64, 64 trials per tau_q, tau_q = 250, 500, 1000, 200 warm-up sweeps at T=6"
        extra = {"rho": [round(p[1], 6) for p in pts]}
    elif w == "coarsening":
        n, trials, total, warm = 128, 16, 2000, 200
        steps, rho = ising_hip.run_coarsening(n, 1.0, 10.0, 0.5, warm, total, 100, args.seed, trials)
        updates = n ** 3 * trials * (warm + total)
        desc = "coarsening: L=128, 16 trials, 200 sweeps at T=10, quench to T=0.5, 2000 steps, sample every 100"
        extra = {"rho_last": round(float(rho[-1].mean()), 6)}
    else:
        sizes, tsteps, warm, samp = [16, 32, 64, 128], 8, 200, 200
        res = ising_hip.run_fss(sizes, 1.0, 0.0, 4.3, 4.7, tsteps, warm, samp, args.seed)
        updates = sum(L ** 3 for L in sizes) * tsteps * (warm + samp)
        desc = "annealed FSS: L=16,32,64,128 together, 8 temperatures 4.7 -> 4.3, 200 warm-up + 200 measured sweeps each"
        extra = {"binder_L128": [round(1 - o.m4 / (3 * o.m2 * o.m2), 4) for o in res[-1][1]]}
    wall = time.perf_counter() - t0
    line = {"metric": "spin updates/ns", "value": round(updates / (wall * 1e9), 3), "unit": "updates/ns",
            "n_gpus": 1, "steps": None, "warmup": None, "ms_per_step": None, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
            "config": dict({"workload": desc, "updates": updates, "seconds": round(wall, 3)}, **extra)}
    if args.cpu_seconds > 0:
        line["cpu_baseline"] = cpu_baseline(args.cpu_seconds, "ising")
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps: exchange intervals of 10 sweeps for the PT workloads (default 300), "
                         "sweeps for the learned, cluster and driver workloads (default 3000)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed warm-up steps, same unit (default 30, or 300 sweeps)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--workload", default="ising", choices=sorted(WORKLOADS) + ["batched", "learned", "kz", "coarsening", "fss",
                                                                  "cluster"])
    ap.add_argument("--hidden", type=int, default=16,
                    help="learned proposal: MLP hidden width (8, 16, 32 or 64; 16 = 707 parameters, 64 = 8963; "
                         "16 gives the most independent samples/s: the wider networks reach the same acceptance "
                         "at a costlier pass)")
    ap.add_argument("--train-steps", type=int, default=400, help="learned workload: Adam steps")
    ap.add_argument("--learned-or", type=int, default=2,
                    help="learned workload: overrelaxation passes per sweep of the learned chain (plain keeps 2)")
    ap.add_argument("--fused", type=int, default=0, choices=[0, 8, 16, 32],
                    help="int8 Ising: one launch per sweep, blocks streaming this many z planes (0: one launch per colour)")
    ap.add_argument("--cluster-part", default="all", choices=["all", "throughput"],
                    help="cluster workload: 'throughput' times only the Swendsen-Wang updates of the L=8..128 x 32 ensemble")
    ap.add_argument("--layout", default="auto", metavar="W,R,S",
                    help="int8 thread-step layout (run width, runs per step, steps per tile); "
                         "'auto' times all layouts during warm-up and keeps the fastest")
    ap.add_argument("--sizes", default=None, metavar="L,L,...",
                    help="single-model workloads: lattice sizes in place of the default ladder (profiling runs)")
    ap.add_argument("--batch-mode", default="branches", choices=["branches", "worklist", "persistent"],
                    help="batched workload: each member's own launches on a graph branch per member, one "
                         "work-list launch per pass for all members, or that work list on a persistent grid")
    ap.add_argument("--no-beyond-mall", action="store_true",
                    help="ising workload: skip the L=256 x 32 sub-result (working set above the Infinity Cache)")
    ap.add_argument("--no-learned", action="store_true",
                    help="ising workload: skip the learned-vs-plain sub-result (XY L=32,64 and Heisenberg L=32 "
                         "x 32 replicas, 1000 + 4000 sweeps per chain)")
    ap.add_argument("--stream-calibration", action="store_true",
                    help="only time a read-two / write-one byte stream and a copy at the int8 pass's sizes")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (PMC byte-counter calibration)")
    args = ap.parse_args()
    pt_workload = args.workload in WORKLOADS or args.workload == "batched"
    if args.steps is None:
        args.steps = 300 if pt_workload else 3000
    if args.warmup is None:
        args.warmup = 30 if pt_workload else 300
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.stream_calibration:
        stream_calibration()
        return
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: the library has no multi-GPU path (an ensemble would be "
              "sharded by replica across GPUs, with the exchange energies gathered every exchange round)",
              file=sys.stderr)
        sys.exit(2)

    import ising_hip
    if ising_hip.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)

    if args.workload == "batched":
        run_batched(args, ising_hip)
        return
    if args.workload == "learned":
        run_learned(args, ising_hip)
        return
    if args.workload in ("kz", "coarsening", "fss"):
        run_driver(args, ising_hip)
        return
    if args.workload == "cluster":
        run_cluster(args, ising_hip)
        return
    model, sizes, (tlo, thi), bpu, dtype, opts, desc = WORKLOADS[args.workload]
    if args.sizes:
        sizes = [int(v) for v in args.sizes.split(",")]
        desc = (desc.replace("L=8,16,32,64,128", "L=" + args.sizes).replace("L=8..128", "L=" + args.sizes)
                .replace("L=16,32,64", "L=" + args.sizes))
    temps = linspace(tlo, thi, R)
    me = opts["measure_every"]
    if model == "ising" and args.layout != "auto":
        w, r, s = (int(v) for v in args.layout.split(","))
        opts = dict(opts, ising_width=w, ising_runs=r, ising_spt=s)
    if model == "ising" and args.fused:
        opts = dict(opts, ising_fused=args.fused)
    sweeps = args.steps * SWEEPS_PER_STEP
    assert opts["exchange_every"] == SWEEPS_PER_STEP
    pt = ising_hip.PTEnsemble(model, sizes, R, temps, args.seed, record_capacity=sweeps // me + 8, **opts)
    layout = None
    if model == "ising":
        if args.layout == "auto":
            layout = pt.autotune(20)  # warm-up sweeps that also pick the thread-step layout of every size
        else:
            layout = tuple(int(v) for v in args.layout.split(","))
    elif model in ("xy", "heisenberg"):
        layout = pt.autotune(20)  # the overrelaxation pass's sites per thread
    # the timed sweeps replay one captured graph of G sweeps (production runs
    # replay graphs too), captured here; the warm-up sweeps are whole replays of
    # the same graph, at least one, since the first replay of a newly
    # instantiated graph runs ~1% slower than the later ones
    G = graph_length(sweeps)
    warm_sweeps = warmup_sweeps(args.warmup, G)
    pt.prepare_graph(G)
    pt.warmup(warm_sweeps)
    pt.sample(0)  # enters the sampling phase (a control-block round trip), outside the timed region
    pt.sync()

    t0 = time.perf_counter()
    pt.sample(sweeps)
    ev_ms = pt.sync()
    wall = time.perf_counter() - t0

    recs = pt.read_ising() if model.startswith("ising") else pt.read_cont()
    assert recs[0].shape[0] == sweeps // me
    per_size = [pt.get_layout(g) for g in range(len(sizes))] if model == "ising" else None
    pt.close()

    sites = sum(L ** 3 for L in sizes) * R
    passes = 1 + opts.get("n_overrelax", 0) if not model.startswith("ising") else 1
    updates = sites * sweeps                       # Metropolis updates (the reported metric)
    value = updates / (wall * 1e9)
    fused = model == "ising" and args.fused > 0
    cont = not model.startswith("ising")
    # colour passes (or whole sweeps when fused); a continuous sweep adds one closing launch
    launches = ((1 if fused else 2) * passes + (1 if cont else 0)) * sweeps
    avg_launch_s = ev_ms * 1e-3 / launches
    if cont:
        # whole sweep: every pass moves bpu bytes per update; the closing launch moves ~nothing
        alg_bytes_per_launch = None
        alg_bytes_per_sweep = bpu * sites * passes
        achieved = alg_bytes_per_sweep / (ev_ms * 1e-3 / sweeps) / 1e9
    else:
        alg_bytes_per_launch = int(bpu * sites / (1 if fused else 2))  # bytes per update x updates per launch
        achieved = alg_bytes_per_launch / avg_launch_s / 1e9
    kernel = {"ising": "ising_fused_kernel" if fused else "ising_half_kernel", "ising_msc": "msc_half_kernel",
              "xy": "whole sweep: cont_or_kernel<2> x 4 + cont_metro_pair_kernel<2> x 2 + cont_close_kernel",
              "heisenberg": "whole sweep: cont_or_kernel<3> x 4 + cont_metro_kernel<3> x 2 + cont_close_kernel"}[model]
    line = {
        "metric": "spin updates/ns",
        "value": round(value, 3),
        "unit": "updates/ns",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": desc, "sizes": sizes, "replicas": R,
                   "step": "one exchange interval: %d sweeps of every lattice" % SWEEPS_PER_STEP,
                   "sweeps_per_step": SWEEPS_PER_STEP, "sweeps": sweeps, "warmup_sweeps": warm_sweeps,
                   "updates_per_step": sites * SWEEPS_PER_STEP, "site_passes_per_sweep": passes},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": load_traffic() if args.workload == "ising" and not fused and not args.sizes else None,
                     "kernel": kernel, "alg_bytes_per_launch": alg_bytes_per_launch,
                     "avg_launch_us": round(avg_launch_s * 1e6, 3), "launches": launches,
                     "device_updates_per_ns": round(updates / (ev_ms * 1e6), 3)},
    }
    line["config"]["graph_sweeps"] = G
    if cont:
        line["roofline"]["alg_bytes_per_update"] = bpu
        line["roofline"]["avg_sweep_us"] = round(ev_ms * 1e3 / sweeps, 3)
        if model == "xy":
            # stored as two FP32 components (24 B per update); one FP32 angle would be 12 B
            line["roofline"]["frac_at_12B_angle"] = round(achieved / 2 / HBM_PEAK_GBS, 4)
    if args.workload == "ising" and not fused and not args.no_beyond_mall and not args.sizes:
        line["config"]["beyond_mall"] = beyond_mall(args, ising_hip, opts, layout)
    if args.workload == "ising" and not fused and not args.no_learned and not args.sizes:
        lsteps, lwarm = 4000, 1000
        line["config"]["learned_vs_plain"] = learned_digest(
            learned_compare(lsteps, lwarm, args.seed, args.hidden, args.train_steps, args.learned_or, cluster=False),
            lsteps, lwarm, args.hidden, args.train_steps)
    if layout is not None and model == "ising":
        line["config"]["layout"] = {"run_width": layout[0], "runs_per_step": layout[1], "steps_per_tile": layout[2],
                                    "per_size": {L: list(lay) for L, lay in zip(sizes, per_size)}}
    elif layout is not None:
        line["config"]["overrelaxation_sites_per_thread"] = layout[0]
    if passes > 1:
        line["config"]["or_updates_per_ns"] = round(updates * (passes - 1) / (wall * 1e9), 3)
        line["config"]["combined_updates_per_ns"] = round(updates * passes / (wall * 1e9), 3)
    if args.cpu_seconds > 0:
        line["cpu_baseline"] = cpu_baseline(args.cpu_seconds, model)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
