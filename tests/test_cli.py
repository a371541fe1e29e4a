"""gpu_fss command line: flags, defaults and validation messages of
src/bin/gpu_fss.rs:20-110 and src/cli.rs:10-107 (CPU), and the output files of
the reference's GPU smoke test (tests/cli.rs:677-729), byte-compared against
files rebuilt from the oracle (GPU)."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as o

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GPU_FSS = os.path.join(ROOT, "ising-rs_amd", "bin", "gpu_fss")


def run(args, **kw):
    return subprocess.run([GPU_FSS] + args, capture_output=True, text=True, timeout=600, **kw)


@pytest.mark.parametrize("args,msg", [
    (["--tmin", "5", "--tmax", "4"], "Error: --tmin (5) must be less than --tmax (4)"),
    (["--tmin", "-1"], "Error: --tmin must be positive (got -1)"),
    (["--tmax", "0"], "Error: --tmax must be positive (got 0)"),
    (["--samples", "0"], "Error: --samples must be at least 1"),
    (["--warmup", "0"], "Error: --warmup must be at least 1"),
    (["--measure-every", "0"], "Error: --measure-every must be at least 1"),
    (["--exchange-every", "0"], "Error: --exchange-every must be at least 1"),
    (["--samples", "abc"], "Error: invalid value 'abc' for --samples: invalid digit found in string"),
    (["--samples", ""], "Error: invalid value '' for --samples: cannot parse integer from empty string"),
    (["--tmin", "x"], "Error: invalid value 'x' for --tmin: invalid float literal"),
    (["--seed"], "Error: --seed requires a value"),
])
def test_validation_messages(args, msg, tmp_path):
    r = run(["--outdir", str(tmp_path)] + args)
    assert r.returncode == 1
    assert msg in r.stderr


def test_unknown_model_and_unknown_flags(tmp_path):
    r = run(["--bogus", "--model", "potts", "--outdir", str(tmp_path)])
    assert r.returncode == 1
    assert "GPU FSS: model=potts, sizes=[8, 16, 32, 64], T=4.4..4.6, replicas=20" in r.stderr
    assert "Error: --model must be ising, xy, or heisenberg" in r.stderr


@pytest.mark.parametrize("args,msg", [
    (["--algorithm", "annealing"], "Error: --algorithm must be metropolis, msc, sw, or wolff"),
    (["--algorithm", "msc", "--model", "xy"], "Error: --algorithm msc needs --model ising"),
])
def test_algorithm_flag_validation(args, msg, tmp_path):
    r = run(["--outdir", str(tmp_path)] + args)
    assert r.returncode == 1 and msg in r.stderr


def test_invalid_sizes_are_dropped_like_the_reference(tmp_path):
    # filter_map(parse().ok()) drops "x" and " " entries; an empty list runs nothing
    r = run(["--sizes", "x,,y", "--outdir", str(tmp_path)])
    assert r.returncode == 0 and "sizes=[]" in r.stderr


def expected_ising_csvs(L, R, tmin, tmax, warmup, samples, ee, me, seed, msc=False):
    temps = o.linspace(tmin, tmax, R)
    run_pt = o.msc_pt_run if msc else o.ising_pt_run
    tidx, bond, mag, _ = run_pt(L, R, temps, seed, warmup, samples, ee, me)
    n3 = float(L ** 3)
    ts = ["temp_idx,E,M"]
    series = {k: ([], []) for k in range(R)}
    for t, b, m in zip(tidx, bond, mag):
        e_per = -1.0 * float(b) / n3
        m_per = abs(float(m)) / n3
        ts.append("%d,%.8f,%.8f" % (t, e_per, m_per))
        series[int(t)][0].append(e_per)
        series[int(t)][1].append(m_per)
    summ = ["T,E,E_err,M,M_err,M2,M2_err,M4,M4_err,Cv,Cv_err,chi,chi_err"]
    for k in range(R):
        e, m = series[k]
        if not e:
            continue
        obs = o.jackknife(np.array(e), np.array(m), 1.0 / temps[k], n3, min(20, len(e)))
        summ.append(("%.6f," % temps[k]) + ",".join("%.6f" % v for v in obs))
    return "\n".join(ts) + "\n", "\n".join(summ) + "\n"


@pytest.mark.gpu
@pytest.mark.parametrize("batched", ["1", "0"])
def test_reference_smoke_invocation_writes_identical_files(tmp_path, batched):
    # tests/cli.rs:684-711: --model ising --sizes 4 --tmin 4.4 --tmax 4.6 --replicas 4 --warmup 50 --samples 100
    r = run(["--model", "ising", "--sizes", "4,6", "--tmin", "4.4", "--tmax", "4.6", "--replicas", "4", "--warmup", "50",
             "--samples", "100", "--exchange-every", "10", "--outdir", str(tmp_path), "--batched", batched])
    assert r.returncode == 0, r.stderr
    for L in (4, 6):
        ts_ref, sum_ref = expected_ising_csvs(L, 4, 4.4, 4.6, 50, 100, 10, 1, 42)
        ts = open(tmp_path / ("gpu_fss_ising_N%d_timeseries.csv" % L)).read()
        summ = open(tmp_path / ("gpu_fss_ising_N%d_summary.csv" % L)).read()
        assert ts == ts_ref
        assert summ == sum_ref
        assert summ.startswith("T,E,") and len(summ.splitlines()) == 5


@pytest.mark.gpu
@pytest.mark.parametrize("model", ["xy", "heisenberg"])
def test_continuous_smoke_outputs(tmp_path, model):
    # scripts/run_gpu_windows_pipeline.py:71-108: L=4, --overrelax 2; CSVs have data and no NaN
    r = run(["--model", model, "--sizes", "4", "--tmin", "1.4", "--tmax", "2.2", "--replicas", "4", "--warmup", "50",
             "--samples", "100", "--overrelax", "2", "--measure-every", "5", "--outdir", str(tmp_path)])
    assert r.returncode == 0, r.stderr
    ts = open(tmp_path / ("gpu_fss_%s_N4_timeseries.csv" % model)).read().splitlines()
    summ = open(tmp_path / ("gpu_fss_%s_N4_summary.csv" % model)).read().splitlines()
    assert ts[0] == "temp_idx,E,M" and len(ts) == 1 + 4 * 20
    assert summ[0].startswith("T,E,E_err") and len(summ) == 5
    assert "nan" not in "".join(ts + summ).lower()


@pytest.mark.gpu
@pytest.mark.parametrize("model", ["ising", "xy"])
def test_binary_timeseries_holds_the_exact_records(tmp_path, model):
    """--binary 1 writes the unrounded records; rounding them as the CSV does
    reproduces the CSV rows, and the Ising records equal the oracle's."""
    import ising_hip
    args = ["--model", model, "--sizes", "4,6", "--tmin", "4.4" if model == "ising" else "1.4", "--tmax",
            "4.6" if model == "ising" else "2.2", "--replicas", "4", "--warmup", "50", "--samples", "60",
            "--exchange-every", "10", "--outdir", str(tmp_path), "--binary", "1"]
    if model != "ising":
        args += ["--overrelax", "2", "--measure-every", "3"]
    r = run(args)
    assert r.returncode == 0, r.stderr
    for L in (4, 6):
        ts = ising_hip.read_timeseries(str(tmp_path / ("gpu_fss_%s_N%d_timeseries.bin" % (model, L))))
        rows = open(tmp_path / ("gpu_fss_%s_N%d_timeseries.csv" % (model, L))).read().splitlines()[1:]
        n3 = float(L ** 3)
        assert ts["L"] == L and ts["replicas"] == 4 and ts["model"] == model
        assert len(rows) == ts["temp_idx"].size
        if model == "ising":
            e = -ts["bond"] / n3
            m = np.abs(ts["mag"]) / n3
            rt, rb, rm, _ = o.ising_pt_run(L, 4, o.linspace(4.4, 4.6, 4), 42, 50, 60, 10, 1)
            assert np.array_equal(ts["bond"].reshape(-1), rb) and np.array_equal(ts["mag"].reshape(-1), rm)
            assert np.array_equal(ts["temp_idx"].reshape(-1), rt)
        else:
            e = ts["e"] / n3
            m = np.sqrt(ts["mx"] ** 2 + ts["my"] ** 2 + ts["mz"] ** 2) / n3
        got = ["%d,%.8f,%.8f" % (t, a, b) for t, a, b in zip(ts["temp_idx"].reshape(-1), e.reshape(-1), m.reshape(-1))]
        assert got == rows
        assert np.allclose(ts["temps"], np.linspace(float(args[5]), float(args[7]), 4))


@pytest.mark.gpu
def test_multispin_algorithm_writes_the_oracle_files(tmp_path):
    # --algorithm msc: the same files from the multispin-coded chain, record for record
    r = run(["--model", "ising", "--sizes", "4,8", "--tmin", "4.4", "--tmax", "4.6", "--replicas", "4", "--warmup",
             "50", "--samples", "100", "--exchange-every", "10", "--outdir", str(tmp_path), "--algorithm", "msc"])
    assert r.returncode == 0, r.stderr
    assert "algorithm: msc" in r.stderr
    for L in (4, 8):
        ts_ref, sum_ref = expected_ising_csvs(L, 4, 4.4, 4.6, 50, 100, 10, 1, 42, msc=True)
        assert open(tmp_path / ("gpu_fss_ising_N%d_timeseries.csv" % L)).read() == ts_ref
        assert open(tmp_path / ("gpu_fss_ising_N%d_summary.csv" % L)).read() == sum_ref


@pytest.mark.gpu
@pytest.mark.parametrize("algorithm,model", [("sw", "ising"), ("wolff", "ising"), ("sw", "xy")])
def test_cluster_algorithms_run_the_hybrid_chain(tmp_path, algorithm, model):
    # a cluster update after every Metropolis sweep: same files, one record per measured sweep
    r = run(["--model", model, "--sizes", "6", "--tmin", "4.4" if model == "ising" else "1.8", "--tmax",
             "4.6" if model == "ising" else "2.4", "--replicas", "4", "--warmup", "40", "--samples", "60",
             "--outdir", str(tmp_path), "--algorithm", algorithm, "--overrelax", "1"])
    assert r.returncode == 0, r.stderr
    ts = open(tmp_path / ("gpu_fss_%s_N6_timeseries.csv" % model)).read().splitlines()
    summ = open(tmp_path / ("gpu_fss_%s_N6_summary.csv" % model)).read().splitlines()
    assert ts[0] == "temp_idx,E,M" and len(ts) == 1 + 4 * 60
    assert len(summ) == 5 and "nan" not in "".join(ts + summ).lower()
    e = np.array([float(x.split(",")[1]) for x in ts[1:]])
    assert np.all(e < 0) and np.all(e >= -3.0)
