"""CPU-side checks of the C-ABI library: it loads, exports every declared
symbol, its host-only helpers agree with the oracle, and device entry points
fail cleanly (error code + message, no crash) where no GPU is visible."""
import os
import re

import numpy as np
import pytest

import ising_hip
import oracle_lib as o

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if not hdr.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(ising_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(ising_hip.lib, n), n


def test_version_and_device_count_do_not_fail():
    assert "gfx950" in ising_hip.version()
    assert ising_hip.device_count() >= 0


def test_linspace_matches_oracle():
    for (a, b, n) in ((4.4, 4.62, 32), (1.4, 1.5, 32), (2.15, 2.25, 7), (4.0, 5.0, 1)):
        assert np.array_equal(ising_hip.linspace_temperatures(a, b, n), o.linspace(a, b, n))


def test_jackknife_matches_oracle_bitwise():
    rng = np.random.default_rng(3)
    for n in (20, 401, 5000):
        e = rng.normal(-1.2, 0.05, n)
        m = np.abs(rng.normal(0.4, 0.1, n))
        for nb in (2, 20):
            assert np.array_equal(ising_hip.jackknife_observables(e, m, 0.22, 4096.0, nb),
                                  o.jackknife(e, m, 0.22, 4096.0, nb))
    assert np.all(ising_hip.jackknife_observables(np.ones(3), np.ones(3), 1.0, 1.0, 20) == 0)


def test_host_replica_exchange_follows_reference_rule():
    temps = np.linspace(4.4, 4.6, 6)
    e = np.array([-100.0, -120.0, -90.0, -95.0, -200.0, -10.0])
    r2t = np.arange(6, dtype=np.int32)
    t2r = np.arange(6, dtype=np.int32)
    # uniforms of 0 accept every pair with a finite delta; 1 accepts only delta < 0
    acc = ising_hip.replica_exchange(temps, e, r2t, t2r, np.ones(3), 0)
    b = 1.0 / temps
    want = sum(1 for i in (0, 2, 4) if (b[i] - b[i + 1]) * (e[i + 1] - e[i]) < 0)
    assert acc == want
    assert np.array_equal(t2r[r2t], np.arange(6))
    acc = ising_hip.replica_exchange(temps, e, r2t, t2r, np.zeros(2), 1)
    assert acc == 2 and np.array_equal(t2r[r2t], np.arange(6))


@pytest.mark.skipif(ising_hip.device_count() > 0, reason="checks the no-device error path")
def test_device_entry_points_fail_loudly_without_a_gpu():
    with pytest.raises(ising_hip.IsingError) as ei:
        ising_hip.LatticeGpu(8, 1)
    assert ei.value.code == -2
    with pytest.raises(ising_hip.IsingError):
        ising_hip.PTEnsemble("ising", [8], 4, np.linspace(4.4, 4.6, 4), 1)
    # reduce_gpu::* checks for a device before it touches the (here fake) device address
    for call in (ising_hip.load_reduce_kernels, lambda: ising_hip.reduce_ising(0x1000, 8),
                 lambda: ising_hip.measure_continuous_gpu(0x1000, 8, 3),
                 lambda: ising_hip.reduce_continuous_mag(0x1000, 100, 2)):
        with pytest.raises(ising_hip.IsingError) as ei:
            call()
        assert ei.value.code == -2


def test_invalid_arguments_are_rejected_before_touching_the_device():
    with pytest.raises(ising_hip.IsingError) as ei:
        ising_hip.LatticeGpu(7, 1)  # odd L
    assert ei.value.code == -1
    with pytest.raises(ising_hip.IsingError) as ei:
        ising_hip.PTEnsemble("ising", [8], 4, np.array([4.4, -1.0, 4.5, 4.6]), 1)
    assert ei.value.code == -1
    with pytest.raises(ising_hip.IsingError):
        ising_hip.PTEnsemble("ising", [8], 2, np.array([4.4, 4.6]), 1, exchange_every=0)
    bad = [
        lambda: ising_hip.LatticeGpu(0, 1),
        lambda: ising_hip.LatticeGpu(4096, 1),  # beyond 2..2046
        lambda: ising_hip.ContinuousGpuLattice(8, 4, 1),
        lambda: ising_hip.PTEnsemble("ising", [], 2, [4.4, 4.5], 1),
        lambda: ising_hip.PTEnsemble("ising", [8], 0, [], 1),
        lambda: ising_hip.PTEnsemble("ising", [8], 2, [4.4, float("nan")], 1),
        lambda: ising_hip.PTEnsemble("ising", [8], 2, [4.4, float("inf")], 1),
        lambda: ising_hip.PTEnsemble("ising", [8, 9], 2, [4.4, 4.5], 1),
        lambda: ising_hip.PTEnsemble("ising", [8], 2, [4.4, 4.5], 1, measure_every=0),
        lambda: ising_hip.PTEnsemble("ising", [8], 2, [4.4, 4.5], 1, record_capacity=0),
        lambda: ising_hip.PTEnsemble("ising", [8], 2, [4.4, 4.5], 1, ising_spt=9),
        lambda: ising_hip.PTEnsemble("ising", [8], 2, [4.4, 4.5], 1, ising_fused=12),
    ]
    for make in bad:
        with pytest.raises(ising_hip.IsingError) as ei:
            make()
        assert ei.value.code == -1


def test_binary_timeseries_reader_parses_the_documented_layout(tmp_path):
    """The reader of gpu_fss --binary output against a file written field by
    field in the documented layout (no GPU needed)."""
    import ising_hip
    R, n, L = 3, 2, 4
    temps = np.array([4.4, 4.5, 4.6])
    recs = np.array([[[10, -4, 2], [12, 6, 0], [-2, 8, 1]], [[14, 2, 0], [16, -6, 2], [0, 0, 1]]], np.int64)
    raw = (b"ISINGTS1" + np.array([1, 0, L, R, 5, 0], np.int32).tobytes() + np.array([n], np.int64).tobytes() +
           np.array([1.0], np.float64).tobytes() + temps.tobytes() + recs.tobytes())
    p = tmp_path / "x.bin"
    p.write_bytes(raw)
    ts = ising_hip.read_timeseries(str(p))
    assert ts["model"] == "ising" and ts["L"] == L and ts["replicas"] == R and ts["measure_every"] == 5
    assert np.array_equal(ts["bond"], recs[:, :, 0]) and np.array_equal(ts["mag"], recs[:, :, 1])
    assert np.array_equal(ts["temp_idx"], recs[:, :, 2]) and np.array_equal(ts["temps"], temps)
    p.write_bytes(raw[:-8])
    with pytest.raises(ValueError):
        ising_hip.read_timeseries(str(p))
