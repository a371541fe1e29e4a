"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure only: the oracle is the checker, never the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
ROCRAND_CHECK = os.path.join(ROOT, "oracle", "_build", "philox_rocrand_check")

MODEL_ISING, MODEL_XY, MODEL_HEIS = 0, 1, 2
P_INIT, P_ISING, P_TIE, P_EXCH, P_CONT, P_MSC = 1, 2, 3, 4, 5, 6


def _load():
    if not os.path.exists(_LIB):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return C.CDLL(_LIB)


lib = _load()

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
_i8p = np.ctypeslib.ndpointer(np.int8, flags="C")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
_i32p_or_null = C.c_void_p

lib.o_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
lib.o_tag.restype = C.c_uint32
lib.o_tag.argtypes = [C.c_int, C.c_int]
lib.o_stream.restype = C.c_uint32
lib.o_stream.argtypes = [C.c_int, C.c_int]
lib.o_det_exp.restype = C.c_double
lib.o_det_exp.argtypes = [C.c_double]
lib.o_u01.restype = C.c_double
lib.o_u01.argtypes = [C.c_uint32, C.c_uint32]
lib.o_ising_thresholds.argtypes = [C.c_double, C.c_double, C.c_double, _u32p, _u32p]
lib.o_ising2d_thresholds.argtypes = [C.c_double, C.c_double, C.c_double, _u32p, _u32p]
lib.o_ising_init.argtypes = [C.c_int, C.c_uint64, C.c_uint32, _i8p]
lib.o_ising_halfsweep.argtypes = [C.c_int, _i8p, C.c_int, C.c_int64, C.c_uint64, C.c_uint32, _u32p, _u32p]
lib.o_ising2d_halfsweep.argtypes = [C.c_int, _i8p, C.c_int, C.c_int64, C.c_uint64, C.c_uint32, _u32p, _u32p]
lib.o_ising_measure.argtypes = [C.c_int, _i8p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
lib.o_ising2d_measure.argtypes = [C.c_int, _i8p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
lib.o_exchange.restype = C.c_int
lib.o_exchange.argtypes = [C.c_int, _f64p, _f64p, _i32p, _i32p, C.c_int64, C.c_uint64, C.c_uint32, C.c_int]
lib.o_ising_pt_run.restype = C.c_int64
lib.o_ising_pt_run.argtypes = [C.c_int, C.c_int, _f64p, C.c_uint64, C.c_double, C.c_double, C.c_int64, C.c_int64,
                               C.c_int, C.c_int, _i32p, _i64p, _i64p, _i8p]
lib.o_jackknife.argtypes = [_f64p, _f64p, C.c_int64, C.c_double, C.c_double, C.c_int, _f64p]
lib.o_linspace.argtypes = [C.c_double, C.c_double, C.c_int, _f64p]
lib.o_cont_init.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_uint32, _f64p]
lib.o_cont_metropolis_half.argtypes = [C.c_int, C.c_int, _f64p, C.c_int, C.c_int64, C.c_uint64, C.c_uint32,
                                       C.c_double, C.c_double, C.c_double, C.c_int, C.c_void_p, C.c_void_p]
lib.o_cont_overrelax_half.argtypes = [C.c_int, C.c_int, _f64p, C.c_int, C.c_double]
lib.o_cont_measure.argtypes = [C.c_int, C.c_int, _f64p, C.c_double, C.POINTER(C.c_double), _f64p]
lib.o_cont_site_energy.argtypes = [C.c_int, C.c_int, _f64p, C.c_double, _f64p]


class Rng(C.Structure):
    _fields_ = [("s", C.c_uint64 * 4)]


lib.o_rng_seed.argtypes = [C.POINTER(Rng), C.c_uint64]
lib.o_rng_f64.restype = C.c_double
lib.o_rng_f64.argtypes = [C.POINTER(Rng)]
lib.o_ref_metropolis_sweep.argtypes = [C.c_int, _i8p, C.c_double, C.c_double, C.c_double, C.POINTER(Rng)]
lib.o_ref_wolff_step.restype = C.c_int64
lib.o_ref_wolff_step.argtypes = [C.c_int, _i8p, C.c_double, C.c_double, C.POINTER(Rng), _i32p,
                                 np.ctypeslib.ndpointer(np.uint8, flags="C")]
lib.o_ref_heis_metropolis_sweep.argtypes = [C.c_int, _f64p, C.c_double, C.c_double, C.c_double, C.POINTER(Rng),
                                            C.POINTER(C.c_int64)]
lib.o_ref_heis_overrelax_sweep.argtypes = [C.c_int, _f64p, C.c_double]
lib.o_ref_xy_wolff_sweep.restype = C.c_int64
lib.o_ref_xy_wolff_sweep.argtypes = [C.c_int, _f64p, C.c_double, C.c_double, C.POINTER(Rng), _i32p,
                                     np.ctypeslib.ndpointer(np.uint8, flags="C")]
lib.o_bench_ref_metropolis.restype = C.c_int64
lib.o_bench_ref_metropolis.argtypes = [C.c_int, C.c_double, C.c_uint64, C.c_double, C.POINTER(C.c_double)]
lib.o_bench_ref_heis.restype = C.c_int64
lib.o_bench_ref_heis.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, C.c_uint64, C.c_double,
                                 C.POINTER(C.c_double)]
lib.o_bench_ref_xy_wolff.restype = C.c_int64
lib.o_bench_ref_xy_wolff.argtypes = [C.c_int, C.c_double, C.c_uint64, C.c_double, C.POINTER(C.c_double),
                                     C.POINTER(C.c_int64)]


def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib.o_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def thresholds(beta, J=1.0, h=0.0):
    hi = np.zeros(14, np.uint32)
    lo = np.zeros(14, np.uint32)
    lib.o_ising_thresholds(beta, J, h, hi, lo)
    return hi, lo


def thresholds2d(beta, J=1.0, h=0.0):
    hi = np.zeros(10, np.uint32)
    lo = np.zeros(10, np.uint32)
    lib.o_ising2d_thresholds(beta, J, h, hi, lo)
    return hi, lo


def ising_init(L, seed, replica=0):
    sp = np.zeros(L ** 3, np.int8)
    lib.o_ising_init(L, seed, lib.o_stream(L, replica), sp)
    return sp


def ising_sweep(L, sp, t, seed, beta, J=1.0, h=0.0, replica=0):
    hi, lo = thresholds(beta, J, h)
    st = lib.o_stream(L, replica)
    for c in (0, 1):
        lib.o_ising_halfsweep(L, sp, c, t, seed, st, hi, lo)


def ising_measure(L, sp):
    b, m = C.c_int64(), C.c_int64()
    lib.o_ising_measure(L, np.ascontiguousarray(sp, np.int8), C.byref(b), C.byref(m))
    return b.value, m.value


def ising2d_measure(L, sp):
    b, m = C.c_int64(), C.c_int64()
    lib.o_ising2d_measure(L, np.ascontiguousarray(sp, np.int8), C.byref(b), C.byref(m))
    return b.value, m.value


def ising_pt_run(L, R, temps, seed, warmup, samples, exchange_every, measure_every, J=1.0, h=0.0):
    n_meas = samples // measure_every
    tidx = np.zeros(max(1, n_meas * R), np.int32)
    bond = np.zeros(max(1, n_meas * R), np.int64)
    mag = np.zeros(max(1, n_meas * R), np.int64)
    final = np.zeros(L ** 3 * R, np.int8)
    n = lib.o_ising_pt_run(L, R, np.asarray(temps, np.float64), seed, J, h, warmup, samples, exchange_every,
                           measure_every, tidx, bond, mag, final)
    return tidx[:n], bond[:n], mag[:n], final.reshape(R, L ** 3)


def jackknife(e, m, beta, n3, n_blocks):
    out = np.zeros(12)
    lib.o_jackknife(np.ascontiguousarray(e, np.float64), np.ascontiguousarray(m, np.float64), len(e), beta, n3,
                    n_blocks, out)
    return out


def linspace(tmin, tmax, n):
    out = np.zeros(n)
    lib.o_linspace(tmin, tmax, n, out)
    return out


def rng(seed):
    r = Rng()
    lib.o_rng_seed(C.byref(r), seed)
    return r


def rocrand_philox(rows):
    """rows: iterable of (c0,c1,c2,c3,k0,k1) -> np.uint32 array (n,4) from rocRAND's engine."""
    if not os.path.exists(ROCRAND_CHECK):
        return None
    inp = "".join("%x %x %x %x %x %x\n" % tuple(int(v) for v in r) for r in rows)
    out = subprocess.run([ROCRAND_CHECK], input=inp, capture_output=True, text=True, check=True).stdout
    return np.array([[int(w, 16) for w in line.split()] for line in out.strip().splitlines()], np.uint32)

lib.o_ising_chain.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_int64,
                              C.c_int64, _i64p, _i64p]


def ising_chain(dim, L, seed, beta, warmup, n, J=1.0, h=0.0):
    b = np.zeros(n, np.int64)
    m = np.zeros(n, np.int64)
    lib.o_ising_chain(dim, L, seed, beta, J, h, warmup, n, b, m)
    return b, m

lib.o_tau_int.restype = C.c_double
lib.o_tau_int.argtypes = [_f64p, C.c_int64, C.c_double, C.POINTER(C.c_int64)]


def tau_int(x, c=6.0):
    w = C.c_int64()
    t = lib.o_tau_int(np.ascontiguousarray(x, np.float64), len(x), c, C.byref(w))
    return t, w.value

lib.o_ref_ising_chain.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_double, C.c_int64, C.c_int64, _i64p, _i64p]
lib.o_ref_heis_chain.argtypes = [C.c_int, C.c_uint64, C.c_double, C.c_double, C.c_int, C.c_int64, C.c_int64, _f64p,
                                 _f64p]
lib.o_ref_xy_chain.argtypes = [C.c_int, C.c_uint64, C.c_double, C.c_int64, C.c_int64, _f64p, _f64p]


def ref_ising_chain(kind, L, seed, beta, warmup, n):
    """kind 'metropolis' (random-site sweeps) or 'wolff' (cluster steps): (bond, mag) per step."""
    b = np.zeros(n, np.int64)
    m = np.zeros(n, np.int64)
    lib.o_ref_ising_chain(0 if kind == "metropolis" else 1, L, seed, beta, warmup, n, b, m)
    return b, m


def ref_heis_chain(L, seed, beta, delta, n_or, warmup, n):
    e = np.zeros(n)
    m = np.zeros(n)
    lib.o_ref_heis_chain(L, seed, beta, delta, n_or, warmup, n, e, m)
    return e, m


def ref_xy_chain(L, seed, beta, warmup, n):
    e = np.zeros(n)
    m = np.zeros(n)
    lib.o_ref_xy_chain(L, seed, beta, warmup, n, e, m)
    return e, m

lib.o_msc_thresholds.argtypes = [C.c_double, C.c_double, _u32p]
lib.o_msc_halfsweep.argtypes = [C.c_int, C.c_int, _i8p, C.c_int, C.c_int64, C.c_uint64, _u32p]
lib.o_msc_pt_run.restype = C.c_int64
lib.o_msc_pt_run.argtypes = [C.c_int, C.c_int, _f64p, C.c_uint64, C.c_double, C.c_int64, C.c_int64, C.c_int, C.c_int,
                             _i32p, _i64p, _i64p, _i8p]


def msc_thresholds(beta, J=1.0):
    out = np.zeros(3, np.uint32)
    lib.o_msc_thresholds(beta, J, out)
    return out


def msc_pt_run(L, R, temps, seed, warmup, samples, exchange_every, measure_every, J=1.0):
    n_meas = samples // measure_every
    tidx = np.zeros(max(1, n_meas * R), np.int32)
    bond = np.zeros(max(1, n_meas * R), np.int64)
    mag = np.zeros(max(1, n_meas * R), np.int64)
    final = np.zeros(L ** 3 * R, np.int8)
    n = lib.o_msc_pt_run(L, R, np.asarray(temps, np.float64), seed, J, warmup, samples, exchange_every,
                         measure_every, tidx, bond, mag, final)
    return tidx[:n], bond[:n], mag[:n], final.reshape(R, L ** 3)

lib.o_ising_pt_run_dim.restype = C.c_int64
lib.o_ising_pt_run_dim.argtypes = [C.c_int, C.c_int, C.c_int, _f64p, C.c_uint64, C.c_double, C.c_double, C.c_int64,
                                   C.c_int64, C.c_int, C.c_int, _i32p, _i64p, _i64p, _i8p]


def ising2d_pt_run(L, R, temps, seed, warmup, samples, exchange_every, measure_every, J=1.0, h=0.0):
    n_meas = samples // measure_every
    tidx = np.zeros(max(1, n_meas * R), np.int32)
    bond = np.zeros(max(1, n_meas * R), np.int64)
    mag = np.zeros(max(1, n_meas * R), np.int64)
    final = np.zeros(L ** 2 * R, np.int8)
    n = lib.o_ising_pt_run_dim(2, L, R, np.asarray(temps, np.float64), seed, J, h, warmup, samples, exchange_every,
                               measure_every, tidx, bond, mag, final)
    return tidx[:n], bond[:n], mag[:n], final.reshape(R, L ** 2)

_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C")
lib.o_bond_threshold.restype = C.c_uint32
lib.o_bond_threshold.argtypes = [C.c_double, C.c_double]
lib.o_ising_cluster.restype = C.c_int64
lib.o_ising_cluster.argtypes = [C.c_int, C.c_int, _i8p, C.c_int, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32,
                                C.c_double, _i32p, _u8p]
lib.o_cont_cluster.restype = C.c_int64
lib.o_cont_cluster.argtypes = [C.c_int, C.c_int, _f64p, C.c_int, C.c_uint32, C.c_uint64, C.c_uint32, C.c_int,
                               C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int64), _i32p, _u8p]
lib.o_ising_cluster_chain.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_int, C.c_double, C.c_double, C.c_int64,
                                      C.c_int64, _i64p, _i64p, _i64p]


def bond_threshold(beta, J=1.0):
    return int(lib.o_bond_threshold(beta, J))


def ising_cluster(dim, L, spins, kind, u, seed, stream, beta=None, J=1.0, thr=None):
    """One cluster update of `spins` (int8 +-1, modified in place); returns (flipped, bonds)."""
    n = spins.size
    if thr is None:
        thr = bond_threshold(beta, J)
    lab = np.zeros(n, np.int32)
    bonds = np.zeros(n, np.uint8)
    f = lib.o_ising_cluster(dim, L, spins, kind, u, seed, stream, thr, J, lab, bonds)
    return int(f), bonds


def cont_cluster(L, ncomp, spins, kind, u, seed, stream, model, beta, J=1.0, margin=1e-5):
    """One embedded cluster update of `spins` ([L^3 * ncomp] float64, in place):
    returns (flipped, bonds, bonds within `margin` of their probability)."""
    n = L ** 3
    lab = np.zeros(n, np.int32)
    bonds = np.zeros(n, np.uint8)
    marg = C.c_int64()
    f = lib.o_cont_cluster(L, ncomp, spins, kind, u, seed, stream, model, beta, J, margin, C.byref(marg), lab, bonds)
    return int(f), bonds, marg.value


def ising_cluster_chain(dim, L, seed, kind, beta, warmup, n, J=1.0):
    b = np.zeros(n, np.int64)
    m = np.zeros(n, np.int64)
    f = np.zeros(n, np.int64)
    lib.o_ising_cluster_chain(dim, L, seed, kind, beta, J, warmup, n, b, m, f)
    return b, m, f
