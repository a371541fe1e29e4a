"""Python mirror of the reference GPU API, over libising_hip.so (C ABI).

Classes and functions follow faulknco/ising-rs's `ising::cuda` surface:

* ``LatticeGpu``            -- src/cuda/lattice_gpu.rs:9-306
* ``ContinuousGpuLattice``  -- src/cuda/gpu_lattice_continuous.rs:9-250
* ``replica_exchange``, ``linspace_temperatures`` -- src/parallel_tempering.rs:5-46
* ``jackknife_observables`` -- src/bin/gpu_fss.rs:327-412
* ``PTEnsemble``            -- the device-resident PT loop of src/bin/gpu_fss.rs

Every compute call goes through the HIP library; there is no CPU fallback.  If
the library is not built, importing this module raises.
"""
import ctypes as C
import os
from collections import namedtuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libising_hip.so")

MODEL_ISING, MODEL_XY, MODEL_HEISENBERG, MODEL_ISING_MSC, MODEL_ISING2D = 0, 1, 2, 3, 4
CLUSTER_SW, CLUSTER_WOLFF = 0, 1
_CLUSTER = {"sw": CLUSTER_SW, "wolff": CLUSTER_WOLFF, CLUSTER_SW: CLUSTER_SW, CLUSTER_WOLFF: CLUSTER_WOLFF}
_MODELS = {"ising": MODEL_ISING, "xy": MODEL_XY, "heisenberg": MODEL_HEISENBERG, "ising_msc": MODEL_ISING_MSC,
           "ising2d": MODEL_ISING2D}

if not os.path.exists(LIB_PATH):
    raise ImportError("libising_hip.so is not built (%s); run `python -c 'import __graft_entry__ as g; g.build()'`"
                      % LIB_PATH)
lib = C.CDLL(LIB_PATH)


class IsingError(RuntimeError):
    """A libising_hip call failed; `code` is the ISING_E_* value."""

    def __init__(self, code, msg):
        super().__init__("libising_hip error %d: %s" % (code, msg))
        self.code = code


class Opts(C.Structure):
    _fields_ = [("j", C.c_double), ("h", C.c_double), ("exchange_every", C.c_int32), ("measure_every", C.c_int32),
                ("delta", C.c_double), ("n_overrelax", C.c_int32), ("renorm_every", C.c_int32),
                ("record_capacity", C.c_int32), ("adapt_every", C.c_int32), ("adapt_target", C.c_double),
                ("graph_sweeps", C.c_int32), ("ising_spt", C.c_int32), ("ising_width", C.c_int32),
                ("ising_runs", C.c_int32), ("ising_fused", C.c_int32)]


class Observables(C.Structure):
    _fields_ = [("temperature", C.c_double), ("energy", C.c_double), ("magnetisation", C.c_double),
                ("heat_capacity", C.c_double), ("susceptibility", C.c_double), ("m2", C.c_double),
                ("m4", C.c_double)]


_vp = C.c_void_p
_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)
_SIGS = {
    "ising_last_error": (C.c_char_p, []),
    "ising_version": (C.c_char_p, []),
    "ising_device_count": (C.c_int, []),
    "ising_lattice_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.c_uint64]),
    "ising_lattice_destroy": (None, [_vp]),
    "ising_lattice_size": (C.c_int, [_vp]),
    "ising_lattice_step": (C.c_int, [_vp, C.c_float, C.c_float, C.c_float]),
    "ising_lattice_warm_up": (C.c_int, [_vp, C.c_float, C.c_float, C.c_float, C.c_size_t]),
    "ising_lattice_get_spins": (C.c_int, [_vp, _vp]),
    "ising_lattice_set_spins": (C.c_int, [_vp, _vp]),
    "ising_lattice_measure_gpu": (C.c_int, [_vp, C.c_float, _f64p, _f64p]),
    "ising_lattice_energy_gpu": (C.c_int, [_vp, C.c_float, _f64p]),
    "ising_lattice_measure_raw": (C.c_int, [_vp, _i64p, _i64p]),
    "ising_lattice_measure": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_size_t,
                                        C.POINTER(Observables)]),
    "ising_lattice_domain_wall_density": (C.c_int, [_vp, _f64p]),
    "ising_lattice_sweeps": (C.c_int, [_vp, _i64p]),
    "ising_cont_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.c_int, C.c_uint64]),
    "ising_cont_destroy": (None, [_vp]),
    "ising_cont_sweep": (C.c_int, [_vp, C.c_float, C.c_float, C.c_float, C.c_size_t]),
    "ising_cont_measure_gpu": (C.c_int, [_vp, C.c_float, _f64p, _f64p, _f64p, _f64p]),
    "ising_cont_measure_raw": (C.c_int, [_vp, _f64p, _f64p, _f64p, _f64p]),
    "ising_cont_get_spins": (C.c_int, [_vp, _vp]),
    "ising_cont_set_spins": (C.c_int, [_vp, _vp]),
    "ising_cont_metropolis_pass": (C.c_int, [_vp, C.c_int, C.c_float, C.c_float, C.c_float]),
    "ising_cont_overrelax_pass": (C.c_int, [_vp, C.c_int, C.c_float]),
    "ising_cont_sweeps": (C.c_int, [_vp, _i64p]),
    "ising_pt_default_opts": (None, [C.POINTER(Opts)]),
    "ising_pt_create": (C.c_int, [C.POINTER(_vp), C.c_int, _vp, C.c_int32, C.c_int32, _vp, C.c_uint64,
                                  C.POINTER(Opts)]),
    "ising_pt_destroy": (None, [_vp]),
    "ising_pt_warmup": (C.c_int, [_vp, C.c_int64]),
    "ising_pt_sample": (C.c_int, [_vp, C.c_int64]),
    "ising_pt_sync": (C.c_int, [_vp, _f64p]),
    "ising_pt_pending": (C.c_int, [_vp, _i64p]),
    "ising_pt_read_ising": (C.c_int, [_vp, C.c_int64, _vp, _vp, _vp, _i64p]),
    "ising_pt_read_cont": (C.c_int, [_vp, C.c_int64, _vp, _vp, _vp, _vp, _vp, _i64p]),
    "ising_pt_get_perm": (C.c_int, [_vp, C.c_int32, _vp, _vp]),
    "ising_pt_exchange_stats": (C.c_int, [_vp, _i64p, _i64p]),
    "ising_pt_get_spins_i8": (C.c_int, [_vp, C.c_int32, C.c_int32, _vp]),
    "ising_pt_set_spins_i8": (C.c_int, [_vp, C.c_int32, C.c_int32, _vp]),
    "ising_pt_get_spins_f32": (C.c_int, [_vp, C.c_int32, C.c_int32, _vp]),
    "ising_pt_set_spins_f32": (C.c_int, [_vp, C.c_int32, C.c_int32, _vp]),
    "ising_pt_get_deltas": (C.c_int, [_vp, C.c_int32, _vp]),
    "ising_pt_learn_init": (C.c_int, [_vp, C.c_int32, C.c_uint64]),
    "ising_pt_learn_size": (C.c_int, [_vp, _i64p]),
    "ising_pt_learn_get_params": (C.c_int, [_vp, _vp, C.c_int64]),
    "ising_pt_learn_set_params": (C.c_int, [_vp, _vp, C.c_int64]),
    "ising_pt_learn_train": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_double, _f64p]),
    "ising_pt_learn_eval": (C.c_int, [_vp, C.c_int64, _vp, _vp, _f64p, _vp]),
    "ising_pt_learn_enable": (C.c_int, [_vp, C.c_int32]),
    "ising_pt_learn_infer": (C.c_int, [_vp, C.c_int64, _vp, _vp]),
    "ising_pt_mh_stats": (C.c_int, [_vp, _i64p, _i64p]),
    "ising_pt_heat_bath": (C.c_int, [_vp, C.c_int32]),
    "ising_pt_metropolis_pass": (C.c_int, [_vp, C.c_int32, C.c_int64]),
    "ising_pt_sweeps": (C.c_int, [_vp, _i64p]),
    "ising_pt_prepare_graph": (C.c_int, [_vp, C.c_int32]),
    "ising_pt_device_bytes": (C.c_int, [_vp, _i64p]),
    "ising_pt_autotune": (C.c_int, [_vp, C.c_int32] + [C.POINTER(C.c_int32)] * 3),
    "ising_pt_get_layout": (C.c_int, [_vp, C.c_int32] + [C.POINTER(C.c_int32)] * 3),
    "ising_pt_helicity": (C.c_int, [_vp, _vp]),
    "ising_pt_anneal": (C.c_int, [_vp, _vp, C.c_int64, _vp, _vp]),
    "ising_pt_cluster": (C.c_int, [_vp, C.c_int32, C.c_int64, _vp, _vp, _vp]),
    "ising_pt_cluster_cont": (C.c_int, [_vp, C.c_int32, C.c_int64, _vp, _vp]),
    "ising_pt_cluster_count": (C.c_int, [_vp, _i64p]),
    "ising_lattice_cluster_step": (C.c_int, [_vp, C.c_int32, C.c_float, C.c_float, _i64p]),
    "ising_cont_cluster_step": (C.c_int, [_vp, C.c_int32, C.c_float, C.c_float, _i64p]),
    "ising_kz_run": (C.c_int, [C.c_int32, C.c_double, C.c_double, C.c_double, C.c_int64, C.c_int32, C.c_int64,
                               C.c_int64, C.c_double, C.c_uint64, _vp]),
    "ising_kz_sweep": (C.c_int, [C.c_int32, C.c_double, C.c_double, C.c_double, _vp, C.c_int32, C.c_int32,
                                 C.c_int64, C.c_int64, C.c_double, C.c_uint64, _vp, _vp]),
    "ising_coarsening_run": (C.c_int, [C.c_int32, C.c_double, C.c_double, C.c_double, C.c_int64, C.c_int64,
                                       C.c_int64, C.c_uint64, C.c_int32, _vp, _vp, _i64p]),
    "ising_fss_anneal": (C.c_int, [_vp, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int32,
                                   C.c_int64, C.c_int64, C.c_uint64, C.POINTER(Observables)]),
    "ising_batch_create": (C.c_int, [C.POINTER(_vp), _vp, C.c_int32, C.c_int32]),
    "ising_batch_destroy": (None, [_vp]),
    "ising_batch_warmup": (C.c_int, [_vp, C.c_int64]),
    "ising_batch_sample": (C.c_int, [_vp, C.c_int64]),
    "ising_batch_sync": (C.c_int, [_vp, _f64p]),
    "ising_batch_set_mode": (C.c_int, [_vp, C.c_int32]),
    "ising_load_reduce_kernels": (C.c_int, []),
    "ising_reduce_ising": (C.c_int, [_vp, C.c_int32, C.c_float, _f64p, _f64p]),
    "ising_measure_ising_gpu": (C.c_int, [_vp, C.c_int32, C.c_float, _f64p, _f64p]),
    "ising_reduce_ising_noalloc": (C.c_int, [_vp, C.c_int32, C.c_float, _vp, _vp]),
    "ising_reduce_continuous_mag": (C.c_int, [_vp, C.c_int64, C.c_int32, _f64p, _f64p, _f64p]),
    "ising_reduce_continuous_energy": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_float, _f64p]),
    "ising_measure_continuous_gpu": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_float, _f64p, _f64p, _f64p,
                                               _f64p]),
    "ising_linspace_temperatures": (C.c_int, [C.c_double, C.c_double, C.c_int32, _vp]),
    "ising_replica_exchange": (C.c_int, [_vp, _vp, C.c_int32, _vp, _vp, _vp, C.c_int32, C.POINTER(C.c_int32)]),
    "ising_jackknife": (C.c_int, [_vp, _vp, C.c_int64, C.c_double, C.c_double, C.c_int32, _vp]),
    "ising_tau_int": (C.c_int, [_vp, C.c_int64, C.c_double, _f64p, _i64p]),
    "ising_binder_crossing": (C.c_int, [_vp, _vp, C.c_int32, _vp, _vp, C.c_int32, _f64p, C.POINTER(C.c_int32)]),
    "ising_wham": (C.c_int, [C.c_int32, _vp, _vp, _vp, C.c_int32, _vp, C.c_int32, C.c_double, _vp,
                             C.POINTER(C.c_int32), C.c_int32, _vp, _vp]),
}
for _name, (_res, _args) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc != 0:
        raise IsingError(rc, lib.ising_last_error().decode())


def version():
    return lib.ising_version().decode()


def device_count():
    return lib.ising_device_count()


ObservablesT = namedtuple("ObservablesT", "temperature energy magnetisation heat_capacity susceptibility m2 m4")


class LatticeGpu:
    """One 3D Ising lattice on the GPU (mirror of LatticeGpu, lattice_gpu.rs:9-306)."""

    def __init__(self, n, seed):
        self._h = _vp()
        _check(lib.ising_lattice_create(C.byref(self._h), int(n), int(seed) & (2 ** 64 - 1)))
        self.n = int(n)

    def close(self):
        if getattr(self, "_h", None):
            lib.ising_lattice_destroy(self._h)
            self._h = None

    __del__ = close

    def step(self, beta, j=1.0, h=0.0):
        _check(lib.ising_lattice_step(self._h, beta, j, h))

    def warm_up(self, beta, j, h, sweeps):
        _check(lib.ising_lattice_warm_up(self._h, beta, j, h, int(sweeps)))

    def get_spins(self):
        out = np.empty(self.n ** 3, np.int8)
        _check(lib.ising_lattice_get_spins(self._h, _ptr(out)))
        return out

    def set_spins(self, spins):
        a = np.ascontiguousarray(spins, np.int8).reshape(-1)
        if a.size != self.n ** 3:
            raise ValueError("need n^3 spins")
        _check(lib.ising_lattice_set_spins(self._h, _ptr(a)))

    def measure_gpu(self, j=1.0):
        e, m = C.c_double(), C.c_double()
        _check(lib.ising_lattice_measure_gpu(self._h, j, C.byref(e), C.byref(m)))
        return e.value, m.value

    def energy_gpu(self, j=1.0):
        e = C.c_double()
        _check(lib.ising_lattice_energy_gpu(self._h, j, C.byref(e)))
        return e.value

    def measure_raw(self):
        b, m = C.c_int64(), C.c_int64()
        _check(lib.ising_lattice_measure_raw(self._h, C.byref(b), C.byref(m)))
        return b.value, m.value

    def measure(self, temperature, j, h, samples):
        o = Observables()
        _check(lib.ising_lattice_measure(self._h, temperature, j, h, int(samples), C.byref(o)))
        return ObservablesT(o.temperature, o.energy, o.magnetisation, o.heat_capacity, o.susceptibility, o.m2, o.m4)

    def domain_wall_density(self):
        d = C.c_double()
        _check(lib.ising_lattice_domain_wall_density(self._h, C.byref(d)))
        return d.value

    def sweeps(self):
        v = C.c_int64()
        _check(lib.ising_lattice_sweeps(self._h, C.byref(v)))
        return v.value

    def cluster_step(self, kind, beta, j=1.0):
        """One Swendsen-Wang ("sw") or Wolff ("wolff") update; returns the sites flipped."""
        f = C.c_int64()
        _check(lib.ising_lattice_cluster_step(self._h, _CLUSTER[kind], beta, j, C.byref(f)))
        return f.value


class WolffGpuLattice(LatticeGpu):
    """GPU Wolff lattice the reference plans (WolffGpuLattice,
    docs/plans/2026-03-09-gpu-compute-optimizations-impl.md:1041-1330): `step`
    flips one cluster and returns its size; measurement and spins are the
    LatticeGpu calls on the same int8 lattice."""

    def step(self, beta, j=1.0):
        return self.cluster_step("wolff", beta, j)

    def warm_up(self, beta, j, steps):
        for _ in range(int(steps)):
            self.step(beta, j)


class ContinuousGpuLattice:
    """One XY (n_comp=2) or Heisenberg (n_comp=3) lattice on the GPU
    (mirror of ContinuousGpuLattice, gpu_lattice_continuous.rs:9-250)."""

    def __init__(self, n, n_comp, seed):
        self._h = _vp()
        _check(lib.ising_cont_create(C.byref(self._h), int(n), int(n_comp), int(seed) & (2 ** 64 - 1)))
        self.n = int(n)
        self.n_comp = int(n_comp)

    def close(self):
        if getattr(self, "_h", None):
            lib.ising_cont_destroy(self._h)
            self._h = None

    __del__ = close

    def sweep(self, beta, j, delta, n_overrelax):
        _check(lib.ising_cont_sweep(self._h, beta, j, delta, int(n_overrelax)))

    def measure_gpu(self, j=1.0):
        v = [C.c_double() for _ in range(4)]
        _check(lib.ising_cont_measure_gpu(self._h, j, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def measure_raw(self):
        v = [C.c_double() for _ in range(4)]
        _check(lib.ising_cont_measure_raw(self._h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def get_spins(self):
        out = np.empty(self.n ** 3 * self.n_comp, np.float32)
        _check(lib.ising_cont_get_spins(self._h, _ptr(out)))
        return out

    def set_spins(self, spins):
        a = np.ascontiguousarray(spins, np.float32).reshape(-1)
        if a.size != self.n ** 3 * self.n_comp:
            raise ValueError("need n^3 * n_comp components")
        _check(lib.ising_cont_set_spins(self._h, _ptr(a)))

    def metropolis_pass(self, color, beta, j, delta):
        _check(lib.ising_cont_metropolis_pass(self._h, int(color), beta, j, delta))

    def cluster_step(self, kind, beta, j=1.0):
        """One embedded Swendsen-Wang ("sw") or Wolff ("wolff") update; returns the sites flipped."""
        f = C.c_int64()
        _check(lib.ising_cont_cluster_step(self._h, _CLUSTER[kind], beta, j, C.byref(f)))
        return f.value

    def overrelax_pass(self, color, j=1.0):
        _check(lib.ising_cont_overrelax_pass(self._h, int(color), j))

    def sweeps(self):
        v = C.c_int64()
        _check(lib.ising_cont_sweeps(self._h, C.byref(v)))
        return v.value


def default_opts():
    o = Opts()
    lib.ising_pt_default_opts(C.byref(o))
    return o


class PTEnsemble:
    """Every lattice size and replica of a PT run, resident on one GPU.

    Records come back per measured sweep as arrays of shape (n_sweeps, n_sizes, R).
    """

    def __init__(self, model, sizes, replicas, temps, seed, **opts):
        o = default_opts()
        for k, v in opts.items():
            if not hasattr(o, k):
                raise TypeError("unknown option %r" % k)
            setattr(o, k, v)
        self.model = _MODELS[model] if isinstance(model, str) else int(model)
        self.sizes = [int(s) for s in sizes]
        self.R = int(replicas)
        self.temps = np.ascontiguousarray(temps, np.float64)
        if self.temps.size != self.R:
            raise ValueError("need one temperature per replica")
        self.opts = o
        self._h = _vp()
        sz = np.ascontiguousarray(self.sizes, np.int32)
        _check(lib.ising_pt_create(C.byref(self._h), self.model, _ptr(sz), len(self.sizes), self.R,
                                   _ptr(self.temps), int(seed) & (2 ** 64 - 1), C.byref(o)))

    def close(self):
        if getattr(self, "_h", None):
            lib.ising_pt_destroy(self._h)
            self._h = None

    __del__ = close

    @property
    def total_replicas(self):
        return len(self.sizes) * self.R

    def warmup(self, sweeps):
        _check(lib.ising_pt_warmup(self._h, int(sweeps)))

    def sample(self, sweeps):
        _check(lib.ising_pt_sample(self._h, int(sweeps)))

    def sync(self):
        ms = C.c_double()
        _check(lib.ising_pt_sync(self._h, C.byref(ms)))
        return ms.value

    def pending(self):
        n = C.c_int64()
        _check(lib.ising_pt_pending(self._h, C.byref(n)))
        return n.value

    def read_ising(self, max_sweeps=None):
        if max_sweeps is None:
            max_sweeps = self.pending()
        n_tot = max(1, max_sweeps) * self.total_replicas
        tidx = np.empty(n_tot, np.int32)
        bond = np.empty(n_tot, np.int64)
        mag = np.empty(n_tot, np.int64)
        n = C.c_int64()
        _check(lib.ising_pt_read_ising(self._h, int(max_sweeps), _ptr(tidx), _ptr(bond), _ptr(mag), C.byref(n)))
        k = n.value * self.total_replicas
        shape = (n.value, len(self.sizes), self.R)
        return tidx[:k].reshape(shape), bond[:k].reshape(shape), mag[:k].reshape(shape)

    def read_cont(self, max_sweeps=None):
        if max_sweeps is None:
            max_sweeps = self.pending()
        n_tot = max(1, max_sweeps) * self.total_replicas
        tidx = np.empty(n_tot, np.int32)
        e, mx, my, mz = (np.empty(n_tot, np.float64) for _ in range(4))
        n = C.c_int64()
        _check(lib.ising_pt_read_cont(self._h, int(max_sweeps), _ptr(tidx), _ptr(e), _ptr(mx), _ptr(my), _ptr(mz),
                                      C.byref(n)))
        k = n.value * self.total_replicas
        shape = (n.value, len(self.sizes), self.R)
        return (tidx[:k].reshape(shape), e[:k].reshape(shape), mx[:k].reshape(shape), my[:k].reshape(shape),
                mz[:k].reshape(shape))

    def get_perm(self, size_index=0):
        r2t = np.empty(self.R, np.int32)
        t2r = np.empty(self.R, np.int32)
        _check(lib.ising_pt_get_perm(self._h, size_index, _ptr(r2t), _ptr(t2r)))
        return r2t, t2r

    def exchange_stats(self):
        a, b = C.c_int64(), C.c_int64()
        _check(lib.ising_pt_exchange_stats(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def get_spins(self, size_index, replica):
        L = self.sizes[size_index]
        if self.model == MODEL_ISING2D:
            out = np.empty(L ** 2, np.int8)
            _check(lib.ising_pt_get_spins_i8(self._h, size_index, replica, _ptr(out)))
        elif self.model in (MODEL_ISING, MODEL_ISING_MSC):
            out = np.empty(L ** 3, np.int8)
            _check(lib.ising_pt_get_spins_i8(self._h, size_index, replica, _ptr(out)))
        else:
            nc = 2 if self.model == MODEL_XY else 3
            out = np.empty(L ** 3 * nc, np.float32)
            _check(lib.ising_pt_get_spins_f32(self._h, size_index, replica, _ptr(out)))
        return out

    def set_spins(self, size_index, replica, spins):
        if self.model in (MODEL_ISING, MODEL_ISING_MSC, MODEL_ISING2D):
            a = np.ascontiguousarray(spins, np.int8).reshape(-1)
            _check(lib.ising_pt_set_spins_i8(self._h, size_index, replica, _ptr(a)))
        else:
            a = np.ascontiguousarray(spins, np.float32).reshape(-1)
            _check(lib.ising_pt_set_spins_f32(self._h, size_index, replica, _ptr(a)))

    def get_deltas(self, size_index=0):
        out = np.empty(self.R, np.float32)
        _check(lib.ising_pt_get_deltas(self._h, size_index, _ptr(out)))
        return out

    # ---- learned proposal ----
    def learn_init(self, hidden=64, seed=0):
        _check(lib.ising_pt_learn_init(self._h, int(hidden), int(seed)))

    def learn_params(self):
        n = C.c_int64()
        _check(lib.ising_pt_learn_size(self._h, C.byref(n)))
        out = np.empty(n.value, np.float32)
        _check(lib.ising_pt_learn_get_params(self._h, _ptr(out), n.value))
        return out

    def learn_set_params(self, params):
        a = np.ascontiguousarray(params, np.float32)
        _check(lib.ising_pt_learn_set_params(self._h, _ptr(a), a.size))

    def learn_train(self, steps, batch=4096, lr=1e-3):
        loss = C.c_double()
        _check(lib.ising_pt_learn_train(self._h, int(steps), int(batch), float(lr), C.byref(loss)))
        return loss.value

    def learn_eval(self, x, c):
        x = np.ascontiguousarray(x, np.float32)
        c = np.ascontiguousarray(c, np.float32)
        n = C.c_int64()
        _check(lib.ising_pt_learn_size(self._h, C.byref(n)))
        g = np.empty(n.value, np.float32)
        loss = C.c_double()
        _check(lib.ising_pt_learn_eval(self._h, len(c), _ptr(x), _ptr(c), C.byref(loss), _ptr(g)))
        return loss.value, g

    def learn_infer(self, x):
        """Network outputs [n, 3] for feature rows x [n, 6], as the fused sweep evaluates them."""
        xa = np.ascontiguousarray(x, np.float32).reshape(-1, 6)
        out = np.empty((xa.shape[0], 3), np.float32)
        _check(lib.ising_pt_learn_infer(self._h, xa.shape[0], _ptr(xa), _ptr(out)))
        return out

    def learn_enable(self, on=True):
        _check(lib.ising_pt_learn_enable(self._h, 1 if on else 0))

    def heat_bath(self, on=True):
        """Heat-bath passes (the exact single-site conditional) in place of the Metropolis passes."""
        _check(lib.ising_pt_heat_bath(self._h, 1 if on else 0))

    def mh_stats(self):
        a, b = C.c_int64(), C.c_int64()
        _check(lib.ising_pt_mh_stats(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def metropolis_pass(self, color, t):
        _check(lib.ising_pt_metropolis_pass(self._h, int(color), int(t)))

    def helicity(self):
        """[total replicas, 6] XY stiffness sums C_x, C_y, C_z, J_x, J_y, J_z of the current state."""
        out = np.empty((self.total_replicas, 6), np.float64)
        _check(lib.ising_pt_helicity(self._h, _ptr(out)))
        return out

    def anneal(self, betas, measure=False):
        """Warm-up sweeps with inverse temperature betas[k] at sweep k for every
        replica; with measure, returns ([n, total replicas] bond sums, mags)."""
        b = np.ascontiguousarray(betas, np.float64)
        if not measure:
            _check(lib.ising_pt_anneal(self._h, _ptr(b), b.size, None, None))
            return None
        bond = np.empty((b.size, self.total_replicas), np.int64)
        mag = np.empty_like(bond)
        _check(lib.ising_pt_anneal(self._h, _ptr(b), b.size, _ptr(bond), _ptr(mag)))
        return bond, mag

    def cluster(self, kind, n=1, measure=False):
        """n Swendsen-Wang ("sw") or Wolff ("wolff") updates of every replica at
        its current temperature.  Returns [n, total replicas] sites flipped and,
        with measure, the state after each update: (bond sums, magnetisations)
        for Ising ensembles, [n, total replicas, 4] E, Mx, My, Mz otherwise."""
        n = int(n)
        k = _CLUSTER[kind]
        flipped = np.empty((n, self.total_replicas), np.int64)
        if self.model in (MODEL_XY, MODEL_HEISENBERG):
            meas = np.empty((n, self.total_replicas, 4), np.float64) if measure else None
            _check(lib.ising_pt_cluster_cont(self._h, k, n, _ptr(flipped), _ptr(meas) if measure else None))
            return (flipped, meas) if measure else flipped
        if not measure:
            _check(lib.ising_pt_cluster(self._h, k, n, _ptr(flipped), None, None))
            return flipped
        bond = np.empty((n, self.total_replicas), np.int64)
        mag = np.empty_like(bond)
        _check(lib.ising_pt_cluster(self._h, k, n, _ptr(flipped), _ptr(bond), _ptr(mag)))
        return flipped, bond, mag

    def cluster_count(self):
        v = C.c_int64()
        _check(lib.ising_pt_cluster_count(self._h, C.byref(v)))
        return v.value

    def autotune(self, sweeps=20):
        """Times the int8 thread-step layouts (every size at once, then size by
        size) or the overrelaxation sites per thread over warm-up sweeps and
        keeps the fastest; returns the largest size's (run width, runs per step,
        steps per tile), or (sites per thread, 0, 0) for XY / Heisenberg."""
        w, r, s = C.c_int32(), C.c_int32(), C.c_int32()
        _check(lib.ising_pt_autotune(self._h, int(sweeps), C.byref(w), C.byref(r), C.byref(s)))
        return w.value, r.value, s.value

    def get_layout(self, size_index=0):
        w, r, s = C.c_int32(), C.c_int32(), C.c_int32()
        _check(lib.ising_pt_get_layout(self._h, int(size_index), C.byref(w), C.byref(r), C.byref(s)))
        return w.value, r.value, s.value

    def prepare_graph(self, graph_sweeps):
        """Capture `graph_sweeps` sweeps as a HIP graph now (nothing runs);
        warmup/sample calls then replay it.  0 turns graph replay off."""
        _check(lib.ising_pt_prepare_graph(self._h, int(graph_sweeps)))

    def sweeps(self):
        v = C.c_int64()
        _check(lib.ising_pt_sweeps(self._h, C.byref(v)))
        return v.value

    def device_bytes(self):
        v = C.c_int64()
        _check(lib.ising_pt_device_bytes(self._h, C.byref(v)))
        return v.value


class Batch:
    """Several PTEnsembles advanced together by one captured graph per block of sweeps."""

    BRANCHES, WORKLIST, PERSISTENT = 0, 1, 2

    def __init__(self, members, graph_sweeps=32, mode=0):
        """mode: BRANCHES (each member's own launches on a graph branch per
        member), WORKLIST (launches over work lists of the members' blocks: one
        per pass for the continuous members, one per colour pass for the Ising
        member) or PERSISTENT (those work lists walked by persistent grids)."""
        self.members = list(members)
        arr = (_vp * len(self.members))(*[m._h for m in self.members])
        self._arr = arr
        self._h = _vp()
        _check(lib.ising_batch_create(C.byref(self._h), arr, len(self.members), int(graph_sweeps)))
        if mode:
            self.set_mode(mode)

    def set_mode(self, mode):
        _check(lib.ising_batch_set_mode(self._h, int(mode)))

    def close(self):
        if getattr(self, "_h", None):
            lib.ising_batch_destroy(self._h)
            self._h = None

    __del__ = close

    def warmup(self, sweeps):
        _check(lib.ising_batch_warmup(self._h, int(sweeps)))

    def sample(self, sweeps):
        _check(lib.ising_batch_sample(self._h, int(sweeps)))

    def sync(self):
        ms = C.c_double()
        _check(lib.ising_batch_sync(self._h, C.byref(ms)))
        return ms.value


def kz_run(n, j, t_start, t_end, tau_q, n_trials, warmup_sweeps=200, freeze_sweeps=0, freeze_temperature=0.01,
           seed=42):
    """One Kibble-Zurek quench of n_trials lattices (kz_gpu.rs:9-45) -> domain-wall density per trial."""
    rho = np.empty(n_trials, np.float64)
    _check(lib.ising_kz_run(n, j, t_start, t_end, tau_q, n_trials, warmup_sweeps, freeze_sweeps,
                            freeze_temperature, seed, _ptr(rho)))
    return rho


def run_kz_sweep(n, j, t_start, t_end, tau_q_values, n_trials, warmup_sweeps=200, freeze_sweeps=0,
                 freeze_temperature=0.01, base_seed=42):
    """run_kz_sweep_gpu (kz_gpu.rs:49-95) -> list of (tau_q, rho, rho_err, n_trials)."""
    tq = np.ascontiguousarray(tau_q_values, np.int64)
    rho, err = np.empty(tq.size), np.empty(tq.size)
    _check(lib.ising_kz_sweep(n, j, t_start, t_end, _ptr(tq), tq.size, n_trials, warmup_sweeps, freeze_sweeps,
                              freeze_temperature, base_seed, _ptr(rho), _ptr(err)))
    return [(int(t), float(r), float(e), n_trials) for t, r, e in zip(tq, rho, err)]


def run_coarsening(n, j=1.0, t_high=10.0, t_quench=0.5, warmup_sweeps=200, total_steps=50000, sample_every=100,
                   seed=42, n_trials=1):
    """run_coarsening_gpu (coarsening_gpu.rs:6-31) -> (steps, rho[sample, trial])."""
    cap = (total_steps + sample_every - 1) // sample_every
    steps = np.empty(max(cap, 1), np.int64)
    rho = np.empty((max(cap, 1), n_trials), np.float64)
    ns = C.c_int64()
    _check(lib.ising_coarsening_run(n, j, t_high, t_quench, warmup_sweeps, total_steps, sample_every, seed, n_trials,
                                    _ptr(steps), _ptr(rho), C.byref(ns)))
    return steps[:ns.value], rho[:ns.value]


def run_fss(sizes, j=1.0, h=0.0, t_min=4.0, t_max=5.0, t_steps=11, warmup_sweeps=1000, sample_sweeps=1000, seed=42):
    """run_fss_gpu (fss_gpu.rs:7-39) -> list of (n, [Observables ascending in T])."""
    sz = np.ascontiguousarray(sizes, np.int32)
    out = (Observables * (len(sizes) * t_steps))()
    _check(lib.ising_fss_anneal(_ptr(sz), sz.size, j, h, t_min, t_max, t_steps, warmup_sweeps, sample_sweeps, seed,
                                out))
    return [(int(L), [out[g * t_steps + k] for k in range(t_steps)]) for g, L in enumerate(sizes)]


def read_timeseries(path):
    """Reads a binary timeseries written by gpu_fss --binary 1.

    Returns a dict: model, L, replicas, measure_every, J, temps (by
    temperature index) and arrays shaped [measured sweep, replica]: temp_idx
    plus bond and mag (int64, Ising) or e, mx, my, mz (FP64 totals)."""
    raw = open(path, "rb").read()
    if raw[:8] != b"ISINGTS1":
        raise ValueError("not an ISINGTS1 file: %s" % path)
    version, model, L, R, me, kind = np.frombuffer(raw, np.int32, 6, 8)
    if version != 1:
        raise ValueError("unsupported timeseries version %d" % version)
    n = int(np.frombuffer(raw, np.int64, 1, 32)[0])
    J = float(np.frombuffer(raw, np.float64, 1, 40)[0])
    temps = np.frombuffer(raw, np.float64, R, 48).copy()
    off = 48 + 8 * R
    if kind == 0:
        rec = np.frombuffer(raw, np.int64, 3 * n * R, off).reshape(n, R, 3)
        fields = {"bond": rec[:, :, 0].copy(), "mag": rec[:, :, 1].copy(), "temp_idx": rec[:, :, 2].astype(np.int32)}
    else:
        rec = np.frombuffer(raw, np.float64, 5 * n * R, off).reshape(n, R, 5)
        fields = {"e": rec[:, :, 0].copy(), "mx": rec[:, :, 1].copy(), "my": rec[:, :, 2].copy(),
                  "mz": rec[:, :, 3].copy(),
                  "temp_idx": np.frombuffer(raw, np.int64, 5 * n * R, off).reshape(n, R, 5)[:, :, 4].astype(np.int32)}
    if len(raw) != off + (24 if kind == 0 else 40) * n * R:
        raise ValueError("truncated timeseries file: %s" % path)
    out = {"model": ["ising", "xy", "heisenberg", "ising_msc", "ising2d"][model], "L": int(L), "replicas": int(R),
           "measure_every": int(me), "J": J, "temps": temps}
    out.update(fields)
    return out


# reduce_gpu::* (src/cuda/reduce_gpu.rs:8-178).  `spins` is a lattice in DEVICE
# memory in the reference's layout: an integer device address, or any object
# with a data_ptr() method (a torch tensor on the GPU, for instance).
def _dptr(spins):
    return C.c_void_p(spins.data_ptr() if hasattr(spins, "data_ptr") else int(spins))


def load_reduce_kernels():
    _check(lib.ising_load_reduce_kernels())


def reduce_ising(spins, n, j=1.0):
    """(E total, M total) of an n^3 int8 +-1 device lattice (reduce_gpu.rs:23-66)."""
    e, m = C.c_double(), C.c_double()
    _check(lib.ising_reduce_ising(_dptr(spins), int(n), float(j), C.byref(e), C.byref(m)))
    return e.value, m.value


def measure_ising_gpu(spins, n, j=1.0):
    """(E per spin, |M| per spin) (reduce_gpu.rs:70-79)."""
    e, m = C.c_double(), C.c_double()
    _check(lib.ising_measure_ising_gpu(_dptr(spins), int(n), float(j), C.byref(e), C.byref(m)))
    return e.value, m.value


def reduce_continuous_mag(spins, n_sites, n_comp):
    """(Mx, My, Mz) totals of n_sites component-interleaved float spins (reduce_gpu.rs:83-135)."""
    v = [C.c_double() for _ in range(3)]
    _check(lib.ising_reduce_continuous_mag(_dptr(spins), int(n_sites), int(n_comp), *[C.byref(x) for x in v]))
    return tuple(x.value for x in v)


def reduce_continuous_energy(spins, n, n_comp, j=1.0):
    """Total energy -j sum S_i . S_fwd of an n^3 lattice (reduce_gpu.rs:138-163)."""
    e = C.c_double()
    _check(lib.ising_reduce_continuous_energy(_dptr(spins), int(n), int(n_comp), float(j), C.byref(e)))
    return e.value


def measure_continuous_gpu(spins, n, n_comp, j=1.0):
    """(E, Mx, My, Mz) totals (reduce_gpu.rs:167-178)."""
    v = [C.c_double() for _ in range(4)]
    _check(lib.ising_measure_continuous_gpu(_dptr(spins), int(n), int(n_comp), float(j), *[C.byref(x) for x in v]))
    return tuple(x.value for x in v)


def linspace_temperatures(t_min, t_max, n):
    out = np.empty(int(n), np.float64)
    _check(lib.ising_linspace_temperatures(t_min, t_max, int(n), _ptr(out)))
    return out


def replica_exchange(temperatures, energies, replica_to_temp, temp_to_replica, uniforms, even_odd):
    """In-place exchange on int32 permutation arrays; one uniform per attempted pair."""
    t = np.ascontiguousarray(temperatures, np.float64)
    e = np.ascontiguousarray(energies, np.float64)
    u = np.ascontiguousarray(uniforms, np.float64)
    acc = C.c_int32()
    for a in (replica_to_temp, temp_to_replica):
        if a.dtype != np.int32 or not a.flags.c_contiguous:
            raise TypeError("permutation arrays must be contiguous int32")
    _check(lib.ising_replica_exchange(_ptr(t), _ptr(e), len(t), _ptr(replica_to_temp), _ptr(temp_to_replica),
                                      _ptr(u), int(even_odd), C.byref(acc)))
    return acc.value


def jackknife_observables(e, m, beta, n3, n_blocks):
    e = np.ascontiguousarray(e, np.float64)
    m = np.ascontiguousarray(m, np.float64)
    out = np.empty(12, np.float64)
    _check(lib.ising_jackknife(_ptr(e), _ptr(m), len(e), beta, n3, int(n_blocks), _ptr(out)))
    return out


def tau_int(x, c=6.0):
    """Integrated autocorrelation time (Sokal automatic windowing) -> (tau, window)."""
    a = np.ascontiguousarray(x, np.float64)
    t, w = C.c_double(), C.c_int64()
    _check(lib.ising_tau_int(_ptr(a), len(a), c, C.byref(t), C.byref(w)))
    return t.value, w.value


def series_by_temperature(tidx, bond, mag, L, J=1.0, dims=3):
    """Per-temperature series of one lattice size from ensemble records
    (read_ising()[k][:, g, :] for size g): [total energies], [|m| per spin],
    the inputs of the reference's reweighting (analyze_gpu_fss.py:60-84:
    E = -J * bond, |M| / L^dims), indexed by temperature slot."""
    t = np.asarray(tidx).reshape(-1)
    e = -J * np.asarray(bond, np.float64).reshape(-1)
    m = np.abs(np.asarray(mag, np.float64).reshape(-1)) / float(L) ** dims
    k_max = int(t.max()) + 1 if t.size else 0
    return [e[t == k] for k in range(k_max)], [m[t == k] for k in range(k_max)]


def _closest_run_errors(e_sim, m_sim, delta_beta, beta_t, V, n_blocks):
    # single-histogram jackknife from one run (analyze_gpu_fss.py:118-141, 212-235)
    bs = len(e_sim) // n_blocks
    jk = np.zeros((3, n_blocks))
    for b in range(n_blocks):
        idx = np.concatenate([np.arange(0, b * bs), np.arange((b + 1) * bs, n_blocks * bs)])
        lw = -delta_beta * e_sim[idx]
        lw -= lw.max()
        w = np.exp(lw)
        w /= w.sum()
        mj = m_sim[idx]
        jm, jm2, jm4 = w @ mj, w @ mj ** 2, w @ mj ** 4
        jk[:, b] = (beta_t * V * (jm2 - jm ** 2), jm, 1.0 - jm4 / (3.0 * jm2 ** 2) if jm2 > 1e-15 else 0.0)
    return jk


def reweight_fine(energy_lists, mag_lists, betas, temps_fine, n, method="wham", n_blocks=20, max_iter=100,
                  tol=1e-8):
    """Observables on a fine temperature grid (analyze_gpu_fss.py:86-253):
    method "wham" (wham_reweight_fine: all runs, the WHAM solve on the GPU) or
    "single" (reweight_single_histogram: the run closest in temperature).
    energy_lists[k] / mag_lists[k]: total energies and |m| per spin of run k at
    betas[k]; n: linear size (V = n^3).  Returns a dict of arrays T, M, M2,
    M4, chi_conn, chi_err, Cv, U, E (per spin), M_err, U_err; the errors are the
    reference's single-histogram jackknife over n_blocks blocks of the closest
    run."""
    V = float(n) ** 3
    betas = np.asarray(betas, np.float64)
    T = np.asarray(temps_fine, np.float64)
    e_l = [np.asarray(e, np.float64) for e in energy_lists]
    m_l = [np.asarray(m, np.float64) for m in mag_lists]
    sim_T = 1.0 / betas
    if method == "wham":
        obs = [m_l, [m ** 2 for m in m_l], [m ** 4 for m in m_l], e_l, [e ** 2 for e in e_l]]
        _, _, means = wham(e_l, betas, obs, 1.0 / T, max_iter, tol)
        M, M2, M4, E, E2 = (means[:, j] for j in range(5))
    elif method == "single":
        M, M2, M4, E, E2 = (np.zeros(T.size) for _ in range(5))
        for i, t in enumerate(T):
            k = int(np.argmin(np.abs(sim_T - t)))
            lw = -(1.0 / t - betas[k]) * e_l[k]
            lw -= lw.max()
            w = np.exp(lw)
            w /= w.sum()
            M[i], M2[i], M4[i] = w @ m_l[k], w @ m_l[k] ** 2, w @ m_l[k] ** 4
            E[i], E2[i] = w @ e_l[k], w @ e_l[k] ** 2
    else:
        raise ValueError("method must be 'wham' or 'single'")
    bt = 1.0 / T
    chi = bt * V * (M2 - M ** 2)
    cv = bt ** 2 * (E2 - E ** 2) / V
    U = np.where(M2 > 1e-15, 1.0 - M4 / (3.0 * np.where(M2 > 1e-15, M2, 1.0) ** 2), 0.0)
    chi_err, M_err, U_err = (np.zeros(T.size) for _ in range(3))
    f = (n_blocks - 1.0) / n_blocks
    for i, t in enumerate(T):
        k = int(np.argmin(np.abs(sim_T - t)))
        jk = _closest_run_errors(e_l[k], m_l[k], 1.0 / t - betas[k], 1.0 / t, V, n_blocks)
        chi_err[i] = np.sqrt(f * np.sum((jk[0] - chi[i]) ** 2))
        M_err[i] = np.sqrt(f * np.sum((jk[1] - M[i]) ** 2))
        U_err[i] = np.sqrt(f * np.sum((jk[2] - U[i]) ** 2))
    return {"T": T, "M": M, "M2": M2, "M4": M4, "chi_conn": chi, "chi_err": chi_err, "Cv": cv, "U": U,
            "E": E / V, "M_err": M_err, "U_err": U_err}


def binder_crossing(t1, u1, t2, u2):
    """Crossing temperature of two Binder-cumulant curves (the rule of
    analyze_gpu_fss.py:256-277: not-a-knot cubic splines on a 2000-point common
    grid, first sign change, linear root); None when they do not cross."""
    a1, b1 = np.ascontiguousarray(t1, np.float64), np.ascontiguousarray(u1, np.float64)
    a2, b2 = np.ascontiguousarray(t2, np.float64), np.ascontiguousarray(u2, np.float64)
    if a1.shape != b1.shape or a2.shape != b2.shape:
        raise ValueError("each curve needs as many cumulants as temperatures")
    tc, found = C.c_double(), C.c_int32()
    _check(lib.ising_binder_crossing(_ptr(a1), _ptr(b1), len(a1), _ptr(a2), _ptr(b2), len(a2), C.byref(tc),
                                     C.byref(found)))
    return tc.value if found.value else None


def binder_crossings(curves):
    """find_binder_crossings (analyze_gpu_fss.py:256-277): curves maps L to
    (temperatures, U4); returns [(L1, L2, T_c)] for adjacent sizes that cross."""
    sizes = sorted(curves)
    out = []
    for n1, n2 in zip(sizes, sizes[1:]):
        tc = binder_crossing(curves[n1][0], curves[n1][1], curves[n2][0], curves[n2][1])
        if tc is not None:
            out.append((n1, n2, tc))
    return out


def wham(energy_lists, betas, observables=(), beta_targets=(), max_iter=100, tol=1e-8):
    """Multiple-histogram reweighting on the GPU (reweighting.py:45-108).

    energy_lists[k]: total energies of run k at betas[k]; observables: list of
    lists shaped like energy_lists.  Returns (f, iterations, means[n_targets, n_obs]).
    """
    K = len(energy_lists)
    n_k = np.array([len(e) for e in energy_lists], np.int64)
    E = np.ascontiguousarray(np.concatenate(energy_lists), np.float64)
    obs = (np.ascontiguousarray(np.concatenate([np.concatenate(o) for o in observables]), np.float64)
           if len(observables) else np.zeros(1))
    b = np.ascontiguousarray(betas, np.float64)
    bt = np.ascontiguousarray(beta_targets, np.float64) if len(beta_targets) else np.zeros(1)
    f = np.empty(K)
    out = np.empty(max(1, len(beta_targets) * len(observables)))
    it = C.c_int32()
    _check(lib.ising_wham(K, _ptr(n_k), _ptr(b), _ptr(E), len(observables), _ptr(obs), max_iter, tol, _ptr(f),
                          C.byref(it), len(beta_targets), _ptr(bt), _ptr(out)))
    return f, it.value, out[: len(beta_targets) * len(observables)].reshape(len(beta_targets), len(observables))
