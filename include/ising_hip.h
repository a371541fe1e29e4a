/*
 * libising_hip -- MI355X (gfx950) parallel-tempering engine for 3D Ising, XY
 * and Heisenberg spins.  C ABI: plain pointers and sizes, no C++ types.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative ISING_E_* code on
 *    failure; ising_last_error() then gives a message (thread-local).
 *  - No C++ exception crosses the ABI.
 *  - The library owns device memory; callers own every host buffer.
 *  - One HIP stream per handle.  A handle is not thread-safe; distinct
 *    handles are.
 *  - Spin arrays exchanged with the caller use the reference's layout:
 *    z-major, index (z*L + y)*L + x; Ising as int8 +-1, continuous spins
 *    component-interleaved float (x, y[, z]) per site.
 *
 * Reference interfaces replaced (faulknco/ising-rs):
 *  - LatticeGpu                     src/cuda/lattice_gpu.rs:9-306
 *  - ContinuousGpuLattice           src/cuda/gpu_lattice_continuous.rs:9-250
 *  - reduce_gpu::*                  src/cuda/reduce_gpu.rs:8-178
 *  - parallel_tempering::*          src/parallel_tempering.rs:5-46
 *  - the PT loop of gpu_fss         src/bin/gpu_fss.rs:166-582 (ensemble API)
 *  - jackknife_observables          src/bin/gpu_fss.rs:327-412
 */
#ifndef ISING_HIP_H
#define ISING_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ISING_OK 0
#define ISING_E_ARG (-1)     /* invalid argument */
#define ISING_E_HIP (-2)     /* HIP runtime error (device, allocation, launch) */
#define ISING_E_STATE (-3)   /* call not valid in the handle's current state */
#define ISING_E_FULL (-4)    /* measurement buffer full: drain records first */

#define ISING_MODEL_ISING 0
#define ISING_MODEL_XY 1
#define ISING_MODEL_HEISENBERG 2
/* 3D Ising with the replicas of a size packed across the bits of a 32-bit
 * word per site (multispin coding, 1 bit per spin).  Requires h = 0, J > 0.
 * Same records and spin API as ISING_MODEL_ISING; its own random stream. */
#define ISING_MODEL_ISING_MSC 3
/* 2D square-lattice Ising (L x L, 4 neighbours): the same checkerboard,
 * random-number and record scheme with z removed; for the exact 2D
 * (Onsager / enumeration) checks.  Spin arrays hold L^2 int8. */
#define ISING_MODEL_ISING2D 4

/* Message for the last failing call on this thread ("" if none). */
const char* ising_last_error(void);
/* Library version string. */
const char* ising_version(void);
/* Number of visible HIP devices (0 when none); never fails. */
int ising_device_count(void);

/* ------------------------------------------------------------------------
 * Single Ising lattice: mirror of LatticeGpu (src/cuda/lattice_gpu.rs).
 * ------------------------------------------------------------------------ */
typedef struct ising_lattice ising_lattice_t;

/* LatticeGpu::new(n, seed) (lattice_gpu.rs:22).  n even, 2 <= n <= 2046.
 * Hot start from the library's counter-based stream (the reference draws an
 * unseeded host start, lattice_gpu.rs:38-40). */
int ising_lattice_create(ising_lattice_t** out, int n, uint64_t seed);
void ising_lattice_destroy(ising_lattice_t* lat);
int ising_lattice_size(const ising_lattice_t* lat);
/* LatticeGpu::step(beta, j, h) (lattice_gpu.rs:85): one sweep, colour 0 then 1. Async. */
int ising_lattice_step(ising_lattice_t* lat, float beta, float j, float h);
/* LatticeGpu::warm_up (lattice_gpu.rs:184). */
int ising_lattice_warm_up(ising_lattice_t* lat, float beta, float j, float h, size_t sweeps);
/* LatticeGpu::get_spins (lattice_gpu.rs:110): out holds n^3 int8 (+-1). */
int ising_lattice_get_spins(ising_lattice_t* lat, int8_t* out);
/* Not in the reference: upload a configuration (n^3 int8 +-1). */
int ising_lattice_set_spins(ising_lattice_t* lat, const int8_t* in);
/* LatticeGpu::measure_gpu(j) (lattice_gpu.rs:116): E per spin (forward bonds,
 * h excluded as in reduce_energy_ising) and |M| per spin. */
int ising_lattice_measure_gpu(ising_lattice_t* lat, float j, double* e_per_spin, double* m_abs_per_spin);
/* LatticeGpu::energy_gpu(j) (lattice_gpu.rs:159): total energy. */
int ising_lattice_energy_gpu(ising_lattice_t* lat, float j, double* e_total);
/* Exact integer observables: forward-bond sum B (E = -J B) and sum of spins. */
int ising_lattice_measure_raw(ising_lattice_t* lat, int64_t* bond, int64_t* mag);

typedef struct {
    double temperature;
    double energy;          /* <E> per spin */
    double magnetisation;   /* <|M|> per spin */
    double heat_capacity;
    double susceptibility;  /* from the signed magnetisation, lattice_gpu.rs:240 */
    double m2;
    double m4;
} ising_observables_t;

/* LatticeGpu::measure(T, j, h, samples) (lattice_gpu.rs:193-251); E includes -h M
 * as energy_magnetisation_host does (lattice_gpu.rs:285-306). */
int ising_lattice_measure(ising_lattice_t* lat, double temperature, double j, double h, size_t samples,
                          ising_observables_t* out);
/* LatticeGpu::domain_wall_density (lattice_gpu.rs:254-280). */
int ising_lattice_domain_wall_density(ising_lattice_t* lat, double* out);
/* Global sweep counter (number of sweeps applied so far). */
int ising_lattice_sweeps(const ising_lattice_t* lat, int64_t* out);

/* ------------------------------------------------------------------------
 * Single continuous-spin lattice: mirror of ContinuousGpuLattice
 * (src/cuda/gpu_lattice_continuous.rs).  n_comp = 2 (XY) or 3 (Heisenberg).
 * ------------------------------------------------------------------------ */
typedef struct ising_cont ising_cont_t;

/* ContinuousGpuLattice::new(n, n_comp, seed, device) (gpu_lattice_continuous.rs:24):
 * uniform random start from the library's counter-based stream. */
int ising_cont_create(ising_cont_t** out, int n, int n_comp, uint64_t seed);
void ising_cont_destroy(ising_cont_t* lat);
/* ContinuousGpuLattice::sweep(beta, j, delta, n_overrelax) (:112): n_overrelax
 * overrelaxation sweeps, then one Metropolis sweep (colour 0 then 1). Async.
 * delta: Heisenberg cap half-angle / XY maximum rotation (radians). */
int ising_cont_sweep(ising_cont_t* lat, float beta, float j, float delta, size_t n_overrelax);
/* ContinuousGpuLattice::measure_gpu(j) (:177): totals E, Mx, My, Mz (FP64 sums). */
int ising_cont_measure_gpu(ising_cont_t* lat, float j, double* e, double* mx, double* my, double* mz);
/* ContinuousGpuLattice::measure_raw() (:243) = measure_gpu(1.0). */
int ising_cont_measure_raw(ising_cont_t* lat, double* e, double* mx, double* my, double* mz);
/* Spins as n^3 * n_comp floats, component-interleaved (the `spins` field, :13). */
int ising_cont_get_spins(ising_cont_t* lat, float* out);
int ising_cont_set_spins(ising_cont_t* lat, const float* in);
/* Single colour passes (not in the reference; used by parity tests): one
 * Metropolis or one overrelaxation pass over `color` at the current sweep
 * index, without advancing it. */
int ising_cont_metropolis_pass(ising_cont_t* lat, int color, float beta, float j, float delta);
int ising_cont_overrelax_pass(ising_cont_t* lat, int color, float j);
int ising_cont_sweeps(const ising_cont_t* lat, int64_t* out);

/* ------------------------------------------------------------------------
 * reduce_gpu::* (src/cuda/reduce_gpu.rs:8-178): totals of a lattice the
 * caller holds in DEVICE memory, in the reference's layout (z-major; Ising
 * int8 +-1, continuous spins component-interleaved float).  Each call
 * allocates its partials, runs on the device's default stream (the caller
 * orders any work on other streams that writes d_spins) and returns after
 * copying the totals to the host.  Ising sums are exact (int64); continuous
 * sums are FP64 in a fixed order (the reference adds FP32 block partials).
 * Energies count the three forward bonds of each site, without a field.
 * ------------------------------------------------------------------------ */
/* load_reduce_kernels (reduce_gpu.rs:8-21): fails with ISING_E_HIP when no device is visible. */
int ising_load_reduce_kernels(void);
/* reduce_ising (reduce_gpu.rs:23-66): E total = -j * sum s_i (s_x+1 + s_y+1 + s_z+1), M total.  1 <= n <= 1024. */
int ising_reduce_ising(const int8_t* d_spins, int32_t n, float j, double* e_total, double* m_total);
/* measure_ising_gpu (reduce_gpu.rs:70-79): E per spin and |M| per spin. */
int ising_measure_ising_gpu(const int8_t* d_spins, int32_t n, float j, double* e_per_spin, double* m_abs_per_spin);
/* reduce_continuous_mag (reduce_gpu.rs:83-135): Mx, My, Mz totals over n_sites
 * sites (mz = 0 for n_comp = 2).  0 <= n_sites <= 2^30. */
int ising_reduce_continuous_mag(const float* d_spins, int64_t n_sites, int32_t n_comp, double* mx, double* my,
                                double* mz);
/* reduce_continuous_energy (reduce_gpu.rs:138-163): -j * sum S_i . S_fwd over an n^3 lattice. */
int ising_reduce_continuous_energy(const float* d_spins, int32_t n, int32_t n_comp, float j, double* e_total);
/* measure_continuous_gpu (reduce_gpu.rs:167-178): (E, Mx, My, Mz) totals. */
int ising_measure_continuous_gpu(const float* d_spins, int32_t n, int32_t n_comp, float j, double* e, double* mx,
                                 double* my, double* mz);
/* No counterpart in the reference: ising_reduce_ising for callers that cannot
 * allocate (code under HIP graph capture, tight loops, harnesses that forbid
 * hipMalloc).  Allocates nothing: the partial sums live in a fixed 1 MiB
 * device array, one per device, held until the process ends.  Writes
 * d_out[0] = E total and d_out[1] = M total (caller-provided DEVICE memory, 2
 * doubles) asynchronously on `stream` (a hipStream_t; NULL = the default
 * stream).  n^3 <= 2^28 (n <= 645), else ISING_E_ARG.  Calls that overlap on
 * one device race on the array: serialise them. */
int ising_reduce_ising_noalloc(const int8_t* d_spins, int32_t n, float j, double* d_out, void* stream);

/* ------------------------------------------------------------------------
 * Parallel-tempering ensemble: every lattice size and replica of a run in one
 * set of launches; exchange, measurement and the sweep counter on the device.
 * Replaces the host loop of gpu_fss (gpu_fss.rs:186-286 / 441-545).
 * ------------------------------------------------------------------------ */
typedef struct ising_pt ising_pt_t;

typedef struct {
    double j;                /* coupling J (gpu_fss uses 1.0) */
    double h;                /* field, Ising only (gpu_fss uses 0.0) */
    int32_t exchange_every;  /* --exchange-every, >= 1 */
    int32_t measure_every;   /* --measure-every, >= 1 */
    double delta;            /* --delta: continuous proposal size (radians): Heisenberg cap
                                half-angle, XY maximum rotation */
    int32_t n_overrelax;     /* --overrelax: OR passes per sweep (continuous) */
    int32_t renorm_every;    /* continuous: renormalise every k sweeps (0 = never) */
    int32_t record_capacity; /* measured sweeps buffered on the device before a drain */
    int32_t adapt_every;     /* continuous: during warm-up, rescale each temperature's
                                delta every k sweeps towards adapt_target (0 = fixed) */
    double adapt_target;     /* target Metropolis acceptance for the adaptation */
    int32_t graph_sweeps;    /* sweeps per captured HIP graph replayed by warmup/sample
                                (0 = launch every kernel directly) */
    int32_t ising_spt;       /* Ising: thread steps per block tile, 1..4 (0 = automatic);
                                results do not depend on it */
    int32_t ising_width;     /* Ising: run width where L % 64 == 0, 16 or 32 sites
                                (0 = automatic); results do not depend on it */
    int32_t ising_runs;      /* Ising: 16-site runs per thread step, 1 or 2 (0 = automatic);
                                results do not depend on it */
    /* int8 Ising: 0 = two launches per sweep (one per colour); 8, 16 or 32 =
     * one launch per sweep, each block streaming that many z planes (at most
     * L) of its replicas through LDS and updating both colours (colour 0 of
     * the two planes beyond its ends is recomputed); lattices with L < 32 are
     * staged whole.  Results are identical either way.  Needs every size to
     * be 3D with L < 32 (L % 8 == 0) or L % 32 == 0 and L <= 128 (ignored
     * otherwise). */
    int32_t ising_fused;
} ising_pt_opts_t;

/* Defaults: j=1, h=0, exchange_every=10, measure_every=1, delta=0.5,
 * n_overrelax=5, renorm_every=100, record_capacity=4096, adapt_every=0,
 * adapt_target=0.5, graph_sweeps=32, ising_spt=0, ising_width=0, ising_runs=0, ising_fused=0. */
void ising_pt_default_opts(ising_pt_opts_t* opts);

/* model: ISING_MODEL_*.  sizes: n_sizes even L values.  temps: `replicas`
 * temperatures shared by every size (replica r starts at temperature r).
 * Replica r of size L uses stream (L << 16) | r under key `seed`. */
int ising_pt_create(ising_pt_t** out, int model, const int32_t* sizes, int32_t n_sizes, int32_t replicas,
                    const double* temps, uint64_t seed, const ising_pt_opts_t* opts);
void ising_pt_destroy(ising_pt_t* pt);
/* Warm-up sweeps: no exchange, no measurement (gpu_fss.rs:199-210). */
int ising_pt_warmup(ising_pt_t* pt, int64_t sweeps);
/* Sample sweeps: measure when (s+1) % measure_every == 0, exchange when
 * (s+1) % exchange_every == 0 at round s / exchange_every (gpu_fss.rs:235-280).
 * Fails with ISING_E_FULL if the records would not fit. */
int ising_pt_sample(ising_pt_t* pt, int64_t sweeps);
/* Wait for queued work; return the device time of the last warmup/sample call (ms, HIP events). */
int ising_pt_sync(ising_pt_t* pt, double* last_ms);
/* Number of measured sweeps waiting in the device buffer. */
int ising_pt_pending(ising_pt_t* pt, int64_t* n_measured_sweeps);
/* Drain Ising records: for each measured sweep, one record per replica in
 * (size, replica) order: temperature index, bond sum B, magnetisation sum M.
 * Arrays hold max_sweeps * total_replicas entries; *n_out = sweeps drained. */
int ising_pt_read_ising(ising_pt_t* pt, int64_t max_sweeps, int32_t* tidx, int64_t* bond, int64_t* mag,
                        int64_t* n_out);
/* Drain continuous records: E total (FP64), Mx, My, Mz totals, temperature index. */
int ising_pt_read_cont(ising_pt_t* pt, int64_t max_sweeps, int32_t* tidx, double* e, double* mx, double* my,
                       double* mz, int64_t* n_out);
/* Permutations of one size: r2t[replica] -> temperature index, t2r[temp] -> replica. */
int ising_pt_get_perm(ising_pt_t* pt, int32_t size_index, int32_t* r2t, int32_t* t2r);
/* Exchange statistics since creation. */
int ising_pt_exchange_stats(ising_pt_t* pt, int64_t* tries, int64_t* accepts);
/* Configuration of one replica: Ising int8 +-1 (L^3) or float (L^3 * ncomp). */
int ising_pt_get_spins_i8(ising_pt_t* pt, int32_t size_index, int32_t replica, int8_t* out);
int ising_pt_set_spins_i8(ising_pt_t* pt, int32_t size_index, int32_t replica, const int8_t* in);
int ising_pt_get_spins_f32(ising_pt_t* pt, int32_t size_index, int32_t replica, float* out);
int ising_pt_set_spins_f32(ising_pt_t* pt, int32_t size_index, int32_t replica, const float* in);
/* Current proposal size per temperature of one size (continuous models). */
int ising_pt_get_deltas(ising_pt_t* pt, int32_t size_index, float* deltas);
/* ---- Learned proposal (continuous models; no counterpart in the reference) ----
 * An MLP (6 rotation-invariant neighbour features -> H -> H -> H -> 3, ReLU,
 * BF16 weights, FP32 accumulation, H in {8, 16, 32, 64}) gives a two-component
 * von Mises (XY) / Fisher (Heisenberg) mixture around the local field; moves
 * are accepted by exact Metropolis-Hastings including the proposal ratio.
 * Parameters: flat FP32 vector W1[H][6], b1[H], W2[H][H], b2[H], W3[H][H],
 * b3[H], W4[3][H], b4[3] (row-major). */
int ising_pt_learn_init(ising_pt_t* pt, int32_t hidden, uint64_t seed);
/* Number of network parameters (after ising_pt_learn_init). */
int ising_pt_learn_size(ising_pt_t* pt, int64_t* n_params);
int ising_pt_learn_get_params(ising_pt_t* pt, float* out, int64_t n);
int ising_pt_learn_set_params(ising_pt_t* pt, const float* in, int64_t n);
/* Adam steps on batches of sites drawn from the ensemble's current
 * configurations; *mean_loss = mean -log q of the last batch. */
int ising_pt_learn_train(ising_pt_t* pt, int32_t steps, int32_t batch, double lr, double* mean_loss);
/* Loss and mean gradient on explicit samples (x: n * 6 features, c: n values of s.h^). */
int ising_pt_learn_eval(ising_pt_t* pt, int64_t n, const float* x, const float* c, double* mean_loss, float* grad);
/* Network outputs (n * 3 floats: two concentration logits and the mixture
 * logit) for n feature rows, evaluated as the fused sweep does: BF16 weights,
 * inputs and activations on the matrix cores, FP32 accumulation. */
int ising_pt_learn_infer(ising_pt_t* pt, int64_t n, const float* x, float* out);
/* 1: sweeps use the learned proposal for the Metropolis passes; 0: the cap/rotation proposal. */
int ising_pt_learn_enable(ising_pt_t* pt, int32_t on);
/* 1: sweeps use heat-bath passes (no counterpart in the reference): every
 * site draws its spin from the exact single-site conditional exp(beta J s.h)
 * -- Fisher on S^2, von Mises on S^1 (Best-Fisher rejection, at most 7
 * attempts; when all fail the site keeps its spin) -- in place of the
 * Metropolis passes; overrelaxation passes stay.  The learned proposal, which
 * depends on the neighbours only, cannot mix faster per pass.  Turns the
 * learned proposal off; 0 returns to the cap/rotation Metropolis passes. */
int ising_pt_heat_bath(ising_pt_t* pt, int32_t on);
/* Learned Metropolis-Hastings moves attempted / accepted since creation. */
int ising_pt_mh_stats(ising_pt_t* pt, int64_t* tries, int64_t* accepts);
/* One learned (or, with ising_pt_heat_bath on, heat-bath) pass over one
 * colour at sweep index t, without measurement, exchange or advancing the
 * sweep counter (used by exactness tests). */
int ising_pt_metropolis_pass(ising_pt_t* pt, int32_t color, int64_t t);
/* XY spin-stiffness sums of every replica's current state (no counterpart in
 * the reference): out[6 * replica + k] = C_x, C_y, C_z, J_x, J_y, J_z with, over
 * the forward bonds (i, i + d), C_d = sum cos(theta_i - theta_j) and
 * J_d = sum sin(theta_i - theta_j).  The helicity modulus along d at inverse
 * temperature beta is (<C_d> - beta <J_d^2>) / N. */
int ising_pt_helicity(ising_pt_t* pt, double* out);

/* Annealing schedule (int8 Ising ensembles, before sampling): n sweeps, sweep
 * k with inverse temperature betas[k] for every replica (acceptance tables
 * built on the host in FP64 as for the ensemble).  When bond / mag are not
 * NULL they receive [n][total replicas] int64 bond sums and magnetisations
 * measured on the device after every sweep.  Serves the drivers below. */
int ising_pt_anneal(ising_pt_t* pt, const double* betas, int64_t n, int64_t* bond, int64_t* mag);

/* ------------------------------------------------------------------------
 * Cluster updates: the GPU Wolff the reference plans as a label-propagation
 * step (docs/plans/2026-03-09-gpu-compute-optimizations-design.md:82-116;
 * WolffGpuLattice, gpu-compute-optimizations-impl.md:1041-1330), with
 * the CPU algorithms of src/wolff.rs:27-77 (Ising) and src/xy/wolff.rs:16-73
 * (XY embedding).  One bond configuration per update (Fortuin-Kasteleyn bonds
 * with probability 1 - exp(-2 beta |J|) between satisfied neighbours; for
 * continuous spins 1 - exp(-2 beta J a_i a_j) with a = s . r along a random
 * unit vector r per replica and update), labelled on the device by union-find;
 * then either every cluster flips with probability 1/2 (Swendsen-Wang) or the
 * cluster of one random seed site flips (Wolff).  Random draws are Philox,
 * keyed by (seed, update index, size, replica): results are identical for any
 * tiling.  Ising needs h = 0 and int8 storage (not ISING_MODEL_ISING_MSC).
 * ------------------------------------------------------------------------ */
#define ISING_CLUSTER_SW 0
#define ISING_CLUSTER_WOLFF 1

/* n cluster updates of every replica of an int8 Ising ensemble (3D or 2D) at
 * its current temperature; may be interleaved with warm-up / sample calls.
 * When not NULL, flipped / bond / mag receive [n][total replicas] int64:
 * sites flipped by each update, then the bond sum and magnetisation after it. */
int ising_pt_cluster(ising_pt_t* pt, int32_t kind, int64_t n, int64_t* flipped, int64_t* bond, int64_t* mag);
/* The same for XY / Heisenberg ensembles; meas (when not NULL) receives
 * [n][total replicas][4] doubles E, Mx, My, Mz after each update. */
int ising_pt_cluster_cont(ising_pt_t* pt, int32_t kind, int64_t n, int64_t* flipped, double* meas);
/* Cluster updates applied so far (the update index of the next one). */
int ising_pt_cluster_count(const ising_pt_t* pt, int64_t* out);
/* WolffGpuLattice::step(beta, j, rng) (gpu-compute-optimizations-impl.md:1150-1248)
 * on a single lattice: one update of `kind`; *flipped (when not NULL) receives
 * the cluster size (Wolff) or the number of sites flipped (Swendsen-Wang).
 * Measurement, warm-up and spins use the lattice's other calls. */
int ising_lattice_cluster_step(ising_lattice_t* lat, int32_t kind, float beta, float j, int64_t* flipped);
/* The same for a single XY / Heisenberg lattice. */
int ising_cont_cluster_step(ising_cont_t* lat, int32_t kind, float beta, float j, int64_t* flipped);

/* ------------------------------------------------------------------------
 * The other LatticeGpu callers as batched drivers: every trial / size runs as
 * a replica / group of one ensemble, so one launch per colour advances all.
 * ------------------------------------------------------------------------ */

/* One Kibble-Zurek quench of n_trials independent L = n lattices
 * (src/cuda/kz_gpu.rs:9-45): warm up at t_start, ramp linearly to t_end over
 * tau_q sweeps (kibble_zurek.rs:78-84), optionally freeze, and return each
 * trial's domain-wall density (fraction of forward bonds across a wall,
 * lattice_gpu.rs:254-280).  Trials are the Philox streams of one seed. */
int ising_kz_run(int32_t n, double j, double t_start, double t_end, int64_t tau_q, int32_t n_trials,
                 int64_t warmup_sweeps, int64_t freeze_sweeps, double freeze_temperature, uint64_t seed,
                 double* rho);

/* run_kz_sweep_gpu (src/cuda/kz_gpu.rs:49-95): for each tau_q, n_trials
 * quenches keyed by base_seed + tau_q (kibble_zurek.rs:102-106); mean and
 * standard error of the density (kibble_zurek.rs:86-100). */
int ising_kz_sweep(int32_t n, double j, double t_start, double t_end, const int64_t* tau_q, int32_t n_tau,
                   int32_t n_trials, int64_t warmup_sweeps, int64_t freeze_sweeps, double freeze_temperature,
                   uint64_t base_seed, double* rho, double* rho_err);

/* run_coarsening_gpu (src/cuda/coarsening_gpu.rs:6-31): warm up at t_high,
 * quench to t_quench, sample the domain-wall density after steps 0,
 * sample_every, ... < total_steps.  rho_out is [sample][trial]; n_trials = 1 is
 * the reference's single lattice.  steps_out and rho_out hold at least
 * ceil(total_steps / sample_every) (x n_trials) entries. */
int ising_coarsening_run(int32_t n, double j, double t_high, double t_quench, int64_t warmup_sweeps,
                         int64_t total_steps, int64_t sample_every, uint64_t seed, int32_t n_trials,
                         int64_t* steps_out, double* rho_out, int64_t* n_samples);

/* run_fss_gpu (src/cuda/fss_gpu.rs:7-39): annealed temperature sweep from
 * t_max down to t_min in t_steps points; at each, warm_up then measure
 * (LatticeGpu::measure, lattice_gpu.rs:193-252) with every size advancing
 * together.  out is [size][temperature index, ascending T]. */
int ising_fss_anneal(const int32_t* sizes, int32_t n_sizes, double j, double h, double t_min, double t_max,
                     int32_t t_steps, int64_t warmup_sweeps, int64_t sample_sweeps, uint64_t seed,
                     ising_observables_t* out);

/* Autotuner (no counterpart in the reference; the reference plans one,
 * docs/plans/2026-03-08-gpu-phase2-performance.md:601-655).  int8 Ising
 * (3D and 2D): times each thread-step layout (16-site runs, one or two per
 * step, or 32-site runs, x 1/2/4 steps per block tile) for every size at once,
 * then refines size by size, largest first, keeping a size's candidate when it
 * is faster by more than 0.5%; the outputs are the largest size's layout
 * (ising_pt_get_layout gives every size's).  XY / Heisenberg: times the
 * overrelaxation pass with one and two sites per thread (*width = the choice,
 * *runs = *spt = 0).  Results do not depend on any of these, so the timed
 * sweeps are ordinary warm-up sweeps (each candidate's first sweep is not
 * timed; about 50 * (sweeps + 1) sweeps for an int8 ensemble of five sizes).
 * When the environment variable ISING_HIP_TUNE_CACHE names a file, a choice
 * stored there for the same model, replica count, sizes and device
 * architecture is applied without timing anything, and a new choice is
 * appended.  Before sampling only; not for multispin-coded ensembles.
 * Outputs may be NULL. */
int ising_pt_autotune(ising_pt_t* pt, int32_t sweeps, int32_t* width, int32_t* runs, int32_t* spt);
/* Thread-step layout of one size (int8 Ising) or the overrelaxation sites per
 * thread in *width (XY / Heisenberg). */
int ising_pt_get_layout(ising_pt_t* pt, int32_t size_index, int32_t* width, int32_t* runs, int32_t* spt);
/* Sweeps per captured HIP graph (no counterpart in the reference): sets
 * graph_sweeps and captures the graph now, without running a sweep, so that
 * a later warmup/sample call of k sweeps replays it k / graph_sweeps times
 * (and launches the remaining k % graph_sweeps sweeps directly).  0 turns
 * graph replay off.  Results do not depend on it. */
int ising_pt_prepare_graph(ising_pt_t* pt, int32_t graph_sweeps);
/* Sweeps applied so far (warm-up + sample). */
int ising_pt_sweeps(const ising_pt_t* pt, int64_t* out);
/* Bytes of spin storage on the device. */
int ising_pt_device_bytes(const ising_pt_t* pt, int64_t* out);

/* ------------------------------------------------------------------------
 * Batch: several ensembles (e.g. Ising, XY and Heisenberg of one batched run)
 * advanced together.  One captured HIP graph per `graph_sweeps` sweeps holds
 * a branch per member on the member's own stream, so one graph launch runs
 * every model concurrently.  Members keep their own records, permutations
 * and read functions; results are identical to running them separately.
 * Not in the reference (it runs one model and one size at a time).
 * ------------------------------------------------------------------------ */
typedef struct ising_batch ising_batch_t;
int ising_batch_create(ising_batch_t** out, ising_pt_t* const* members, int32_t n, int32_t graph_sweeps);
void ising_batch_destroy(ising_batch_t* b);
int ising_batch_warmup(ising_batch_t* b, int64_t sweeps);
int ising_batch_sample(ising_batch_t* b, int64_t sweeps);
int ising_batch_sync(ising_batch_t* b, double* last_ms);

/* How a batch launches its members' passes.  No counterpart in the reference.
 *   ISING_BATCH_BRANCHES (default): each member's own launches, on a graph
 *     branch per member.
 *   ISING_BATCH_WORKLIST: launches over work lists of (member, size,
 *     replica, tile) blocks: one per pass for the continuous members
 *     (overrelaxation, Metropolis, the closing step) and one per colour pass
 *     of the Ising member (2 n_overrelax + 5 launches per sweep with all
 *     three models).  At most one int8 Ising ensemble (two-launch sweeps) and
 *     two continuous ensembles (XY, Heisenberg, Metropolis passes) with equal
 *     n_overrelax.
 *   ISING_BATCH_PERSISTENT: the same work list walked by a persistent grid
 *     (the blocks the device holds at once).
 * Lattices, records and permutations equal the branch mode's bit for bit. */
#define ISING_BATCH_BRANCHES 0
#define ISING_BATCH_WORKLIST 1
#define ISING_BATCH_PERSISTENT 2
int ising_batch_set_mode(ising_batch_t* b, int32_t mode);

/* ------------------------------------------------------------------------
 * Host-side statistics and PT helpers (no GPU needed).
 * ------------------------------------------------------------------------ */
/* linspace_temperatures (parallel_tempering.rs:39-46). */
int ising_linspace_temperatures(double t_min, double t_max, int32_t n, double* out);
/* replica_exchange (parallel_tempering.rs:5-36) on host arrays; one caller-
 * supplied uniform per attempted pair (uniforms[k] for the k-th pair).
 * Returns accepted swaps in *accepted. */
int ising_replica_exchange(const double* temps, const double* energies, int32_t n, int32_t* r2t, int32_t* t2r,
                           const double* uniforms, int32_t even_odd, int32_t* accepted);
/* jackknife_observables (gpu_fss.rs:327-412): out[12] = E, E_err, M, M_err, M2,
 * M2_err, M4, M4_err, Cv, Cv_err, chi, chi_err. */
int ising_jackknife(const double* e, const double* m, int64_t n, double beta, double n3, int32_t n_blocks,
                    double* out);

/* Multiple-histogram (WHAM / Ferrenberg-Swendsen) reweighting on the device in
 * FP64, restating analysis/scripts/reweighting.py:45-108 (wham_reweight):
 * K runs at betas[k] with n_k[k] samples each; `energies` holds the total
 * energies of all runs concatenated (N = sum n_k), `obs` holds n_obs
 * observables, each an array of N values (observable-major).  Iterates the
 * free energies until max |df| < tol or max_iter (f_out[K], f_out[0] = 0,
 * *iters = iterations used), then writes the reweighted means
 * out[t * n_obs + j] = <obs_j> at beta_targets[t], with the per-sample
 * denominators of the last iteration (computed from the free energies before
 * its update), as the reference does.  K <= 1024, n_obs <= 8. */
int ising_wham(int32_t K, const int64_t* n_k, const double* betas, const double* energies, int32_t n_obs,
               const double* obs, int32_t max_iter, double tol, double* f_out, int32_t* iters, int32_t n_targets,
               const double* beta_targets, double* out);

/* Integrated autocorrelation time of a series by Sokal's automatic windowing:
 * tau(W) = 1/2 + sum_{t=1..W} rho(t), W = smallest window with W >= c * tau(W)
 * (c = 6 is customary).  Not in the reference (its plans name it); used to
 * report independent samples per second.  *window receives W. */
int ising_tau_int(const double* x, int64_t n, double c, double* tau, int64_t* window);

/* Crossing temperature of two Binder-cumulant curves U1(T1), U2(T2)
 * (analysis/scripts/analyze_gpu_fss.py:256-277, find_binder_crossings): both
 * curves are interpolated by not-a-knot cubic splines (the interpolant of
 * scipy's interp1d(kind="cubic")) on 2000 equally spaced temperatures of the
 * common range [max(min T1, min T2), min(max T1, max T2)]; the first sign
 * change of U1 - U2 between grid points i, i + 1 gives the linear root
 * T_i - d_i (T_{i+1} - T_i) / (d_{i+1} - d_i).  Temperatures may come in any
 * order (they are sorted); each curve needs >= 4 distinct temperatures.
 * *found = 0 (and *tc untouched) when the curves do not cross. */
int ising_binder_crossing(const double* t1, const double* u1, int32_t n1, const double* t2, const double* u2,
                          int32_t n2, double* tc, int32_t* found);

#ifdef __cplusplus
}
#endif
#endif
