/*
 * CPU oracle for the Ising / XY / Heisenberg parallel-tempering path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (libising_hip.so, the
 * gpu_fss command, the Python bindings) links, loads or calls this code.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
 * and there only as the checker or the timed CPU baseline.
 *
 * Two families of functions live here:
 *
 *  1. Checkerboard replays ("o_*"): the same algorithm the GPU library runs,
 *     restated on the full z-major lattice (index s = (z*L + y)*L + x) with the
 *     same Philox4x32-10 counter assignment, so Ising lattices, int64
 *     observables and the temperature permutation can be compared bit for bit,
 *     and continuous-spin half-sweeps can be replayed in FP64.
 *
 *  2. Reference-faithful CPU algorithms ("o_ref_*"): restatements of the
 *     reference's own CPU paths (random-site Metropolis, Wolff, Heisenberg cap
 *     Metropolis + sequential overrelaxation, XY Wolff).  They use their own
 *     xoshiro256++ stream, so they are algorithm-faithful, not stream-identical
 *     to the reference's Rust `rand` draws.  They serve as the equilibrium
 *     oracle and as the timed CPU baseline.
 *
 * Parity pins (what checks each function) are listed next to each group.
 */
#ifndef ISING_ORACLE_H
#define ISING_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- Philox4x32-10 (Salmon et al., SC'11) ----------------------
 * Pinned by: Random123 known-answer vectors and rocRAND's open-source
 * philox4x32_10 engine header (tests/test_oracle_philox.py). */
void o_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Counter word 3 = (purpose << 8) | model.  Counter word 2 = stream id. */
enum { O_MODEL_ISING = 0, O_MODEL_XY = 1, O_MODEL_HEIS = 2 };
enum { O_P_INIT = 1, O_P_ISING = 2, O_P_TIE = 3, O_P_EXCH = 4, O_P_CONT = 5, O_P_MSC = 6 };
uint32_t o_tag(int purpose, int model);
uint32_t o_stream(int L, int replica);         /* (L << 16) | replica */

/* FP64 exp built only from explicit fma, rint and ldexp, so the host and the
 * device produce identical bits.  Pinned by comparison with libm exp. */
double o_det_exp(double x);
/* 53-bit uniform in [0,1) from two Philox words. */
double o_u01(uint32_t a, uint32_t b);

/* ---------------- Ising checkerboard replay --------------------------------
 * Spins: int8 +-1, full z-major layout, even L.  Colour c = (x+y+z)&1.
 * Half-sweep colour 0 then colour 1 (metropolis_sweep_kernel parity order,
 * src/cuda/lattice_gpu.rs:88).  Acceptance follows src/cuda/kernels.cu:47-53:
 * dE = 2 s (J sum_nb + h); accept iff dE <= 0 or u < exp(-beta dE), with u a
 * 32-bit uniform compared against a host-built threshold table.
 * Pinned by: brute-force stationary distribution on a 2x2x2 lattice,
 * T->0 / T->inf limits (metropolis.rs:70-100), the 2D 4x4 exact-enumeration
 * fixture and CPU Wolff equilibrium comparisons. */
/* thr[k] with k = b*7 + n (b = spin up, n = up neighbours) as a 33-bit value
 * split into hi (0..65536) and lo (0..65535). */
void o_ising_thresholds(double beta, double J, double h, uint32_t thr_hi[14], uint32_t thr_lo[14]);
void o_ising_init(int L, uint64_t seed, uint32_t stream, int8_t* spins);
void o_ising_halfsweep(int L, int8_t* spins, int color, int64_t t, uint64_t seed, uint32_t stream,
                       const uint32_t thr_hi[14], const uint32_t thr_lo[14]);
/* Forward-bond sum B = sum_s s_s (s_{x+1} + s_{y+1} + s_{z+1}) and M = sum_s s_s
 * (src/cuda/reduce_kernel.cu:7-66).  E_total = -J * B. */
void o_ising_measure(int L, const int8_t* spins, int64_t* bond, int64_t* mag);
/* 2D square-lattice variant (Lz = 1), used against the exact-enumeration and
 * Onsager fixtures: same colour/RNG scheme with z removed. */
void o_ising2d_halfsweep(int L, int8_t* spins, int color, int64_t t, uint64_t seed, uint32_t stream,
                         const uint32_t thr_hi[10], const uint32_t thr_lo[10]);
void o_ising2d_thresholds(double beta, double J, double h, uint32_t thr_hi[10], uint32_t thr_lo[10]);
void o_ising2d_measure(int L, const int8_t* spins, int64_t* bond, int64_t* mag);

/* ---------------- Replica exchange (parallel_tempering.rs:5-36) -------------
 * Even/odd adjacent pairs on the temperature->replica permutation; one Philox
 * uniform per pair is always drawn.  Returns the number of accepted swaps.
 * Pinned by: bijectivity (parallel_tempering.rs:53-67), closed-form
 * acceptance frequencies. */
int o_exchange(int R, const double* betas, const double* energies, int* r2t, int* t2r, int64_t round,
               uint64_t seed, uint32_t stream, int model);

/* Full Ising PT run of one lattice size, exactly as the GPU ensemble runs it:
 * W warm-up sweeps (no exchange), then S sample sweeps with measurement when
 * (s+1) % measure_every == 0 and exchange when (s+1) % exchange_every == 0 at
 * round s / exchange_every (src/bin/gpu_fss.rs:199-286).
 * Records (written in replica order within a measured sweep): tidx, B, M.
 * Returns the number of records written. */
int64_t o_ising_pt_run(int L, int R, const double* temps, uint64_t seed, double J, double h,
                       int64_t warmup, int64_t samples, int exchange_every, int measure_every,
                       int32_t* rec_tidx, int64_t* rec_bond, int64_t* rec_mag, int8_t* final_spins);

/* Same for the 2D square lattice (dim = 2) or the cube (dim = 3). */
int64_t o_ising_pt_run_dim(int dim, int L, int R, const double* temps, uint64_t seed, double J, double h,
                           int64_t warmup, int64_t samples, int exchange_every, int measure_every,
                           int32_t* rec_tidx, int64_t* rec_bond, int64_t* rec_mag, int8_t* final_spins);

/* ---------------- Jackknife (src/bin/gpu_fss.rs:327-412) --------------------
 * out = E, E_err, M, M_err, M2, M2_err, M4, M4_err, Cv, Cv_err, chi, chi_err.
 * Pinned by closed-form series (constant, alternating, linear). */
void o_jackknife(const double* e, const double* m, int64_t n, double beta, double n3, int n_blocks,
                 double out[12]);
void o_linspace(double tmin, double tmax, int n, double* out);

/* ---------------- Continuous spins (XY n=2, Heisenberg n=3) ----------------
 * Spins stored component-interleaved (AoS, as src/cuda/gpu_lattice_continuous.rs
 * keeps them), full z-major layout, FP64.
 * Metropolis half-sweep: one Philox call per site (counter = (s>>1, 2t+c,
 * stream, tag(O_P_CONT, model))): words 0,1 drive the proposal, word 2 the
 * acceptance.  Heisenberg: cap proposal around the current spin, cos(theta)
 * uniform in [cos(delta), 1], phi uniform in [0, 2pi) (heisenberg/metropolis.rs:54-92).
 * XY: rotation by alpha uniform in [-delta, delta).  Accept iff dE <= 0 or
 * u < exp(-beta dE).
 * Overrelaxation: S' = 2 (S.h) h / |h|^2 - S, skipped when |h|^2 < 1e-30
 * (heisenberg/overrelax.rs:15-48), no renormalisation.
 * When `accept_u`/`accept_p` are non-null they receive, per site of the colour,
 * the uniform and the acceptance probability (for near-threshold masks).
 * Pinned by: OR energy conservation, unit norm, single-site conditional
 * histograms against the exact Boltzmann conditional, and equilibrium against
 * the o_ref_* algorithms. */
void o_cont_init(int L, int ncomp, uint64_t seed, uint32_t stream, double* spins);
void o_cont_metropolis_half(int L, int ncomp, double* spins, int color, int64_t t, uint64_t seed,
                            uint32_t stream, double beta, double J, double delta, int renorm,
                            double* accept_u, double* accept_p);
void o_cont_overrelax_half(int L, int ncomp, double* spins, int color, double J);
void o_cont_measure(int L, int ncomp, const double* spins, double J, double* energy, double m[3]);
/* Per-site local energy e_s = -J S_s . sum_nb S_nb (used by the 1e-5 parity check). */
void o_cont_site_energy(int L, int ncomp, const double* spins, double J, double* out);

/* ---------------- Reference-faithful CPU algorithms -------------------------
 * RNG: xoshiro256++ seeded through splitmix64 (as rand_xoshiro's
 * seed_from_u64 does), integer ranges by Lemire's method, f64 = (x>>11)*2^-53. */
typedef struct { uint64_t s[4]; } o_rng;
void o_rng_seed(o_rng* r, uint64_t seed);
uint64_t o_rng_next(o_rng* r);
double o_rng_f64(o_rng* r);
uint64_t o_rng_below(o_rng* r, uint64_t n);

/* Ising, 3D cubic, +-1 spins, full layout. */
void o_ref_metropolis_sweep(int L, int8_t* spins, double beta, double J, double h, o_rng* rng); /* metropolis.rs:8-27 */
int64_t o_ref_wolff_step(int L, int8_t* spins, double beta, double J, o_rng* rng, int32_t* stack, uint8_t* mark); /* wolff.rs:27-65 */
/* Heisenberg, 3D cubic, AoS doubles. */
void o_ref_heis_metropolis_sweep(int L, double* spins, double beta, double J, double delta, o_rng* rng, int64_t* accepted); /* heisenberg/metropolis.rs:10-36 */
void o_ref_heis_overrelax_sweep(int L, double* spins, double J);  /* heisenberg/overrelax.rs:15-48 */
/* XY Wolff, 3D cubic, AoS doubles. */
int64_t o_ref_xy_wolff_sweep(int L, double* spins, double beta, double J, o_rng* rng, int32_t* stack, uint8_t* mark); /* xy/wolff.rs:16-73 */

/* Timed baseline: random-site Metropolis sweeps until `seconds` elapse;
 * returns updates performed and wall seconds. */
int64_t o_bench_ref_metropolis(int L, double beta, uint64_t seed, double seconds, double* wall);
int64_t o_bench_ref_heis(int L, double beta, double delta, int n_or, uint64_t seed, double seconds, double* wall);
/* xy/wolff.rs:16-73 timed: spins flipped in `seconds` (*steps = Wolff steps). */
int64_t o_bench_ref_xy_wolff(int L, double beta, uint64_t seed, double seconds, double* wall, int64_t* steps);
/* ---------------- Multispin-coded Ising (replicas across word bits) --------
 * Pinned by: the 2x2x2 brute-force Boltzmann distribution per replica, the
 * replica-independence of the bit lanes, and equality of the lazy bit-plane
 * comparison with the full 32-bit one (tests/test_oracle_msc.py). */
void o_msc_thresholds(double beta, double J, uint32_t out[3]);
void o_msc_halfsweep(int L, int R, int8_t* spins, int color, int64_t t, uint64_t seed, const uint32_t* thr);
int64_t o_msc_pt_run(int L, int R, const double* temps, uint64_t seed, double J, int64_t warmup, int64_t samples,
                     int exchange_every, int measure_every, int32_t* rec_tidx, int64_t* rec_bond, int64_t* rec_mag,
                     int8_t* final_spins);
void o_ref_ising_chain(int kind, int L, uint64_t seed, double beta, int64_t warmup, int64_t n,
                       int64_t* out_bond, int64_t* out_mag);
void o_ref_heis_chain(int L, uint64_t seed, double beta, double delta, int n_or, int64_t warmup, int64_t n,
                      double* out_e, double* out_m);
void o_ref_xy_chain(int L, uint64_t seed, double beta, int64_t warmup, int64_t n, double* out_e, double* out_m);
double o_tau_int(const double* x, int64_t n, double c, int64_t* window);
void o_ising_chain(int dim, int L, uint64_t seed, double beta, double J, double h, int64_t warmup,
                   int64_t n_sweeps, int64_t* out_bond, int64_t* out_mag);

/* ---------------- Cluster updates (Swendsen-Wang / Wolff) ------------------
 * One bond configuration per update, as the library draws it: Fortuin-
 * Kasteleyn bonds on the forward bonds (+x, +y[, +z]) of every site, active
 * between satisfied neighbours (aligned for J > 0, anti-aligned for J < 0)
 * with probability floor(2^32 p) / 2^32, p = 1 - exp(-2 beta |J|)
 * (src/wolff.rs:27-77 grows the same clusters from one seed).  Draws:
 * w = Philox(c0 = site / 2, u, stream, tag(7, 0)); the site's words are
 * (wa, wb) = (w0, w1) for even x, (w2, w3) for odd x; the 16-bit draws of
 * bonds +x, +y, +z are wa & 0xffff, wa >> 16, wb & 0xffff, compared with the
 * high 16 bits of the threshold; on equality the 16-bit draw
 * Philox(site, u, stream, tag(8, d)) word 0 >> 16 against the low 16 bits.
 * Bit 31 of wb is the site's cluster coin.  Clusters by
 * union-find (root = smallest site index); Swendsen-Wang (kind 0) flips a
 * cluster when its root's coin is set, Wolff (kind 1) flips the cluster of
 * the site floor(N ((w0 << 32) | w1) / 2^64), w = Philox(0, u, stream, tag(10, 0)).
 * Returns the number of sites flipped; bonds[i] (when not NULL) receives the
 * site's bond bits (bit d), coin (bit 3) and, in bit 5 + d, whether bond d
 * was decided by the tie draw.
 * Pinned by: exact stationarity of the Boltzmann distribution on 2x2x2 and
 * 4x4 lattices (full transition matrices from enumerated bond draws, not from
 * this code), clusters = connected components found by breadth-first search,
 * Wolff mean cluster size = <M^2> / N, and equilibrium against o_ref_*
 * (tests/test_oracle_cluster.py). */
uint32_t o_bond_threshold(double beta, double J);
int64_t o_ising_cluster(int dim, int L, int8_t* spins, int kind, uint32_t u, uint64_t seed, uint32_t stream,
                        uint32_t thr, double J, int32_t* lab, uint8_t* bonds);
/* Continuous spins, Wolff embedding (src/xy/wolff.rs:16-73): r from
 * Philox(0, u, stream, tag(9, model)) (XY: phi = 2 pi u01(w0, w1); Heisenberg:
 * z = 2 u01(w0, w1) - 1, phi = 2 pi u01(w2, w3)), a_i = s_i . r, bond d of
 * site i active iff x = 2 beta J a_i a_j > 0 and u24(w_d) < 1 - exp(-x) with
 * w = Philox(i, u, stream, tag(7, model)); bit 31 of w_3 is the coin; a flip
 * is s -> s - 2 a r.  FP64 throughout; *n_marginal counts bonds whose uniform
 * lies within `margin` of the probability (where an FP32 device may differ).
 * Pinned by: energy conservation of every cluster-reflected configuration's
 * embedded Ising weights (reflection is an isometry), unit norm, and
 * equilibrium against o_ref_xy_chain. */
int64_t o_cont_cluster(int L, int ncomp, double* spins, int kind, uint32_t u, uint64_t seed, uint32_t stream,
                       int model, double beta, double J, double margin, int64_t* n_marginal, int32_t* lab,
                       uint8_t* bonds);
/* Cluster-update chains for equilibrium checks: `warmup` then `n` updates of
 * one lattice, recording the bond sum and magnetisation after each. */
void o_ising_cluster_chain(int dim, int L, uint64_t seed, int kind, double beta, double J, int64_t warmup, int64_t n,
                           int64_t* out_bond, int64_t* out_mag, int64_t* out_flipped);

#ifdef __cplusplus
}
#endif
#endif
