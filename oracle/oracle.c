/*
 * CPU oracle for the parallel-tempering path -- TEST INFRASTRUCTURE ONLY.
 * See oracle.h for what each function restates and what pins it.
 * Build: oracle/Makefile (gcc, -ffp-contract=off so FP64 replays round the
 * way the source reads).
 */
#define _POSIX_C_SOURCE 200809L
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------ */
/* Philox4x32-10                                                             */
/* ------------------------------------------------------------------------ */
#define PH_M0 0xD2511F53u
#define PH_M1 0xCD9E8D57u
#define PH_W0 0x9E3779B9u
#define PH_W1 0xBB67AE85u

void o_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)PH_M0 * x0;
        uint64_t p1 = (uint64_t)PH_M1 * x2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ x1 ^ k0;
        uint32_t n2 = hi0 ^ x3 ^ k1;
        x0 = n0; x1 = lo1; x2 = n2; x3 = lo0;
        k0 += PH_W0; k1 += PH_W1;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

uint32_t o_tag(int purpose, int model) { return ((uint32_t)purpose << 8) | (uint32_t)model; }
uint32_t o_stream(int L, int replica) { return ((uint32_t)L << 16) | (uint32_t)replica; }

static void ph(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint64_t seed, uint32_t out[4]) {
    uint32_t ctr[4] = {c0, c1, c2, c3};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    o_philox4x32_10(ctr, key, out);
}

/* ------------------------------------------------------------------------ */
/* Deterministic exp: Cody-Waite reduction by ln2, degree-13 Taylor in       */
/* Horner form with explicit fma, exact scaling by ldexp.                    */
/* ------------------------------------------------------------------------ */
double o_det_exp(double x) {
    if (x != x) return x;
    if (x < -700.0) return 0.0;
    if (x > 700.0) return INFINITY;
    const double inv_ln2 = 1.4426950408889634074;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    double kd = rint(x * inv_ln2);
    double r = fma(-kd, ln2_hi, x);
    r = fma(-kd, ln2_lo, r);
    double p = 1.0 / 6227020800.0;          /* 1/13! */
    p = fma(p, r, 1.0 / 479001600.0);       /* 1/12! */
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)kd);
}

double o_u01(uint32_t a, uint32_t b) {
    uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
    return (double)m * (1.0 / 9007199254740992.0);
}

/* 24-bit uniform in [0,1), exactly representable in FP32 and FP64. */
static double u24(uint32_t w) { return (double)(w >> 8) * (1.0 / 16777216.0); }

/* ------------------------------------------------------------------------ */
/* Ising checkerboard                                                        */
/* ------------------------------------------------------------------------ */
static uint32_t thr_from_p(double p, uint32_t* lo) {
    /* 33-bit threshold T = round-down(p * 2^32), T = 2^32 when p >= 1, split
     * as T = hi * 2^25 + lo with hi in 0..128 and lo < 2^25. */
    uint64_t t;
    if (!(p < 1.0)) t = 1ull << 32;
    else if (p <= 0.0) t = 0;
    else t = (uint64_t)ldexp(p, 32);
    *lo = (uint32_t)(t & 0x1ffffffu);
    return (uint32_t)(t >> 25);
}

static void thresholds_z(int z, double beta, double J, double h, uint32_t* hi, uint32_t* lo) {
    for (int b = 0; b <= 1; ++b) {
        double s = b ? 1.0 : -1.0;
        for (int n = 0; n <= z; ++n) {
            double nb = (double)(2 * n - z);
            double dE = 2.0 * s * (J * nb + h);
            double p = (dE <= 0.0) ? 1.0 : exp(-beta * dE);
            hi[b * (z + 1) + n] = thr_from_p(p, &lo[b * (z + 1) + n]);
        }
    }
}

void o_ising_thresholds(double beta, double J, double h, uint32_t thr_hi[14], uint32_t thr_lo[14]) {
    thresholds_z(6, beta, J, h, thr_hi, thr_lo);
}
void o_ising2d_thresholds(double beta, double J, double h, uint32_t thr_hi[10], uint32_t thr_lo[10]) {
    thresholds_z(4, beta, J, h, thr_hi, thr_lo);
}

void o_ising_init(int L, uint64_t seed, uint32_t stream, int8_t* spins) {
    int64_t n = (int64_t)L * L * L;
    for (int64_t s = 0; s < n; ++s) {
        uint32_t w[4];
        ph((uint32_t)s, 0u, stream, o_tag(O_P_INIT, O_MODEL_ISING), seed, w);
        spins[s] = (w[0] >> 31) ? 1 : -1;
    }
}

/* Accept decision for one site, u < T for the 32-bit uniform
 * u = r7 * 2^25 + u25: r7 is the low 7 bits of the site's byte of the group's
 * Philox word (16 sites per call, site j of group h >> 4 takes byte j), and
 * only when r7 equals T's high part is u25 drawn, as the top 25 bits of a
 * second Philox call keyed by the site's half index. */
static int ising_accept(uint32_t r7, uint32_t thi, uint32_t tlo, uint32_t h, uint32_t hs, uint32_t stream,
                        int model, uint64_t seed) {
    if (r7 < thi) return 1;
    if (r7 > thi) return 0;
    uint32_t w[4];
    ph(h, hs, stream, o_tag(O_P_TIE, model), seed, w);
    return (w[0] >> 7) < tlo;
}

static uint32_t site_r7(const uint32_t w[4], uint32_t h) {
    const uint32_t j = h & 15u;
    return (w[j >> 2] >> (8 * (j & 3))) & 0x7fu;
}

void o_ising_halfsweep(int L, int8_t* spins, int color, int64_t t, uint64_t seed, uint32_t stream,
                       const uint32_t thr_hi[14], const uint32_t thr_lo[14]) {
    const uint32_t hs = (uint32_t)(2 * t + color);
    uint32_t cur_g = 0xffffffffu, w[4] = {0, 0, 0, 0};
    for (int z = 0; z < L; ++z)
        for (int y = 0; y < L; ++y)
            for (int x = 0; x < L; ++x) {
                if (((x + y + z) & 1) != color) continue;
                int64_t s = ((int64_t)z * L + y) * L + x;
                int xp = (x + 1) % L, xm = (x + L - 1) % L;
                int yp = (y + 1) % L, ym = (y + L - 1) % L;
                int zp = (z + 1) % L, zm = (z + L - 1) % L;
                int nb = spins[((int64_t)z * L + y) * L + xp] + spins[((int64_t)z * L + y) * L + xm] +
                         spins[((int64_t)z * L + yp) * L + x] + spins[((int64_t)z * L + ym) * L + x] +
                         spins[((int64_t)zp * L + y) * L + x] + spins[((int64_t)zm * L + y) * L + x];
                int n_up = (nb + 6) / 2;
                int b = spins[s] > 0;
                int k = b * 7 + n_up;
                uint32_t h = (uint32_t)(s >> 1);
                uint32_t g = h >> 4;
                if (g != cur_g) {
                    ph(g, hs, stream, o_tag(O_P_ISING, O_MODEL_ISING), seed, w);
                    cur_g = g;
                }
                if (ising_accept(site_r7(w, h), thr_hi[k], thr_lo[k], h, hs, stream, O_MODEL_ISING, seed))
                    spins[s] = (int8_t)-spins[s];
            }
}

void o_ising_measure(int L, const int8_t* spins, int64_t* bond, int64_t* mag) {
    int64_t B = 0, M = 0;
    for (int z = 0; z < L; ++z)
        for (int y = 0; y < L; ++y)
            for (int x = 0; x < L; ++x) {
                int64_t s = ((int64_t)z * L + y) * L + x;
                int v = spins[s];
                int f = spins[((int64_t)z * L + y) * L + (x + 1) % L] + spins[((int64_t)z * L + (y + 1) % L) * L + x] +
                        spins[((int64_t)((z + 1) % L) * L + y) * L + x];
                B += v * f;
                M += v;
            }
    *bond = B;
    *mag = M;
}

/* 2D: site s = y*L + x, colour (x+y)&1, half index s>>1, same group/tie rule. */
void o_ising2d_halfsweep(int L, int8_t* spins, int color, int64_t t, uint64_t seed, uint32_t stream,
                         const uint32_t thr_hi[10], const uint32_t thr_lo[10]) {
    const uint32_t hs = (uint32_t)(2 * t + color);
    uint32_t cur_g = 0xffffffffu, w[4] = {0, 0, 0, 0};
    for (int y = 0; y < L; ++y)
        for (int x = 0; x < L; ++x) {
            if (((x + y) & 1) != color) continue;
            int64_t s = (int64_t)y * L + x;
            int nb = spins[(int64_t)y * L + (x + 1) % L] + spins[(int64_t)y * L + (x + L - 1) % L] +
                     spins[(int64_t)((y + 1) % L) * L + x] + spins[(int64_t)((y + L - 1) % L) * L + x];
            int n_up = (nb + 4) / 2;
            int b = spins[s] > 0;
            int k = b * 5 + n_up;
            uint32_t h = (uint32_t)(s >> 1);
            uint32_t g = h >> 4;
            if (g != cur_g) {
                ph(g, hs, stream, o_tag(O_P_ISING, O_MODEL_ISING), seed, w);
                cur_g = g;
            }
            if (ising_accept(site_r7(w, h), thr_hi[k], thr_lo[k], h, hs, stream, O_MODEL_ISING, seed))
                spins[s] = (int8_t)-spins[s];
        }
}

void o_ising2d_measure(int L, const int8_t* spins, int64_t* bond, int64_t* mag) {
    int64_t B = 0, M = 0;
    for (int y = 0; y < L; ++y)
        for (int x = 0; x < L; ++x) {
            int v = spins[(int64_t)y * L + x];
            B += v * (spins[(int64_t)y * L + (x + 1) % L] + spins[(int64_t)((y + 1) % L) * L + x]);
            M += v;
        }
    *bond = B;
    *mag = M;
}

/* ------------------------------------------------------------------------ */
/* Replica exchange                                                          */
/* ------------------------------------------------------------------------ */
int o_exchange(int R, const double* betas, const double* energies, int* r2t, int* t2r, int64_t round,
               uint64_t seed, uint32_t stream, int model) {
    int acc = 0;
    for (int i = (int)(round & 1); i + 1 < R; i += 2) {
        int ri = t2r[i], rj = t2r[i + 1];
        double delta = (betas[i] - betas[i + 1]) * (energies[rj] - energies[ri]);
        uint32_t w[4];
        ph((uint32_t)i, (uint32_t)round, stream, o_tag(O_P_EXCH, model), seed, w);
        double u = o_u01(w[0], w[1]);
        if (delta <= 0.0 || u < o_det_exp(-delta)) {
            t2r[i] = rj;
            t2r[i + 1] = ri;
            r2t[ri] = i + 1;
            r2t[rj] = i;
            ++acc;
        }
    }
    return acc;
}

void o_linspace(double tmin, double tmax, int n, double* out) {
    if (n == 1) {
        out[0] = tmin;
        return;
    }
    for (int i = 0; i < n; ++i) out[i] = tmin + (tmax - tmin) * (double)i / (double)(n - 1);
}

static void half(int dim, int L, int8_t* sp, int c, int64_t t, uint64_t seed, uint32_t st, const uint32_t* hi,
                 const uint32_t* lo) {
    if (dim == 2) o_ising2d_halfsweep(L, sp, c, t, seed, st, hi, lo);
    else o_ising_halfsweep(L, sp, c, t, seed, st, hi, lo);
}

int64_t o_ising_pt_run_dim(int dim, int L, int R, const double* temps, uint64_t seed, double J, double h,
                       int64_t warmup, int64_t samples, int exchange_every, int measure_every,
                       int32_t* rec_tidx, int64_t* rec_bond, int64_t* rec_mag, int8_t* final_spins) {
    int64_t n3 = dim == 2 ? (int64_t)L * L : (int64_t)L * L * L;
    int8_t* sp = (int8_t*)malloc((size_t)(n3 * R));
    int* r2t = (int*)malloc(sizeof(int) * R);
    int* t2r = (int*)malloc(sizeof(int) * R);
    double* betas = (double*)malloc(sizeof(double) * R);
    double* en = (double*)malloc(sizeof(double) * R);
    uint32_t (*thi)[14] = malloc(sizeof(uint32_t) * 14 * R);
    uint32_t (*tlo)[14] = malloc(sizeof(uint32_t) * 14 * R);
    for (int i = 0; i < R; ++i) {
        r2t[i] = t2r[i] = i;
        betas[i] = 1.0 / temps[i];
        if (dim == 2) {
            o_ising2d_thresholds(betas[i], J, h, thi[i], tlo[i]);
            for (int64_t s = 0; s < n3; ++s) {
                uint32_t w[4];
                ph((uint32_t)s, 0u, o_stream(L, i), o_tag(O_P_INIT, O_MODEL_ISING), seed, w);
                sp[n3 * i + s] = (w[0] >> 31) ? 1 : -1;
            }
        } else {
            o_ising_thresholds(betas[i], J, h, thi[i], tlo[i]);
            o_ising_init(L, seed, o_stream(L, i), sp + n3 * i);
        }
    }
    int64_t t = 0, nrec = 0;
    for (int64_t w = 0; w < warmup; ++w, ++t)
        for (int r = 0; r < R; ++r)
            for (int c = 0; c < 2; ++c)
                half(dim, L, sp + n3 * r, c, t, seed, o_stream(L, r), thi[r2t[r]], tlo[r2t[r]]);
    for (int64_t s = 0; s < samples; ++s, ++t) {
        for (int r = 0; r < R; ++r)
            for (int c = 0; c < 2; ++c)
                half(dim, L, sp + n3 * r, c, t, seed, o_stream(L, r), thi[r2t[r]], tlo[r2t[r]]);
        int do_m = ((s + 1) % measure_every) == 0;
        int do_x = ((s + 1) % exchange_every) == 0;
        if (do_m || do_x) {
            for (int r = 0; r < R; ++r) {
                int64_t B, M;
                if (dim == 2) o_ising2d_measure(L, sp + n3 * r, &B, &M);
                else o_ising_measure(L, sp + n3 * r, &B, &M);
                en[r] = -J * (double)B;
                if (do_m) {
                    if (rec_tidx) rec_tidx[nrec] = r2t[r];
                    if (rec_bond) rec_bond[nrec] = B;
                    if (rec_mag) rec_mag[nrec] = M;
                    ++nrec;
                }
            }
        }
        if (do_x)
            o_exchange(R, betas, en, r2t, t2r, s / exchange_every, seed, o_stream(L, 0), O_MODEL_ISING);
    }
    if (final_spins) memcpy(final_spins, sp, (size_t)(n3 * R));
    free(sp); free(r2t); free(t2r); free(betas); free(en); free(thi); free(tlo);
    return nrec;
}

int64_t o_ising_pt_run(int L, int R, const double* temps, uint64_t seed, double J, double h,
                       int64_t warmup, int64_t samples, int exchange_every, int measure_every,
                       int32_t* rec_tidx, int64_t* rec_bond, int64_t* rec_mag, int8_t* final_spins) {
    return o_ising_pt_run_dim(3, L, R, temps, seed, J, h, warmup, samples, exchange_every, measure_every, rec_tidx,
                              rec_bond, rec_mag, final_spins);
}

/* ------------------------------------------------------------------------ */
/* Jackknife                                                                 */
/* ------------------------------------------------------------------------ */
void o_jackknife(const double* e, const double* m, int64_t n, double beta, double n3, int n_blocks,
                 double out[12]) {
    if (n < n_blocks || n_blocks < 2) {
        for (int i = 0; i < 12; ++i) out[i] = 0.0;
        return;
    }
    int64_t bs = n / n_blocks, nu = bs * n_blocks;
    double se = 0, se2 = 0, sm = 0, sm2 = 0, sm4 = 0;
    for (int64_t i = 0; i < nu; ++i) {
        double mm = m[i] * m[i];
        se += e[i]; se2 += e[i] * e[i]; sm += m[i]; sm2 += mm; sm4 += mm * mm;
    }
    double s = (double)nu;
    double ae = se / s, ae2 = se2 / s, am = sm / s, am2 = sm2 / s, am4 = sm4 / s;
    double cv = beta * beta * (ae2 - ae * ae) * n3;
    double chi = beta * (am2 - am * am) * n3;
    double full[6] = {ae, am, am2, am4, cv, chi};
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int b = 0; b < n_blocks; ++b) {
        int64_t st = b * bs, en = st + bs;
        double be = 0, be2 = 0, bm = 0, bm2 = 0, bm4 = 0;
        for (int64_t i = 0; i < nu; ++i) {
            if (i >= st && i < en) continue;
            double mm = m[i] * m[i];
            be += e[i]; be2 += e[i] * e[i]; bm += m[i]; bm2 += mm; bm4 += mm * mm;
        }
        double sj = (double)(nu - bs);
        double je = be / sj, je2 = be2 / sj, jm = bm / sj, jm2 = bm2 / sj, jm4 = bm4 / sj;
        double jk[6] = {je, jm, jm2, jm4, beta * beta * (je2 - je * je) * n3, beta * (jm2 - jm * jm) * n3};
        for (int k = 0; k < 6; ++k) acc[k] += (jk[k] - full[k]) * (jk[k] - full[k]);
    }
    double nb = (double)n_blocks;
    for (int k = 0; k < 6; ++k) {
        out[2 * k] = full[k];
        out[2 * k + 1] = sqrt(acc[k] * (nb - 1.0) / nb);
    }
}

/* ------------------------------------------------------------------------ */
/* Continuous spins                                                          */
/* ------------------------------------------------------------------------ */
static const double TWO_PI = 6.283185307179586476925286766559;

void o_cont_init(int L, int ncomp, uint64_t seed, uint32_t stream, double* spins) {
    int64_t n = (int64_t)L * L * L;
    int model = ncomp == 2 ? O_MODEL_XY : O_MODEL_HEIS;
    for (int64_t s = 0; s < n; ++s) {
        uint32_t w[4];
        ph((uint32_t)s, 0u, stream, o_tag(O_P_INIT, model), seed, w);
        if (ncomp == 2) {
            double th = TWO_PI * u24(w[0]);
            spins[2 * s] = cos(th);
            spins[2 * s + 1] = sin(th);
        } else {
            double z = 2.0 * u24(w[0]) - 1.0;
            double phi = TWO_PI * u24(w[1]);
            double r = sqrt(fmax(0.0, 1.0 - z * z));
            spins[3 * s] = r * cos(phi);
            spins[3 * s + 1] = r * sin(phi);
            spins[3 * s + 2] = z;
        }
    }
}

static void local_field(int L, int ncomp, const double* sp, int x, int y, int z, double hf[3]) {
    int64_t nb[6] = {
        ((int64_t)z * L + y) * L + (x + 1) % L,     ((int64_t)z * L + y) * L + (x + L - 1) % L,
        ((int64_t)z * L + (y + 1) % L) * L + x,     ((int64_t)z * L + (y + L - 1) % L) * L + x,
        ((int64_t)((z + 1) % L) * L + y) * L + x,   ((int64_t)((z + L - 1) % L) * L + y) * L + x};
    hf[0] = hf[1] = hf[2] = 0.0;
    for (int k = 0; k < 6; ++k)
        for (int c = 0; c < ncomp; ++c) hf[c] += sp[nb[k] * ncomp + c];
}

void o_cont_metropolis_half(int L, int ncomp, double* spins, int color, int64_t t, uint64_t seed,
                            uint32_t stream, double beta, double J, double delta, int renorm,
                            double* accept_u, double* accept_p) {
    int model = ncomp == 2 ? O_MODEL_XY : O_MODEL_HEIS;
    uint32_t hs = (uint32_t)(2 * t + color);
    int64_t idx = 0;
    for (int z = 0; z < L; ++z)
        for (int y = 0; y < L; ++y)
            for (int x = 0; x < L; ++x) {
                if (((x + y + z) & 1) != color) continue;
                int64_t s = ((int64_t)z * L + y) * L + x;
                double hf[3];
                local_field(L, ncomp, spins, x, y, z, hf);
                double* S = spins + s * ncomp;
                uint32_t w[4];
                ph((uint32_t)(s >> 1), hs, stream, o_tag(O_P_CONT, model), seed, w);
                double nw[3] = {0, 0, 0};
                if (ncomp == 2) {
                    double a = delta * (2.0 * u24(w[0]) - 1.0);
                    double ca = cos(a), sa = sin(a);
                    nw[0] = S[0] * ca - S[1] * sa;
                    nw[1] = S[1] * ca + S[0] * sa;
                } else {
                    double ct = cos(delta) + (1.0 - cos(delta)) * u24(w[0]);
                    double st = sqrt(fmax(0.0, 1.0 - ct * ct));
                    double phi = TWO_PI * u24(w[1]);
                    /* frame perpendicular to S (heisenberg/metropolis.rs:72-92) */
                    double tv[3] = {0, 0, 0};
                    if (fabs(S[0]) < 0.9) tv[0] = 1.0; else tv[1] = 1.0;
                    double ts = tv[0] * S[0] + tv[1] * S[1] + tv[2] * S[2];
                    double u[3] = {tv[0] - ts * S[0], tv[1] - ts * S[1], tv[2] - ts * S[2]};
                    double un = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
                    u[0] /= un; u[1] /= un; u[2] /= un;
                    double v[3] = {S[1] * u[2] - S[2] * u[1], S[2] * u[0] - S[0] * u[2], S[0] * u[1] - S[1] * u[0]};
                    double cp = cos(phi), sp = sin(phi);
                    for (int c = 0; c < 3; ++c) nw[c] = st * cp * u[c] + st * sp * v[c] + ct * S[c];
                    double nn = sqrt(nw[0] * nw[0] + nw[1] * nw[1] + nw[2] * nw[2]);
                    nw[0] /= nn; nw[1] /= nn; nw[2] /= nn;
                }
                double dE = 0.0;
                for (int c = 0; c < ncomp; ++c) dE += (nw[c] - S[c]) * hf[c];
                dE *= -J;
                double uacc = u24(w[2]);
                double p = dE <= 0.0 ? 1.0 : exp(-beta * dE);
                if (accept_u) accept_u[idx] = uacc;
                if (accept_p) accept_p[idx] = p;
                ++idx;
                if (dE <= 0.0 || uacc < p)
                    for (int c = 0; c < ncomp; ++c) S[c] = nw[c];
                if (renorm) {
                    double nn = 0.0;
                    for (int c = 0; c < ncomp; ++c) nn += S[c] * S[c];
                    nn = sqrt(nn);
                    for (int c = 0; c < ncomp; ++c) S[c] /= nn;
                }
            }
}

void o_cont_overrelax_half(int L, int ncomp, double* spins, int color, double J) {
    for (int z = 0; z < L; ++z)
        for (int y = 0; y < L; ++y)
            for (int x = 0; x < L; ++x) {
                if (((x + y + z) & 1) != color) continue;
                int64_t s = ((int64_t)z * L + y) * L + x;
                double hf[3];
                local_field(L, ncomp, spins, x, y, z, hf);
                for (int c = 0; c < ncomp; ++c) hf[c] *= J;
                double h2 = 0.0, sh = 0.0;
                double* S = spins + s * ncomp;
                for (int c = 0; c < ncomp; ++c) {
                    h2 += hf[c] * hf[c];
                    sh += S[c] * hf[c];
                }
                if (h2 < 1e-30) continue;
                double sc = 2.0 * sh / h2;
                for (int c = 0; c < ncomp; ++c) S[c] = sc * hf[c] - S[c];
            }
}

void o_cont_measure(int L, int ncomp, const double* spins, double J, double* energy, double m[3]) {
    double e = 0.0;
    m[0] = m[1] = m[2] = 0.0;
    for (int z = 0; z < L; ++z)
        for (int y = 0; y < L; ++y)
            for (int x = 0; x < L; ++x) {
                int64_t s = ((int64_t)z * L + y) * L + x;
                int64_t f[3] = {((int64_t)z * L + y) * L + (x + 1) % L, ((int64_t)z * L + (y + 1) % L) * L + x,
                                ((int64_t)((z + 1) % L) * L + y) * L + x};
                for (int c = 0; c < ncomp; ++c) m[c] += spins[s * ncomp + c];
                for (int k = 0; k < 3; ++k) {
                    double d = 0.0;
                    for (int c = 0; c < ncomp; ++c) d += spins[s * ncomp + c] * spins[f[k] * ncomp + c];
                    e -= J * d;
                }
            }
    *energy = e;
}

void o_cont_site_energy(int L, int ncomp, const double* spins, double J, double* out) {
    for (int z = 0; z < L; ++z)
        for (int y = 0; y < L; ++y)
            for (int x = 0; x < L; ++x) {
                int64_t s = ((int64_t)z * L + y) * L + x;
                double hf[3];
                local_field(L, ncomp, spins, x, y, z, hf);
                double d = 0.0;
                for (int c = 0; c < ncomp; ++c) d += spins[s * ncomp + c] * hf[c];
                out[s] = -J * d;
            }
}

/* ------------------------------------------------------------------------ */
/* Reference-faithful CPU algorithms                                          */
/* ------------------------------------------------------------------------ */
static uint64_t splitmix64(uint64_t* x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
void o_rng_seed(o_rng* r, uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) r->s[i] = splitmix64(&x);
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
uint64_t o_rng_next(o_rng* r) {
    uint64_t* s = r->s;
    uint64_t res = rotl(s[0] + s[3], 23) + s[0];
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return res;
}
double o_rng_f64(o_rng* r) { return (double)(o_rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }
uint64_t o_rng_below(o_rng* r, uint64_t n) {
    /* Lemire's multiply-and-reject on 64-bit words */
    unsigned __int128 m = (unsigned __int128)o_rng_next(r) * n;
    uint64_t l = (uint64_t)m;
    if (l < n) {
        uint64_t th = (0 - n) % n;
        while (l < th) {
            m = (unsigned __int128)o_rng_next(r) * n;
            l = (uint64_t)m;
        }
    }
    return (uint64_t)(m >> 64);
}

static inline void nbr6(int L, int64_t s, int64_t nb[6]) {
    int64_t L2 = (int64_t)L * L;
    int x = (int)(s % L), y = (int)((s / L) % L), z = (int)(s / L2);
    nb[0] = ((int64_t)z * L + y) * L + (x + L - 1) % L;
    nb[1] = ((int64_t)z * L + y) * L + (x + 1) % L;
    nb[2] = ((int64_t)z * L + (y + L - 1) % L) * L + x;
    nb[3] = ((int64_t)z * L + (y + 1) % L) * L + x;
    nb[4] = ((int64_t)((z + L - 1) % L) * L + y) * L + x;
    nb[5] = ((int64_t)((z + 1) % L) * L + y) * L + x;
}

void o_ref_metropolis_sweep(int L, int8_t* spins, double beta, double J, double h, o_rng* rng) {
    int64_t n = (int64_t)L * L * L;
    for (int64_t it = 0; it < n; ++it) {
        int64_t s = (int64_t)o_rng_below(rng, (uint64_t)n);
        int64_t nb[6];
        nbr6(L, s, nb);
        double sum = 0.0;
        for (int k = 0; k < 6; ++k) sum += spins[nb[k]];
        double dE = 2.0 * spins[s] * (J * sum + h);
        if (dE < 0.0 || o_rng_f64(rng) < exp(-beta * dE)) spins[s] = (int8_t)-spins[s];
    }
}

int64_t o_ref_wolff_step(int L, int8_t* spins, double beta, double J, o_rng* rng, int32_t* stack, uint8_t* mark) {
    if (J <= 0.0) return 0;
    int64_t n = (int64_t)L * L * L;
    double padd = 1.0 - exp(-2.0 * beta * J);
    int64_t seed = (int64_t)o_rng_below(rng, (uint64_t)n);
    int8_t target = spins[seed];
    memset(mark, 0, (size_t)n);
    int64_t top = 0, count = 0;
    mark[seed] = 1;
    stack[top++] = (int32_t)seed;
    while (top > 0) {
        int64_t i = stack[--top];
        int64_t nb[6];
        nbr6(L, i, nb);
        for (int k = 0; k < 6; ++k) {
            int64_t j = nb[k];
            if (!mark[j] && spins[j] == target && o_rng_f64(rng) < padd) {
                mark[j] = 1;
                stack[top++] = (int32_t)j;
            }
        }
    }
    for (int64_t i = 0; i < n; ++i)
        if (mark[i]) {
            spins[i] = (int8_t)-spins[i];
            ++count;
        }
    return count;
}

void o_ref_heis_metropolis_sweep(int L, double* spins, double beta, double J, double delta, o_rng* rng,
                                 int64_t* accepted) {
    int64_t n = (int64_t)L * L * L;
    double cd = cos(delta);
    for (int64_t it = 0; it < n; ++it) {
        int64_t s = (int64_t)o_rng_below(rng, (uint64_t)n);
        int64_t nb[6];
        nbr6(L, s, nb);
        double hf[3] = {0, 0, 0};
        for (int k = 0; k < 6; ++k)
            for (int c = 0; c < 3; ++c) hf[c] += spins[nb[k] * 3 + c];
        double* S = spins + s * 3;
        double eo = -J * (S[0] * hf[0] + S[1] * hf[1] + S[2] * hf[2]);
        double ct = cd + (1.0 - cd) * o_rng_f64(rng);
        double st = sqrt(1.0 - ct * ct);
        double phi = TWO_PI * o_rng_f64(rng);
        double tv[3] = {0, 0, 0};
        if (fabs(S[0]) < 0.9) tv[0] = 1.0; else tv[1] = 1.0;
        double ts = tv[0] * S[0] + tv[1] * S[1] + tv[2] * S[2];
        double u[3] = {tv[0] - ts * S[0], tv[1] - ts * S[1], tv[2] - ts * S[2]};
        double un = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        u[0] /= un; u[1] /= un; u[2] /= un;
        double v[3] = {S[1] * u[2] - S[2] * u[1], S[2] * u[0] - S[0] * u[2], S[0] * u[1] - S[1] * u[0]};
        double cp = cos(phi), sp = sin(phi), nw[3];
        for (int c = 0; c < 3; ++c) nw[c] = st * cp * u[c] + st * sp * v[c] + ct * S[c];
        double nn = sqrt(nw[0] * nw[0] + nw[1] * nw[1] + nw[2] * nw[2]);
        for (int c = 0; c < 3; ++c) nw[c] /= nn;
        double en = -J * (nw[0] * hf[0] + nw[1] * hf[1] + nw[2] * hf[2]);
        double dE = en - eo;
        if (dE < 0.0 || o_rng_f64(rng) < exp(-beta * dE)) {
            S[0] = nw[0]; S[1] = nw[1]; S[2] = nw[2];
            if (accepted) ++*accepted;
        }
    }
}

void o_ref_heis_overrelax_sweep(int L, double* spins, double J) {
    int64_t n = (int64_t)L * L * L;
    for (int64_t s = 0; s < n; ++s) {
        int64_t nb[6];
        nbr6(L, s, nb);
        double hf[3] = {0, 0, 0};
        for (int k = 0; k < 6; ++k)
            for (int c = 0; c < 3; ++c) hf[c] += J * spins[nb[k] * 3 + c];
        double h2 = hf[0] * hf[0] + hf[1] * hf[1] + hf[2] * hf[2];
        if (h2 < 1e-30) continue;
        double* S = spins + s * 3;
        double sc = 2.0 * (S[0] * hf[0] + S[1] * hf[1] + S[2] * hf[2]) / h2;
        double nw[3] = {sc * hf[0] - S[0], sc * hf[1] - S[1], sc * hf[2] - S[2]};
        double nn = sqrt(nw[0] * nw[0] + nw[1] * nw[1] + nw[2] * nw[2]);
        S[0] = nw[0] / nn; S[1] = nw[1] / nn; S[2] = nw[2] / nn;
    }
}

int64_t o_ref_xy_wolff_sweep(int L, double* spins, double beta, double J, o_rng* rng, int32_t* stack, uint8_t* mark) {
    int64_t n = (int64_t)L * L * L;
    double phi = TWO_PI * o_rng_f64(rng);
    double r0 = cos(phi), r1 = sin(phi);
    int64_t seed = (int64_t)o_rng_below(rng, (uint64_t)n);
    memset(mark, 0, (size_t)n);
    int64_t top = 0, count = 0;
    mark[seed] = 1;
    stack[top++] = (int32_t)seed;
    while (top > 0) {
        int64_t i = stack[--top];
        double pi = spins[2 * i] * r0 + spins[2 * i + 1] * r1;
        int64_t nb[6];
        nbr6(L, i, nb);
        for (int k = 0; k < 6; ++k) {
            int64_t j = nb[k];
            if (mark[j]) continue;
            double pj = spins[2 * j] * r0 + spins[2 * j + 1] * r1;
            if (pi * pj <= 0.0) continue;
            double pb = 1.0 - exp(-2.0 * beta * J * fabs(pi) * fabs(pj));
            if (o_rng_f64(rng) < pb) {
                mark[j] = 1;
                stack[top++] = (int32_t)j;
            }
        }
    }
    for (int64_t i = 0; i < n; ++i) {
        if (!mark[i]) continue;
        double d = spins[2 * i] * r0 + spins[2 * i + 1] * r1;
        double a = spins[2 * i] - 2.0 * d * r0, b = spins[2 * i + 1] - 2.0 * d * r1;
        double nn = sqrt(a * a + b * b);
        spins[2 * i] = a / nn;
        spins[2 * i + 1] = b / nn;
        ++count;
    }
    return count;
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int64_t o_bench_ref_metropolis(int L, double beta, uint64_t seed, double seconds, double* wall) {
    int64_t n = (int64_t)L * L * L, upd = 0;
    int8_t* sp = (int8_t*)malloc((size_t)n);
    o_rng rng;
    o_rng_seed(&rng, seed);
    for (int64_t i = 0; i < n; ++i) sp[i] = (o_rng_next(&rng) >> 63) ? 1 : -1;
    double t0 = now_s(), t1 = t0;
    while (t1 - t0 < seconds) {
        o_ref_metropolis_sweep(L, sp, beta, 1.0, 0.0, &rng);
        upd += n;
        t1 = now_s();
    }
    *wall = t1 - t0;
    free(sp);
    return upd;
}

int64_t o_bench_ref_heis(int L, double beta, double delta, int n_or, uint64_t seed, double seconds, double* wall) {
    int64_t n = (int64_t)L * L * L, upd = 0;
    double* sp = (double*)malloc(sizeof(double) * 3 * (size_t)n);
    o_rng rng;
    o_rng_seed(&rng, seed);
    for (int64_t i = 0; i < n; ++i) {
        double z = 2.0 * o_rng_f64(&rng) - 1.0, phi = TWO_PI * o_rng_f64(&rng), r = sqrt(1.0 - z * z);
        sp[3 * i] = r * cos(phi); sp[3 * i + 1] = r * sin(phi); sp[3 * i + 2] = z;
    }
    double t0 = now_s(), t1 = t0;
    while (t1 - t0 < seconds) {
        o_ref_heis_metropolis_sweep(L, sp, beta, 1.0, delta, &rng, NULL);
        for (int k = 0; k < n_or; ++k) o_ref_heis_overrelax_sweep(L, sp, 1.0);
        upd += n * (1 + n_or);
        t1 = now_s();
    }
    *wall = t1 - t0;
    free(sp);
    return upd;
}

/* The reference's only XY CPU path, Wolff (xy/wolff.rs:16-73), from a random
 * start: returns the spins flipped (cluster sizes summed) in `seconds`. */
int64_t o_bench_ref_xy_wolff(int L, double beta, uint64_t seed, double seconds, double* wall, int64_t* steps) {
    int64_t n = (int64_t)L * L * L, flipped = 0, k = 0;
    double* sp = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    uint8_t* mark = (uint8_t*)malloc((size_t)n);
    o_rng rng;
    o_rng_seed(&rng, seed);
    for (int64_t i = 0; i < n; ++i) {
        double phi = TWO_PI * o_rng_f64(&rng);
        sp[2 * i] = cos(phi); sp[2 * i + 1] = sin(phi);
    }
    double t0 = now_s(), t1 = t0;
    while (t1 - t0 < seconds) {
        flipped += o_ref_xy_wolff_sweep(L, sp, beta, 1.0, &rng, stack, mark);
        ++k;
        t1 = now_s();
    }
    *wall = t1 - t0;
    *steps = k;
    free(sp); free(stack); free(mark);
    return flipped;
}

/* Single-chain helpers for the statistical tests (one lattice, fixed beta,
 * measured every sweep).  dim = 2 uses the square lattice, dim = 3 the cube. */
void o_ising_chain(int dim, int L, uint64_t seed, double beta, double J, double h, int64_t warmup,
                   int64_t n_sweeps, int64_t* out_bond, int64_t* out_mag) {
    int64_t n = dim == 2 ? (int64_t)L * L : (int64_t)L * L * L;
    int8_t* sp = (int8_t*)malloc((size_t)n);
    uint32_t hi[14], lo[14];
    uint32_t st = o_stream(L, 0);
    if (dim == 2) {
        for (int64_t s = 0; s < n; ++s) {
            uint32_t w[4];
            ph((uint32_t)s, 0u, st, o_tag(O_P_INIT, O_MODEL_ISING), seed, w);
            sp[s] = (w[0] >> 31) ? 1 : -1;
        }
        o_ising2d_thresholds(beta, J, h, hi, lo);
    } else {
        o_ising_init(L, seed, st, sp);
        o_ising_thresholds(beta, J, h, hi, lo);
    }
    for (int64_t t = 0; t < warmup + n_sweeps; ++t) {
        for (int c = 0; c < 2; ++c) {
            if (dim == 2) o_ising2d_halfsweep(L, sp, c, t, seed, st, hi, lo);
            else o_ising_halfsweep(L, sp, c, t, seed, st, hi, lo);
        }
        if (t >= warmup) {
            int64_t B, M;
            if (dim == 2) o_ising2d_measure(L, sp, &B, &M);
            else o_ising_measure(L, sp, &B, &M);
            out_bond[t - warmup] = B;
            out_mag[t - warmup] = M;
        }
    }
    free(sp);
}

/* Sokal automatic windowing (test restatement): tau = 1/2 + sum_{t<=W} rho(t),
 * W the first lag with W >= c tau(W). */
double o_tau_int(const double* x, int64_t n, double c, int64_t* window) {
    double mean = 0.0, c0 = 0.0, t = 0.5;
    int64_t W = 0;
    for (int64_t i = 0; i < n; ++i) mean += x[i];
    mean /= (double)n;
    for (int64_t i = 0; i < n; ++i) c0 += (x[i] - mean) * (x[i] - mean);
    c0 /= (double)n;
    if (!(c0 > 0.0)) {
        if (window) *window = 0;
        return 0.5;
    }
    for (int64_t lag = 1; lag < n; ++lag) {
        double ct = 0.0;
        for (int64_t i = 0; i + lag < n; ++i) ct += (x[i] - mean) * (x[i + lag] - mean);
        t += ct / (double)n / c0;
        W = lag;
        if ((double)lag >= c * t) break;
    }
    if (window) *window = W;
    return t;
}

/* Reference-faithful Ising chain (kind 0: random-site Metropolis sweeps,
 * metropolis.rs:8-27; kind 1: Wolff steps, wolff.rs:27-65) from a hot start,
 * measured (forward-bond sum, magnetisation) after every step. */
void o_ref_ising_chain(int kind, int L, uint64_t seed, double beta, int64_t warmup, int64_t n,
                       int64_t* out_bond, int64_t* out_mag) {
    int64_t N = (int64_t)L * L * L;
    int8_t* sp = (int8_t*)malloc((size_t)N);
    int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * (size_t)N);
    uint8_t* mark = (uint8_t*)malloc((size_t)N);
    o_rng rng;
    o_rng_seed(&rng, seed);
    for (int64_t i = 0; i < N; ++i) sp[i] = (o_rng_next(&rng) >> 63) ? 1 : -1;
    for (int64_t s = 0; s < warmup + n; ++s) {
        if (kind == 0) o_ref_metropolis_sweep(L, sp, beta, 1.0, 0.0, &rng);
        else o_ref_wolff_step(L, sp, beta, 1.0, &rng, stack, mark);
        if (s >= warmup) o_ising_measure(L, sp, &out_bond[s - warmup], &out_mag[s - warmup]);
    }
    free(sp); free(stack); free(mark);
}

/* Reference-faithful Heisenberg chain: cap Metropolis + n_or sequential OR
 * sweeps (heisenberg/sweep.rs:50-62), measured (E, |M|) after every sweep. */
void o_ref_heis_chain(int L, uint64_t seed, double beta, double delta, int n_or, int64_t warmup, int64_t n,
                      double* out_e, double* out_m) {
    int64_t N = (int64_t)L * L * L;
    double* sp = (double*)malloc(sizeof(double) * 3 * (size_t)N);
    o_rng rng;
    o_rng_seed(&rng, seed);
    for (int64_t i = 0; i < N; ++i) {
        double z = 2.0 * o_rng_f64(&rng) - 1.0, phi = TWO_PI * o_rng_f64(&rng), r = sqrt(1.0 - z * z);
        sp[3 * i] = r * cos(phi); sp[3 * i + 1] = r * sin(phi); sp[3 * i + 2] = z;
    }
    for (int64_t s = 0; s < warmup + n; ++s) {
        o_ref_heis_metropolis_sweep(L, sp, beta, 1.0, delta, &rng, NULL);
        for (int k = 0; k < n_or; ++k) o_ref_heis_overrelax_sweep(L, sp, 1.0);
        if (s >= warmup) {
            double e, m[3];
            o_cont_measure(L, 3, sp, 1.0, &e, m);
            out_e[s - warmup] = e;
            out_m[s - warmup] = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        }
    }
    free(sp);
}

/* Reference-faithful XY chain: Wolff sweeps (xy/wolff.rs:16-73), measured (E, |M|). */
void o_ref_xy_chain(int L, uint64_t seed, double beta, int64_t warmup, int64_t n, double* out_e, double* out_m) {
    int64_t N = (int64_t)L * L * L;
    double* sp = (double*)malloc(sizeof(double) * 2 * (size_t)N);
    int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * (size_t)N);
    uint8_t* mark = (uint8_t*)malloc((size_t)N);
    o_rng rng;
    o_rng_seed(&rng, seed);
    for (int64_t i = 0; i < N; ++i) {
        double phi = TWO_PI * o_rng_f64(&rng);
        sp[2 * i] = cos(phi); sp[2 * i + 1] = sin(phi);
    }
    for (int64_t s = 0; s < warmup + n; ++s) {
        o_ref_xy_wolff_sweep(L, sp, beta, 1.0, &rng, stack, mark);
        if (s >= warmup) {
            double e, m[3];
            o_cont_measure(L, 2, sp, 1.0, &e, m);
            out_e[s - warmup] = e;
            out_m[s - warmup] = sqrt(m[0] * m[0] + m[1] * m[1]);
        }
    }
    free(sp); free(stack); free(mark);
}

/* ------------------------------------------------------------------------ */
/* Multispin-coded Ising (replicas packed across the bits of a site word)   */
/* ------------------------------------------------------------------------ */
/* Replica r of a size-L group lives in bit r % 32 of word w = r / 32 of each
 * site.  A colour-c half-sweep of word w: aligned-neighbour count a_r per
 * replica; a_r <= 3 (dE <= 0) flips; otherwise accept iff u_r < T(a_r, beta_r)
 * with u_r the 32-bit uniform whose bit (31 - j) is bit (r % 32) of random
 * plane j, plane j = word (j & 3) of Philox(h, 2t + c, (L << 16) | w,
 * (O_P_MSC << 8) | (j >> 2)).  T(a, beta) = floor(2^32 exp(-beta J (4a - 12))).
 * Spins: R arrays of L^3 int8 +-1 (full layout).  thr[r][3] = T(4), T(5), T(6)
 * for replica r's current temperature. */
void o_msc_halfsweep(int L, int R, int8_t* spins, int color, int64_t t, uint64_t seed, const uint32_t* thr) {
    const uint32_t hs = (uint32_t)(2 * t + color);
    int64_t n3 = (int64_t)L * L * L;
    for (int z = 0; z < L; ++z)
        for (int y = 0; y < L; ++y)
            for (int x = 0; x < L; ++x) {
                if (((x + y + z) & 1) != color) continue;
                int64_t s = ((int64_t)z * L + y) * L + x;
                int64_t nb[6] = {((int64_t)z * L + y) * L + (x + 1) % L, ((int64_t)z * L + y) * L + (x + L - 1) % L,
                                 ((int64_t)z * L + (y + 1) % L) * L + x, ((int64_t)z * L + (y + L - 1) % L) * L + x,
                                 ((int64_t)((z + 1) % L) * L + y) * L + x,
                                 ((int64_t)((z + L - 1) % L) * L + y) * L + x};
                uint32_t h = (uint32_t)(s >> 1);
                for (int w = 0; w * 32 < R; ++w) {
                    uint32_t planes[32];
                    for (int q = 0; q < 8; ++q) {
                        uint32_t o[4];
                        ph(h, hs, ((uint32_t)L << 16) | (uint32_t)w, (O_P_MSC << 8) | (uint32_t)q, seed, o);
                        for (int k = 0; k < 4; ++k) planes[4 * q + k] = o[k];
                    }
                    for (int b = 0; b < 32 && w * 32 + b < R; ++b) {
                        int r = w * 32 + b;
                        int8_t* sp = spins + n3 * r;
                        int a = 0;
                        for (int k = 0; k < 6; ++k) a += sp[nb[k]] == sp[s];
                        int flip = 1;
                        if (a >= 4) {
                            uint32_t u = 0;
                            for (int j = 0; j < 32; ++j) u |= ((planes[j] >> b) & 1u) << (31 - j);
                            flip = u < thr[3 * r + (a - 4)];
                        }
                        if (flip) sp[s] = (int8_t)-sp[s];
                    }
                }
            }
}

void o_msc_thresholds(double beta, double J, uint32_t out[3]) {
    for (int a = 4; a <= 6; ++a) {
        double p = exp(-beta * J * (4.0 * a - 12.0));
        out[a - 4] = (uint32_t)ldexp(p, 32); /* p < 1 for a >= 4, J > 0 */
    }
}

int64_t o_msc_pt_run(int L, int R, const double* temps, uint64_t seed, double J, int64_t warmup, int64_t samples,
                     int exchange_every, int measure_every, int32_t* rec_tidx, int64_t* rec_bond, int64_t* rec_mag,
                     int8_t* final_spins) {
    int64_t n3 = (int64_t)L * L * L;
    int8_t* sp = (int8_t*)malloc((size_t)(n3 * R));
    int* r2t = (int*)malloc(sizeof(int) * R);
    int* t2r = (int*)malloc(sizeof(int) * R);
    double* betas = (double*)malloc(sizeof(double) * R);
    double* en = (double*)malloc(sizeof(double) * R);
    uint32_t* tt = (uint32_t*)malloc(sizeof(uint32_t) * 3 * R);   /* per temperature */
    uint32_t* tr = (uint32_t*)malloc(sizeof(uint32_t) * 3 * R);   /* per replica */
    for (int i = 0; i < R; ++i) {
        r2t[i] = t2r[i] = i;
        betas[i] = 1.0 / temps[i];
        o_msc_thresholds(betas[i], J, tt + 3 * i);
        o_ising_init(L, seed, o_stream(L, i), sp + n3 * i);
    }
    int64_t t = 0, nrec = 0;
    for (int64_t s = -warmup; s < samples; ++s, ++t) {
        for (int r = 0; r < R; ++r)
            for (int k = 0; k < 3; ++k) tr[3 * r + k] = tt[3 * r2t[r] + k];
        for (int c = 0; c < 2; ++c) o_msc_halfsweep(L, R, sp, c, t, seed, tr);
        if (s < 0) continue;
        int do_m = ((s + 1) % measure_every) == 0;
        int do_x = ((s + 1) % exchange_every) == 0;
        if (do_m || do_x) {
            for (int r = 0; r < R; ++r) {
                int64_t B, M;
                o_ising_measure(L, sp + n3 * r, &B, &M);
                en[r] = -J * (double)B;
                if (do_m) {
                    if (rec_tidx) rec_tidx[nrec] = r2t[r];
                    if (rec_bond) rec_bond[nrec] = B;
                    if (rec_mag) rec_mag[nrec] = M;
                    ++nrec;
                }
            }
        }
        if (do_x)
            o_exchange(R, betas, en, r2t, t2r, s / exchange_every, seed, o_stream(L, 0), O_MODEL_ISING);
    }
    if (final_spins) memcpy(final_spins, sp, (size_t)(n3 * R));
    free(sp); free(r2t); free(t2r); free(betas); free(en); free(tt); free(tr);
    return nrec;
}

/* ------------------------------------------------------------------------ */
/* Cluster updates (library: csrc/cluster.hip)                              */
/* ------------------------------------------------------------------------ */
static int64_t uf_find(int32_t* lab, int64_t x) {
    while (lab[x] != x) {
        lab[x] = lab[lab[x]];
        x = lab[x];
    }
    return x;
}

static void uf_union(int32_t* lab, int64_t a, int64_t b) {
    a = uf_find(lab, a);
    b = uf_find(lab, b);
    if (a == b) return;
    if (a < b) lab[b] = (int32_t)a;
    else lab[a] = (int32_t)b;
}

uint32_t o_bond_threshold(double beta, double J) {
    double p = -expm1(-2.0 * beta * fabs(J));
    if (!(p > 0.0)) return 0u;
    double t = floor(ldexp(p, 32));
    return t >= 4294967295.0 ? 0xffffffffu : (uint32_t)t;
}

static int64_t fwd(int dim, int L, int x, int y, int z, int d) {
    if (d == 0) x = (x + 1) % L;
    else if (d == 1) y = (y + 1) % L;
    else z = (z + 1) % L;
    (void)dim;
    return ((int64_t)z * L + y) * L + x;
}

static uint64_t mulhi64(uint64_t a, uint64_t b) {
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
}

/* Flip step shared by both spin kinds: returns the sites whose cluster flips in mark[]. */
static int64_t choose_flips(int64_t N, int kind, uint32_t u, uint64_t seed, uint32_t stream, int model,
                            int32_t* lab, const uint8_t* bonds, uint8_t* mark) {
    int64_t rs = -1, cnt = 0;
    if (kind == 1) {
        uint32_t w[4];
        ph(0u, u, stream, o_tag(10, model), seed, w);
        const uint64_t v = ((uint64_t)w[0] << 32) | w[1];
        rs = uf_find(lab, (int64_t)mulhi64(v, (uint64_t)N));
    }
    for (int64_t i = 0; i < N; ++i) {
        const int64_t r = uf_find(lab, i);
        mark[i] = (uint8_t)(kind == 1 ? r == rs : ((bonds[r] >> 3) & 1u));
        cnt += mark[i];
    }
    return cnt;
}

int64_t o_ising_cluster(int dim, int L, int8_t* sp, int kind, uint32_t u, uint64_t seed, uint32_t stream,
                        uint32_t thr, double J, int32_t* lab, uint8_t* bonds) {
    const int Lz = dim == 3 ? L : 1;
    const int64_t N = (int64_t)Lz * L * L;
    const uint32_t hi = thr >> 16, lo = thr & 0xffffu;
    uint8_t* own = NULL;
    if (!bonds) bonds = own = (uint8_t*)malloc((size_t)N);
    for (int64_t i = 0; i < N; ++i) lab[i] = (int32_t)i;
    for (int z = 0; z < Lz; ++z)
        for (int y = 0; y < L; ++y)
            for (int x = 0; x < L; ++x) {
                const int64_t i = ((int64_t)z * L + y) * L + x;
                uint32_t w[4];
                ph((uint32_t)(i >> 1), u, stream, o_tag(7, 0), seed, w);
                const uint32_t wa = w[2 * (x & 1)], wb = w[2 * (x & 1) + 1];
                uint32_t bits = (wb >> 31) << 3;
                for (int d = 0; d < dim; ++d) {
                    const int64_t j = fwd(dim, L, x, y, z, d);
                    const int sat = J > 0.0 ? sp[i] == sp[j] : sp[i] != sp[j];
                    if (!sat) continue;
                    const uint32_t r16 = d == 0 ? wa & 0xffffu : (d == 1 ? wa >> 16 : wb & 0xffffu);
                    int act = r16 < hi;
                    if (r16 == hi && lo != 0u) {
                        uint32_t t[4];
                        ph((uint32_t)i, u, stream, o_tag(8, d), seed, t);
                        act = (t[0] >> 16) < lo;
                        bits |= 32u << d; /* decided by the tie draw (reported, not used) */
                    }
                    if (act) {
                        bits |= 1u << d;
                        uf_union(lab, i, j);
                    }
                }
                bonds[i] = (uint8_t)bits;
            }
    uint8_t* mark = (uint8_t*)malloc((size_t)N);
    const int64_t cnt = choose_flips(N, kind, u, seed, stream, 0, lab, bonds, mark);
    for (int64_t i = 0; i < N; ++i)
        if (mark[i]) sp[i] = (int8_t)-sp[i];
    free(mark);
    free(own);
    return cnt;
}

int64_t o_cont_cluster(int L, int nc, double* sp, int kind, uint32_t u, uint64_t seed, uint32_t stream, int model,
                       double beta, double J, double margin, int64_t* n_marginal, int32_t* lab, uint8_t* bonds) {
    const int64_t N = (int64_t)L * L * L;
    uint8_t* own = NULL;
    if (!bonds) bonds = own = (uint8_t*)malloc((size_t)N);
    double r[3] = {0.0, 0.0, 0.0};
    {
        uint32_t w[4];
        ph(0u, u, stream, o_tag(9, model), seed, w);
        const double two_pi = 6.283185307179586;
        if (nc == 2) {
            const double phi = two_pi * o_u01(w[0], w[1]);
            r[0] = cos(phi);
            r[1] = sin(phi);
        } else {
            const double zc = 2.0 * o_u01(w[0], w[1]) - 1.0, phi = two_pi * o_u01(w[2], w[3]);
            const double s = sqrt(fmax(0.0, 1.0 - zc * zc));
            r[0] = s * cos(phi);
            r[1] = s * sin(phi);
            r[2] = zc;
        }
    }
    double* a = (double*)malloc(sizeof(double) * (size_t)N);
    for (int64_t i = 0; i < N; ++i) {
        double v = 0.0;
        for (int k = 0; k < nc; ++k) v += sp[i * nc + k] * r[k];
        a[i] = v;
        lab[i] = (int32_t)i;
    }
    int64_t marg = 0;
    for (int z = 0; z < L; ++z)
        for (int y = 0; y < L; ++y)
            for (int x = 0; x < L; ++x) {
                const int64_t i = ((int64_t)z * L + y) * L + x;
                uint32_t w[4];
                ph((uint32_t)i, u, stream, o_tag(7, model), seed, w);
                uint32_t bits = (w[3] >> 31) << 3;
                for (int d = 0; d < 3; ++d) {
                    const int64_t j = fwd(3, L, x, y, z, d);
                    const double xx = 2.0 * beta * J * a[i] * a[j];
                    const double uu = u24(w[d]);
                    const double p = xx > 0.0 ? -expm1(-xx) : 0.0;
                    if (fabs(uu - p) < margin) ++marg;
                    if (xx > 0.0 && uu < p) {
                        bits |= 1u << d;
                        uf_union(lab, i, j);
                    }
                }
                bonds[i] = (uint8_t)bits;
            }
    uint8_t* mark = (uint8_t*)malloc((size_t)N);
    const int64_t cnt = choose_flips(N, kind, u, seed, stream, model, lab, bonds, mark);
    for (int64_t i = 0; i < N; ++i)
        if (mark[i])
            for (int k = 0; k < nc; ++k) sp[i * nc + k] -= 2.0 * a[i] * r[k];
    if (n_marginal) *n_marginal = marg;
    free(mark);
    free(a);
    free(own);
    return cnt;
}

void o_ising_cluster_chain(int dim, int L, uint64_t seed, int kind, double beta, double J, int64_t warmup, int64_t n,
                           int64_t* out_bond, int64_t* out_mag, int64_t* out_flipped) {
    const int64_t N = dim == 3 ? (int64_t)L * L * L : (int64_t)L * L;
    /* the hot start is keyed by site index: a square lattice takes the first L^2 of the cube's */
    int8_t* sp = (int8_t*)malloc((size_t)L * L * L);
    int32_t* lab = (int32_t*)malloc(sizeof(int32_t) * (size_t)N);
    const uint32_t stream = o_stream(L, 0), thr = o_bond_threshold(beta, J);
    o_ising_init(L, seed, stream, sp);
    for (int64_t k = 0; k < warmup + n; ++k) {
        const int64_t f = o_ising_cluster(dim, L, sp, kind, (uint32_t)k, seed, stream, thr, J, lab, NULL);
        if (k < warmup) continue;
        int64_t b, m;
        if (dim == 3) o_ising_measure(L, sp, &b, &m);
        else o_ising2d_measure(L, sp, &b, &m);
        if (out_bond) out_bond[k - warmup] = b;
        if (out_mag) out_mag[k - warmup] = m;
        if (out_flipped) out_flipped[k - warmup] = f;
    }
    free(lab);
    free(sp);
}
