// Sweep control shared by the int8 and multispin-coded Ising kernels.
//
// A sweep is two launches, colour 0 then colour 1; there is no closing launch
// between sweeps.  The closing step of sweep t (records, replica exchange) is
// done at the start of the colour-0 launch of sweep t + 1, and a closing
// kernel runs only after the last sweep of a host call, so that what the host
// reads is complete.  Every field that a launch reads is written only by
// other launches (a kernel boundary orders them):
//
//   colour 0 reads  t, flags, closed, flp, round, perm[round & 1], acc[(t-1) & 1]
//            writes (one block) t1, fl1, round1, records;  (first block of each
//            group) perm[(round + 1) & 1] when it exchanges;  acc[t & 1] by atomics
//   colour 1 reads  t1, fl1, round1, perm[round1 & 1]
//            writes (one block) t, flags (next sweep), flp, round, phases,
//            zero acc[(t1 + 1) & 1];  acc[t1 & 1] by atomics
//   closing  reads and writes what colour 0 would, for sweep t - 1, and sets
//            closed = t - 1 so that the next colour-0 launch does not redo it.
//
// Every colour-0 block that closes an exchange sweep decides its own
// replica's pair itself (the decision is a pure function of the committed
// energies and counters), so all blocks agree with the permutation that the
// group's first block writes.
#pragma once
#include "kernels.h"

namespace ising {

// Accumulators of the sweep with parity (t & 1): [copy][global replica][2].
__device__ __forceinline__ int64_t* acc_of(const IsingArgs& a, int64_t t) { return a.acc + (t & 1) * a.acc_stride; }

// Quantity q (0: bond sum, 1: magnetisation) of global replica gr, summed over copies.
__device__ __forceinline__ int64_t acc_total(const IsingArgs& a, const int64_t* acc, int gr, int q) {
    int64_t v = 0;
    for (int c = 0; c < a.acc_copies; ++c)
        v += __hip_atomic_load(&acc[2 * ((int64_t)c * a.total_reps + gr) + q], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    return v;
}

// Metropolis decision of exchange round `round` for the temperature pair
// (i, i + 1) of group G (parallel_tempering.rs:5-36, with Delta on the
// committed bond sums and one Philox draw per pair and round).
// Takes the bond sums of the replicas at temperatures i and i + 1.
__device__ __forceinline__ bool exch_accept(const IsingArgs& a, const DevGroup& G, int i, int64_t round, int64_t bi,
                                            int64_t bj) {
    const double Ei = -a.J * (double)bi;
    const double Ej = -a.J * (double)bj;
    const double delta = (a.betas[G.rep0 + i] - a.betas[G.rep0 + i + 1]) * (Ej - Ei);
    const u32x4 w = philox((uint32_t)i, (uint32_t)round, stream_id((uint32_t)G.L, 0u), tag(P_EXCH, MODEL_ISING), a.k0,
                           a.k1);
    return delta <= 0.0 || u01(w.x, w.y) < det_exp(-delta);
}

__device__ __forceinline__ bool exch_pair(const IsingArgs& a, const DevGroup& G, int i, int64_t round,
                                          const int32_t* t2r_old, const int64_t* acc) {
    const int ri = t2r_old[G.rep0 + i], rj = t2r_old[G.rep0 + i + 1];
    return exch_accept(a, G, i, round, acc_total(a, acc, G.rep0 + ri, 0), acc_total(a, acc, G.rep0 + rj, 0));
}

// Temperature index of replica r of group G after exchange round `round`.
__device__ __forceinline__ int exch_tidx(const IsingArgs& a, const DevGroup& G, int r, int64_t round,
                                         const int32_t* r2t_old, const int32_t* t2r_old, const int64_t* acc) {
    const int i = r2t_old[G.rep0 + r];
    const int par = (int)(round & 1);
    if (i < par) return i;
    const int i0 = par + ((i - par) & ~1);
    if (i0 + 1 >= G.R) return i;
    if (!exch_pair(a, G, i0, round, t2r_old, acc)) return i;
    return i == i0 ? i0 + 1 : i0;
}

// The same from a block's copies of the group's permutation (temperature ->
// replica, t2r[k] for k < G.R) and bond sums by replica (bond[r]), gathered
// in one round of loads; i is the replica's temperature before the round.
__device__ __forceinline__ int exch_tidx_gathered(const IsingArgs& a, const DevGroup& G, int i, int64_t round,
                                                  const int32_t* t2r, const int64_t* bond) {
    const int par = (int)(round & 1);
    if (i < par) return i;
    const int i0 = par + ((i - par) & ~1);
    if (i0 + 1 >= G.R) return i;
    if (!exch_accept(a, G, i0, round, bond[t2r[i0]], bond[t2r[i0 + 1]])) return i;
    return i == i0 ? i0 + 1 : i0;
}

// Writes the permutation after exchange round `round` for group G into the
// other parity buffer (copy, then the accepted swaps) and counts the tries.
// All threads of one block.
__device__ __forceinline__ void exch_write_group(const IsingArgs& a, const DevGroup& G, int64_t round,
                                                 const int64_t* acc) {
    const int po = (int)(round & 1);
    const int32_t* r2t_o = a.r2t + (int64_t)po * a.total_reps;
    const int32_t* t2r_o = a.t2r + (int64_t)po * a.total_reps;
    int32_t* r2t_n = a.r2t + (int64_t)(po ^ 1) * a.total_reps;
    int32_t* t2r_n = a.t2r + (int64_t)(po ^ 1) * a.total_reps;
    for (int k = threadIdx.x; k < G.R; k += blockDim.x) {
        r2t_n[G.rep0 + k] = r2t_o[G.rep0 + k];
        t2r_n[G.rep0 + k] = t2r_o[G.rep0 + k];
    }
    __syncthreads();
    const int npairs = (G.R - po) / 2;
    for (int q = threadIdx.x; q < npairs; q += blockDim.x) {
        const int i = po + 2 * q;
        if (exch_pair(a, G, i, round, t2r_o, acc)) {
            const int ri = t2r_o[G.rep0 + i], rj = t2r_o[G.rep0 + i + 1];
            t2r_n[G.rep0 + i] = rj;
            t2r_n[G.rep0 + i + 1] = ri;
            r2t_n[G.rep0 + ri] = i + 1;
            r2t_n[G.rep0 + rj] = i;
            atomicAdd((unsigned long long*)&a.ctl->exch_accepts, 1ull);
        }
    }
    if (threadIdx.x == 0) atomicAdd((unsigned long long*)&a.ctl->exch_tries, (unsigned long long)npairs);
    if (a.thr_rep) {
        // thresholds by replica under the new permutation
        __syncthreads();
        for (int k = threadIdx.x; k < G.R * kIsingTab; k += blockDim.x) {
            const int rr = k / kIsingTab, e = k - rr * kIsingTab;
            a.thr_rep[(int64_t)(G.rep0 + rr) * kIsingTab + e] =
                a.thr[(int64_t)(G.rep0 + r2t_n[G.rep0 + rr]) * kIsingTab + e];
        }
    }
}

// exch_write_group for every group at once (one block; the closing kernel):
// the same permutations, counts and tables, each phase over all groups, so
// the block waits at three barriers instead of three per group.
__device__ __forceinline__ void exch_write_all(const IsingArgs& a, int64_t round, const int64_t* acc) {
    const int po = (int)(round & 1);
    const int32_t* r2t_o = a.r2t + (int64_t)po * a.total_reps;
    const int32_t* t2r_o = a.t2r + (int64_t)po * a.total_reps;
    int32_t* r2t_n = a.r2t + (int64_t)(po ^ 1) * a.total_reps;
    int32_t* t2r_n = a.t2r + (int64_t)(po ^ 1) * a.total_reps;
    for (int k = threadIdx.x; k < a.total_reps; k += blockDim.x) {
        r2t_n[k] = r2t_o[k];
        t2r_n[k] = t2r_o[k];
    }
    __syncthreads();
    int base = 0;  // pairs of the groups before g
    int64_t tries = 0;
    for (int g = 0; g < a.ngroups; ++g) {
        const DevGroup G = a.groups[g];
        const int npairs = (G.R - po) / 2;
        for (int q = (int)threadIdx.x - base; q < npairs; q += blockDim.x) {
            if (q < 0) continue;
            const int i = po + 2 * q;
            if (exch_pair(a, G, i, round, t2r_o, acc)) {
                const int ri = t2r_o[G.rep0 + i], rj = t2r_o[G.rep0 + i + 1];
                t2r_n[G.rep0 + i] = rj;
                t2r_n[G.rep0 + i + 1] = ri;
                r2t_n[G.rep0 + ri] = i + 1;
                r2t_n[G.rep0 + rj] = i;
                atomicAdd((unsigned long long*)&a.ctl->exch_accepts, 1ull);
            }
        }
        base = (base + npairs) % (int)blockDim.x;
        tries += npairs;
    }
    if (threadIdx.x == 0) atomicAdd((unsigned long long*)&a.ctl->exch_tries, (unsigned long long)tries);
    if (a.thr_rep) {
        __syncthreads();
        for (int k = threadIdx.x; k < a.total_reps * kIsingTab; k += blockDim.x) {
            const int gr = k / kIsingTab, e = k - gr * kIsingTab;
            int g = 0;
            while (g + 1 < a.ngroups && gr >= a.groups[g + 1].rep0) ++g;
            a.thr_rep[(int64_t)gr * kIsingTab + e] = a.thr[(int64_t)(a.groups[g].rep0 + r2t_n[gr]) * kIsingTab + e];
        }
    }
}

// Measurement records of sweep `ts` (one per replica, in global replica
// order, temperature index from the permutation in use during that sweep).
// All threads of one block; returns nothing, updates n_rec / overflow.
__device__ __forceinline__ void write_records(const IsingArgs& a, const int64_t* acc, const int32_t* r2t) {
    DevCtl* ctl = a.ctl;
    const int nrec = ctl->n_rec;
    const bool room = nrec < ctl->rec_cap;
    if (room)
        for (int gr = threadIdx.x; gr < a.total_reps; gr += blockDim.x) {
            IsingRec r;
            r.bond = acc_total(a, acc, gr, 0);
            r.mag = acc_total(a, acc, gr, 1);
            r.tidx = r2t[gr];
            r.pad = 0;
            a.rec[(int64_t)nrec * a.total_reps + gr] = r;
        }
    if (threadIdx.x == 0) {
        if (room) ctl->n_rec = nrec + 1;
        else ctl->overflow = 1;
    }
}

// What a colour launch needs to know about its sweep.
struct SweepView {
    int64_t t;       // sweep index
    int32_t flags;   // F_* of sweep t
    bool closing;    // colour 0 only: this launch closes sweep t - 1
    int32_t pflags;  // F_* of sweep t - 1 when closing
    int64_t round;   // exchanges committed before sweep t - 1 closes (colour 0) / before sweep t (colour 1)
    bool exch;       // colour 0 only: closing performs exchange round `round`
    int par;         // permutation parity this launch runs with
};

__device__ __forceinline__ SweepView view_of(const DevCtl* c, int color) {
    SweepView v;
    if (color == 0) {
        v.t = c->t;
        v.flags = c->flags;
        v.closing = c->closed < v.t - 1;
        v.pflags = v.closing ? c->flp : 0;
        v.round = c->round;
        v.exch = (v.pflags & F_EXCH) != 0;
        v.par = (int)((v.round + (v.exch ? 1 : 0)) & 1);
    } else {
        v.t = c->t1;
        v.flags = c->fl1;
        v.closing = false;
        v.pflags = 0;
        v.round = c->round1;
        v.exch = false;
        v.par = (int)(v.round & 1);
    }
    return v;
}

// Per-sweep bookkeeping, one block of each colour launch (all its threads).
__device__ __forceinline__ void sweep_bookkeeping(const IsingArgs& a, const SweepView& v, int color) {
    DevCtl* ctl = a.ctl;
    if (color == 0) {
        if (v.closing && (v.pflags & F_MEAS))
            write_records(a, acc_of(a, v.t - 1), a.r2t + (int64_t)(v.round & 1) * a.total_reps);
        if (threadIdx.x == 0) {
            ctl->t1 = v.t;
            ctl->fl1 = v.flags;
            ctl->round1 = v.round + (v.exch ? 1 : 0);
        }
    } else {
        int64_t* nxt = acc_of(a, v.t + 1);  // consumed by the closing of sweep t - 1
        for (int64_t i = threadIdx.x; i < a.acc_stride; i += blockDim.x) nxt[i] = 0;
        if (threadIdx.x == 0) {
            ctl->flp = v.flags;
            ctl->round = v.round;
            ctl->t = v.t + 1;
            advance_phases(ctl);
        }
    }
}

}  // namespace ising
