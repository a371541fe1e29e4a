// Host side of libising_hip.so: the C ABI declared in include/ising_hip.h.
//
// The ensemble object owns every lattice of a parallel-tempering run on one
// device: spin pool, per-temperature acceptance tables, temperature
// permutations, measurement records and the sweep control block.  A sweep is
// two launches (colour 0, colour 1); measurement, replica exchange and the
// sweep counter live on the device, so the host only enqueues launches and
// drains records.  Host code never computes spins or observables: there is no
// CPU fallback, and every entry point that needs a device fails loudly without one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/ising_hip.h"
#include "kernels.h"

using namespace ising;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

}  // namespace

// Error reporting for entry points defined in other source files of the library.
int ising::set_error(int code, const char* msg) { return fail(code, msg); }

namespace {

#define HIP_TRY(expr)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(ISING_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
    } while (0)

#define SYNC_ALL(p)                        \
    do {                                   \
        const int rs_ = (p)->sync_all();   \
        if (rs_ != ISING_OK) return rs_;   \
    } while (0)

#define ORDER(p)                           \
    do {                                   \
        const int ro_ = (p)->order();      \
        if (ro_ != ISING_OK) return ro_;   \
    } while (0)

// Transfers to and from an ensemble's buffers are ordered on the ensemble's
// own stream and waited for: that stream is non-blocking, so the legacy null
// stream (plain hipMemcpy / hipMemset) would not be ordered against its kernels.
#define ST_COPY(p, dst, src, n, kind)                                        \
    do {                                                                     \
        HIP_TRY(hipMemcpyAsync((dst), (src), (n), (kind), (p)->st));         \
        HIP_TRY(hipStreamSynchronize((p)->st));                              \
    } while (0)

#define ST_FILL(p, dst, v, n)                                                \
    do {                                                                     \
        HIP_TRY(hipMemsetAsync((dst), (v), (n), (p)->st));                   \
        HIP_TRY(hipStreamSynchronize((p)->st));                              \
    } while (0)

// 33-bit threshold T = floor(p * 2^32) (T = 2^32 when p >= 1), split in halves.
void threshold_pair(double p, uint32_t* hi, uint32_t* lo) {
    uint64_t t;
    if (!(p < 1.0)) t = 1ull << 32;
    else if (p <= 0.0) t = 0;
    else t = (uint64_t)std::ldexp(p, 32);
    *hi = (uint32_t)(t >> 25);         // 0..128, compared with a 7-bit draw
    *lo = (uint32_t)(t & 0x1ffffffu);  // compared with 25 more bits on a tie
}

// Acceptance table for k = b*(z+1) + n_up, z = 2 * dims neighbours:
// dE = 2 s (J (2 n_up - z) + h), T(k) = floor(2^32 min(1, exp(-beta dE))) =
// hi * 2^25 + lo.  out[0..13] = hi, out[14..27] = lo, and out[28..31] the hi
// bytes packed for byte-permute lookups: out[28] = hi[0..3], out[29] =
// hi[4..6] (b = 0), out[30] = hi[7..10], out[31] = hi[11..13] (b = 1).
void ising_table(double beta, double J, double h, uint32_t out[kIsingTab], int dims = 3) {
    const int z = 2 * dims;
    for (int i = 0; i < kIsingTab; ++i) out[i] = 0;
    for (int b = 0; b <= 1; ++b) {
        const double s = b ? 1.0 : -1.0;
        for (int n = 0; n <= z; ++n) {
            const double dE = 2.0 * s * (J * (double)(2 * n - z) + h);
            const double p = dE <= 0.0 ? 1.0 : std::exp(-beta * dE);
            threshold_pair(p, &out[b * (z + 1) + n], &out[14 + b * (z + 1) + n]);
        }
    }
    if (dims == 3)
        for (int b = 0; b <= 1; ++b)
            for (int n = 0; n <= 6; ++n) out[28 + 2 * b + (n >> 2)] |= out[7 * b + n] << (8 * (n & 3));
}

bool valid_L(int L) { return L >= 2 && L <= 2046 && (L % 2) == 0; }

}  // namespace

struct ising_pt {
    int model = ISING_MODEL_ISING;
    int ncomp = 0;
    int dims = 3;
    int R = 0;
    int total_reps = 0;
    int nblocks = 0;
    uint64_t seed = 0;
    ising_pt_opts_t opts{};
    std::vector<DevGroup> groups;
    std::vector<double> temps;
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    DevGroup* d_groups = nullptr;
    uint8_t* d_pool = nullptr;
    int64_t pool_bytes = 0;
    uint32_t* d_thr = nullptr;
    uint32_t* d_thr_rep = nullptr;  // int8 Ising: thresholds by replica under the permutation in force
    uint32_t* d_mthr = nullptr;  // multispin coding thresholds
    uint32_t* d_mplanes = nullptr;  // multispin coding threshold bit planes per word
    int32_t* d_r2t = nullptr;
    int32_t* d_t2r = nullptr;
    double* d_betas = nullptr;
    DevCtl* d_ctl = nullptr;
    int64_t* d_acc = nullptr;
    void* d_rec = nullptr;
    // continuous models
    float* d_fpool = nullptr;
    float* d_deltas = nullptr;
    float* d_omcd = nullptr;  // 1 - cos(delta) per temperature slot (Heisenberg cap)
    float* d_betas_f = nullptr;
    double* d_part0 = nullptr;
    double* d_part1 = nullptr;
    double* d_dacc = nullptr;
    int64_t* d_cnt = nullptr;
    bool inline_cont = false;  // single-lattice continuous mode
    // learned proposal
    int learn_H = 0, learn_P = 0;
    bool learned_on = false;
    bool heatbath_on = false;  // continuous: heat-bath passes in place of the Metropolis passes
    int64_t learn_step = 0;
    float* d_lmaster = nullptr;
    uint16_t* d_lnet = nullptr;
    uint4* d_lfrag = nullptr;  // d_lnet as per-lane matrix-core fragments (rebuilt after every parameter change)
    float *d_lm1 = nullptr, *d_lm2 = nullptr, *d_lgpart = nullptr, *d_lgout = nullptr;
    double* d_llpart = nullptr;
    int learn_cap_blocks = 0;
    int64_t* d_mh = nullptr;  // [kMhSlots][8] learned-move counters
    float beta_inline = 0.f, delta_inline = 0.f, j_inline = 1.f;
    int layout_width = 0, layout_runs = 0, layout_spt = 0;  // Ising thread-step layout of the first size
    std::vector<std::array<int, 3>> glayout;  // Ising thread-step layout per size: run width, runs, steps per tile
    int or_spt = 0;  // continuous: sites per thread of the overrelaxation pass (0: the model's default)
    // continuous: 64-bit offsets in every pass (ISING_HIP_OFFSETS64 set at
    // creation; otherwise only lattices whose half-lattices reach 2^30 sites
    // use them), so tests can run that path at sizes the oracle replays
    bool offsets64 = std::getenv("ISING_HIP_OFFSETS64") != nullptr;
    int fused_z = 0;         // one-launch sweeps: z planes per block asked for (0: two launches per sweep)
    bool fused_on = false;   // one-launch sweeps in use (every size qualifies)
    int fnblocks = 0, fused_lds = 0;
    uint8_t* d_pool2 = nullptr;  // one-launch sweeps: the second lattice buffer
    uint32_t* d_sched = nullptr;   // annealing schedule: table rows per sweep
    int64_t sched_cap = 0;         // rows allocated
    bool d_sched_on = false;
    int64_t sched_t0 = 0;
    int64_t* d_smeas = nullptr;    // annealing: per-sweep B, M of every replica
    int64_t smeas_cap = 0;         // int64s allocated
    // cluster updates (allocated at the first call)
    std::vector<ClusterGroup> cgroups;
    ClusterGroup* d_cg = nullptr;
    int cl_blocks = 0;
    int32_t* d_clabel = nullptr;
    uint8_t* d_cbond = nullptr;
    uint32_t* d_cbthr = nullptr;
    int64_t* d_cflip = nullptr;   // [update][global replica] sites flipped
    int64_t cflip_cap = 0;
    double* d_cmeas = nullptr;    // continuous: [update][global replica][4] E, Mx, My, Mz
    int64_t cmeas_cap = 0;
    int64_t c_host = 0;           // cluster updates enqueued (Philox c1 of the next one)
    int64_t t_host = 0;        // sweeps enqueued
    int64_t s_host = -1;       // sample sweeps enqueued (-1 before sampling starts)
    int64_t pending_host = 0;  // measured sweeps the device will have written
    bool inline_thr = false;   // single-lattice mode
    uint32_t thr_inline[kIsingTab];
    hipGraphExec_t graph_exec = nullptr;  // `opts.graph_sweeps` sweeps, captured once

    ~ising_pt() {
        if (st) (void)hipStreamSynchronize(st);
        (void)hipFree(d_groups);
        (void)hipFree(d_pool);
        (void)hipFree(d_thr);
        (void)hipFree(d_thr_rep);
        (void)hipFree(d_pool2);
        (void)hipFree(d_mthr);
        (void)hipFree(d_mplanes);
        (void)hipFree(d_sched);
        (void)hipFree(d_smeas);
        (void)hipFree(d_r2t);
        (void)hipFree(d_t2r);
        (void)hipFree(d_betas);
        (void)hipFree(d_ctl);
        (void)hipFree(d_acc);
        (void)hipFree(d_rec);
        (void)hipFree(d_fpool);
        (void)hipFree(d_deltas);
        (void)hipFree(d_omcd);
        (void)hipFree(d_betas_f);
        (void)hipFree(d_part0);
        (void)hipFree(d_part1);
        (void)hipFree(d_dacc);
        (void)hipFree(d_cnt);
        (void)hipFree(d_lmaster);
        (void)hipFree(d_lnet);
        (void)hipFree(d_lfrag);
        (void)hipFree(d_lm1);
        (void)hipFree(d_lm2);
        (void)hipFree(d_lgpart);
        (void)hipFree(d_lgout);
        (void)hipFree(d_llpart);
        (void)hipFree(d_mh);
        (void)hipFree(d_cg);
        (void)hipFree(d_clabel);
        (void)hipFree(d_cbond);
        (void)hipFree(d_cbthr);
        (void)hipFree(d_cflip);
        (void)hipFree(d_cmeas);
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (ext_ev) (void)hipEventDestroy(ext_ev);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (st) (void)hipStreamDestroy(st);
    }

    IsingArgs ising_args() const {
        IsingArgs a{};
        a.groups = d_groups;
        a.ngroups = (int32_t)groups.size();
        a.total_reps = total_reps;
        a.nblocks = nblocks;
        a.pool = d_pool;
        a.thr = d_thr;
        a.thr_rep = d_thr_rep;
        a.fnblocks = fnblocks;
        a.gk_on = groups.size() <= (size_t)kArgGroups ? 1 : 0;
        if (a.gk_on) std::memcpy(a.gk, groups.data(), sizeof(DevGroup) * groups.size());
        a.r2t = d_r2t;
        a.t2r = d_t2r;
        a.betas = d_betas;
        a.ctl = d_ctl;
        a.acc = d_acc;
        a.rec = (IsingRec*)d_rec;
        a.k0 = (uint32_t)seed;
        a.k1 = (uint32_t)(seed >> 32);
        a.J = opts.j;
        a.inline_thr = inline_thr ? 1 : 0;
        if (inline_thr) std::memcpy(a.thr_inline, thr_inline, sizeof(thr_inline));
        a.mthr = d_mthr;
        a.mplanes = d_mplanes;
        a.acc_copies = model == ISING_MODEL_ISING_MSC ? kMscAccCopies : kIsingAccCopies;
        a.acc_stride = 2 * (int64_t)a.acc_copies * total_reps;
        a.sched = d_sched_on ? d_sched : nullptr;
        a.sched_t0 = sched_t0;
        return a;
    }

    ContArgs cont_args() const {
        ContArgs a{};
        a.groups = d_groups;
        a.ngroups = (int32_t)groups.size();
        a.total_reps = total_reps;
        a.ncomp = ncomp;
        a.model = model == ISING_MODEL_XY ? MODEL_XY : MODEL_HEIS;
        a.pool = d_fpool;
        a.deltas = d_deltas;
        a.omcd = d_omcd;
        a.betas_f = d_betas_f;
        a.betas = d_betas;
        a.r2t = d_r2t;
        a.t2r = d_t2r;
        a.ctl = d_ctl;
        a.part0 = d_part0;
        a.part1 = d_part1;
        a.dacc = d_dacc;
        a.acc_cnt = d_cnt;
        a.rec = (ContRec*)d_rec;
        a.k0 = (uint32_t)seed;
        a.k1 = (uint32_t)(seed >> 32);
        a.J = inline_cont ? (double)j_inline : opts.j;
        a.Jf = inline_cont ? j_inline : (float)opts.j;
        a.renorm_every = opts.renorm_every;
        a.adapt_every = opts.adapt_every;
        a.adapt_target = (float)opts.adapt_target;
        a.inline_params = inline_cont ? 1 : 0;
        a.beta_inline = beta_inline;
        a.delta_inline = delta_inline;
        a.omcd_inline = (float)(1.0 - std::cos((double)delta_inline));
        a.net = d_lnet;
        a.lfrag = d_lfrag;
        a.mh = d_mh;
        a.learn_H = learn_H;
        a.or_spt = or_spt;
        a.off32 = offsets64 ? 0 : 1;
        for (const DevGroup& g : groups)
            if (g.stride >= (int64_t(1) << 30)) a.off32 = 0;
        return a;
    }

    // Sweeps through a captured graph where possible.  Kernel arguments of an
    // ensemble never change after creation and every per-sweep decision is read
    // from the device control block, so one captured block of sweeps replays
    // unchanged.  Single-lattice handles pass per-call parameters by value and
    // always launch directly.
    bool graph_capable() const { return opts.graph_sweeps > 0 && !inline_thr && !inline_cont; }

    // Captures `opts.graph_sweeps` sweeps into graph_exec (nothing runs).
    int capture_graph() {
        const int G = opts.graph_sweeps;
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_sweeps(G);
        hipGraph_t captured = nullptr;
        const hipError_t ec = hipStreamEndCapture(st, &captured);
        t_host -= G;  // capture enqueued nothing
        if (rc) {
            if (captured) (void)hipGraphDestroy(captured);
            return rc;
        }
        HIP_TRY(ec);
        graph = captured;
        const hipError_t ei = hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIP_TRY(ei);
        HIP_TRY(hipGraphUpload(graph_exec, st));  // the first replay then starts without the upload
        return ISING_OK;
    }

    int run_sweeps(int64_t n) {
        const int G = opts.graph_sweeps;
        if (graph_capable() && n >= G) {
            if (!graph_exec) {
                const int rc = capture_graph();
                if (rc) return rc;
            }
            const int64_t reps = n / G;
            for (int64_t i = 0; i < reps; ++i) HIP_TRY(hipGraphLaunch(graph_exec, st));
            t_host += reps * G;
            n -= reps * G;
        }
        return n > 0 ? enqueue_sweeps(n) : ISING_OK;
    }

    int enqueue_sweeps(int64_t n, int n_or = -1, bool close_last = true) {
        if (model == ISING_MODEL_ISING && fused_on) {
            // one launch per sweep, alternating the two lattice buffers; it
            // closes the sweep before it
            if (!d_pool2) return fail(ISING_E_STATE, "second lattice buffer missing");
            IsingArgs a = ising_args();
            for (int64_t i = 0; i < n; ++i) {
                a.pool = (i & 1) ? d_pool2 : d_pool;
                a.pool_out = (i & 1) ? d_pool : d_pool2;
                HIP_TRY(launch_ising_fused(a, fused_lds, st));
            }
            if (n & 1) HIP_TRY(hipMemcpyAsync(d_pool, d_pool2, pool_bytes, hipMemcpyDeviceToDevice, st));
            a.pool = d_pool;
            if (n > 0 && close_last) HIP_TRY(launch_ising_close(a, st));
        } else if (model == ISING_MODEL_ISING) {
            // each colour-0 launch closes the sweep before it; the closing
            // kernel closes the last one (see sweepctl.h)
            const IsingArgs a = ising_args();
            for (int64_t i = 0; i < n; ++i) {
                HIP_TRY(launch_ising_half(a, nblocks, 0, st));
                HIP_TRY(launch_ising_half(a, nblocks, 1, st));
            }
            if (n > 0 && close_last) HIP_TRY(launch_ising_close(a, st));
        } else if (model == ISING_MODEL_ISING_MSC) {
            const IsingArgs a = ising_args();
            for (int64_t i = 0; i < n; ++i) {
                HIP_TRY(launch_msc_half(a, nblocks, 0, st));
                HIP_TRY(launch_msc_half(a, nblocks, 1, st));
            }
            if (n > 0 && close_last) HIP_TRY(launch_ising_close(a, st));
        } else {
            // overrelaxation passes first, then one Metropolis sweep
            // (the order of ContinuousGpuLattice::sweep, gpu_lattice_continuous.rs:124-168)
            const ContArgs a = cont_args();
            const int nor = n_or >= 0 ? n_or : opts.n_overrelax;
            for (int64_t i = 0; i < n; ++i) {
                for (int k = 0; k < nor; ++k) {
                    HIP_TRY(launch_cont_or(a, nblocks, 0, st));
                    HIP_TRY(launch_cont_or(a, nblocks, 1, st));
                }
                if (heatbath_on) {
                    HIP_TRY(launch_cont_heatbath(a, nblocks, 0, 1, -1, st));
                    HIP_TRY(launch_cont_heatbath(a, nblocks, 1, 1, -1, st));
                } else if (learned_on) {
                    HIP_TRY(launch_cont_learned(a, nblocks, 0, 1, -1, st));
                    HIP_TRY(launch_cont_learned(a, nblocks, 1, 1, -1, st));
                } else {
                    HIP_TRY(launch_cont_metro(a, nblocks, 0, 1, st));
                    HIP_TRY(launch_cont_metro(a, nblocks, 1, 1, st));
                }
                HIP_TRY(launch_cont_close(a, st));
            }
        }
        t_host += n;
        return ISING_OK;
    }

    hipStream_t ext_st = nullptr;  // stream of a batch that also drives this ensemble
    hipEvent_t ext_ev = nullptr;
    int graph_gen = 0;             // bumped whenever captured launches go stale

    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        graph_exec = nullptr;
        ++graph_gen;
    }

    // Work enqueued on this ensemble's stream follows what a batch queued before.
    int order() {
        if (!ext_st) return ISING_OK;
        if (!ext_ev) HIP_TRY(hipEventCreateWithFlags(&ext_ev, hipEventDisableTiming));
        HIP_TRY(hipEventRecord(ext_ev, ext_st));
        HIP_TRY(hipStreamWaitEvent(st, ext_ev, 0));
        return ISING_OK;
    }

    int sync_all() {
        if (ext_st) HIP_TRY(hipStreamSynchronize(ext_st));
        HIP_TRY(hipStreamSynchronize(st));
        return ISING_OK;
    }

    // The device's control block, once every enqueued sweep has run.  Every
    // host read of device state goes through here, and fails if a sweep the
    // host enqueued was not closed on the device or a launch left a block
    // count behind.
    int read_ctl(DevCtl* out) {
        const int rc = sync_all();
        if (rc) return rc;
        ST_COPY(this, out, d_ctl, sizeof(DevCtl), hipMemcpyDeviceToHost);
        if (out->t != t_host)
            return fail(ISING_E_STATE, "device sweep counter " + std::to_string(out->t) + " differs from the " +
                                           std::to_string(t_host) + " sweeps enqueued");
        if (out->cdone != 0 || out->fdone != 0) return fail(ISING_E_STATE, "a sweep was left unclosed on the device");
        return ISING_OK;
    }

    // First sample sweep: fix sample_start and the phase counters on the device.
    int begin_sampling() {
        if (s_host >= 0) return ISING_OK;
        DevCtl c;
        const int rc = read_ctl(&c);
        if (rc) return rc;
        c.sample_start = t_host;
        set_flags(&c, t_host);
        ST_COPY(this, d_ctl, &c, sizeof(DevCtl), hipMemcpyHostToDevice);
        s_host = 0;
        return ISING_OK;
    }
};

struct ising_lattice {
    ising_pt* pt = nullptr;
    int L = 0;
    float beta = NAN, j = NAN, h = NAN;
    ~ising_lattice() { delete pt; }
};

namespace {

int64_t measured_in(int64_t s0, int64_t n, int every) {
    // count s in [s0, s0+n) with (s+1) % every == 0
    return (s0 + n) / every - s0 / every;
}

// Accumulator words: two sweep parities of [copies][replica][2], then 8
// slots for standalone measurements.
static int64_t acc_words(int total_reps) { return 2 * 2 * (int64_t)kMscAccCopies * total_reps + 8; }

// Ising thread-step layout of every group: 3 = 32 sites in one row
// (L % 64 == 0), 2 = 16-site runs in one row (L % 32 == 0; `runs` of them per
// step, a block row apart), 1 = two 8-site runs (L % 16 == 0), 0 = 16 sites
// addressed one by one (other L, and 2D).  width (16 or 32; 0 = 16) caps the
// run where L % 64 == 0; runs (1 or 2; 0 = 2) applies to 16-site runs; spt
// (1..4; 0 = 1) is the number of thread steps per block tile.  Results do not
// depend on any of them: only the grid changes.
// Per-size layouts: lw[g], lr[g], ls[g] as above for group g (results do not depend on them).
static int layout_ising_each(ising_pt* p, const int* lw, const int* lr, const int* ls) {
    p->glayout.resize(p->groups.size());
    int blk = 0;
    for (size_t gi = 0; gi < p->groups.size(); ++gi) {
        DevGroup& G = p->groups[gi];
        const int width = lw[gi] == 0 ? 16 : lw[gi], runs = lr[gi] == 0 ? 2 : lr[gi], spt = ls[gi] == 0 ? 1 : ls[gi];
        G.fast = G.dims == 2 ? 0 : ((G.Lh % 32) == 0 ? 3 : ((G.Lh % 16) == 0 ? 2 : ((G.Lh % 8) == 0 ? 1 : 0)));
        if (G.fast == 3 && width == 16) G.fast = 2;
        G.runs = G.fast == 2 ? runs : 1;
        G.spt = G.fast ? spt : 1;
        p->glayout[gi][0] = width;
        p->glayout[gi][1] = runs;
        p->glayout[gi][2] = spt;
        G.tsz = (int32_t)((G.fast == 3 || G.runs == 2 ? 2 : 1) * kIsingTile * G.spt);
        G.tiles = (int32_t)((G.n_half + G.tsz - 1) / G.tsz);
        G.mag_T = G.sh_T = 0;
        if (G.tiles >= 2) fastdiv_magic((uint32_t)G.tiles, &G.mag_T, &G.sh_T);
        G.blk0 = blk;
        if ((int64_t)blk + (int64_t)G.tiles * G.R > INT32_MAX / 2) return fail(ISING_E_ARG, "ensemble too large");
        blk += G.tiles * G.R;
    }
    p->nblocks = blk;
    // one-launch sweeps: sizes with L % 32 == 0 (L <= 128) streamed in z,
    // fused_z planes (at most L) of P = 16384 / L^2 replicas per block; sizes
    // with L < 32 whole, one replica per block; every size must qualify.  The
    // largest size takes the first blocks (its blocks run longest).
    {
        int64_t lds = 0;
        bool ok = p->fused_z > 0;
        for (DevGroup& G : p->groups) {
            G.fz = 0;
            G.frep = 1;
            G.fblocks = G.R;
            if (G.dims != 3) {
                ok = false;
            } else if (G.L < 32) {
                if ((G.Lh % 4) != 0 || ((int64_t)G.L * G.L * G.Lh) % 16 != 0) ok = false;
                lds = std::max<int64_t>(lds, (int64_t)G.L * G.L * G.L);
            } else if (ok && G.L % 32 == 0 && G.L <= 128) {
                G.fz = std::min(p->fused_z, G.L);
                G.frep = std::min(G.R, 16384 / (G.L * G.L));
                G.fblocks = (G.R + G.frep - 1) / G.frep * (G.L / G.fz);
                lds = std::max<int64_t>(lds, 9 * (int64_t)G.frep * G.L * G.L / 2);  // 4 + 5 colour-plane slots
            } else {
                ok = false;
            }
        }
        std::vector<int> order(p->groups.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return p->groups[x].L > p->groups[y].L; });
        int fb = 0;
        for (int i : order) {
            p->groups[i].fblk0 = fb;
            fb += p->groups[i].fblocks;
        }
        p->fused_on = ok && lds <= 73728;
        p->fnblocks = fb;
        p->fused_lds = (int)lds;
        if (p->fused_on && p->d_pool && !p->d_pool2) HIP_TRY(hipMalloc(&p->d_pool2, p->pool_bytes));
    }
    p->layout_width = p->glayout.empty() ? 16 : p->glayout[0][0];
    p->layout_runs = p->glayout.empty() ? 2 : p->glayout[0][1];
    p->layout_spt = p->glayout.empty() ? 1 : p->glayout[0][2];
    if (p->d_groups) {
        ST_COPY(p, p->d_groups, p->groups.data(), sizeof(DevGroup) * p->groups.size(), hipMemcpyHostToDevice);
        p->drop_graph();
    }
    return ISING_OK;
}

static int layout_ising(ising_pt* p, int width, int runs, int spt) {
    const size_t n = p->groups.size();
    std::vector<int> w(n, width), r(n, runs), t(n, spt);
    return layout_ising_each(p, w.data(), r.data(), t.data());
}

int build_ensemble(ising_pt* p, int model, const int32_t* sizes, int n_sizes, int R, const double* temps,
                   uint64_t seed, const ising_pt_opts_t* opts) {
    if (model != ISING_MODEL_ISING && model != ISING_MODEL_XY && model != ISING_MODEL_HEISENBERG &&
        model != ISING_MODEL_ISING_MSC && model != ISING_MODEL_ISING2D)
        return fail(ISING_E_ARG, "model must be ISING_MODEL_ISING, _ISING_MSC, _ISING2D, _XY or _HEISENBERG");
    if (n_sizes < 1 || n_sizes > 64 || !sizes) return fail(ISING_E_ARG, "need 1..64 lattice sizes");
    if (R < 1 || R > 65535 || !temps) return fail(ISING_E_ARG, "need 1..65535 replicas and their temperatures");
    for (int i = 0; i < n_sizes; ++i)
        if (!valid_L(sizes[i])) return fail(ISING_E_ARG, "lattice sizes must be even, 2..2046");
    for (int i = 0; i < R; ++i)
        if (!(temps[i] > 0.0) || !std::isfinite(temps[i])) return fail(ISING_E_ARG, "temperatures must be positive and finite");
    ising_pt_opts_t o;
    if (opts) o = *opts;
    else ising_pt_default_opts(&o);
    if (o.exchange_every < 1 || o.measure_every < 1) return fail(ISING_E_ARG, "exchange_every and measure_every must be >= 1");
    if (o.record_capacity < 1) return fail(ISING_E_ARG, "record_capacity must be >= 1");
    if (o.graph_sweeps < 0 || o.graph_sweeps > 100000) return fail(ISING_E_ARG, "graph_sweeps must be in 0..100000");
    if (o.ising_spt < 0 || o.ising_spt > 4) return fail(ISING_E_ARG, "ising_spt must be in 0..4");
    if (o.ising_fused != 0 && o.ising_fused != 8 && o.ising_fused != 16 && o.ising_fused != 32)
        return fail(ISING_E_ARG, "ising_fused must be 0, 8, 16 or 32");
    if (o.ising_width != 0 && o.ising_width != 16 && o.ising_width != 32)
        return fail(ISING_E_ARG, "ising_width must be 0, 16 or 32");
    if (o.ising_runs < 0 || o.ising_runs > 2) return fail(ISING_E_ARG, "ising_runs must be 0, 1 or 2");
    if (model == ISING_MODEL_ISING_MSC && !(o.h == 0.0 && o.j > 0.0))
        return fail(ISING_E_ARG, "multispin coding needs h = 0 and J > 0");
    if (model == ISING_MODEL_XY || model == ISING_MODEL_HEISENBERG) {
        if (!(o.delta > 0.0) || !std::isfinite(o.delta)) return fail(ISING_E_ARG, "delta must be > 0");
        if (o.n_overrelax < 0 || o.renorm_every < 0 || o.adapt_every < 0)
            return fail(ISING_E_ARG, "n_overrelax, renorm_every and adapt_every must be >= 0");
        if (o.adapt_every > 0 && !(o.adapt_target > 0.0 && o.adapt_target < 1.0))
            return fail(ISING_E_ARG, "adapt_target must be in (0, 1)");
    }

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(ISING_E_HIP, "no HIP device visible");

    const bool two_d = model == ISING_MODEL_ISING2D;
    p->model = two_d ? ISING_MODEL_ISING : model;  // the 2D lattice runs the Ising kernels
    p->dims = two_d ? 2 : 3;
    const bool msc = model == ISING_MODEL_ISING_MSC;
    p->ncomp = (model == ISING_MODEL_ISING || two_d || msc) ? 0 : (model == ISING_MODEL_XY ? 2 : 3);
    p->R = R;
    p->seed = seed;
    p->opts = o;
    p->temps.assign(temps, temps + R);
    p->total_reps = n_sizes * R;

    const bool ising = model == ISING_MODEL_ISING || two_d;
    const int64_t words = (R + 31) / 32;          // multispin coding: words per site
    const int64_t ncomp = ising ? 1 : (msc ? words : p->ncomp);  // arrays per colour half-lattice
    const int64_t elem = ising ? 1 : 4;           // bytes per element
    const int64_t tile = ising ? kIsingTile : (msc ? kMscTile : kContTile);
    const int64_t units = msc ? words : R;        // work units per group: replicas or replica words
    int64_t off = 0;
    int blk = 0;
    for (int i = 0; i < n_sizes; ++i) {
        DevGroup G{};
        G.L = sizes[i];
        G.Lh = sizes[i] / 2;
        G.R = R;
        G.dims = p->dims;
        G.spt = 1;
        if (G.Lh >= 2) fastdiv_magic((uint32_t)G.Lh, &G.mag_Lh, &G.sh_Lh);
        fastdiv_magic((uint32_t)G.L, &G.mag_L, &G.sh_L);
        G.rep0 = i * R;
        G.n_half = (int64_t)sizes[i] * sizes[i] * (two_d ? 1 : sizes[i]) / 2;
        G.stride = ((G.n_half * elem + 255) / 256) * 256 / elem;
        G.off = off;
        G.tsz = (int32_t)tile;
        G.tiles = (int32_t)((G.n_half + G.tsz - 1) / G.tsz);
        if (G.tiles >= 2) fastdiv_magic((uint32_t)G.tiles, &G.mag_T, &G.sh_T);
        G.blk0 = blk;
        if ((int64_t)blk + (int64_t)G.tiles * units > INT32_MAX / 2) return fail(ISING_E_ARG, "ensemble too large");
        blk += G.tiles * (int32_t)units;
        off += 2 * (msc ? words : (int64_t)R * ncomp) * G.stride;
        p->groups.push_back(G);
    }
    p->nblocks = blk;
    p->pool_bytes = off * elem;
    if (ising) {
        p->fused_z = model == ISING_MODEL_ISING ? o.ising_fused : 0;
        const int rc = layout_ising(p, o.ising_width, o.ising_runs, o.ising_spt);
        if (rc) return rc;
    }

    HIP_TRY(hipStreamCreateWithFlags(&p->st, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&p->ev0));
    HIP_TRY(hipEventCreate(&p->ev1));
    HIP_TRY(hipMalloc(&p->d_groups, sizeof(DevGroup) * p->groups.size()));
    if (ising || msc) HIP_TRY(hipMalloc(&p->d_pool, p->pool_bytes));
    if (p->fused_on) HIP_TRY(hipMalloc(&p->d_pool2, p->pool_bytes));
    else HIP_TRY(hipMalloc(&p->d_fpool, p->pool_bytes));
    if (msc) HIP_TRY(hipMalloc(&p->d_mthr, sizeof(uint32_t) * 3 * p->total_reps));
    if (msc) HIP_TRY(hipMalloc(&p->d_mplanes, sizeof(uint32_t) * 96 * p->total_reps));
    HIP_TRY(hipMalloc(&p->d_thr, sizeof(uint32_t) * kIsingTab * p->total_reps));
    if (ising) HIP_TRY(hipMalloc(&p->d_thr_rep, sizeof(uint32_t) * kIsingTab * p->total_reps));
    HIP_TRY(hipMalloc(&p->d_r2t, sizeof(int32_t) * 2 * p->total_reps));  // two permutation parities
    HIP_TRY(hipMalloc(&p->d_t2r, sizeof(int32_t) * 2 * p->total_reps));
    HIP_TRY(hipMalloc(&p->d_betas, sizeof(double) * p->total_reps));
    HIP_TRY(hipMalloc(&p->d_ctl, sizeof(DevCtl)));
    HIP_TRY(hipMalloc(&p->d_acc, sizeof(int64_t) * acc_words(p->total_reps)));
    HIP_TRY(hipMalloc(&p->d_rec, ((ising || msc) ? sizeof(IsingRec) : sizeof(ContRec)) *
                                     (size_t)o.record_capacity * p->total_reps));
    if (!ising && !msc) {
        HIP_TRY(hipMalloc(&p->d_deltas, sizeof(float) * p->total_reps));
        HIP_TRY(hipMalloc(&p->d_omcd, sizeof(float) * p->total_reps));
        HIP_TRY(hipMalloc(&p->d_betas_f, sizeof(float) * p->total_reps));
        HIP_TRY(hipMalloc(&p->d_part0, sizeof(double) * 4 * (size_t)p->nblocks));
        HIP_TRY(hipMalloc(&p->d_part1, sizeof(double) * 4 * (size_t)p->nblocks));
        HIP_TRY(hipMalloc(&p->d_dacc, sizeof(double) * 4 * (size_t)p->total_reps));
        HIP_TRY(hipMalloc(&p->d_cnt, sizeof(int64_t) * kCntSlots * (size_t)p->total_reps));
    }

    std::vector<uint32_t> thr(kIsingTab * (size_t)p->total_reps);
    std::vector<int32_t> ident(2 * (size_t)p->total_reps);
    std::vector<double> betas(p->total_reps);
    std::vector<float> betas_f(p->total_reps), deltas(p->total_reps), omcd(p->total_reps);
    std::vector<uint32_t> mthr(3 * (size_t)p->total_reps);
    for (int g = 0; g < n_sizes; ++g)
        for (int i = 0; i < R; ++i) {
            const double beta = 1.0 / temps[i];
            betas[g * R + i] = beta;
            betas_f[g * R + i] = (float)beta;
            deltas[g * R + i] = (float)o.delta;
            omcd[g * R + i] = (float)(1.0 - std::cos((double)(float)o.delta));
            ident[g * R + i] = ident[p->total_reps + g * R + i] = i;
            ising_table(beta, o.j, o.h, &thr[kIsingTab * (size_t)(g * R + i)], p->dims);
            for (int k = 0; k < 3; ++k)  // multispin coding: T(a) for a = 4, 5, 6 aligned neighbours
                mthr[3 * (size_t)(g * R + i) + k] = (uint32_t)std::ldexp(std::exp(-beta * o.j * (4.0 * (4 + k) - 12.0)), 32);
        }
    DevCtl ctl{};
    ctl.t = 0;
    ctl.sample_start = INT64_MAX;
    ctl.measure_every = o.measure_every;
    ctl.exchange_every = o.exchange_every;
    ctl.rec_cap = o.record_capacity;
    ctl.renorm_every = (model == ISING_MODEL_XY || model == ISING_MODEL_HEISENBERG) ? o.renorm_every : 0;
    ctl.adapt_every = (model == ISING_MODEL_XY || model == ISING_MODEL_HEISENBERG) ? o.adapt_every : 0;
    ctl.closed = -1;
    set_flags(&ctl, 0);
    ST_COPY(p, p->d_groups, p->groups.data(), sizeof(DevGroup) * p->groups.size(), hipMemcpyHostToDevice);
    ST_COPY(p, p->d_thr, thr.data(), sizeof(uint32_t) * thr.size(), hipMemcpyHostToDevice);
    if (p->d_thr_rep)  // the starting permutation is the identity
        ST_COPY(p, p->d_thr_rep, thr.data(), sizeof(uint32_t) * thr.size(), hipMemcpyHostToDevice);
    ST_COPY(p, p->d_r2t, ident.data(), sizeof(int32_t) * ident.size(), hipMemcpyHostToDevice);
    ST_COPY(p, p->d_t2r, ident.data(), sizeof(int32_t) * ident.size(), hipMemcpyHostToDevice);
    ST_COPY(p, p->d_betas, betas.data(), sizeof(double) * betas.size(), hipMemcpyHostToDevice);
    ST_COPY(p, p->d_ctl, &ctl, sizeof(DevCtl), hipMemcpyHostToDevice);
    ST_FILL(p, p->d_acc, 0, sizeof(int64_t) * acc_words(p->total_reps));
    if (msc) {
        ST_COPY(p, p->d_mthr, mthr.data(), sizeof(uint32_t) * mthr.size(), hipMemcpyHostToDevice);
        // threshold bit planes per 32-replica word under the identity permutation
        std::vector<uint32_t> planes(96 * (size_t)p->total_reps, 0u);
        for (int g = 0; g < n_sizes; ++g)
            for (int wi = 0; 32 * wi < R; ++wi)
                for (int kj = 0; kj < 96; ++kj) {
                    uint32_t v = 0;
                    for (int bit = 0; bit < std::min(32, R - 32 * wi); ++bit)
                        v |= ((mthr[3 * (size_t)(g * R + 32 * wi + bit) + (kj >> 5)] >> (31 - (kj & 31))) & 1u) << bit;
                    planes[(size_t)(g * R + 32 * wi) * 96 + kj] = v;
                }
        ST_COPY(p, p->d_mplanes, planes.data(), sizeof(uint32_t) * planes.size(), hipMemcpyHostToDevice);
        ST_FILL(p, p->d_pool, 0, p->pool_bytes);
        HIP_TRY(launch_msc_init(p->d_groups, (int)p->groups.size(), (uint32_t*)p->d_pool, p->nblocks, (uint32_t)seed,
                                (uint32_t)(seed >> 32), p->st));
    } else if (ising) {
        ST_FILL(p, p->d_pool, 0, p->pool_bytes);
        HIP_TRY(launch_ising_init(p->d_groups, (int)p->groups.size(), p->d_pool, p->nblocks, (uint32_t)seed,
                                  (uint32_t)(seed >> 32), p->st));
    } else {
        ST_COPY(p, p->d_deltas, deltas.data(), sizeof(float) * deltas.size(), hipMemcpyHostToDevice);
        ST_COPY(p, p->d_omcd, omcd.data(), sizeof(float) * omcd.size(), hipMemcpyHostToDevice);
        ST_COPY(p, p->d_betas_f, betas_f.data(), sizeof(float) * betas_f.size(), hipMemcpyHostToDevice);
        ST_FILL(p, p->d_fpool, 0, p->pool_bytes);
        ST_FILL(p, p->d_part0, 0, sizeof(double) * 4 * (size_t)p->nblocks);
        ST_FILL(p, p->d_part1, 0, sizeof(double) * 4 * (size_t)p->nblocks);
        ST_FILL(p, p->d_dacc, 0, sizeof(double) * 4 * (size_t)p->total_reps);
        ST_FILL(p, p->d_cnt, 0, sizeof(int64_t) * kCntSlots * (size_t)p->total_reps);
        HIP_TRY(launch_cont_init(p->d_groups, (int)p->groups.size(), p->d_fpool, p->ncomp,
                                 model == ISING_MODEL_XY ? MODEL_XY : MODEL_HEIS, p->nblocks, (uint32_t)seed,
                                 (uint32_t)(seed >> 32), p->st));
    }
    SYNC_ALL(p);
    return ISING_OK;
}

const DevGroup* group_of(ising_pt* p, int size_index) {
    if (size_index < 0 || size_index >= (int)p->groups.size()) return nullptr;
    return &p->groups[size_index];
}

}  // namespace

extern "C" {

const char* ising_last_error(void) { return g_err.c_str(); }
const char* ising_version(void) { return "libising_hip 0.1 (gfx950)"; }
int ising_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

void ising_pt_default_opts(ising_pt_opts_t* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->j = 1.0;
    o->h = 0.0;
    o->exchange_every = 10;
    o->measure_every = 1;
    o->delta = 0.5;
    o->n_overrelax = 5;
    o->renorm_every = 100;
    o->record_capacity = 4096;
    o->adapt_every = 0;
    o->adapt_target = 0.5;
    o->graph_sweeps = 32;
    o->ising_spt = 0;
    o->ising_width = 0;
    o->ising_runs = 0;
    o->ising_fused = 0;
}

int ising_pt_create(ising_pt_t** out, int model, const int32_t* sizes, int32_t n_sizes, int32_t replicas,
                    const double* temps, uint64_t seed, const ising_pt_opts_t* opts) {
    if (!out) return fail(ISING_E_ARG, "out is null");
    *out = nullptr;
    ising_pt* p = new (std::nothrow) ising_pt();
    if (!p) return fail(ISING_E_HIP, "host allocation failed");
    const int rc = build_ensemble(p, model, sizes, n_sizes, replicas, temps, seed, opts);
    if (rc != ISING_OK) {
        const std::string keep = g_err;
        delete p;
        g_err = keep;
        return rc;
    }
    *out = p;
    return ISING_OK;
}

void ising_pt_destroy(ising_pt_t* pt) { delete pt; }

int ising_pt_warmup(ising_pt_t* p, int64_t sweeps) {
    if (!p) return fail(ISING_E_ARG, "null handle");
    if (sweeps < 0) return fail(ISING_E_ARG, "sweeps must be >= 0");
    if (p->s_host >= 0) return fail(ISING_E_STATE, "warm-up after sampling started");
    ORDER(p);
    HIP_TRY(hipEventRecord(p->ev0, p->st));
    const int rc = p->run_sweeps(sweeps);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(p->ev1, p->st));
    p->timed = true;
    return ISING_OK;
}

int ising_pt_sample(ising_pt_t* p, int64_t sweeps) {
    if (!p) return fail(ISING_E_ARG, "null handle");
    if (sweeps < 0) return fail(ISING_E_ARG, "sweeps must be >= 0");
    const int rc0 = p->begin_sampling();
    if (rc0) return rc0;
    const int64_t add = measured_in(p->s_host, sweeps, p->opts.measure_every);
    if (p->pending_host + add > p->opts.record_capacity)
        return fail(ISING_E_FULL, "measurement buffer would overflow: drain records before sampling more");
    ORDER(p);
    HIP_TRY(hipEventRecord(p->ev0, p->st));
    const int rc = p->run_sweeps(sweeps);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(p->ev1, p->st));
    p->timed = true;
    p->s_host += sweeps;
    p->pending_host += add;
    return ISING_OK;
}

int ising_pt_sync(ising_pt_t* p, double* last_ms) {
    if (!p) return fail(ISING_E_ARG, "null handle");
    const int rc = p->sync_all();
    if (rc) return rc;
    if (last_ms) {
        float ms = 0.f;
        if (p->timed) HIP_TRY(hipEventElapsedTime(&ms, p->ev0, p->ev1));
        *last_ms = ms;
    }
    return ISING_OK;
}

int ising_pt_pending(ising_pt_t* p, int64_t* n) {
    if (!p || !n) return fail(ISING_E_ARG, "null argument");
    DevCtl c;
    const int rc = p->read_ctl(&c);
    if (rc) return rc;
    if (c.overflow) return fail(ISING_E_FULL, "measurement records were dropped (device buffer overflow)");
    *n = c.n_rec;
    return ISING_OK;
}

static int drain(ising_pt* p, int64_t max_sweeps, size_t rec_size, void* host, int64_t* n_out) {
    DevCtl c;
    int rc = p->read_ctl(&c);
    if (rc) return rc;
    if (c.overflow) return fail(ISING_E_FULL, "measurement records were dropped (device buffer overflow)");
    const int64_t n = std::min<int64_t>(c.n_rec, std::max<int64_t>(max_sweeps, 0));
    const size_t per = rec_size * (size_t)p->total_reps;
    if (n > 0) ST_COPY(p, host, p->d_rec, per * n, hipMemcpyDeviceToHost);
    const int64_t rest = c.n_rec - n;
    if (rest > 0 && n > 0)
        ST_COPY(p, p->d_rec, (char*)p->d_rec + per * n, per * rest, hipMemcpyDeviceToDevice);
    const int32_t nr = (int32_t)rest;
    ST_COPY(p, &p->d_ctl->n_rec, &nr, sizeof(int32_t), hipMemcpyHostToDevice);
    p->pending_host -= n;
    *n_out = n;
    return ISING_OK;
}

int ising_pt_read_ising(ising_pt_t* p, int64_t max_sweeps, int32_t* tidx, int64_t* bond, int64_t* mag,
                        int64_t* n_out) {
    if (!p || !n_out) return fail(ISING_E_ARG, "null argument");
    if (p->model != ISING_MODEL_ISING && p->model != ISING_MODEL_ISING_MSC)
        return fail(ISING_E_STATE, "not an Ising ensemble");
    std::vector<IsingRec> buf((size_t)std::max<int64_t>(0, max_sweeps) * p->total_reps);
    int64_t n = 0;
    const int rc = drain(p, max_sweeps, sizeof(IsingRec), buf.data(), &n);
    if (rc) return rc;
    for (int64_t i = 0; i < n * p->total_reps; ++i) {
        if (tidx) tidx[i] = buf[i].tidx;
        if (bond) bond[i] = buf[i].bond;
        if (mag) mag[i] = buf[i].mag;
    }
    *n_out = n;
    return ISING_OK;
}

int ising_pt_read_cont(ising_pt_t* p, int64_t max_sweeps, int32_t* tidx, double* e, double* mx, double* my,
                       double* mz, int64_t* n_out) {
    if (!p || !n_out) return fail(ISING_E_ARG, "null argument");
    if (p->model == ISING_MODEL_ISING || p->model == ISING_MODEL_ISING_MSC)
        return fail(ISING_E_STATE, "not a continuous-spin ensemble");
    std::vector<ContRec> buf((size_t)std::max<int64_t>(0, max_sweeps) * p->total_reps);
    int64_t n = 0;
    const int rc = drain(p, max_sweeps, sizeof(ContRec), buf.data(), &n);
    if (rc) return rc;
    for (int64_t i = 0; i < n * p->total_reps; ++i) {
        if (tidx) tidx[i] = buf[i].tidx;
        if (e) e[i] = buf[i].e;
        if (mx) mx[i] = buf[i].mx;
        if (my) my[i] = buf[i].my;
        if (mz) mz[i] = buf[i].mz;
    }
    *n_out = n;
    return ISING_OK;
}

int ising_pt_get_perm(ising_pt_t* p, int32_t size_index, int32_t* r2t, int32_t* t2r) {
    if (!p) return fail(ISING_E_ARG, "null handle");
    const DevGroup* G = group_of(p, size_index);
    if (!G) return fail(ISING_E_ARG, "size_index out of range");
    DevCtl c;
    const int rc = p->read_ctl(&c);
    if (rc) return rc;
    // Ising ensembles keep two permutation buffers and use the one of the
    // committed exchange count's parity; continuous ensembles update one in place
    const int64_t off = p->ncomp == 0 ? (c.round & 1) * p->total_reps : 0;
    if (r2t) ST_COPY(p, r2t, p->d_r2t + off + G->rep0, sizeof(int32_t) * G->R, hipMemcpyDeviceToHost);
    if (t2r) ST_COPY(p, t2r, p->d_t2r + off + G->rep0, sizeof(int32_t) * G->R, hipMemcpyDeviceToHost);
    return ISING_OK;
}

int ising_pt_exchange_stats(ising_pt_t* p, int64_t* tries, int64_t* accepts) {
    if (!p) return fail(ISING_E_ARG, "null handle");
    DevCtl c;
    const int rc = p->read_ctl(&c);
    if (rc) return rc;
    if (tries) *tries = c.exch_tries;
    if (accepts) *accepts = c.exch_accepts;
    return ISING_OK;
}

// Host copy of both colour halves of one Ising replica as 0/1 bytes (either storage).
static int ising_halves(ising_pt* p, const DevGroup* G, int replica, std::vector<uint8_t>& half,
                        std::vector<uint32_t>* words) {
    half.assign(2 * (size_t)G->n_half, 0);
    SYNC_ALL(p);
    if (p->model == ISING_MODEL_ISING) {
        for (int c = 0; c < 2; ++c)
            ST_COPY(p, half.data() + c * G->n_half,
                              p->d_pool + G->off + (2 * (int64_t)replica + c) * G->stride, G->n_half,
                              hipMemcpyDeviceToHost);
        return ISING_OK;
    }
    const int wi = replica / 32, bit = replica % 32;
    words->assign(2 * (size_t)G->n_half, 0);
    const uint32_t* base = (const uint32_t*)p->d_pool + G->off;
    for (int c = 0; c < 2; ++c)
        ST_COPY(p, words->data() + c * G->n_half, base + (2 * (int64_t)wi + c) * G->stride,
                          sizeof(uint32_t) * G->n_half, hipMemcpyDeviceToHost);
    for (size_t i = 0; i < half.size(); ++i) half[i] = ((*words)[i] >> bit) & 1u;
    return ISING_OK;
}

int ising_pt_get_spins_i8(ising_pt_t* p, int32_t size_index, int32_t replica, int8_t* out) {
    if (!p || !out) return fail(ISING_E_ARG, "null argument");
    if (p->ncomp != 0) return fail(ISING_E_STATE, "not an Ising ensemble");
    const DevGroup* G = group_of(p, size_index);
    if (!G || replica < 0 || replica >= G->R) return fail(ISING_E_ARG, "size_index or replica out of range");
    std::vector<uint8_t> half;
    std::vector<uint32_t> words;
    const int rc = ising_halves(p, G, replica, half, &words);
    if (rc) return rc;
    const int L = G->L, Lh = G->Lh, Lz = G->dims == 3 ? L : 1;
    for (int z = 0; z < Lz; ++z)
        for (int y = 0; y < L; ++y)
            for (int x = 0; x < L; ++x) {
                const int c = (x + y + z) & 1;
                const int64_t h = ((int64_t)z * L + y) * Lh + (x >> 1);
                out[((int64_t)z * L + y) * L + x] = half[c * G->n_half + h] ? 1 : -1;
            }
    return ISING_OK;
}

int ising_pt_set_spins_i8(ising_pt_t* p, int32_t size_index, int32_t replica, const int8_t* in) {
    if (!p || !in) return fail(ISING_E_ARG, "null argument");
    if (p->ncomp != 0) return fail(ISING_E_STATE, "not an Ising ensemble");
    const DevGroup* G = group_of(p, size_index);
    if (!G || replica < 0 || replica >= G->R) return fail(ISING_E_ARG, "size_index or replica out of range");
    const int L = G->L, Lh = G->Lh, Lz = G->dims == 3 ? L : 1;
    std::vector<uint8_t> half(2 * (size_t)G->n_half);
    for (int z = 0; z < Lz; ++z)
        for (int y = 0; y < L; ++y)
            for (int x = 0; x < L; ++x) {
                const int8_t v = in[((int64_t)z * L + y) * L + x];
                if (v != 1 && v != -1) return fail(ISING_E_ARG, "Ising spins must be +1 or -1");
                const int c = (x + y + z) & 1;
                half[c * G->n_half + ((int64_t)z * L + y) * Lh + (x >> 1)] = v > 0 ? 1 : 0;
            }
    SYNC_ALL(p);
    if (p->model == ISING_MODEL_ISING) {
        for (int c = 0; c < 2; ++c)
            ST_COPY(p, p->d_pool + G->off + (2 * (int64_t)replica + c) * G->stride,
                              half.data() + c * G->n_half, G->n_half, hipMemcpyHostToDevice);
        return ISING_OK;
    }
    std::vector<uint8_t> old;
    std::vector<uint32_t> words;
    const int rc = ising_halves(p, G, replica, old, &words);
    if (rc) return rc;
    const int wi = replica / 32, bit = replica % 32;
    for (size_t i = 0; i < words.size(); ++i) words[i] = (words[i] & ~(1u << bit)) | ((uint32_t)half[i] << bit);
    uint32_t* base = (uint32_t*)p->d_pool + G->off;
    for (int c = 0; c < 2; ++c)
        ST_COPY(p, base + (2 * (int64_t)wi + c) * G->stride, words.data() + c * G->n_half,
                          sizeof(uint32_t) * G->n_half, hipMemcpyHostToDevice);
    return ISING_OK;
}

int ising_pt_get_spins_f32(ising_pt_t* p, int32_t size_index, int32_t replica, float* out) {
    if (!p || !out) return fail(ISING_E_ARG, "null argument");
    if (p->ncomp == 0) return fail(ISING_E_STATE, "not a continuous-spin ensemble");
    const DevGroup* G = group_of(p, size_index);
    if (!G || replica < 0 || replica >= G->R) return fail(ISING_E_ARG, "size_index or replica out of range");
    const int nc = p->ncomp;
    SYNC_ALL(p);
    std::vector<float> half(2 * (size_t)nc * G->n_half);
    for (int c = 0; c < 2; ++c)
        for (int k = 0; k < nc; ++k)
            ST_COPY(p, half.data() + (size_t)(c * nc + k) * G->n_half,
                              p->d_fpool + G->off + ((2 * (int64_t)replica + c) * nc + k) * G->stride,
                              sizeof(float) * G->n_half, hipMemcpyDeviceToHost);
    const int L = G->L, Lh = G->Lh;
    for (int z = 0; z < L; ++z)
        for (int y = 0; y < L; ++y)
            for (int x = 0; x < L; ++x) {
                const int c = (x + y + z) & 1;
                const int64_t h = ((int64_t)z * L + y) * Lh + (x >> 1);
                const int64_t s = ((int64_t)z * L + y) * L + x;
                for (int k = 0; k < nc; ++k) out[s * nc + k] = half[(size_t)(c * nc + k) * G->n_half + h];
            }
    return ISING_OK;
}

int ising_pt_set_spins_f32(ising_pt_t* p, int32_t size_index, int32_t replica, const float* in) {
    if (!p || !in) return fail(ISING_E_ARG, "null argument");
    if (p->ncomp == 0) return fail(ISING_E_STATE, "not a continuous-spin ensemble");
    const DevGroup* G = group_of(p, size_index);
    if (!G || replica < 0 || replica >= G->R) return fail(ISING_E_ARG, "size_index or replica out of range");
    const int nc = p->ncomp;
    const int L = G->L, Lh = G->Lh;
    std::vector<float> half(2 * (size_t)nc * G->n_half);
    for (int z = 0; z < L; ++z)
        for (int y = 0; y < L; ++y)
            for (int x = 0; x < L; ++x) {
                const int c = (x + y + z) & 1;
                const int64_t h = ((int64_t)z * L + y) * Lh + (x >> 1);
                const int64_t s = ((int64_t)z * L + y) * L + x;
                for (int k = 0; k < nc; ++k) {
                    const float v = in[s * nc + k];
                    if (!std::isfinite(v)) return fail(ISING_E_ARG, "spin components must be finite");
                    half[(size_t)(c * nc + k) * G->n_half + h] = v;
                }
            }
    SYNC_ALL(p);
    for (int c = 0; c < 2; ++c)
        for (int k = 0; k < nc; ++k)
            ST_COPY(p, p->d_fpool + G->off + ((2 * (int64_t)replica + c) * nc + k) * G->stride,
                              half.data() + (size_t)(c * nc + k) * G->n_half, sizeof(float) * G->n_half,
                              hipMemcpyHostToDevice);
    return ISING_OK;
}

int ising_pt_get_deltas(ising_pt_t* p, int32_t size_index, float* deltas) {
    if (!p || !deltas) return fail(ISING_E_ARG, "null argument");
    if (p->ncomp == 0) return fail(ISING_E_STATE, "not a continuous-spin ensemble");
    const DevGroup* G = group_of(p, size_index);
    if (!G) return fail(ISING_E_ARG, "size_index out of range");
    SYNC_ALL(p);
    ST_COPY(p, deltas, p->d_deltas + G->rep0, sizeof(float) * G->R, hipMemcpyDeviceToHost);
    return ISING_OK;
}

// ---------------------------------------------------------------------------
// Learned proposal
// ---------------------------------------------------------------------------

static int need_learn(ising_pt* p) {
    if (!p) return fail(ISING_E_ARG, "null handle");
    if (p->ncomp == 0) return fail(ISING_E_STATE, "the learned proposal needs an XY or Heisenberg ensemble");
    if (!p->d_lnet) return fail(ISING_E_STATE, "call ising_pt_learn_init first");
    return ISING_OK;
}

int ising_pt_learn_init(ising_pt_t* p, int32_t hidden, uint64_t seed) {
    if (!p) return fail(ISING_E_ARG, "null handle");
    if (p->ncomp == 0) return fail(ISING_E_STATE, "the learned proposal needs an XY or Heisenberg ensemble");
    const int P = learn_params(hidden);
    if (P < 0) return fail(ISING_E_ARG, "hidden width must be 8, 16, 32 or 64");
    SYNC_ALL(p);
    const int cap = 4096;  // training blocks per step (64 samples each)
    (void)hipFree(p->d_lmaster); (void)hipFree(p->d_lnet); (void)hipFree(p->d_lfrag); (void)hipFree(p->d_lm1);
    (void)hipFree(p->d_lm2); (void)hipFree(p->d_lgpart); (void)hipFree(p->d_lgout); (void)hipFree(p->d_llpart);
    p->d_lmaster = nullptr; p->d_lnet = nullptr; p->d_lfrag = nullptr;
    HIP_TRY(hipMalloc(&p->d_lmaster, sizeof(float) * P));
    HIP_TRY(hipMalloc(&p->d_lnet, sizeof(uint16_t) * P));
    HIP_TRY(hipMalloc(&p->d_lfrag, learn_frag_bytes(hidden)));
    HIP_TRY(hipMalloc(&p->d_lm1, sizeof(float) * P));
    HIP_TRY(hipMalloc(&p->d_lm2, sizeof(float) * P));
    HIP_TRY(hipMalloc(&p->d_lgpart, sizeof(float) * (size_t)P * cap));
    HIP_TRY(hipMalloc(&p->d_lgout, sizeof(float) * P));
    HIP_TRY(hipMalloc(&p->d_llpart, sizeof(double) * cap));
    if (!p->d_mh) {
        HIP_TRY(hipMalloc(&p->d_mh, sizeof(int64_t) * 8 * kMhSlots));
        ST_FILL(p, p->d_mh, 0, sizeof(int64_t) * 8 * kMhSlots);
    }
    p->learn_H = hidden;
    p->learn_P = P;
    p->learn_cap_blocks = cap;
    p->learn_step = 0;
    // hidden layers W ~ N(0, 2/fan_in) (He, for ReLU), output layer N(0, 0.01/fan_in), biases 0;
    // Box-Muller on splitmix64
    std::vector<float> w(P, 0.0f);
    uint64_t state = seed ^ 0x5DEECE66DULL;
    auto next = [&]() {
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return ((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
    };
    auto gauss = [&]() {
        const double u1 = next() + 1e-300, u2 = next();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    const int H = hidden;
    int o = 0;
    for (int i = 0; i < H * 6; ++i) w[o++] = (float)(gauss() * std::sqrt(2.0 / 6.0));
    o += H;
    for (int i = 0; i < H * H; ++i) w[o++] = (float)(gauss() * std::sqrt(2.0 / H));
    o += H;
    for (int i = 0; i < H * H; ++i) w[o++] = (float)(gauss() * std::sqrt(2.0 / H));
    o += H;
    for (int i = 0; i < 3 * H; ++i) w[o++] = (float)(0.1 * gauss() / std::sqrt((double)H));
    ST_COPY(p, p->d_lmaster, w.data(), sizeof(float) * P, hipMemcpyHostToDevice);
    ST_FILL(p, p->d_lm1, 0, sizeof(float) * P);
    ST_FILL(p, p->d_lm2, 0, sizeof(float) * P);
    HIP_TRY(launch_learn_pack(p->d_lmaster, p->d_lnet, P, p->st));
    HIP_TRY(launch_learn_frag(p->d_lnet, hidden, p->d_lfrag, p->st));
    SYNC_ALL(p);
    p->drop_graph();  // captured launches carry the old network pointer
    return ISING_OK;
}

int ising_pt_learn_size(ising_pt_t* p, int64_t* n) {
    const int rc = need_learn(p);
    if (rc) return rc;
    if (!n) return fail(ISING_E_ARG, "null argument");
    *n = p->learn_P;
    return ISING_OK;
}

int ising_pt_learn_get_params(ising_pt_t* p, float* out, int64_t n) {
    const int rc = need_learn(p);
    if (rc) return rc;
    if (!out || n != p->learn_P) return fail(ISING_E_ARG, "need exactly ising_pt_learn_size() floats");
    SYNC_ALL(p);
    ST_COPY(p, out, p->d_lmaster, sizeof(float) * n, hipMemcpyDeviceToHost);
    return ISING_OK;
}

int ising_pt_learn_set_params(ising_pt_t* p, const float* in, int64_t n) {
    const int rc = need_learn(p);
    if (rc) return rc;
    if (!in || n != p->learn_P) return fail(ISING_E_ARG, "need exactly ising_pt_learn_size() floats");
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(in[i])) return fail(ISING_E_ARG, "parameters must be finite");
    SYNC_ALL(p);
    ST_COPY(p, p->d_lmaster, in, sizeof(float) * n, hipMemcpyHostToDevice);
    HIP_TRY(launch_learn_pack(p->d_lmaster, p->d_lnet, (int)n, p->st));
    HIP_TRY(launch_learn_frag(p->d_lnet, p->learn_H, p->d_lfrag, p->st));
    SYNC_ALL(p);
    return ISING_OK;
}

int ising_pt_learn_train(ising_pt_t* p, int32_t steps, int32_t batch, double lr, double* mean_loss) {
    int rc = need_learn(p);
    if (rc) return rc;
    if (steps < 0 || batch < kLearnBlock || !(lr > 0.0)) return fail(ISING_E_ARG, "steps >= 0, batch >= 64, lr > 0");
    const int nblk = std::min(p->learn_cap_blocks, batch / kLearnBlock);
    ORDER(p);
    const ContArgs a = p->cont_args();
    for (int s = 0; s < steps; ++s) {
        ++p->learn_step;
        HIP_TRY(launch_learn_grad(a, nblk, p->learn_step, 0, nullptr, nullptr, 0, p->d_lgpart, p->d_llpart, p->st));
        HIP_TRY(launch_learn_adam(p->d_lmaster, p->d_lnet, p->d_lm1, p->d_lm2, p->d_lgpart, nblk, p->learn_P,
                                  1.0f / (float)(nblk * kLearnBlock), (float)lr, p->learn_step, nullptr, p->st));
    }
    if (steps > 0) HIP_TRY(launch_learn_frag(p->d_lnet, p->learn_H, p->d_lfrag, p->st));
    if (mean_loss) {
        std::vector<double> lp(nblk);
        HIP_TRY(hipMemcpyAsync(lp.data(), p->d_llpart, sizeof(double) * nblk, hipMemcpyDeviceToHost, p->st));
        SYNC_ALL(p);
        double sum = 0.0;
        for (double v : lp) sum += v;
        *mean_loss = steps > 0 ? sum / (double)(nblk * kLearnBlock) : NAN;
    }
    return ISING_OK;
}

int ising_pt_learn_eval(ising_pt_t* p, int64_t n, const float* x, const float* c, double* mean_loss, float* grad) {
    int rc = need_learn(p);
    if (rc) return rc;
    if (n < 1 || !x || !c) return fail(ISING_E_ARG, "need n >= 1 samples");
    const int nblk = (int)((n + kLearnBlock - 1) / kLearnBlock);
    if (nblk > p->learn_cap_blocks) return fail(ISING_E_ARG, "too many samples");
    ORDER(p);
    float *dx = nullptr, *dc = nullptr;
    HIP_TRY(hipMalloc(&dx, sizeof(float) * 6 * n));
    HIP_TRY(hipMalloc(&dc, sizeof(float) * n));
    ST_COPY(p, dx, x, sizeof(float) * 6 * n, hipMemcpyHostToDevice);
    ST_COPY(p, dc, c, sizeof(float) * n, hipMemcpyHostToDevice);
    const ContArgs a = p->cont_args();
    hipError_t e = launch_learn_grad(a, nblk, 0, 1, dx, dc, n, p->d_lgpart, p->d_llpart, p->st);
    if (e == hipSuccess)
        e = launch_learn_adam(p->d_lmaster, p->d_lnet, p->d_lm1, p->d_lm2, p->d_lgpart, nblk, p->learn_P,
                              1.0f / (float)n, 0.0f, 1, p->d_lgout, p->st);
    std::vector<double> lp(nblk);
    if (e == hipSuccess) e = hipMemcpyAsync(lp.data(), p->d_llpart, sizeof(double) * nblk, hipMemcpyDeviceToHost, p->st);
    if (e == hipSuccess && grad)
        e = hipMemcpyAsync(grad, p->d_lgout, sizeof(float) * p->learn_P, hipMemcpyDeviceToHost, p->st);
    if (e == hipSuccess) e = hipStreamSynchronize(p->st);
    (void)hipFree(dx);
    (void)hipFree(dc);
    HIP_TRY(e);
    if (mean_loss) {
        double sum = 0.0;
        for (double v : lp) sum += v;
        *mean_loss = sum / (double)n;
    }
    return ISING_OK;
}

int ising_pt_learn_infer(ising_pt_t* p, int64_t n, const float* x, float* out) {
    int rc = need_learn(p);
    if (rc) return rc;
    if (n < 1 || !x || !out) return fail(ISING_E_ARG, "need n >= 1 feature rows and an output buffer");
    if (n > (1ll << 28)) return fail(ISING_E_ARG, "too many feature rows");
    ORDER(p);
    float *dx = nullptr, *dout = nullptr;
    HIP_TRY(hipMalloc(&dx, sizeof(float) * 6 * n));
    hipError_t e = hipMalloc(&dout, sizeof(float) * 3 * n);
    if (e == hipSuccess) e = hipMemcpyAsync(dx, x, sizeof(float) * 6 * n, hipMemcpyHostToDevice, p->st);
    if (e == hipSuccess) e = launch_learn_infer(p->d_lfrag, p->learn_H, dx, n, dout, p->st);
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, p->st);
    if (e == hipSuccess) e = hipStreamSynchronize(p->st);
    (void)hipFree(dx);
    (void)hipFree(dout);
    HIP_TRY(e);
    return ISING_OK;
}

int ising_pt_learn_enable(ising_pt_t* p, int32_t on) {
    if (on) {
        const int rc = need_learn(p);
        if (rc) return rc;
    } else if (!p) {
        return fail(ISING_E_ARG, "null handle");
    }
    if ((bool)on != p->learned_on || (on && p->heatbath_on)) {
        SYNC_ALL(p);
        p->drop_graph();
    }
    p->learned_on = on != 0;
    if (on) p->heatbath_on = false;
    return ISING_OK;
}

int ising_pt_heat_bath(ising_pt_t* p, int32_t on) {
    if (!p) return fail(ISING_E_ARG, "null handle");
    if (p->ncomp == 0) return fail(ISING_E_STATE, "heat-bath passes exist for XY / Heisenberg ensembles");
    if (p->inline_cont) return fail(ISING_E_STATE, "single-lattice handles take per-call parameters");
    if ((bool)on != p->heatbath_on || (on && p->learned_on)) {
        SYNC_ALL(p);
        p->drop_graph();
    }
    p->heatbath_on = on != 0;
    if (on) p->learned_on = false;
    return ISING_OK;
}

int ising_pt_mh_stats(ising_pt_t* p, int64_t* tries, int64_t* accepts) {
    if (!p) return fail(ISING_E_ARG, "null handle");
    DevCtl c;
    const int rc = p->read_ctl(&c);
    if (rc) return rc;
    int64_t t = c.mh_tries, acc = c.mh_accepts;
    if (p->d_mh) {
        std::vector<int64_t> v(8 * kMhSlots);
        ST_COPY(p, v.data(), p->d_mh, sizeof(int64_t) * v.size(), hipMemcpyDeviceToHost);
        for (int k = 0; k < kMhSlots; ++k) {
            t += v[8 * (size_t)k];
            acc += v[8 * (size_t)k + 1];
        }
    }
    if (tries) *tries = t;
    if (accepts) *accepts = acc;
    return ISING_OK;
}

int ising_pt_metropolis_pass(ising_pt_t* p, int32_t color, int64_t t) {
    if (!p) return fail(ISING_E_ARG, "null handle");
    if (p->ncomp == 0) return fail(ISING_E_STATE, "continuous ensembles only");
    if ((color != 0 && color != 1) || t < 0) return fail(ISING_E_ARG, "color 0/1 and t >= 0");
    ORDER(p);
    if (p->heatbath_on) HIP_TRY(launch_cont_heatbath(p->cont_args(), p->nblocks, color, 0, t, p->st));
    else if (p->learned_on) HIP_TRY(launch_cont_learned(p->cont_args(), p->nblocks, color, 0, t, p->st));
    else return fail(ISING_E_STATE, "enable the learned proposal or heat-bath passes first");
    return ISING_OK;
}

int ising_pt_helicity(ising_pt_t* p, double* out) {
    if (!p || !out) return fail(ISING_E_ARG, "null argument");
    if (p->model != ISING_MODEL_XY) return fail(ISING_E_STATE, "the helicity sums are defined for XY ensembles");
    SYNC_ALL(p);
    std::vector<double> part(6 * (size_t)p->nblocks);
    double* d_part = nullptr;
    HIP_TRY(hipMalloc(&d_part, sizeof(double) * part.size()));
    hipError_t e = launch_cont_helicity(p->d_groups, (int)p->groups.size(), p->d_fpool, d_part, p->nblocks, p->st);
    if (e == hipSuccess) e = hipMemcpyAsync(part.data(), d_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, p->st);
    if (e == hipSuccess) e = hipStreamSynchronize(p->st);
    (void)hipFree(d_part);
    if (e != hipSuccess) return fail(ISING_E_HIP, std::string("helicity: ") + hipGetErrorString(e));
    // blocks of a replica are consecutive: sum them in block order
    for (const DevGroup& G : p->groups)
        for (int r = 0; r < G.R; ++r) {
            double* o = out + 6 * (size_t)(G.rep0 + r);
            for (int k = 0; k < 6; ++k) o[k] = 0.0;
            for (int t = 0; t < G.tiles; ++t) {
                const double* q = &part[6 * (size_t)(G.blk0 + (int64_t)r * G.tiles + t)];
                for (int k = 0; k < 6; ++k) o[k] += q[k];
            }
        }
    return ISING_OK;
}

int ising_pt_anneal(ising_pt_t* p, const double* betas, int64_t n, int64_t* bond, int64_t* mag) {
    if (!p || (!betas && n > 0)) return fail(ISING_E_ARG, "null argument");
    if (n < 0) return fail(ISING_E_ARG, "n must be >= 0");
    if (p->model != ISING_MODEL_ISING && p->model != ISING_MODEL_ISING2D)
        return fail(ISING_E_STATE, "annealing schedules exist for int8 Ising ensembles");
    if (p->inline_thr) return fail(ISING_E_STATE, "single-lattice handles take beta per step");
    if (p->s_host >= 0) return fail(ISING_E_STATE, "annealing sweeps are warm-up sweeps: call before sampling");
    for (int64_t k = 0; k < n; ++k)
        if (!(betas[k] > 0.0) || !std::isfinite(betas[k])) return fail(ISING_E_ARG, "betas must be positive and finite");
    if (n == 0) return ISING_OK;
    SYNC_ALL(p);
    if (n > p->sched_cap) {
        (void)hipFree(p->d_sched);
        p->d_sched = nullptr;
        p->sched_cap = 0;
        HIP_TRY(hipMalloc(&p->d_sched, sizeof(uint32_t) * kIsingTab * (size_t)n));
        p->sched_cap = n;
    }
    const bool meas = bond || mag;
    const int64_t per = 2 * (int64_t)p->total_reps;
    if (meas && n * per > p->smeas_cap) {
        (void)hipFree(p->d_smeas);
        p->d_smeas = nullptr;
        p->smeas_cap = 0;
        HIP_TRY(hipMalloc(&p->d_smeas, sizeof(int64_t) * (size_t)(n * per)));
        p->smeas_cap = n * per;
    }
    std::vector<uint32_t> rows(kIsingTab * (size_t)n);
    for (int64_t k = 0; k < n; ++k) ising_table(betas[k], p->opts.j, p->opts.h, &rows[kIsingTab * (size_t)k], p->dims);
    ST_COPY(p, p->d_sched, rows.data(), sizeof(uint32_t) * rows.size(), hipMemcpyHostToDevice);
    if (meas) ST_FILL(p, p->d_smeas, 0, sizeof(int64_t) * (size_t)(n * per));
    p->d_sched_on = true;
    p->sched_t0 = p->t_host;
    int rc = ISING_OK;
    if (!meas) {
        rc = p->enqueue_sweeps(n);
    } else {
        // each colour-0 launch closes the sweep before it; one closing kernel at the end
        for (int64_t k = 0; k < n && rc == ISING_OK; ++k) {
            rc = p->enqueue_sweeps(1, -1, k + 1 == n);
            if (rc == ISING_OK) {
                const hipError_t e = launch_ising_measure(p->d_groups, (int)p->groups.size(), p->d_pool,
                                                          p->d_smeas + k * per, p->nblocks, p->st);
                if (e != hipSuccess) rc = fail(ISING_E_HIP, std::string("measure launch: ") + hipGetErrorString(e));
            }
        }
    }
    p->d_sched_on = false;
    if (rc) return rc;
    SYNC_ALL(p);
    if (meas) {
        std::vector<int64_t> v((size_t)(n * per));
        ST_COPY(p, v.data(), p->d_smeas, sizeof(int64_t) * v.size(), hipMemcpyDeviceToHost);
        for (int64_t k = 0; k < n; ++k)
            for (int r = 0; r < p->total_reps; ++r) {
                if (bond) bond[k * p->total_reps + r] = v[(size_t)(k * per + 2 * r)];
                if (mag) mag[k * p->total_reps + r] = v[(size_t)(k * per + 2 * r + 1)];
            }
    }
    return ISING_OK;
}

// ---------------------------------------------------------------------------
// Cluster updates (Swendsen-Wang, Wolff)
// ---------------------------------------------------------------------------

namespace {

// Tiles, label and bond arrays, allocated at the first cluster update.
int cluster_setup(ising_pt* p) {
    if (p->d_cg) return ISING_OK;
    std::vector<ClusterGroup> cg;
    int64_t lab = 0, blk = 0;
    for (const DevGroup& G : p->groups) {
        ClusterGroup C{};
        C.L = G.L;
        C.dims = G.dims;
        C.R = G.R;
        C.rep0 = G.rep0;
        C.tx = std::min(G.dims == 3 ? 32 : 64, G.L);
        C.ty = std::min(G.dims == 3 ? 8 : 32, G.L);
        C.tz = G.dims == 3 ? std::min(8, G.L) : 1;
        C.ntx = (G.L + C.tx - 1) / C.tx;
        C.nty = (G.L + C.ty - 1) / C.ty;
        C.ntz = G.dims == 3 ? (G.L + C.tz - 1) / C.tz : 1;
        C.tiles = C.ntx * C.nty * C.ntz;
        C.nsites = G.dims == 3 ? (int64_t)G.L * G.L * G.L : (int64_t)G.L * G.L;
        if (C.nsites >= (1ll << 31)) return fail(ISING_E_ARG, "cluster updates need L^d < 2^31 sites");
        C.blk0 = (int32_t)blk;
        blk += (int64_t)G.R * C.tiles;
        C.lab0 = lab;
        lab += (int64_t)G.R * C.nsites;
        C.q4 = (G.L + 3) / 4;
        // every extent is even (L even, tiles of even width), so >= 2 as fastdiv needs
        fastdiv_magic((uint32_t)G.L, &C.mag_L, &C.sh_L);
        fastdiv_magic((uint32_t)C.tx, &C.mag_x[0], &C.sh_x[0]);
        fastdiv_magic((uint32_t)(G.L - (C.ntx - 1) * C.tx), &C.mag_x[1], &C.sh_x[1]);
        fastdiv_magic((uint32_t)C.ty, &C.mag_y[0], &C.sh_y[0]);
        fastdiv_magic((uint32_t)(G.L - (C.nty - 1) * C.ty), &C.mag_y[1], &C.sh_y[1]);
        cg.push_back(C);
    }
    if (blk >= (1ll << 31)) return fail(ISING_E_ARG, "too many cluster tiles");
    HIP_TRY(hipMalloc(&p->d_cg, sizeof(ClusterGroup) * cg.size()));
    HIP_TRY(hipMalloc(&p->d_clabel, sizeof(int32_t) * (size_t)lab));
    HIP_TRY(hipMalloc(&p->d_cbond, (size_t)lab));
    HIP_TRY(hipMalloc(&p->d_cbthr, sizeof(uint32_t) * (size_t)p->total_reps));
    ST_COPY(p, p->d_cg, cg.data(), sizeof(ClusterGroup) * cg.size(), hipMemcpyHostToDevice);
    p->cgroups = cg;
    p->cl_blocks = (int)blk;
    return ISING_OK;
}

// Bond threshold floor(2^32 p), p = 1 - exp(-2 beta |J|), at most 2^32 - 1.
uint32_t bond_threshold(double beta, double J) {
    const double pa = -std::expm1(-2.0 * beta * std::fabs(J));
    if (!(pa > 0.0)) return 0u;
    const double t = std::floor(std::ldexp(pa, 32));
    return t >= 4294967295.0 ? 0xffffffffu : (uint32_t)t;
}

int grow(ising_pt* p, int64_t** buf, int64_t* cap, int64_t need) {
    if (need <= *cap) return ISING_OK;
    (void)hipFree(*buf);
    *buf = nullptr;
    *cap = 0;
    HIP_TRY(hipMalloc(buf, sizeof(int64_t) * (size_t)need));
    *cap = need;
    return ISING_OK;
}

// n cluster updates of every replica (kind 0 Swendsen-Wang, 1 Wolff), at each
// replica's temperature (inl: beta and J of a single-lattice call).  After
// each update, optionally: sites flipped, Ising bond sum and magnetisation,
// continuous E, Mx, My, Mz; all indexed [update][global replica].
int cluster_run(ising_pt* p, int kind, int64_t n, bool inl, double beta, double J, int64_t* flipped, int64_t* bond,
                int64_t* mag, double* cmeas) {
    if (kind != ISING_CLUSTER_SW && kind != ISING_CLUSTER_WOLFF) return fail(ISING_E_ARG, "unknown cluster kind");
    if (n < 0) return fail(ISING_E_ARG, "n must be >= 0");
    if (p->model == ISING_MODEL_ISING_MSC)
        return fail(ISING_E_STATE, "cluster updates need int8 or continuous storage (not multispin coding)");
    const bool ising = p->ncomp == 0;
    if (ising && !inl && p->opts.h != 0.0) return fail(ISING_E_STATE, "cluster updates need h = 0");
    if (n == 0) return ISING_OK;
    SYNC_ALL(p);
    int rc = cluster_setup(p);
    if (rc) return rc;
    const int64_t R = p->total_reps;
    if (flipped) rc = grow(p, &p->d_cflip, &p->cflip_cap, n * R);
    if (!rc && ising && (bond || mag)) rc = grow(p, &p->d_smeas, &p->smeas_cap, n * 2 * R);
    if (!rc && !ising && cmeas) {
        int64_t* q = (int64_t*)p->d_cmeas;
        rc = grow(p, &q, &p->cmeas_cap, n * 4 * R);
        p->d_cmeas = (double*)q;
    }
    if (rc) return rc;
    if (ising && !inl) {
        std::vector<uint32_t> thr((size_t)R);
        // slot rep0 + i of every size holds temperature i
        for (int64_t k = 0; k < R; ++k) thr[(size_t)k] = bond_threshold(1.0 / p->temps[(size_t)(k % p->R)], J);
        ST_COPY(p, p->d_cbthr, thr.data(), sizeof(uint32_t) * thr.size(), hipMemcpyHostToDevice);
    }
    if (flipped) ST_FILL(p, p->d_cflip, 0, sizeof(int64_t) * (size_t)(n * R));
    if (ising && (bond || mag)) ST_FILL(p, p->d_smeas, 0, sizeof(int64_t) * (size_t)(n * 2 * R));

    ClusterArgs a{};
    a.groups = p->d_groups;
    a.cg = p->d_cg;
    a.ngroups = (int32_t)p->groups.size();
    a.total_reps = p->total_reps;
    a.ipool = ising ? p->d_pool : nullptr;
    a.fpool = ising ? nullptr : p->d_fpool;
    a.ncomp = p->ncomp;
    a.model = p->model == ISING_MODEL_XY ? MODEL_XY : (p->model == ISING_MODEL_HEISENBERG ? MODEL_HEIS : MODEL_ISING);
    a.kind = kind;
    a.k0 = (uint32_t)p->seed;
    a.k1 = (uint32_t)(p->seed >> 32);
    a.label = p->d_clabel;
    a.bond = p->d_cbond;
    a.r2t = p->d_r2t;
    a.ctl = ising ? p->d_ctl : nullptr;
    a.bthr = p->d_cbthr;
    a.betas = p->d_betas;
    a.J = J;
    a.inline_p = inl ? 1 : 0;
    a.bthr_inline = bond_threshold(beta, J);
    a.beta_inline = beta;
    for (int64_t k = 0; k < n; ++k) {
        a.u = (uint32_t)(p->c_host + k);
        a.flipped = flipped ? p->d_cflip + k * R : nullptr;
        HIP_TRY(launch_cluster_update(a, p->cl_blocks, p->st));
        if (ising && (bond || mag))
            HIP_TRY(launch_ising_measure(p->d_groups, (int)p->groups.size(), p->d_pool, p->d_smeas + k * 2 * R,
                                         p->nblocks, p->st));
        if (!ising && cmeas) {
            HIP_TRY(launch_cont_measure(p->d_groups, (int)p->groups.size(), p->d_fpool, p->ncomp, J, p->d_part0,
                                        p->nblocks, p->st));
            HIP_TRY(launch_cont_fold(p->d_groups, (int)p->groups.size(), (int)R, p->d_part0, p->d_cmeas + k * 4 * R,
                                     p->st));
        }
    }
    p->c_host += n;
    HIP_TRY(hipStreamSynchronize(p->st));
    if (flipped) ST_COPY(p, flipped, p->d_cflip, sizeof(int64_t) * (size_t)(n * R), hipMemcpyDeviceToHost);
    if (ising && (bond || mag)) {
        std::vector<int64_t> v((size_t)(n * 2 * R));
        ST_COPY(p, v.data(), p->d_smeas, sizeof(int64_t) * v.size(), hipMemcpyDeviceToHost);
        for (int64_t k = 0; k < n; ++k)
            for (int64_t r = 0; r < R; ++r) {
                if (bond) bond[k * R + r] = v[(size_t)(k * 2 * R + 2 * r)];
                if (mag) mag[k * R + r] = v[(size_t)(k * 2 * R + 2 * r + 1)];
            }
    }
    if (!ising && cmeas) ST_COPY(p, cmeas, p->d_cmeas, sizeof(double) * (size_t)(n * 4 * R), hipMemcpyDeviceToHost);
    return ISING_OK;
}

}  // namespace

int ising_pt_cluster(ising_pt_t* p, int32_t kind, int64_t n, int64_t* flipped, int64_t* bond, int64_t* mag) {
    if (!p) return fail(ISING_E_ARG, "null handle");
    if (p->ncomp != 0) return fail(ISING_E_STATE, "not an Ising ensemble: use ising_pt_cluster_cont");
    if (p->inline_thr) return fail(ISING_E_STATE, "single-lattice handles use ising_lattice_cluster_step");
    return cluster_run(p, kind, n, false, 0.0, p->opts.j, flipped, bond, mag, nullptr);
}

int ising_pt_cluster_cont(ising_pt_t* p, int32_t kind, int64_t n, int64_t* flipped, double* meas) {
    if (!p) return fail(ISING_E_ARG, "null handle");
    if (p->ncomp == 0) return fail(ISING_E_STATE, "not a continuous ensemble: use ising_pt_cluster");
    if (p->inline_cont) return fail(ISING_E_STATE, "single-lattice handles use ising_cont_cluster_step");
    return cluster_run(p, kind, n, false, 0.0, p->opts.j, flipped, nullptr, nullptr, meas);
}

int ising_pt_cluster_count(const ising_pt_t* p, int64_t* out) {
    if (!p || !out) return fail(ISING_E_ARG, "null argument");
    *out = p->c_host;
    return ISING_OK;
}

// ---------------------------------------------------------------------------
// The other LatticeGpu callers, as batched ensemble drivers
// ---------------------------------------------------------------------------

namespace {

// kibble_zurek.rs:78-84: inclusive linear ramp; the last sweep lands on t_end.
double kz_ramp_temperature(double t_start, double t_end, int64_t step, int64_t tau_q) {
    if (tau_q == 0) return t_end;
    const double frac = (double)(step + 1) / (double)tau_q;
    return t_start + (t_end - t_start) * frac;
}

// The reference's GPU callers pass beta to the device as f32 (kz_gpu.rs:25-35,
// coarsening_gpu.rs:11-21, fss_gpu.rs:22): the same value is used here.
double f32_beta(double T) { return (double)(float)(1.0 / T); }

struct Handle {
    ising_pt* p = nullptr;
    ~Handle() { delete p; }
};

int one_size_ensemble(Handle& h, int n, int R, double j, double t, uint64_t seed) {
    ising_pt_opts_t o;
    ising_pt_default_opts(&o);
    o.j = (double)(float)j;  // passed to the kernel as f32 in the reference
    o.h = 0.0;
    o.exchange_every = 1 << 30;
    o.graph_sweeps = 0;
    std::vector<double> temps((size_t)R, t);
    std::vector<int32_t> sz(1, n);
    return ising_pt_create(&h.p, ISING_MODEL_ISING, sz.data(), 1, R, temps.data(), seed, &o);
}

// Bond sum of every replica of the current state, on the device.
int measure_bonds(ising_pt* p, int64_t* bond) {
    const int64_t per = 2 * (int64_t)p->total_reps;
    SYNC_ALL(p);
    if (per > p->smeas_cap) {
        (void)hipFree(p->d_smeas);
        p->d_smeas = nullptr;
        p->smeas_cap = 0;
        HIP_TRY(hipMalloc(&p->d_smeas, sizeof(int64_t) * (size_t)per));
        p->smeas_cap = per;
    }
    ST_FILL(p, p->d_smeas, 0, sizeof(int64_t) * (size_t)per);
    HIP_TRY(launch_ising_measure(p->d_groups, (int)p->groups.size(), p->d_pool, p->d_smeas, p->nblocks, p->st));
    std::vector<int64_t> v((size_t)per);
    ST_COPY(p, v.data(), p->d_smeas, sizeof(int64_t) * v.size(), hipMemcpyDeviceToHost);
    for (int r = 0; r < p->total_reps; ++r) bond[r] = v[(size_t)(2 * r)];
    return ISING_OK;
}

int domain_walls(ising_pt* p, int n, const int64_t* bond, double* rho) {
    const double bonds = 3.0 * (double)n * n * n;  // forward bonds (lattice_gpu.rs:254-280)
    for (int r = 0; r < p->total_reps; ++r) rho[r] = (bonds - (double)bond[r]) / (2.0 * bonds);
    return ISING_OK;
}

}  // namespace

int ising_kz_run(int32_t n, double j, double t_start, double t_end, int64_t tau_q, int32_t n_trials,
                 int64_t warmup_sweeps, int64_t freeze_sweeps, double freeze_temperature, uint64_t seed,
                 double* rho) {
    if (!rho || n_trials < 1 || tau_q < 0 || warmup_sweeps < 0 || freeze_sweeps < 0 || !(t_start > 0.0))
        return fail(ISING_E_ARG, "need rho[n_trials], n_trials >= 1, tau_q/warmup/freeze >= 0, t_start > 0");
    Handle h;
    int rc = one_size_ensemble(h, n, n_trials, j, t_start, seed);
    if (rc) return rc;
    std::vector<double> betas;
    betas.reserve((size_t)(warmup_sweeps + tau_q + freeze_sweeps + 1));
    for (int64_t k = 0; k < warmup_sweeps; ++k) betas.push_back(f32_beta(t_start));        // kz_gpu.rs:18-24
    for (int64_t k = 0; k < tau_q; ++k)                                                    // kz_gpu.rs:27-31
        betas.push_back(f32_beta(std::max(kz_ramp_temperature(t_start, t_end, k, tau_q), 0.01)));
    for (int64_t k = 0; k < freeze_sweeps; ++k) betas.push_back(f32_beta(std::max(freeze_temperature, 0.01)));
    std::vector<int64_t> bond((size_t)n_trials);
    rc = ising_pt_anneal(h.p, betas.data(), (int64_t)betas.size(), nullptr, nullptr);
    if (rc) return rc;
    rc = measure_bonds(h.p, bond.data());
    if (rc) return rc;
    return domain_walls(h.p, n, bond.data(), rho);
}

int ising_kz_sweep(int32_t n, double j, double t_start, double t_end, const int64_t* tau_q, int32_t n_tau,
                   int32_t n_trials, int64_t warmup_sweeps, int64_t freeze_sweeps, double freeze_temperature,
                   uint64_t base_seed, double* rho, double* rho_err) {
    if (!tau_q || !rho || !rho_err || n_tau < 1) return fail(ISING_E_ARG, "need tau_q[n_tau], rho, rho_err");
    std::vector<double> r((size_t)n_trials);
    for (int k = 0; k < n_tau; ++k) {
        // kibble_zurek.rs:102-106: trial seeds base + tau_q + trial; here the
        // trials of one tau_q are replicas (Philox streams) of one ensemble keyed by base + tau_q
        const uint64_t seed = base_seed + (uint64_t)tau_q[k];
        const int rc = ising_kz_run(n, j, t_start, t_end, tau_q[k], n_trials, warmup_sweeps, freeze_sweeps,
                                    freeze_temperature, seed, r.data());
        if (rc) return rc;
        // kibble_zurek.rs:86-100: mean and standard error (n - 1 variance)
        double mean = 0.0;
        for (double v : r) mean += v;
        mean /= (double)n_trials;
        double var = 0.0;
        for (double v : r) var += (v - mean) * (v - mean);
        rho[k] = mean;
        rho_err[k] = n_trials > 1 ? std::sqrt(var / (double)(n_trials - 1)) / std::sqrt((double)n_trials) : 0.0;
    }
    return ISING_OK;
}

int ising_coarsening_run(int32_t n, double j, double t_high, double t_quench, int64_t warmup_sweeps,
                         int64_t total_steps, int64_t sample_every, uint64_t seed, int32_t n_trials,
                         int64_t* steps_out, double* rho_out, int64_t* n_samples) {
    if (!steps_out || !rho_out || !n_samples || n_trials < 1 || sample_every < 1 || total_steps < 0 ||
        warmup_sweeps < 0 || !(t_high > 0.0))
        return fail(ISING_E_ARG, "need outputs, n_trials >= 1, sample_every >= 1, t_high > 0");
    Handle h;
    int rc = one_size_ensemble(h, n, n_trials, j, t_high, seed);
    if (rc) return rc;
    std::vector<double> warm((size_t)warmup_sweeps, f32_beta(t_high));  // coarsening_gpu.rs:9-13
    rc = ising_pt_anneal(h.p, warm.data(), warmup_sweeps, nullptr, nullptr);
    if (rc) return rc;
    const double bq = f32_beta(std::max(t_quench, 0.01));               // coarsening_gpu.rs:15-16
    int64_t ns = 0;
    std::vector<double> betas((size_t)sample_every, bq);
    std::vector<int64_t> bond((size_t)n_trials);
    for (int64_t step = 0; step < total_steps; step += sample_every) {
        // sample after step (coarsening_gpu.rs:23-28), then run to the next sample point
        rc = ising_pt_anneal(h.p, betas.data(), 1, bond.data(), nullptr);
        if (rc) return rc;
        steps_out[ns] = step;
        rc = domain_walls(h.p, n, bond.data(), &rho_out[(size_t)(ns * n_trials)]);
        if (rc) return rc;
        ++ns;
        const int64_t rest = std::min(sample_every, total_steps - step) - 1;
        if (rest > 0) {
            rc = ising_pt_anneal(h.p, betas.data(), rest, nullptr, nullptr);
            if (rc) return rc;
        }
    }
    *n_samples = ns;
    return ISING_OK;
}

int ising_fss_anneal(const int32_t* sizes, int32_t n_sizes, double j, double h, double t_min, double t_max,
                     int32_t t_steps, int64_t warmup_sweeps, int64_t sample_sweeps, uint64_t seed,
                     ising_observables_t* out) {
    if (!sizes || !out || n_sizes < 1 || t_steps < 2 || sample_sweeps < 1 || warmup_sweeps < 0 ||
        !(t_min > 0.0) || !(t_max > t_min))
        return fail(ISING_E_ARG, "need sizes, out[n_sizes * t_steps], t_steps >= 2, samples >= 1, 0 < t_min < t_max");
    ising_pt_opts_t o;
    ising_pt_default_opts(&o);
    o.j = (double)(float)j;  // the reference's kernel takes J and h as f32 (lattice_gpu.rs:85)
    o.h = (double)(float)h;
    o.exchange_every = 1 << 30;
    o.graph_sweeps = 0;
    std::vector<double> temps(1, t_max);
    Handle hd;
    int rc = ising_pt_create(&hd.p, ISING_MODEL_ISING, sizes, n_sizes, 1, temps.data(), seed, &o);
    if (rc) return rc;
    ising_pt* p = hd.p;
    std::vector<int64_t> B((size_t)(sample_sweeps * n_sizes)), M((size_t)(sample_sweeps * n_sizes));
    for (int32_t step = t_steps - 1; step >= 0; --step) {                 // fss_gpu.rs:17-33: high T to low T
        const double T = t_min + (t_max - t_min) * (double)step / (double)(t_steps - 1);
        const double beta = f32_beta(T);
        std::vector<double> warm((size_t)warmup_sweeps, beta), meas((size_t)sample_sweeps, beta);
        rc = ising_pt_anneal(p, warm.data(), warmup_sweeps, nullptr, nullptr);
        if (rc) return rc;
        rc = ising_pt_anneal(p, meas.data(), sample_sweeps, B.data(), M.data());
        if (rc) return rc;
        for (int g = 0; g < n_sizes; ++g) {                               // lattice_gpu.rs:193-252
            const double n3 = (double)sizes[g] * sizes[g] * sizes[g], b = 1.0 / T;
            double se = 0, se2 = 0, sm = 0, sm2 = 0, sm4 = 0, sms = 0, sms2 = 0;
            for (int64_t k = 0; k < sample_sweeps; ++k) {
                const double Bk = (double)B[(size_t)(k * n_sizes + g)], Mk = (double)M[(size_t)(k * n_sizes + g)];
                const double e = (-j * Bk - h * Mk) / n3;
                const double ms = Mk / n3, ma = std::fabs(ms);
                se += e;
                se2 += e * e;
                sm += ma;
                sm2 += ma * ma;
                sm4 += ma * ma * ma * ma;
                sms += ms;
                sms2 += ms * ms;
            }
            const double s = (double)sample_sweeps;
            const double ae = se / s, ae2 = se2 / s, ams = sms / s, ams2 = sms2 / s;
            ising_observables_t& ob = out[(size_t)g * t_steps + step];
            ob.temperature = T;
            ob.energy = ae;
            ob.magnetisation = sm / s;
            ob.heat_capacity = b * b * (ae2 - ae * ae) * n3;
            ob.susceptibility = b * (ams2 - ams * ams) * n3;
            ob.m2 = sm2 / s;
            ob.m4 = sm4 / s;
        }
    }
    return ISING_OK;
}

}  // extern "C"

namespace {

// Device time of `sweeps` sweeps (after one untimed sweep: the first launch of a shape).
int time_sweeps(ising_pt* p, int sweeps, double* ms_out) {
    int rc = p->enqueue_sweeps(1);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(p->ev0, p->st));
    rc = p->enqueue_sweeps(sweeps);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(p->ev1, p->st));
    HIP_TRY(hipEventSynchronize(p->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, p->ev0, p->ev1));
    *ms_out = ms;
    return ISING_OK;
}

// Tuning cache: lines "key | value" in the file named by ISING_HIP_TUNE_CACHE
// (none when unset).  key = model, replicas, sizes and device architecture.
std::string tune_key(const ising_pt* p) {
    std::string arch = "unknown";
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) arch = prop.gcnArchName;
    std::string k = "model=" + std::to_string(p->model) + " R=" + std::to_string(p->R) + " sizes=";
    for (size_t i = 0; i < p->groups.size(); ++i) k += (i ? "," : "") + std::to_string(p->groups[i].L);
    return k + " device=" + arch;
}

bool tune_cache_read(const ising_pt* p, std::string* value) {
    const char* path = std::getenv("ISING_HIP_TUNE_CACHE");
    if (!path || !*path) return false;
    std::FILE* f = std::fopen(path, "r");
    if (!f) return false;
    const std::string key = tune_key(p) + " | ";
    bool found = false;
    char line[4096];
    while (std::fgets(line, sizeof(line), f)) {
        std::string l(line);
        while (!l.empty() && (l.back() == '\n' || l.back() == '\r')) l.pop_back();
        if (l.compare(0, key.size(), key) == 0) {
            *value = l.substr(key.size());
            found = true;  // the last entry wins
        }
    }
    std::fclose(f);
    return found;
}

void tune_cache_write(const ising_pt* p, const std::string& value) {
    const char* path = std::getenv("ISING_HIP_TUNE_CACHE");
    if (!path || !*path) return;
    std::FILE* f = std::fopen(path, "a");
    if (!f) return;
    std::fprintf(f, "%s | %s\n", tune_key(p).c_str(), value.c_str());
    std::fclose(f);
}

// Layout of group G for a candidate (width, runs, spt): what layout_ising_each makes of it.
std::array<int, 3> effective_layout(const DevGroup& G, int w, int r, int s) {
    int fast = G.dims == 2 ? 0 : ((G.Lh % 32) == 0 ? 3 : ((G.Lh % 16) == 0 ? 2 : ((G.Lh % 8) == 0 ? 1 : 0)));
    if (fast == 3 && w == 16) fast = 2;
    return {fast, fast == 2 ? r : 1, fast ? s : 1};
}

}  // namespace

extern "C" {

int ising_pt_autotune(ising_pt_t* p, int32_t sweeps, int32_t* width, int32_t* runs, int32_t* spt) {
    if (!p) return fail(ISING_E_ARG, "null handle");
    if (sweeps < 1 || sweeps > 100000) return fail(ISING_E_ARG, "sweeps per candidate must be in 1..100000");
    if (p->model == ISING_MODEL_ISING_MSC) return fail(ISING_E_STATE, "multispin-coded ensembles have no tunable layout");
    if (p->s_host >= 0) return fail(ISING_E_STATE, "autotune runs warm-up sweeps: call it before sampling");
    if (p->inline_thr || p->inline_cont) return fail(ISING_E_STATE, "single-lattice handles are not tuned");
    SYNC_ALL(p);
    const size_t ng = p->groups.size();
    std::string cached;
    const bool hit = tune_cache_read(p, &cached);
    if (p->ncomp > 0) {
        // continuous: sites per thread of the overrelaxation pass
        int best_s = 0;
        if (hit && std::sscanf(cached.c_str(), "or %d", &best_s) == 1 && (best_s == 1 || best_s == 2)) {
            p->or_spt = best_s;
        } else {
            double best = 1e300;
            for (int s2 = 1; s2 <= 2; ++s2) {
                p->or_spt = s2;
                double ms = 0.0;
                const int rc = time_sweeps(p, sweeps, &ms);
                if (rc) return rc;
                if (ms < best) {
                    best = ms;
                    best_s = s2;
                }
            }
            p->or_spt = best_s;
            tune_cache_write(p, "or " + std::to_string(best_s));
        }
        p->drop_graph();
        p->timed = false;
        if (width) *width = p->or_spt;
        if (runs) *runs = 0;
        if (spt) *spt = 0;
        return ISING_OK;
    }
    std::vector<int> cw(ng), cr(ng), cs(ng);
    bool have = false;
    if (hit) {
        // "ising w,r,s w,r,s ..." one triple per size
        const char* q = cached.c_str();
        if (std::strncmp(q, "ising", 5) == 0) {
            q += 5;
            size_t g = 0;
            int n = 0, w = 0, r = 0, s2 = 0;
            while (g < ng && std::sscanf(q, " %d,%d,%d%n", &w, &r, &s2, &n) == 3) {
                cw[g] = w;
                cr[g] = r;
                cs[g] = s2;
                ++g;
                q += n;
            }
            have = g == ng;
        }
    }
    if (!have) {
        // (run width, runs per step) x steps per tile, first for every size at once
        static const int kL[3][2] = {{16, 1}, {16, 2}, {32, 1}}, kS[3] = {1, 2, 4};
        double best = 1e300;
        int bw = 16, br = 2, bs = 1;
        for (int li = 0; li < 3; ++li)
            for (int si = 0; si < 3; ++si) {
                int rc = layout_ising(p, kL[li][0], kL[li][1], kS[si]);
                if (rc) return rc;
                double ms = 0.0;
                rc = time_sweeps(p, sweeps, &ms);
                if (rc) return rc;
                if (ms < best) {
                    best = ms;
                    bw = kL[li][0];
                    br = kL[li][1];
                    bs = kS[si];
                }
            }
        std::fill(cw.begin(), cw.end(), bw);
        std::fill(cr.begin(), cr.end(), br);
        std::fill(cs.begin(), cs.end(), bs);
        // then size by size, largest first (coordinate descent): a size's
        // block count changes how the launch's last round of blocks fills the device
        const std::vector<int> uw = cw, ur = cr, us = cs;
        if (ng > 1) {
            int rc = layout_ising_each(p, cw.data(), cr.data(), cs.data());
            if (rc) return rc;
            double base = 0.0;
            rc = time_sweeps(p, sweeps, &base);
            if (rc) return rc;
            std::vector<int> order(ng);
            for (size_t i = 0; i < ng; ++i) order[i] = (int)i;
            std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return p->groups[x].L > p->groups[y].L; });
            for (int g : order) {
                std::vector<std::array<int, 3>> seen{effective_layout(p->groups[g], cw[g], cr[g], cs[g])};
                for (int li = 0; li < 3; ++li)
                    for (int si = 0; si < 3; ++si) {
                        const std::array<int, 3> e = effective_layout(p->groups[g], kL[li][0], kL[li][1], kS[si]);
                        if (std::find(seen.begin(), seen.end(), e) != seen.end()) continue;
                        seen.push_back(e);
                        const int ow = cw[g], orr = cr[g], os = cs[g];
                        cw[g] = kL[li][0];
                        cr[g] = kL[li][1];
                        cs[g] = kS[si];
                        rc = layout_ising_each(p, cw.data(), cr.data(), cs.data());
                        if (rc) return rc;
                        double ms = 0.0;
                        rc = time_sweeps(p, sweeps, &ms);
                        if (rc) return rc;
                        if (ms < 0.995 * base) {
                            base = ms;  // keep it
                        } else {
                            cw[g] = ow;
                            cr[g] = orr;
                            cs[g] = os;
                        }
                    }
            }
            // keep the per-size layouts only if they beat the common one again
            // in interleaved timings (median of three each): single timings
            // differ by about 1% from run to run
            if (cw != uw || cr != ur || cs != us) {
                double tu[3], tp[3];
                for (int k = 0; k < 3; ++k) {
                    rc = layout_ising_each(p, uw.data(), ur.data(), us.data());
                    if (!rc) rc = time_sweeps(p, sweeps, &tu[k]);
                    if (!rc) rc = layout_ising_each(p, cw.data(), cr.data(), cs.data());
                    if (!rc) rc = time_sweeps(p, sweeps, &tp[k]);
                    if (rc) return rc;
                }
                std::sort(tu, tu + 3);
                std::sort(tp, tp + 3);
                if (!(tp[1] < 0.995 * tu[1])) {
                    cw = uw;
                    cr = ur;
                    cs = us;
                }
            }
        }
        std::string v = "ising";
        for (size_t g = 0; g < ng; ++g)
            v += " " + std::to_string(cw[g]) + "," + std::to_string(cr[g]) + "," + std::to_string(cs[g]);
        tune_cache_write(p, v);
    }
    const int rc = layout_ising_each(p, cw.data(), cr.data(), cs.data());
    if (rc) return rc;
    p->timed = false;
    // the largest size's layout
    size_t gl = 0;
    for (size_t g = 1; g < ng; ++g)
        if (p->groups[g].L > p->groups[gl].L) gl = g;
    if (width) *width = cw[gl];
    if (runs) *runs = cr[gl];
    if (spt) *spt = cs[gl];
    return ISING_OK;
}

int ising_pt_get_layout(ising_pt_t* p, int32_t size_index, int32_t* width, int32_t* runs, int32_t* spt) {
    if (!p) return fail(ISING_E_ARG, "null handle");
    if (size_index < 0 || size_index >= (int)p->groups.size()) return fail(ISING_E_ARG, "size_index out of range");
    if (p->ncomp > 0) {
        if (width) *width = p->or_spt ? p->or_spt : (p->ncomp == 2 ? 2 : 1);
        if (runs) *runs = 0;
        if (spt) *spt = 0;
        return ISING_OK;
    }
    if (p->glayout.empty()) return fail(ISING_E_STATE, "multispin-coded ensembles have no thread-step layout");
    if (width) *width = p->glayout[size_index][0];
    if (runs) *runs = p->glayout[size_index][1];
    if (spt) *spt = p->glayout[size_index][2];
    return ISING_OK;
}

int ising_pt_prepare_graph(ising_pt_t* p, int32_t graph_sweeps) {
    if (!p) return fail(ISING_E_ARG, "null handle");
    if (graph_sweeps < 0 || graph_sweeps > 100000) return fail(ISING_E_ARG, "graph_sweeps must be in 0..100000");
    SYNC_ALL(p);
    p->drop_graph();
    p->opts.graph_sweeps = graph_sweeps;
    if (p->graph_capable()) return p->capture_graph();
    return ISING_OK;
}

int ising_pt_sweeps(const ising_pt_t* p, int64_t* out) {
    if (!p || !out) return fail(ISING_E_ARG, "null argument");
    *out = p->t_host;
    return ISING_OK;
}

int ising_pt_device_bytes(const ising_pt_t* p, int64_t* out) {
    if (!p || !out) return fail(ISING_E_ARG, "null argument");
    *out = p->pool_bytes;
    return ISING_OK;
}

// ---------------------------------------------------------------------------
// Single lattice (LatticeGpu mirror)
// ---------------------------------------------------------------------------

int ising_lattice_create(ising_lattice_t** out, int n, uint64_t seed) {
    if (!out) return fail(ISING_E_ARG, "out is null");
    *out = nullptr;
    if (!valid_L(n)) return fail(ISING_E_ARG, "n must be even, 2..2046");
    ising_lattice* l = new (std::nothrow) ising_lattice();
    if (!l) return fail(ISING_E_HIP, "host allocation failed");
    l->L = n;
    const int32_t sizes[1] = {n};
    const double temps[1] = {1.0};
    ising_pt_opts_t o;
    ising_pt_default_opts(&o);
    o.record_capacity = 1;
    const int rc = ising_pt_create(&l->pt, ISING_MODEL_ISING, sizes, 1, 1, temps, seed, &o);
    if (rc) {
        const std::string keep = g_err;
        delete l;
        g_err = keep;
        return rc;
    }
    l->pt->inline_thr = true;
    *out = l;
    return ISING_OK;
}

void ising_lattice_destroy(ising_lattice_t* lat) { delete lat; }
int ising_lattice_size(const ising_lattice_t* lat) { return lat ? lat->L : 0; }

int ising_lattice_step(ising_lattice_t* l, float beta, float j, float h) {
    if (!l) return fail(ISING_E_ARG, "null handle");
    if (!(beta >= 0.0f) || !std::isfinite(beta) || !std::isfinite(j) || !std::isfinite(h))
        return fail(ISING_E_ARG, "beta must be finite and >= 0; j, h finite");
    if (!(beta == l->beta && j == l->j && h == l->h)) {
        ising_table((double)beta, (double)j, (double)h, l->pt->thr_inline);
        l->beta = beta;
        l->j = j;
        l->h = h;
    }
    return l->pt->enqueue_sweeps(1);
}

int ising_lattice_cluster_step(ising_lattice_t* l, int32_t kind, float beta, float j, int64_t* flipped) {
    if (!l) return fail(ISING_E_ARG, "null handle");
    if (!(beta >= 0.0f) || !std::isfinite(beta) || !std::isfinite(j))
        return fail(ISING_E_ARG, "beta must be finite and >= 0; j finite");
    return cluster_run(l->pt, kind, 1, true, (double)beta, (double)j, flipped, nullptr, nullptr, nullptr);
}

int ising_lattice_warm_up(ising_lattice_t* l, float beta, float j, float h, size_t sweeps) {
    for (size_t i = 0; i < sweeps; ++i) {
        const int rc = ising_lattice_step(l, beta, j, h);
        if (rc) return rc;
    }
    return l ? ISING_OK : fail(ISING_E_ARG, "null handle");
}

int ising_lattice_get_spins(ising_lattice_t* l, int8_t* out) {
    if (!l) return fail(ISING_E_ARG, "null handle");
    return ising_pt_get_spins_i8(l->pt, 0, 0, out);
}

int ising_lattice_set_spins(ising_lattice_t* l, const int8_t* in) {
    if (!l) return fail(ISING_E_ARG, "null handle");
    return ising_pt_set_spins_i8(l->pt, 0, 0, in);
}

int ising_lattice_measure_raw(ising_lattice_t* l, int64_t* bond, int64_t* mag) {
    if (!l) return fail(ISING_E_ARG, "null handle");
    ising_pt* p = l->pt;
    int64_t* acc = p->d_acc + acc_words(p->total_reps) - 8;  // standalone-measurement slots
    HIP_TRY(hipMemsetAsync(acc, 0, 2 * sizeof(int64_t), p->st));
    HIP_TRY(launch_ising_measure(p->d_groups, 1, p->d_pool, acc - 0, p->nblocks, p->st));
    int64_t v[2];
    HIP_TRY(hipMemcpyAsync(v, acc, sizeof(v), hipMemcpyDeviceToHost, p->st));
    SYNC_ALL(p);
    if (bond) *bond = v[0];
    if (mag) *mag = v[1];
    return ISING_OK;
}

int ising_lattice_measure_gpu(ising_lattice_t* l, float j, double* e, double* m) {
    int64_t B, M;
    const int rc = ising_lattice_measure_raw(l, &B, &M);
    if (rc) return rc;
    const double n3 = (double)l->L * l->L * l->L;
    if (e) *e = -(double)j * (double)B / n3;
    if (m) *m = std::fabs((double)M) / n3;
    return ISING_OK;
}

int ising_lattice_energy_gpu(ising_lattice_t* l, float j, double* e_total) {
    int64_t B;
    const int rc = ising_lattice_measure_raw(l, &B, nullptr);
    if (rc) return rc;
    if (e_total) *e_total = -(double)j * (double)B;
    return ISING_OK;
}

int ising_lattice_measure(ising_lattice_t* l, double T, double j, double h, size_t samples,
                          ising_observables_t* out) {
    if (!l || !out) return fail(ISING_E_ARG, "null argument");
    if (!(T > 0.0) || samples == 0) return fail(ISING_E_ARG, "temperature must be > 0 and samples >= 1");
    const double beta = 1.0 / T, n3 = (double)l->L * l->L * l->L;
    double se = 0, se2 = 0, sm = 0, sm2 = 0, sm4 = 0, sms = 0, sms2 = 0;
    for (size_t i = 0; i < samples; ++i) {
        int rc = ising_lattice_step(l, (float)beta, (float)j, (float)h);
        if (rc) return rc;
        int64_t B, M;
        rc = ising_lattice_measure_raw(l, &B, &M);
        if (rc) return rc;
        const double e = (-j * (double)B - h * (double)M) / n3;
        const double ms = (double)M / n3, ma = std::fabs(ms);
        se += e;
        se2 += e * e;
        sm += ma;
        sm2 += ma * ma;
        sm4 += ma * ma * ma * ma;
        sms += ms;
        sms2 += ms * ms;
    }
    const double s = (double)samples;
    const double ae = se / s, ae2 = se2 / s, am = sm / s, am2 = sm2 / s, am4 = sm4 / s, ams = sms / s,
                 ams2 = sms2 / s;
    out->temperature = T;
    out->energy = ae;
    out->magnetisation = am;
    out->heat_capacity = beta * beta * (ae2 - ae * ae) * n3;
    out->susceptibility = beta * (ams2 - ams * ams) * n3;
    out->m2 = am2;
    out->m4 = am4;
    return ISING_OK;
}

int ising_lattice_domain_wall_density(ising_lattice_t* l, double* out) {
    int64_t B;
    const int rc = ising_lattice_measure_raw(l, &B, nullptr);
    if (rc) return rc;
    // 3N forward bonds; B = (#equal - #opposite) = 3N - 2 #walls
    const double bonds = 3.0 * (double)l->L * l->L * l->L;
    if (out) *out = (bonds - (double)B) / (2.0 * bonds);
    return ISING_OK;
}

int ising_lattice_sweeps(const ising_lattice_t* l, int64_t* out) {
    if (!l || !out) return fail(ISING_E_ARG, "null argument");
    *out = l->pt->t_host;
    return ISING_OK;
}

// ---------------------------------------------------------------------------
// Single continuous lattice (ContinuousGpuLattice mirror)
// ---------------------------------------------------------------------------

}  // extern "C"

struct ising_cont {
    ising_pt* pt = nullptr;
    int L = 0, ncomp = 0;
    ~ising_cont() { delete pt; }
};

extern "C" {

int ising_cont_create(ising_cont_t** out, int n, int n_comp, uint64_t seed) {
    if (!out) return fail(ISING_E_ARG, "out is null");
    *out = nullptr;
    if (!valid_L(n)) return fail(ISING_E_ARG, "n must be even, 2..2046");
    if (n_comp != 2 && n_comp != 3) return fail(ISING_E_ARG, "n_comp must be 2 (XY) or 3 (Heisenberg)");
    ising_cont* l = new (std::nothrow) ising_cont();
    if (!l) return fail(ISING_E_HIP, "host allocation failed");
    l->L = n;
    l->ncomp = n_comp;
    const int32_t sizes[1] = {n};
    const double temps[1] = {1.0};
    ising_pt_opts_t o;
    ising_pt_default_opts(&o);
    o.record_capacity = 1;
    o.renorm_every = 0;  // the reference lattice never renormalises (continuous_spin_kernel.cu:85-142)
    const int rc = ising_pt_create(&l->pt, n_comp == 2 ? ISING_MODEL_XY : ISING_MODEL_HEISENBERG, sizes, 1, 1,
                                   temps, seed, &o);
    if (rc) {
        const std::string keep = g_err;
        delete l;
        g_err = keep;
        return rc;
    }
    l->pt->inline_cont = true;
    *out = l;
    return ISING_OK;
}

void ising_cont_destroy(ising_cont_t* lat) { delete lat; }

static int cont_params(ising_cont* l, float beta, float j, float delta) {
    if (!l) return fail(ISING_E_ARG, "null handle");
    if (!(beta >= 0.0f) || !std::isfinite(beta) || !std::isfinite(j) || !(delta > 0.0f) || !std::isfinite(delta))
        return fail(ISING_E_ARG, "beta >= 0, finite j and delta > 0 required");
    l->pt->beta_inline = beta;
    l->pt->j_inline = j;
    l->pt->delta_inline = delta;
    return ISING_OK;
}

int ising_cont_sweep(ising_cont_t* l, float beta, float j, float delta, size_t n_overrelax) {
    const int rc = cont_params(l, beta, j, delta);
    if (rc) return rc;
    if (n_overrelax > 1000000) return fail(ISING_E_ARG, "n_overrelax too large");
    return l->pt->enqueue_sweeps(1, (int)n_overrelax);
}

int ising_cont_cluster_step(ising_cont_t* l, int32_t kind, float beta, float j, int64_t* flipped) {
    if (!l) return fail(ISING_E_ARG, "null handle");
    if (!(beta >= 0.0f) || !std::isfinite(beta) || !std::isfinite(j))
        return fail(ISING_E_ARG, "beta must be finite and >= 0; j finite");
    return cluster_run(l->pt, kind, 1, true, (double)beta, (double)j, flipped, nullptr, nullptr, nullptr);
}

int ising_cont_metropolis_pass(ising_cont_t* l, int color, float beta, float j, float delta) {
    const int rc = cont_params(l, beta, j, delta);
    if (rc) return rc;
    if (color != 0 && color != 1) return fail(ISING_E_ARG, "color must be 0 or 1");
    HIP_TRY(launch_cont_metro(l->pt->cont_args(), l->pt->nblocks, color, 0, l->pt->st));
    return ISING_OK;
}

int ising_cont_overrelax_pass(ising_cont_t* l, int color, float j) {
    const int rc = cont_params(l, 1.0f, j, 1.0f);
    if (rc) return rc;
    if (color != 0 && color != 1) return fail(ISING_E_ARG, "color must be 0 or 1");
    HIP_TRY(launch_cont_or(l->pt->cont_args(), l->pt->nblocks, color, l->pt->st));
    return ISING_OK;
}

int ising_cont_measure_gpu(ising_cont_t* l, float j, double* e, double* mx, double* my, double* mz) {
    if (!l) return fail(ISING_E_ARG, "null handle");
    ising_pt* p = l->pt;
    HIP_TRY(launch_cont_measure(p->d_groups, 1, p->d_fpool, p->ncomp, (double)j, p->d_part0, p->nblocks, p->st));
    std::vector<double> part(4 * (size_t)p->nblocks);
    HIP_TRY(hipMemcpyAsync(part.data(), p->d_part0, sizeof(double) * part.size(), hipMemcpyDeviceToHost, p->st));
    SYNC_ALL(p);
    double v[4] = {0, 0, 0, 0};
    for (int b = 0; b < p->nblocks; ++b)
        for (int k = 0; k < 4; ++k) v[k] += part[4 * (size_t)b + k];
    if (e) *e = v[0];
    if (mx) *mx = v[1];
    if (my) *my = v[2];
    if (mz) *mz = v[3];
    return ISING_OK;
}

int ising_cont_measure_raw(ising_cont_t* l, double* e, double* mx, double* my, double* mz) {
    return ising_cont_measure_gpu(l, 1.0f, e, mx, my, mz);
}

int ising_cont_get_spins(ising_cont_t* l, float* out) {
    if (!l) return fail(ISING_E_ARG, "null handle");
    return ising_pt_get_spins_f32(l->pt, 0, 0, out);
}

int ising_cont_set_spins(ising_cont_t* l, const float* in) {
    if (!l) return fail(ISING_E_ARG, "null handle");
    return ising_pt_set_spins_f32(l->pt, 0, 0, in);
}

int ising_cont_sweeps(const ising_cont_t* l, int64_t* out) {
    if (!l || !out) return fail(ISING_E_ARG, "null argument");
    *out = l->pt->t_host;
    return ISING_OK;
}

// ---------------------------------------------------------------------------
// reduce_gpu::* (src/cuda/reduce_gpu.rs:8-178): reductions of a caller's
// device lattice in the reference's layout.  Like the reference, each call
// allocates its partials, runs on the device's default stream and copies the
// totals back synchronously.
// ---------------------------------------------------------------------------

}  // extern "C"

namespace {

int reduce_device_ready() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return fail(ISING_E_HIP, "no HIP device visible");
    return ISING_OK;
}

// Totals of an int8 +-1 lattice: bond sum B (forward neighbours) and spin sum.
int reduce_ising_totals(const int8_t* d_spins, int32_t n, int64_t* bm) {
    if (!d_spins || !bm) return fail(ISING_E_ARG, "null argument");
    if (n < 1 || n > 1024) return fail(ISING_E_ARG, "n must be in 1..1024");
    int rc = reduce_device_ready();
    if (rc) return rc;
    const int nblk = reduce_blocks((int64_t)n * n * n);
    int64_t* buf = nullptr;
    HIP_TRY(hipMalloc(&buf, sizeof(int64_t) * (2 * (size_t)nblk + 2)));
    hipError_t e = launch_reduce_ising(d_spins, n, buf + 2, buf, nullptr);
    if (e == hipSuccess) e = hipMemcpy(bm, buf, 2 * sizeof(int64_t), hipMemcpyDeviceToHost);
    (void)hipFree(buf);
    HIP_TRY(e);
    return ISING_OK;
}

// FP64 totals of a continuous lattice: [0] = sum of S_i . S_fwd (when n > 0), [1..3] = Sx, Sy, Sz.
int reduce_cont_totals(const float* d_spins, int64_t n_sites, int32_t n, int32_t n_comp, double* out) {
    if (!d_spins || !out) return fail(ISING_E_ARG, "null argument");
    if (n_comp != 2 && n_comp != 3) return fail(ISING_E_ARG, "n_comp must be 2 or 3");
    if (n_sites < 0 || n_sites > (int64_t)1 << 30) return fail(ISING_E_ARG, "n_sites must be in 0..2^30");
    int rc = reduce_device_ready();
    if (rc) return rc;
    if (n_sites == 0) {
        for (int q = 0; q < 4; ++q) out[q] = 0.0;
        return ISING_OK;
    }
    const int nblk = reduce_blocks(n_sites);
    double* buf = nullptr;
    HIP_TRY(hipMalloc(&buf, sizeof(double) * (4 * (size_t)nblk + 4)));
    hipError_t e = launch_reduce_cont(d_spins, n_sites, n, n_comp, buf + 4, buf, nullptr);
    if (e == hipSuccess) e = hipMemcpy(out, buf, 4 * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(buf);
    HIP_TRY(e);
    return ISING_OK;
}

}  // namespace

extern "C" {

int ising_load_reduce_kernels(void) { return reduce_device_ready(); }

int ising_reduce_ising(const int8_t* d_spins, int32_t n, float j, double* e_total, double* m_total) {
    int64_t bm[2];
    const int rc = reduce_ising_totals(d_spins, n, bm);
    if (rc) return rc;
    if (e_total) *e_total = -(double)j * (double)bm[0];
    if (m_total) *m_total = (double)bm[1];
    return ISING_OK;
}

int ising_measure_ising_gpu(const int8_t* d_spins, int32_t n, float j, double* e_per_spin, double* m_abs_per_spin) {
    double e = 0.0, m = 0.0;
    const int rc = ising_reduce_ising(d_spins, n, j, &e, &m);
    if (rc) return rc;
    const double n3 = (double)n * n * n;
    if (e_per_spin) *e_per_spin = e / n3;
    if (m_abs_per_spin) *m_abs_per_spin = std::fabs(m) / n3;
    return ISING_OK;
}

int ising_reduce_continuous_mag(const float* d_spins, int64_t n_sites, int32_t n_comp, double* mx, double* my,
                                double* mz) {
    double v[4];
    const int rc = reduce_cont_totals(d_spins, n_sites, 0, n_comp, v);
    if (rc) return rc;
    if (mx) *mx = v[1];
    if (my) *my = v[2];
    if (mz) *mz = v[3];
    return ISING_OK;
}

int ising_reduce_continuous_energy(const float* d_spins, int32_t n, int32_t n_comp, float j, double* e_total) {
    if (n < 1 || n > 1024) return fail(ISING_E_ARG, "n must be in 1..1024");
    double v[4];
    const int rc = reduce_cont_totals(d_spins, (int64_t)n * n * n, n, n_comp, v);
    if (rc) return rc;
    if (e_total) *e_total = -(double)j * v[0];
    return ISING_OK;
}

int ising_measure_continuous_gpu(const float* d_spins, int32_t n, int32_t n_comp, float j, double* e, double* mx,
                                 double* my, double* mz) {
    if (n < 1 || n > 1024) return fail(ISING_E_ARG, "n must be in 1..1024");
    double v[4];
    const int rc = reduce_cont_totals(d_spins, (int64_t)n * n * n, n, n_comp, v);
    if (rc) return rc;
    if (e) *e = -(double)j * v[0];
    if (mx) *mx = v[1];
    if (my) *my = v[2];
    if (mz) *mz = v[3];
    return ISING_OK;
}

// ---------------------------------------------------------------------------
// Batch of ensembles advanced together (one graph with a branch per member)
// ---------------------------------------------------------------------------

}  // extern "C"

struct ising_batch {
    std::vector<ising_pt*> m;
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, fork = nullptr;
    std::vector<hipEvent_t> join;
    std::vector<int> gens;  // members' graph generations the captured graph was built for
    hipGraphExec_t exec = nullptr;
    int G = 32;
    bool timed = false;
    int mode = ISING_BATCH_BRANCHES;  // ISING_BATCH_*
    int resident = 0;                 // persistent work list: blocks the device holds at once
    ~ising_batch() {
        if (st) (void)hipStreamSynchronize(st);
        for (ising_pt* p : m) p->ext_st = nullptr;
        if (exec) (void)hipGraphExecDestroy(exec);
        for (hipEvent_t e : join) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (st) (void)hipStreamDestroy(st);
    }

    // order: members' own streams wait for everything queued on the batch stream
    int members_after_batch() {
        HIP_TRY(hipEventRecord(fork, st));
        for (ising_pt* p : m) HIP_TRY(hipStreamWaitEvent(p->st, fork, 0));
        return ISING_OK;
    }
    int batch_after_members() {
        for (size_t i = 0; i < m.size(); ++i) {
            HIP_TRY(hipEventRecord(join[i], m[i]->st));
            HIP_TRY(hipStreamWaitEvent(st, join[i], 0));
        }
        return ISING_OK;
    }

    // Work-list mode: every pass of every member in one launch of the work-list
    // kernel on the batch stream (batch.hip); n sweeps, then the Ising
    // member's closing kernel.
    int enqueue_worklist(int64_t n) {
        ising_pt* pi = nullptr;
        std::vector<ising_pt*> pc;
        for (ising_pt* p : m) {
            // a member switched to one-launch, learned or heat-bath sweeps after the
            // mode was set would silently run the work list's passes instead
            if (p->fused_on || p->learned_on || p->heatbath_on)
                return fail(ISING_E_STATE, "a work-list batch member runs one-launch, learned or heat-bath sweeps");
            if (p->model == ISING_MODEL_ISING) pi = p;
            else pc.push_back(p);
        }
        BatchPass base{};
        if (pi) base.ia = pi->ising_args();
        for (size_t i = 0; i < pc.size(); ++i) base.ca[i] = pc[i]->cont_args();
        const int nor = pc.empty() ? 0 : pc[0]->opts.n_overrelax;
        const bool persistent = mode == ISING_BATCH_PERSISTENT;
        // Kind W_ISING: the Ising member's colour pass, in launches of its own (in
        // a launch with the continuous Metropolis tiles, the Ising body's
        // registers set the occupancy of every item: 5.26 against 4.92 ms per
        // config-4 sweep).
        auto pass = [&](int kind, int color) -> hipError_t {
            BatchPass bp = base;
            bp.color = color;
            auto add = [&](int k, int member, int count, int tiles) {
                if (bp.nseg < kBatchSegs) bp.seg[bp.nseg++] = WorkSeg{k, member, count, 0, 0, tiles};
            };  // at most two continuous members (ising_batch_set_mode)
            if (kind == W_ISING) {
                add(W_ISING, 0, pi->nblocks, pi->nblocks);
                return launch_batch_pass(bp, persistent, st);
            }
            for (size_t i = 0; i < pc.size(); ++i) {
                const int tiles = pc[i]->nblocks;
                if (kind == W_CLOSE)
                    add(kind, (int)i, (pc[i]->total_reps + kContClosePerBlock - 1) / kContClosePerBlock, 0);
                else  // an XY item holds two tiles
                    add(kind, (int)i, pc[i]->model == ISING_MODEL_XY ? (tiles + 1) / 2 : tiles, tiles);
            }
            return launch_batch_pass(bp, persistent, st);
        };
        for (int64_t i = 0; i < n; ++i) {
            for (int k = 0; k < nor; ++k) {
                HIP_TRY(pass(W_OR, 0));
                HIP_TRY(pass(W_OR, 1));
            }
            for (int c = 0; c < 2; ++c) {
                if (pi) HIP_TRY(pass(W_ISING, c));
                if (!pc.empty()) HIP_TRY(pass(W_METRO, c));
            }
            if (!pc.empty()) HIP_TRY(pass(W_CLOSE, 0));
        }
        if (n > 0 && pi) HIP_TRY(launch_ising_close(pi->ising_args(), st));
        for (ising_pt* p : m) p->t_host += n;
        return ISING_OK;
    }

    int run(int64_t n) {
        int rc = members_after_batch();
        if (rc) return rc;
        rc = batch_after_members();  // the batch stream now follows any eager member work
        if (rc) return rc;
        HIP_TRY(hipEventRecord(ev0, st));
        if (n >= G) {
            bool stale = gens.size() != m.size();
            for (size_t i = 0; !stale && i < m.size(); ++i) stale = gens[i] != m[i]->graph_gen;
            if (exec && stale) {
                (void)hipGraphExecDestroy(exec);
                exec = nullptr;
            }
            if (!exec) {
                HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                hipError_t e = hipEventRecord(fork, st);
                int rc2 = ISING_OK;
                if (mode != ISING_BATCH_BRANCHES) {
                    rc2 = enqueue_worklist(G);
                    if (rc2 == ISING_OK)
                        for (ising_pt* p : m) p->t_host -= G;  // capture enqueued nothing
                }
                for (size_t i = 0; mode == ISING_BATCH_BRANCHES && i < m.size() && e == hipSuccess && rc2 == ISING_OK;
                     ++i) {
                    e = hipStreamWaitEvent(m[i]->st, fork, 0);
                    if (e != hipSuccess) break;
                    rc2 = m[i]->enqueue_sweeps(G);
                    if (rc2 == ISING_OK) m[i]->t_host -= G;  // capture enqueued nothing
                    if (rc2 == ISING_OK) e = hipEventRecord(join[i], m[i]->st);
                    if (e == hipSuccess && rc2 == ISING_OK) e = hipStreamWaitEvent(st, join[i], 0);
                }
                hipGraph_t graph = nullptr;
                const hipError_t ec = hipStreamEndCapture(st, &graph);
                if (rc2 != ISING_OK) {
                    if (graph) (void)hipGraphDestroy(graph);
                    return rc2;
                }
                HIP_TRY(e);
                HIP_TRY(ec);
                const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                HIP_TRY(ei);
                gens.clear();
                for (ising_pt* p : m) gens.push_back(p->graph_gen);
            }
            const int64_t reps = n / G;
            for (int64_t i = 0; i < reps; ++i) HIP_TRY(hipGraphLaunch(exec, st));
            for (ising_pt* p : m) p->t_host += reps * G;
            n -= reps * G;
        }
        if (n > 0 && mode != ISING_BATCH_BRANCHES) {
            rc = enqueue_worklist(n);
            if (rc) return rc;
        } else if (n > 0) {
            rc = members_after_batch();
            if (rc) return rc;
            for (ising_pt* p : m) {
                rc = p->enqueue_sweeps(n);
                if (rc) return rc;
            }
            rc = batch_after_members();
            if (rc) return rc;
        }
        HIP_TRY(hipEventRecord(ev1, st));
        timed = true;
        return ISING_OK;
    }
};

extern "C" {

int ising_batch_create(ising_batch_t** out, ising_pt_t* const* members, int32_t n, int32_t graph_sweeps) {
    if (!out || !members || n < 1 || n > 64 || graph_sweeps < 1 || graph_sweeps > 100000)
        return fail(ISING_E_ARG, "need 1..64 ensembles and graph_sweeps in 1..100000");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (!members[i]) return fail(ISING_E_ARG, "null member");
        if (members[i]->ext_st) return fail(ISING_E_STATE, "an ensemble can belong to one batch only");
        if (members[i]->inline_thr || members[i]->inline_cont) return fail(ISING_E_ARG, "single-lattice handles cannot be batched");
        for (int j = 0; j < i; ++j)
            if (members[j] == members[i]) return fail(ISING_E_ARG, "duplicate member");
    }
    ising_batch* b = new (std::nothrow) ising_batch();
    if (!b) return fail(ISING_E_HIP, "host allocation failed");
    auto bad = [&](hipError_t e, const char* what) {
        delete b;
        return fail(ISING_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking)) != hipSuccess) return bad(e, "stream");
    if ((e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming)) != hipSuccess) return bad(e, "event");
    if ((e = hipEventCreate(&b->ev0)) != hipSuccess) return bad(e, "event");
    if ((e = hipEventCreate(&b->ev1)) != hipSuccess) return bad(e, "event");
    b->join.resize(n, nullptr);
    for (int i = 0; i < n; ++i)
        if ((e = hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming)) != hipSuccess) return bad(e, "event");
    b->G = graph_sweeps;
    for (int i = 0; i < n; ++i) {
        b->m.push_back(members[i]);
        members[i]->ext_st = b->st;
    }
    *out = b;
    return ISING_OK;
}

void ising_batch_destroy(ising_batch_t* b) { delete b; }

int ising_batch_set_mode(ising_batch_t* b, int32_t mode) {
    if (!b) return fail(ISING_E_ARG, "null handle");
    if (mode != ISING_BATCH_BRANCHES && mode != ISING_BATCH_WORKLIST && mode != ISING_BATCH_PERSISTENT)
        return fail(ISING_E_ARG, "mode must be ISING_BATCH_BRANCHES, ISING_BATCH_WORKLIST or ISING_BATCH_PERSISTENT");
    if (mode != ISING_BATCH_BRANCHES) {
        int n_ising = 0, n_cont = 0, nor = -1;
        for (ising_pt* p : b->m) {
            if (p->model == ISING_MODEL_ISING) {
                if (p->fused_on) return fail(ISING_E_ARG, "work-list batches need two-launch Ising sweeps (ising_fused = 0)");
                ++n_ising;
            } else if (p->model == ISING_MODEL_XY || p->model == ISING_MODEL_HEISENBERG) {
                if (p->learned_on || p->heatbath_on)
                    return fail(ISING_E_ARG, "work-list batches run Metropolis passes (no learned or heat-bath passes)");
                if (nor >= 0 && p->opts.n_overrelax != nor)
                    return fail(ISING_E_ARG, "the continuous members of a work-list batch need the same n_overrelax");
                nor = p->opts.n_overrelax;
                ++n_cont;
            } else {
                return fail(ISING_E_ARG, "work-list batches hold int8 Ising, XY and Heisenberg ensembles");
            }
        }
        if (n_ising > 1 || n_cont > kBatchCont)
            return fail(ISING_E_ARG, "a work-list batch holds at most one Ising and two continuous ensembles");
        if (mode == ISING_BATCH_PERSISTENT) {
            b->resident = batch_pass_resident_blocks();
            if (b->resident < 8) return fail(ISING_E_HIP, "cannot size the persistent grid on this device");
        }
    }
    HIP_TRY(hipStreamSynchronize(b->st));
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
    b->mode = mode;
    return ISING_OK;
}

int ising_batch_warmup(ising_batch_t* b, int64_t sweeps) {
    if (!b || sweeps < 0) return fail(ISING_E_ARG, "null handle or negative sweeps");
    for (ising_pt* p : b->m)
        if (p->s_host >= 0) return fail(ISING_E_STATE, "warm-up after sampling started");
    return b->run(sweeps);
}

int ising_batch_sample(ising_batch_t* b, int64_t sweeps) {
    if (!b || sweeps < 0) return fail(ISING_E_ARG, "null handle or negative sweeps");
    for (ising_pt* p : b->m) {
        const int rc = p->begin_sampling();
        if (rc) return rc;
        if (p->pending_host + measured_in(p->s_host, sweeps, p->opts.measure_every) > p->opts.record_capacity)
            return fail(ISING_E_FULL, "a member's measurement buffer would overflow: drain records first");
    }
    const int rc = b->run(sweeps);
    if (rc) return rc;
    for (ising_pt* p : b->m) {
        p->pending_host += measured_in(p->s_host, sweeps, p->opts.measure_every);
        p->s_host += sweeps;
    }
    return ISING_OK;
}

int ising_batch_sync(ising_batch_t* b, double* last_ms) {
    if (!b) return fail(ISING_E_ARG, "null handle");
    for (ising_pt* p : b->m) SYNC_ALL(p);
    HIP_TRY(hipStreamSynchronize(b->st));
    if (last_ms) {
        float ms = 0.f;
        if (b->timed) HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        *last_ms = ms;
    }
    return ISING_OK;
}

// ---------------------------------------------------------------------------
// Host statistics
// ---------------------------------------------------------------------------

int ising_linspace_temperatures(double t_min, double t_max, int32_t n, double* out) {
    if (n < 1 || !out) return fail(ISING_E_ARG, "n must be >= 1 and out non-null");
    if (n == 1) {
        out[0] = t_min;
        return ISING_OK;
    }
    for (int i = 0; i < n; ++i) out[i] = t_min + (t_max - t_min) * (double)i / (double)(n - 1);
    return ISING_OK;
}

int ising_replica_exchange(const double* temps, const double* energies, int32_t n, int32_t* r2t, int32_t* t2r,
                           const double* uniforms, int32_t even_odd, int32_t* accepted) {
    if (!temps || !energies || !r2t || !t2r || !uniforms || n < 0) return fail(ISING_E_ARG, "null argument");
    int acc = 0, k = 0;
    for (int i = even_odd % 2; i + 1 < n; i += 2, ++k) {
        const int ri = t2r[i], rj = t2r[i + 1];
        const double bi = 1.0 / temps[i], bj = 1.0 / temps[i + 1];
        const double delta = (bi - bj) * (energies[rj] - energies[ri]);
        if (delta < 0.0 || uniforms[k] < std::exp(-delta)) {
            t2r[i] = rj;
            t2r[i + 1] = ri;
            r2t[ri] = i + 1;
            r2t[rj] = i;
            ++acc;
        }
    }
    if (accepted) *accepted = acc;
    return ISING_OK;
}

int ising_jackknife(const double* e, const double* m, int64_t n, double beta, double n3, int32_t n_blocks,
                    double* out) {
    if (!out || (n > 0 && (!e || !m))) return fail(ISING_E_ARG, "null argument");
    for (int i = 0; i < 12; ++i) out[i] = 0.0;
    if (n < n_blocks || n_blocks < 2) return ISING_OK;
    const int64_t bs = n / n_blocks, used = bs * n_blocks;
    double se = 0, se2 = 0, sm = 0, sm2 = 0, sm4 = 0;
    for (int64_t i = 0; i < used; ++i) {
        const double q = m[i] * m[i];
        se += e[i];
        se2 += e[i] * e[i];
        sm += m[i];
        sm2 += q;
        sm4 += q * q;
    }
    const double s = (double)used;
    const double full[6] = {se / s, sm / s, sm2 / s, sm4 / s,
                            beta * beta * (se2 / s - (se / s) * (se / s)) * n3,
                            beta * (sm2 / s - (sm / s) * (sm / s)) * n3};
    double var[6] = {0, 0, 0, 0, 0, 0};
    for (int b = 0; b < n_blocks; ++b) {
        // leave block b out, summing the remaining samples in index order
        const int64_t lo = b * bs, hi = lo + bs;
        double be = 0, be2 = 0, bm = 0, bm2 = 0, bm4 = 0;
        for (int64_t i = 0; i < used; ++i) {
            if (i >= lo && i < hi) continue;
            const double q = m[i] * m[i];
            be += e[i];
            be2 += e[i] * e[i];
            bm += m[i];
            bm2 += q;
            bm4 += q * q;
        }
        const double sj = (double)(used - bs);
        const double je = be / sj, je2 = be2 / sj, jm = bm / sj, jm2 = bm2 / sj, jm4 = bm4 / sj;
        const double jk[6] = {je, jm, jm2, jm4, beta * beta * (je2 - je * je) * n3, beta * (jm2 - jm * jm) * n3};
        for (int k = 0; k < 6; ++k) var[k] += (jk[k] - full[k]) * (jk[k] - full[k]);
    }
    const double nb = (double)n_blocks;
    for (int k = 0; k < 6; ++k) {
        out[2 * k] = full[k];
        out[2 * k + 1] = std::sqrt(var[k] * (nb - 1.0) / nb);
    }
    return ISING_OK;
}

int ising_tau_int(const double* x, int64_t n, double c, double* tau, int64_t* window) {
    if (!x || !tau || n < 2 || !(c > 0.0)) return fail(ISING_E_ARG, "need n >= 2 samples and c > 0");
    double mean = 0.0;
    for (int64_t i = 0; i < n; ++i) mean += x[i];
    mean /= (double)n;
    double c0 = 0.0;
    for (int64_t i = 0; i < n; ++i) c0 += (x[i] - mean) * (x[i] - mean);
    c0 /= (double)n;
    if (!(c0 > 0.0)) {
        *tau = 0.5;
        if (window) *window = 0;
        return ISING_OK;
    }
    double t = 0.5;
    int64_t W = 0;
    for (int64_t lag = 1; lag < n; ++lag) {
        double ct = 0.0;
        for (int64_t i = 0; i + lag < n; ++i) ct += (x[i] - mean) * (x[i + lag] - mean);
        ct /= (double)n;
        t += ct / c0;
        W = lag;
        if ((double)lag >= c * t) break;
    }
    *tau = t;
    if (window) *window = W;
    return ISING_OK;
}

// Not-a-knot cubic spline through (x[i], y[i]), x strictly increasing, n >= 4:
// second derivatives m[] from the n x n system (interior continuity of the
// first derivative, third-derivative continuity at x[1] and x[n-2]), solved by
// Gaussian elimination with partial pivoting (n is a temperature count).
static bool spline_not_a_knot(const std::vector<double>& x, const std::vector<double>& y, std::vector<double>* m) {
    const int n = (int)x.size();
    std::vector<double> A((size_t)n * n, 0.0), b(n, 0.0);
    auto a = [&](int r, int c) -> double& { return A[(size_t)r * n + c]; };
    const double h0 = x[1] - x[0], h1 = x[2] - x[1];
    a(0, 0) = h1;
    a(0, 1) = -(h0 + h1);
    a(0, 2) = h0;
    for (int i = 1; i < n - 1; ++i) {
        const double hl = x[i] - x[i - 1], hr = x[i + 1] - x[i];
        a(i, i - 1) = hl;
        a(i, i) = 2.0 * (hl + hr);
        a(i, i + 1) = hr;
        b[i] = 6.0 * ((y[i + 1] - y[i]) / hr - (y[i] - y[i - 1]) / hl);
    }
    const double ha = x[n - 2] - x[n - 3], hb = x[n - 1] - x[n - 2];
    a(n - 1, n - 3) = hb;
    a(n - 1, n - 2) = -(ha + hb);
    a(n - 1, n - 1) = ha;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(a(r, c)) > std::fabs(a(piv, c))) piv = r;
        if (!(std::fabs(a(piv, c)) > 0.0)) return false;
        if (piv != c) {
            for (int k = 0; k < n; ++k) std::swap(a(c, k), a(piv, k));
            std::swap(b[c], b[piv]);
        }
        for (int r = c + 1; r < n; ++r) {
            const double f = a(r, c) / a(c, c);
            if (f == 0.0) continue;
            for (int k = c; k < n; ++k) a(r, k) -= f * a(c, k);
            b[r] -= f * b[c];
        }
    }
    m->assign(n, 0.0);
    for (int r = n - 1; r >= 0; --r) {
        double v = b[r];
        for (int k = r + 1; k < n; ++k) v -= a(r, k) * (*m)[k];
        (*m)[r] = v / a(r, r);
    }
    return true;
}

static double spline_eval(const std::vector<double>& x, const std::vector<double>& y, const std::vector<double>& m,
                          double t) {
    const int n = (int)x.size();
    int i = (int)(std::upper_bound(x.begin(), x.end(), t) - x.begin()) - 1;
    i = std::max(0, std::min(n - 2, i));
    const double h = x[i + 1] - x[i], A = x[i + 1] - t, B = t - x[i];
    return m[i] * A * A * A / (6.0 * h) + m[i + 1] * B * B * B / (6.0 * h) + (y[i] / h - m[i] * h / 6.0) * A +
           (y[i + 1] / h - m[i + 1] * h / 6.0) * B;
}

static bool sorted_curve(const double* t, const double* u, int32_t n, std::vector<double>* x, std::vector<double>* y) {
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int p, int q) { return t[p] < t[q]; });
    x->resize(n);
    y->resize(n);
    for (int i = 0; i < n; ++i) {
        (*x)[i] = t[idx[i]];
        (*y)[i] = u[idx[i]];
        if (!std::isfinite((*x)[i]) || !std::isfinite((*y)[i])) return false;
        if (i > 0 && !((*x)[i] > (*x)[i - 1])) return false;
    }
    return true;
}

int ising_binder_crossing(const double* t1, const double* u1, int32_t n1, const double* t2, const double* u2,
                          int32_t n2, double* tc, int32_t* found) {
    if (!t1 || !u1 || !t2 || !u2 || !tc || !found) return fail(ISING_E_ARG, "null argument");
    if (n1 < 4 || n2 < 4) return fail(ISING_E_ARG, "each curve needs at least 4 temperatures");
    std::vector<double> x1, y1, x2, y2, m1, m2;
    if (!sorted_curve(t1, u1, n1, &x1, &y1) || !sorted_curve(t2, u2, n2, &x2, &y2))
        return fail(ISING_E_ARG, "temperatures must be finite and distinct, cumulants finite");
    if (!spline_not_a_knot(x1, y1, &m1) || !spline_not_a_knot(x2, y2, &m2))
        return fail(ISING_E_ARG, "singular spline system");
    *found = 0;
    const double lo = std::max(x1.front(), x2.front()), hi = std::min(x1.back(), x2.back());
    if (!(hi > lo)) return ISING_OK;
    constexpr int kGrid = 2000;
    double tp = lo, dp = spline_eval(x1, y1, m1, lo) - spline_eval(x2, y2, m2, lo);
    for (int j = 1; j < kGrid; ++j) {
        const double t = lo + (hi - lo) * (double)j / (double)(kGrid - 1);
        const double d = spline_eval(x1, y1, m1, t) - spline_eval(x2, y2, m2, t);
        if (dp * d < 0.0) {
            *tc = tp - dp * (t - tp) / (d - dp);
            *found = 1;
            return ISING_OK;
        }
        tp = t;
        dp = d;
    }
    return ISING_OK;
}

int ising_wham(int32_t K, const int64_t* n_k, const double* betas, const double* energies, int32_t n_obs,
               const double* obs, int32_t max_iter, double tol, double* f_out, int32_t* iters, int32_t n_targets,
               const double* beta_targets, double* out) {
    if (K < 1 || K > 1024 || !n_k || !betas || !energies || n_obs < 0 || n_obs > 8 || (n_obs > 0 && !obs) ||
        max_iter < 1 || n_targets < 0 || (n_targets > 0 && (!beta_targets || !out)))
        return fail(ISING_E_ARG, "ising_wham: invalid arguments");
    int64_t N = 0;
    std::vector<double> logN(K);
    for (int k = 0; k < K; ++k) {
        if (n_k[k] < 1) return fail(ISING_E_ARG, "every run needs at least one sample");
        N += n_k[k];
        logN[k] = std::log((double)n_k[k]);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(ISING_E_HIP, "no HIP device visible");
    hipStream_t st = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    const int nchunk = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (N + 256 * 16 - 1) / (256 * 16)));
    const int maxT = std::max<int>(K, n_targets);
    double *dE = nullptr, *dO = nullptr, *dlogN = nullptr, *dF = nullptr, *dB = nullptr, *dD = nullptr,
           *dT = nullptr, *dP = nullptr, *dOut = nullptr;
    hipError_t e = hipSuccess;
    auto M = [&](double** p, size_t n) {
        if (e == hipSuccess) e = hipMalloc(p, sizeof(double) * std::max<size_t>(n, 1));
    };
    M(&dE, N);
    M(&dO, (size_t)N * n_obs);
    M(&dlogN, K);
    M(&dF, K);
    M(&dB, K);
    M(&dD, N);
    M(&dT, maxT);
    M(&dP, (size_t)maxT * nchunk * 10);
    M(&dOut, (size_t)maxT * (1 + n_obs));
    std::vector<double> f(K, 0.0), res((size_t)maxT * (1 + n_obs));
    int it = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(dE, energies, sizeof(double) * N, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && n_obs) e = hipMemcpyAsync(dO, obs, sizeof(double) * N * n_obs, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dlogN, logN.data(), sizeof(double) * K, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dB, betas, sizeof(double) * K, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dT, betas, sizeof(double) * K, hipMemcpyHostToDevice, st);
    // free-energy iteration (reweighting.py:77-99)
    while (e == hipSuccess && it < max_iter) {
        ++it;
        e = hipMemcpyAsync(dF, f.data(), sizeof(double) * K, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = launch_wham_denom(dE, N, K, dlogN, dF, dB, dD, st);
        if (e == hipSuccess) e = launch_wham_reduce(dE, N, dD, dT, K, 0, dO, dP, nchunk, dOut, st);
        if (e == hipSuccess) e = hipMemcpyAsync(res.data(), dOut, sizeof(double) * K, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) break;
        std::vector<double> fn(K);
        for (int k = 0; k < K; ++k) fn[k] = -res[k];
        const double f0 = fn[0];
        double diff = 0.0;
        for (int k = 0; k < K; ++k) {
            fn[k] -= f0;
            diff = std::max(diff, std::fabs(fn[k] - f[k]));
        }
        f = fn;
        if (diff < tol) break;
    }
    // the weights use log D_n of the last iteration, computed from the free
    // energies before its update, as reweighting.py:101 does (dD holds it);
    // with no iteration at all, D_n of the initial f = 0
    if (e == hipSuccess && it == 0) {
        e = hipMemcpyAsync(dF, f.data(), sizeof(double) * K, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = launch_wham_denom(dE, N, K, dlogN, dF, dB, dD, st);
    }
    if (e == hipSuccess && n_targets > 0) {
        e = hipMemcpyAsync(dT, beta_targets, sizeof(double) * n_targets, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = launch_wham_reduce(dE, N, dD, dT, n_targets, n_obs, dO, dP, nchunk, dOut, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(res.data(), dOut, sizeof(double) * n_targets * (1 + n_obs), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e == hipSuccess)
            for (int t = 0; t < n_targets; ++t)
                for (int j = 0; j < n_obs; ++j) out[(size_t)t * n_obs + j] = res[(size_t)t * (1 + n_obs) + 1 + j];
    }
    for (double* p : {dE, dO, dlogN, dF, dB, dD, dT, dP, dOut}) (void)hipFree(p);
    (void)hipStreamDestroy(st);
    HIP_TRY(e);
    if (f_out)
        for (int k = 0; k < K; ++k) f_out[k] = f[k];
    if (iters) *iters = it;
    return ISING_OK;
}

}  // extern "C"
