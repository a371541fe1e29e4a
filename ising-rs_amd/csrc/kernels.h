// Kernel argument blocks and launchers shared by the kernels and the host API.
#pragma once
#include "common.h"

namespace ising {

constexpr int kArgGroups = 12;  // size descriptors carried in the kernel arguments
struct IsingArgs {
    const DevGroup* groups;
    int32_t ngroups;
    int32_t total_reps;
    int32_t nblocks;      // work blocks of a half-sweep (the grid is padded to a multiple of 8)
    uint8_t* pool;
    const uint32_t* thr;  // [rep0 + tidx][32]: 14 high parts, 14 low parts, 4 packed high bytes
    int32_t* r2t;         // [2][global replica] -> temperature index, by permutation parity
    int32_t* t2r;         // [2][global temperature slot] -> replica
    const double* betas;  // [rep0 + tidx]
    DevCtl* ctl;
    int64_t* acc;         // [2 sweep parities][acc_copies][global replica][2]: bond, mag
    int64_t acc_stride;   // int64s per sweep parity
    IsingRec* rec;
    uint32_t k0, k1;
    double J;
    int32_t inline_thr;           // 1: thresholds come from thr_inline (single-lattice steps)
    uint32_t thr_inline[32];
    const uint32_t* mthr;         // multispin coding: [rep0 + tidx][3] thresholds for a = 4, 5, 6
    uint32_t* mplanes;            // multispin coding: per word, 3 x 32 threshold bit planes of its replicas
    int32_t acc_copies;           // accumulator copies (spread atomics), summed by the closing kernel
    const uint32_t* sched;        // annealing: one table row per sweep for every replica (null: per temperature)
    int64_t sched_t0;             // sweep index of schedule row 0
    // Thresholds by replica: thr_rep[global replica] = thr[rep0 + its temperature index]
    // under the permutation in force, rewritten wherever an exchange writes a new
    // permutation; lets a launch that closes no exchange read its table without
    // first reading the permutation (null: read through the permutation).
    uint32_t* thr_rep;
    // The size descriptors again, in the kernel arguments (gk_on when ngroups <= kArgGroups):
    // a block finds its size and tile without a dependent global load.
    int32_t gk_on;
    int32_t fnblocks;  // one-launch sweeps: boxes over all sizes and replicas
    uint8_t* pool_out; // one-launch sweeps: lattice buffer written (pool is read)
    DevGroup gk[kArgGroups];
};

constexpr int kMscAccCopies = 16;
constexpr int kIsingAccCopies = 2;  // int8 Ising accumulator copies (a power of two, <= kMscAccCopies; 2 measured fastest of 1, 2, 4, 8)
constexpr int kMscTile = 256;  // site words per block in multispin-coded launches
hipError_t launch_msc_half(const IsingArgs& a, int nblocks, int color, hipStream_t st);
hipError_t launch_msc_init(const DevGroup* groups, int ngroups, uint32_t* pool, int nblocks, uint32_t k0, uint32_t k1,
                           hipStream_t st);
hipError_t launch_msc_measure(const DevGroup* groups, int ngroups, const uint32_t* pool, int total_reps, int nblocks,
                              int64_t* acc, hipStream_t st);

constexpr int kMhSlots = 64;  // counter copies, so that waves do not contend on one address
constexpr int kCntSlots = 64;  // acceptance counter copies per replica (same reason), summed at adaptation
struct ContArgs {
    const DevGroup* groups;
    int32_t ngroups;
    int32_t total_reps;
    int32_t ncomp;         // 2 = XY, 3 = Heisenberg
    uint32_t model;        // MODEL_XY or MODEL_HEIS (random-stream tag)
    float* pool;
    float* deltas;         // [rep0 + tidx] proposal size (radians), adapted during warm-up
    float* omcd;           // [rep0 + tidx] 1 - cos(delta) of the FP32 delta, rounded from FP64 (Heisenberg cap)
    const float* betas_f;  // [rep0 + tidx] beta for the Metropolis test
    const double* betas;   // [rep0 + tidx] beta for the exchange test
    int32_t* r2t;
    int32_t* t2r;
    DevCtl* ctl;
    double* part0;         // [block][4] colour-0 partials (M)
    double* part1;         // [block][4] colour-1 partials (E, M)
    double* dacc;          // [global replica][4] E, Mx, My, Mz of the last measured sweep
    int64_t* acc_cnt;      // [global replica][kCntSlots] accepted Metropolis moves (warm-up adaptation)
    ContRec* rec;
    uint32_t k0, k1;
    double J;
    float Jf;
    int32_t renorm_every;
    int32_t adapt_every;   // 0: proposal size fixed
    float adapt_target;    // target acceptance rate
    int32_t inline_params; // 1: beta/delta from the two fields below (single-lattice passes)
    float beta_inline, delta_inline, omcd_inline;
    const uint16_t* net;   // learned proposal: BF16 parameters (null when off)
    const uint4* lfrag;    // the same network as per-lane matrix-core fragments (learn_frag_bytes(H))
    int64_t* mh;           // learned proposal: [kMhSlots][8] moves tried (0) / accepted (1), spread by wave
    int32_t learn_H;       // hidden width of the learned proposal's MLP
    int32_t or_spt;        // sites per thread of the overrelaxation pass (1 or 2; 0: the model's default)
    int32_t off32;         // 1: every group's half-lattice stride < 2^30 (32-bit byte offsets in the passes)
};

int learn_params(int H);  // parameter count of the MLP with hidden width H (8, 16, 32, 64), -1 otherwise
constexpr int kLearnBlock = 64;  // samples per training block
hipError_t launch_cont_learned(const ContArgs& a, int nblocks, int color, int closes, int64_t t_override,
                               hipStream_t st);
hipError_t launch_cont_heatbath(const ContArgs& a, int nblocks, int color, int closes, int64_t t_override,
                                hipStream_t st);
hipError_t launch_learn_grad(const ContArgs& a, int nblk, int64_t step, int mode, const float* xin, const float* cin,
                             int64_t n_in, float* gpart, double* lpart, hipStream_t st);
hipError_t launch_learn_adam(float* master, uint16_t* net, float* m1, float* m2, const float* gpart, int nblk, int P,
                             float inv_n, float lr, int64_t step, float* gout, hipStream_t st);
hipError_t launch_learn_pack(const float* master, uint16_t* net, int P, hipStream_t st);
hipError_t launch_learn_infer(const uint4* frag, int H, const float* x, int64_t n, float* out, hipStream_t st);
size_t learn_frag_bytes(int H);  // bytes of the fragment buffer of hidden width H (0 if unsupported)
hipError_t launch_learn_frag(const uint16_t* net, int H, uint4* frag, hipStream_t st);  // fragments from BF16 parameters

// Work-list launch of several ensembles (batch.hip): the work of one pass of
// every member, as segments of a list of blocks.
enum WorkKind : int32_t { W_ISING = 0, W_OR = 1, W_METRO = 2, W_CLOSE = 3 };
struct WorkSeg {
    int32_t kind;    // WorkKind
    int32_t member;  // continuous member (index into BatchPass::ca) of W_OR, W_METRO, W_CLOSE
    int32_t count;   // work items: the blocks of the member's own launch of this pass
    int32_t chunk;   // items of the segment per XCD, ceil(count / 8) (set by launch_batch_pass)
    int32_t kfirst;  // position of the segment's chunk in every XCD's list (set by launch_batch_pass)
    int32_t tiles;   // the member's tiles (blocks of its own launch); an XY item holds two tiles
};
constexpr int kBatchSegs = 2;  // segments per launch, all of one kind: the Ising member, or the continuous members
constexpr int kBatchCont = 2;  // continuous members of a work-list batch
struct BatchPass {
    IsingArgs ia;             // the int8 Ising member (when a segment is W_ISING)
    ContArgs ca[kBatchCont];  // the continuous members
    int32_t color;
    int32_t nseg;
    int32_t per_xcd;          // items in every XCD's list (set by launch_batch_pass)
    int32_t pad;
    WorkSeg seg[kBatchSegs];
};
hipError_t launch_batch_pass(BatchPass bp, bool persistent, hipStream_t st);  // persistent: grids the device holds at once
int batch_pass_resident_blocks();  // least blocks of a persistent batch kernel the current device holds at once (0 on error)

hipError_t launch_ising_close(const IsingArgs& a, hipStream_t st);
hipError_t launch_cont_metro(const ContArgs& a, int nblocks, int color, int closes, hipStream_t st);
hipError_t launch_cont_close(const ContArgs& a, hipStream_t st);
hipError_t launch_cont_or(const ContArgs& a, int nblocks, int color, hipStream_t st);
hipError_t launch_cont_init(const DevGroup* groups, int ngroups, float* pool, int ncomp, int model, int nblocks,
                            uint32_t k0, uint32_t k1, hipStream_t st);
hipError_t launch_cont_helicity(const DevGroup* groups, int ngroups, const float* pool, double* part, int nblocks,
                                hipStream_t st);
hipError_t launch_cont_measure(const DevGroup* groups, int ngroups, const float* pool, int ncomp, double J,
                               double* part, int nblocks, hipStream_t st);
constexpr int kContTile = 256;  // half-sites per block in continuous launches
constexpr int kContClosePerBlock = 4;  // replicas per block of a continuous closing launch (a wave each)

hipError_t launch_wham_denom(const double* E, int64_t n, int K, const double* logN, const double* f,
                             const double* beta, double* logD, hipStream_t st);
hipError_t launch_wham_reduce(const double* E, int64_t n, const double* logD, const double* tb, int n_targets,
                              int n_obs, const double* obs, double* part, int nchunk, double* out, hipStream_t st);

hipError_t launch_ising_half(const IsingArgs& a, int nblocks, int color, hipStream_t st);
hipError_t launch_ising_fused(const IsingArgs& a, int lds_bytes, hipStream_t st);
hipError_t launch_ising_init(const DevGroup* groups, int ngroups, uint8_t* pool, int nblocks, uint32_t k0,
                             uint32_t k1, hipStream_t st);
hipError_t launch_ising_measure(const DevGroup* groups, int ngroups, const uint8_t* pool, int64_t* acc,
                                int nblocks, hipStream_t st);
// Cluster updates (Swendsen-Wang / Wolff).  One block per (replica, tile);
// tiles are 32 x 8 x 8 sites (3D) or 64 x 32 (2D), clipped to L.
constexpr int kClusterTile = 2048;  // sites per tile
constexpr int kClusterFace = 256;   // largest + face of a tile
struct ClusterGroup {
    int32_t L, dims, R, rep0;
    int32_t tx, ty, tz;       // tile extents
    int32_t ntx, nty, ntz;    // tiles per axis
    int32_t tiles, blk0;      // tiles per replica, first block of this size
    int64_t nsites;           // L^dims
    int64_t lab0;             // element of replica 0 in the label and bond arrays
    int32_t q4;               // ceil(L / 4): Philox groups per lattice row
    int32_t pad;
    // n / d == fastdiv(n, mag, sh) for d = L, and for the x / y extents of a
    // full tile ([0]) and of the last tile along the axis ([1])
    uint32_t mag_L, sh_L;
    uint32_t mag_x[2], sh_x[2];
    uint32_t mag_y[2], sh_y[2];
};
struct ClusterArgs {
    const DevGroup* groups;
    const ClusterGroup* cg;
    int32_t ngroups, total_reps;
    uint8_t* ipool;          // Ising pool (0/1 bytes), or null
    float* fpool;            // continuous pool, or null
    int32_t ncomp;           // 0 Ising, 2 XY, 3 Heisenberg
    uint32_t model;          // random-stream tag of continuous models
    int32_t kind;            // 0 Swendsen-Wang, 1 Wolff
    uint32_t u;              // update index (Philox c1)
    uint32_t k0, k1;
    int32_t* label;          // [lab0 + replica * nsites + site]
    uint8_t* bond;           // same indexing: bits 0-2 bonds +x/+y/+z, bit 3 coin
    const int32_t* r2t;      // [2][total_reps]
    const DevCtl* ctl;       // permutation parity = round & 1 (Ising); null: parity 0
    const uint32_t* bthr;    // Ising: [rep0 + tidx] floor(2^32 (1 - exp(-2 beta |J|)))
    const double* betas;     // continuous: [rep0 + tidx]
    double J;
    int32_t inline_p;        // single-lattice: bthr_inline / beta_inline
    uint32_t bthr_inline;
    double beta_inline;
    int64_t* flipped;        // [global replica] sites flipped, accumulated (null: not counted)
};
hipError_t launch_cluster_update(const ClusterArgs& a, int nblocks, hipStream_t st);
hipError_t launch_cont_fold(const DevGroup* groups, int ngroups, int total_reps, const double* part, double* out,
                            hipStream_t st);

// Reductions of a caller's lattice in the reference's layout (reduce.hip):
// partials per 4096-site tile into `part` (2 int64 / 4 doubles per tile), then
// one block sums them in a fixed order into `out` (2 int64: bond sum, spin sum /
// 4 doubles: bond dot sum, Sx, Sy, Sz).  n <= 1024 (n^3 < 2^32), n_sites < 2^32.
// Sets the calling thread's ising_last_error() message; returns code.
int set_error(int code, const char* msg);

constexpr int kReduceTile = 4096;
inline int reduce_blocks(int64_t sites) { return (int)((sites + kReduceTile - 1) / kReduceTile); }
hipError_t launch_reduce_ising(const int8_t* spins, int32_t n, int64_t* part, int64_t* out, hipStream_t st);
hipError_t launch_reduce_cont(const float* spins, int64_t n_sites, int32_t n, int32_t ncomp, double* part,
                              double* out, hipStream_t st);

constexpr int kIsingTile = 4096;  // half-sites per block step in Ising launches (256 threads x 16)
constexpr int kIsingTab = 32;     // acceptance-table words per temperature

}  // namespace ising
