// Kernel argument blocks and launchers shared by the kernels and the host API.
#pragma once
#include "common.h"

namespace ising {

struct IsingArgs {
    const DevGroup* groups;
    int32_t ngroups;
    int32_t total_reps;
    uint8_t* pool;
    const uint32_t* thr;  // [rep0 + tidx][28]: 14 high halves, then 14 low halves
    int32_t* r2t;         // replica -> temperature index (per group, global replica index)
    int32_t* t2r;         // temperature slot -> replica
    const double* betas;  // [rep0 + tidx]
    DevCtl* ctl;
    int64_t* acc;         // [global replica][2]: bond, mag
    IsingRec* rec;
    uint32_t k0, k1;
    double J;
    int32_t inline_thr;           // 1: thresholds come from thr_inline (single-lattice steps)
    uint32_t thr_inline[28];
};

hipError_t launch_ising_half(const IsingArgs& a, int nblocks, int color, hipStream_t st);
hipError_t launch_ising_init(const DevGroup* groups, int ngroups, uint8_t* pool, int nblocks, uint32_t k0,
                             uint32_t k1, hipStream_t st);
hipError_t launch_ising_measure(const DevGroup* groups, int ngroups, const uint8_t* pool, int64_t* acc,
                                int nblocks, hipStream_t st);
constexpr int kIsingTile = 2048;  // half-sites per block in Ising launches

}  // namespace ising
