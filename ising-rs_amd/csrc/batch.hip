// Passes of several ensembles of different models launched together over work
// lists of their blocks (no counterpart in the reference; it replaces the
// per-model launches of a batched parallel-tempering run, config 4).
//
// A sweep of a batch holding an int8 Ising ensemble and up to two continuous
// ensembles (XY, Heisenberg) with n_or overrelaxation passes is 2 n_or + 5
// launches, each over a work list of the members' blocks for that pass:
//
//   n_or x (OR colour 0, OR colour 1)   continuous members' overrelaxation tiles (batch_or_kernel)
//   Ising colour 0                      Ising colour-0 tiles                (batch_pass_kernel<0, 0, true>)
//   Metropolis colour 0                 continuous members' Metropolis tiles (batch_pass_kernel<N0, N1, false>)
//   Ising colour 1, Metropolis colour 1 the same for colour 1
//   close                               continuous members' closing blocks  (batch_close_kernel)
//
// instead of 2 + 2 (2 n_or + 3) launches on one stream per member.  The Ising
// pass has launches of its own: a kernel holding the Ising body has its
// registers (71 VGPRs), which would set the occupancy of every continuous item
// beside it.  The work list of a launch is a short table of segments in the
// kernel arguments, one per member, item j of a segment being block j of that
// member's own launch of the pass.  A block runs the member's own block body
// (ising_pass.h, cont_pass.h), so every result equals the per-member
// launches'.  The bodies read their block's parameters (size descriptor,
// replica tables, sweep counter) with ld_uniform: as the second of two member
// branches, or inside a persistent grid's item loop, the compiler otherwise
// loaded them with vector loads (XY's Metropolis pass 505 against 318 us).
//
// Items and XCDs: blocks are dealt to the 8 XCDs round-robin (block i runs on
// XCD i % 8).  Each segment is cut into 8 contiguous chunks and XCD x takes
// chunk x of every segment, one after the other, so every member's tiles are
// spread over all XCDs (each XCD draws on its own share of the memory
// bandwidth: a member confined to some XCDs ran at their share) and each XCD
// still walks contiguous tiles of a member, whose z-neighbour planes stay in
// its L2.  With one segment this is the per-member launches' block order.
//
// Persistent grids (ISING_BATCH_PERSISTENT): the same launches with at most
// the blocks the device holds at once of each kernel; block k of XCD x takes
// the XCD's items k, k + (the XCD's blocks), ..., with a barrier between
// items, whose LDS the bodies reuse.
#include "cont_pass.h"
#include "ising_pass.h"

namespace ising {

// Work item of an XCD: its position k in the XCD's list -> segment s and the
// block lb of that segment's member (lb >= the segment's count when the XCD's
// chunk of s is shorter than the longest XCD's).  A macro, not a function:
// dynamic indices into the kernel arguments stay scalar loads only in the
// kernel itself.
#define WL_ITEM(bp, x, k, s, lb)                                      \
    int s = 0;                                                          \
    while (s + 1 < (bp).nseg && (k) >= (bp).seg[s + 1].kfirst) ++s;     \
    const int lb = (x) * (bp).seg[s].chunk + ((k) - (bp).seg[s].kfirst)

// The size descriptor of Ising work block lb (a macro for the same reason).
#define WL_ISING_GROUP(bp, lb, G)                                                \
    DevGroup G;                                                                  \
    if ((bp).ia.gk_on) {                                                         \
        int g_ = 0;                                                              \
        while (g_ + 1 < (bp).ia.ngroups && (lb) >= (bp).ia.gk[g_ + 1].blk0) ++g_; \
        G = (bp).ia.gk[g_];                                                      \
    } else {                                                                     \
        G = (bp).ia.groups[find_group((bp).ia.groups, (bp).ia.ngroups, (lb))];   \
    }

// The passes of continuous member m, whose component count NM is a template
// argument: every instance then has its own body.  (With the member chosen at
// run time, the compiler merged the members' identical bodies into one that
// reads the member's arguments through a computed address: generic loads of
// every argument, and Metropolis passes at half speed.)
//
// An XY item is two tiles (2 lb, 2 lb + 1), one per half of the block, with
// the member's own thread shape: two sites per thread in 128-thread halves.
// A Heisenberg item is one tile, one site per thread.
template <int KIND, int NM, bool O32>
__device__ __forceinline__ void cont_item(const ContArgs& a, const int lb, const int color, const int tiles) {
    if (NM == 2) {
        // the half is uniform in a wave; read as a scalar, it keeps the tile and
        // everything derived from it (size, replica, base pointers) in scalar
        // registers, as in the member's own kernel (per lane: 51 VGPRs instead
        // of 34 in the overrelaxation pass, and 24% slower XY sweeps)
        const int q = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 7));
        const int tid = (int)(threadIdx.x & 127), tile = 2 * lb + q;
        if (KIND == W_OR) {
            if (tile < tiles) cont_or_block<2, 2, O32>(a, tile, color, tid);
        } else {
            cont_metro_pair_block<2, O32>(a, tile, color, 1, tile < tiles, q, tid);
        }
    } else if (KIND == W_OR) {
        cont_or_block<NM, 1, O32>(a, lb, color);
    } else {
        cont_metro_block<NM, O32>(a, lb, color, 1);
    }
}

template <int KIND, int N0, int N1, bool O32>
__device__ __forceinline__ void cont_member_item(const BatchPass& bp, const int member, const int lb, const int tiles) {
    if (N0 && member == 0) cont_item<KIND, N0 ? N0 : 2, O32>(bp.ca[0], lb, bp.color, tiles);
    else if (N1) cont_item<KIND, N1 ? N1 : 2, O32>(bp.ca[1], lb, bp.color, tiles);
}

// A Metropolis launch: the Ising member's colour pass (ISING; no continuous
// member) or the continuous members' Metropolis passes (N0, N1: components of
// continuous members 0 and 1, 0 when absent; O32: 32-bit offsets, see
// comp_at).  At most 80 VGPRs, six waves per SIMD (the Ising body's occupancy
// in its own launch); the continuous instances keep their bodies' registers
// (45-50 VGPRs against 71 with the Ising body beside them).
template <int N0, int N1, bool ISING, bool O32>
__global__ __launch_bounds__(kThreads, 6) void batch_pass_kernel(BatchPass bp) {
    const int x = (int)(blockIdx.x & 7), k = (int)(blockIdx.x >> 3);
    WL_ITEM(bp, x, k, s, lb);
    if (lb >= bp.seg[s].count) return;
    if (ISING) {
        WL_ISING_GROUP(bp, lb, G)
        ising_half_block(bp.ia, lb, bp.color, G);
    } else {
        cont_member_item<W_METRO, N0, N1, O32>(bp, bp.seg[s].member, lb, bp.seg[s].tiles);
    }
}

// The same lists walked by a persistent grid (every pass but the close), one
// kernel per kind of pass, holding only that kind's bodies: with every body
// in one kernel it needed 142 VGPRs (3 waves per SIMD) for passes whose own
// kernels use 34-73.  The continuous kernels keep 40-63 VGPRs; the Ising body
// in the item loop needs 128 (four waves per SIMD; capped at 80 it spilled).
template <int N0, int N1, int KIND, bool O32>
__global__ __launch_bounds__(kThreads, KIND == W_ISING ? 4 : 6) void batch_persistent_kernel(BatchPass bp) {
    const int x = (int)(blockIdx.x & 7), slot = (int)(blockIdx.x >> 3), nslot = (int)(gridDim.x >> 3);
#pragma nounroll
    for (int k = slot; k < bp.per_xcd; k += nslot) {
        // a barrier between items whose bodies reuse LDS (the overrelaxation
        // bodies use none: without it their launches took 0.7 of the time)
        if (KIND != W_OR && k != slot) __syncthreads();
        WL_ITEM(bp, x, k, s, lb);
        if (lb >= bp.seg[s].count) continue;
        if (KIND == W_ISING) {
            WL_ISING_GROUP(bp, lb, G)
            ising_half_block(bp.ia, lb, bp.color, G);
        } else {
            cont_member_item<KIND, N0, N1, O32>(bp, bp.seg[s].member, lb, bp.seg[s].tiles);
        }
    }
}

// An overrelaxation launch: a kernel of its own, with the registers (and
// occupancy) of the overrelaxation bodies.
template <int N0, int N1, bool O32>
__global__ __launch_bounds__(kThreads) void batch_or_kernel(BatchPass bp) {
    const int x = (int)(blockIdx.x & 7), k = (int)(blockIdx.x >> 3);
    WL_ITEM(bp, x, k, s, lb);
    if (lb >= bp.seg[s].count) return;
    cont_member_item<W_OR, N0, N1, O32>(bp, bp.seg[s].member, lb, bp.seg[s].tiles);
}

// The continuous members' closing blocks of one sweep: block i of the launch
// is item i of the concatenated segments (the closing step counts blocks, not
// tiles, and has no XCD order).
__global__ __launch_bounds__(kThreads) void batch_close_kernel(BatchPass bp) {
    const int w = (int)blockIdx.x;
    const int s = bp.nseg > 1 && w >= bp.seg[0].count ? 1 : 0;
    const int lb = s ? w - bp.seg[0].count : w;
    if (lb >= bp.seg[s].count) return;
    if (bp.seg[s].member == 0) cont_close_block(bp.ca[0], lb, bp.seg[s].count);
    else cont_close_block(bp.ca[1], lb, bp.seg[s].count);
}

// Blocks of a persistent kernel the device holds at once (0 if unknown).
template <int N0, int N1, int KIND, bool O32>
int persistent_blocks() {
    static int blocks = -1;
    if (blocks < 0) {
        int dev = 0, per_cu = 0;
        hipDeviceProp_t prop;
        blocks = hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
                         hipOccupancyMaxActiveBlocksPerMultiprocessor(
                             &per_cu, batch_persistent_kernel<N0, N1, KIND, O32>, kThreads, 0) == hipSuccess
                     ? per_cu * prop.multiProcessorCount
                     : 0;
    }
    return blocks;
}

template <int N0, int N1, int KIND, bool O32>
void launch_persistent(const BatchPass& bp, hipStream_t st) {
    const int resident = persistent_blocks<N0, N1, KIND, O32>();
    const int slots = resident / 8 > 0 ? resident / 8 : 1;
    hipLaunchKernelGGL((batch_persistent_kernel<N0, N1, KIND, O32>),
                       dim3(8 * (bp.per_xcd < slots ? bp.per_xcd : slots)), dim3(kThreads), 0, st, bp);
}

// A continuous pass (W_OR or W_METRO) of continuous members with N0, N1 components.
template <int N0, int N1, bool O32>
void launch_cont_kind(const BatchPass& bp, int kind, bool persistent, hipStream_t st) {
    if (persistent && kind == W_OR) {
        launch_persistent<N0, N1, W_OR, O32>(bp, st);
    } else if (persistent) {
        launch_persistent<N0, N1, W_METRO, O32>(bp, st);
    } else if (kind == W_OR) {
        hipLaunchKernelGGL((batch_or_kernel<N0, N1, O32>), dim3(8 * bp.per_xcd), dim3(kThreads), 0, st, bp);
    } else {
        hipLaunchKernelGGL((batch_pass_kernel<N0, N1, false, O32>), dim3(8 * bp.per_xcd), dim3(kThreads), 0, st, bp);
    }
}

template <int N0, int N1>
void launch_cont_o32(const BatchPass& bp, int kind, bool persistent, bool o32, hipStream_t st) {
    if (o32) launch_cont_kind<N0, N1, true>(bp, kind, persistent, st);
    else launch_cont_kind<N0, N1, false>(bp, kind, persistent, st);
}

template <int N0>
void launch_cont_n1(const BatchPass& bp, int n1, int kind, bool persistent, bool o32, hipStream_t st) {
    if (n1 == 2) launch_cont_o32<N0, 2>(bp, kind, persistent, o32, st);
    else if (n1 == 3) launch_cont_o32<N0, 3>(bp, kind, persistent, o32, st);
    else launch_cont_o32<N0, 0>(bp, kind, persistent, o32, st);
}

hipError_t launch_batch_pass(BatchPass bp, bool persistent, hipStream_t st) {
    // chunks per XCD and their prefix sums
    bp.per_xcd = 0;
    int total = 0;
    for (int s = 0; s < bp.nseg; ++s) {
        bp.seg[s].chunk = (bp.seg[s].count + 7) / 8;
        bp.seg[s].kfirst = bp.per_xcd;
        bp.per_xcd += bp.seg[s].chunk;
        total += bp.seg[s].count;
    }
    if (total <= 0) return hipSuccess;
    // the segments of a launch share one kind
    const int kind = bp.seg[0].kind;
    for (int s = 1; s < bp.nseg; ++s)
        if (bp.seg[s].kind != kind) return hipErrorInvalidValue;
    if (kind == W_CLOSE) {
        hipLaunchKernelGGL(batch_close_kernel, dim3(total), dim3(kThreads), 0, st, bp);
        return hipGetLastError();
    }
    if (kind == W_ISING) {
        if (persistent) launch_persistent<0, 0, W_ISING, false>(bp, st);
        else hipLaunchKernelGGL((batch_pass_kernel<0, 0, true, false>), dim3(8 * bp.per_xcd), dim3(kThreads), 0, st, bp);
        return hipGetLastError();
    }
    const int n0 = bp.ca[0].ncomp, n1 = bp.ca[1].ncomp;  // 0 for an absent member
    const bool o32 = (n0 == 0 || bp.ca[0].off32) && (n1 == 0 || bp.ca[1].off32);
    if (n0 == 2) launch_cont_n1<2>(bp, n1, kind, persistent, o32, st);
    else if (n0 == 3) launch_cont_n1<3>(bp, n1, kind, persistent, o32, st);
    else launch_cont_n1<0>(bp, n1, kind, persistent, o32, st);
    return hipGetLastError();
}

int batch_pass_resident_blocks() {
    // the smallest of the persistent kernels' capacities (the Ising kernel's)
    const int a = persistent_blocks<0, 0, W_ISING, false>(), b = persistent_blocks<2, 3, W_METRO, true>();
    return a < b ? a : b;
}

}  // namespace ising
