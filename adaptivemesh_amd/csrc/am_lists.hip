// Device-side compaction of the octree's launch lists for gfx950.
//
// Every kernel of the path launches over a list of slot indices (the
// reference launches over every slot and lets inactive ones return,
// AMR.h:640-644).  After the octree changes, the lists are rebuilt here from
// the device copies of positions, children and boundaries, by stream
// compaction: each slot contributes a count to each list (0 or 1 to the
// tile lists, 0-6 faces to its depth's coarse-face list), and its entries
// land at the exclusive prefix sum of those counts.  The scan is the
// reference's approach for GPU buffers (parallel.h:10-80: block-wise scans
// of 2^9 elements, then the block totals), written as three passes:
//   k_list_counts   per-block totals of every list (and coarse_source
//                   cleared);
//   k_list_offsets  one workgroup: exclusive scan of the block totals per
//                   list, the lists' offsets and counts;
//   k_list_scatter  per block: the in-block exclusive scan of every list,
//                   then each slot writes its entries (and marks the slots
//                   its coarse faces read in coarse_source).
// Entries keep ascending slot order (and face order within a slot), so the
// lists equal a sequential loop over the slots.
#include "am_common.hpp"
#include "am_device.hpp"

namespace am {

constexpr int LIST_THREADS = 256;

struct ListArgs {
    const uint32_t *positions;
    const int32_t *children;
    const uint64_t *boundaries;
    int32_t n_slots;      // slots listed: 0 .. n_slots - 1
    int32_t clear_slots;  // coarse_source entries cleared: 0 .. clear_slots - 1
    int32_t depth_limit;
    int32_t *tiles;       // [all][depth 0]..[depth depth_limit][leaves]
    int32_t *faces;       // [depth 0]..[depth depth_limit]: slot * 6 + face with a coarser target
    uint8_t *coarse_source;
    int32_t *block_totals;  // n_blocks x n_lists, then exclusive block offsets in place
    int32_t *counts;        // am_tile_list_counts
};

// lists: tile list L in [0, depth_limit + 3), face list of depth d at
// depth_limit + 3 + d
__device__ __forceinline__ int n_lists(int depth_limit) { return 2 * depth_limit + 4; }

// The counts of slot i: its depth's tile list, the all-slot list, the leaf
// list and its depth's face list (the only lists it can enter).
struct SlotCounts {
    int depth;       // -1: unused
    bool leaf;
    int face_mask;   // bit f: face f's target is a coarser tile
};
__device__ __forceinline__ SlotCounts slot_counts(const ListArgs &a, int i)
{
    SlotCounts c{-1, false, 0};
    if (i >= a.n_slots) return c;
    c.depth = pos_depth(a.positions[i]);
    if (c.depth < 0) return c;
    bool leaf = true;
#pragma unroll
    for (int k = 0; k < 8; k++) leaf &= a.children[(size_t)i * 8 + k] == -1;
    c.leaf = leaf;
    if (c.depth > 0) {
#pragma unroll
        for (int f = 0; f < 6; f++) {
            const uint32_t nd = (uint32_t)a.boundaries[(size_t)i * 6 + f];
            if (bnd_target(nd) >= 0 && bnd_ddiff(nd) > 0) c.face_mask |= 1 << f;
        }
    }
    return c;
}
__device__ __forceinline__ int count_in(const SlotCounts &c, int i, int n_slots, int L, int depth_limit)
{
    if (L == 0) return i < n_slots ? 1 : 0;
    if (L <= depth_limit + 1) return c.depth == L - 1 ? 1 : 0;
    if (L == depth_limit + 2) return c.leaf ? 1 : 0;
    return c.depth == L - (depth_limit + 3) ? __popc((unsigned)c.face_mask) : 0;
}

// block-wide exclusive scan of v (LDS), returning the block total in *total
__device__ __forceinline__ int block_exclusive_scan(int v, int *buf, int *total)
{
    const int t = (int)threadIdx.x;
    buf[t] = v;
    __syncthreads();
    for (int off = 1; off < LIST_THREADS; off <<= 1) {
        const int add = t >= off ? buf[t - off] : 0;
        __syncthreads();
        buf[t] += add;
        __syncthreads();
    }
    const int incl = buf[t];
    *total = buf[LIST_THREADS - 1];
    __syncthreads();  // buf is reused by the next list
    return incl - v;
}

__global__ void __launch_bounds__(LIST_THREADS) k_list_counts(ListArgs a)
{
    __shared__ int buf[LIST_THREADS];
    const int i = (int)(blockIdx.x * LIST_THREADS + threadIdx.x);
    if (i < a.clear_slots) a.coarse_source[i] = 0;
    const SlotCounts c = slot_counts(a, i);
    const int NL = n_lists(a.depth_limit);
    for (int L = 0; L < NL; L++) {
        int total;
        block_exclusive_scan(count_in(c, i, a.n_slots, L, a.depth_limit), buf, &total);
        if (threadIdx.x == 0) a.block_totals[(size_t)blockIdx.x * NL + L] = total;
    }
}

// one workgroup; thread L scans list L's block totals in block order
__global__ void __launch_bounds__(64) k_list_offsets(ListArgs a, int32_t n_blocks)
{
    const int NL = n_lists(a.depth_limit), L = (int)threadIdx.x;
    __shared__ int tot[64];
    if (L < NL) {
        int run = 0;
        for (int b = 0; b < n_blocks; b++) {
            int *p = a.block_totals + (size_t)b * NL + L;
            const int t = *p;
            *p = run;
            run += t;
        }
        tot[L] = run;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // counts: list_offset[depth_limit + 3], list_count[depth_limit + 3],
        // face_offset[depth_limit + 1], face_count[depth_limit + 1]
        const int NT = a.depth_limit + 3, NF = a.depth_limit + 1;
        int off = 0;
        for (int k = 0; k < NT; k++) {
            a.counts[k] = off;
            a.counts[NT + k] = tot[k];
            off += tot[k];
        }
        off = 0;
        for (int d = 0; d < NF; d++) {
            a.counts[2 * NT + d] = off;
            a.counts[2 * NT + NF + d] = tot[NT + d];
            off += tot[NT + d];
        }
    }
}

__global__ void __launch_bounds__(LIST_THREADS) k_list_scatter(ListArgs a)
{
    __shared__ int buf[LIST_THREADS];
    const int i = (int)(blockIdx.x * LIST_THREADS + threadIdx.x);
    const SlotCounts c = slot_counts(a, i);
    const int NL = n_lists(a.depth_limit), NT = a.depth_limit + 3;
    for (int L = 0; L < NL; L++) {
        int total;
        const int n = count_in(c, i, a.n_slots, L, a.depth_limit);
        const int rank = block_exclusive_scan(n, buf, &total);
        if (n == 0) continue;
        const int base = a.block_totals[(size_t)blockIdx.x * NL + L] + rank;
        if (L < NT) {
            a.tiles[a.counts[L] + base] = i;
        } else {
            int k = a.counts[2 * NT + (L - NT)] + base;
            for (int f = 0; f < 6; f++) {
                if (!((c.face_mask >> f) & 1)) continue;
                a.faces[k++] = i * 6 + f;
                // several faces may mark one slot: they all store 1
                a.coarse_source[bnd_target((uint32_t)a.boundaries[(size_t)i * 6 + f])] = 1;
            }
        }
    }
}

}  // namespace am

extern "C" int32_t am_tile_lists_workspace(int32_t max_slots, int32_t depth_limit)
{
    if (max_slots < 0 || depth_limit < 0 || depth_limit > 29) return -1;
    const int32_t n_blocks = (max_slots + am::LIST_THREADS - 1) / am::LIST_THREADS;
    return n_blocks * (2 * depth_limit + 4);
}

extern "C" int am_tile_lists(const uint32_t *positions, const int32_t *children, const uint64_t *boundaries,
                             int32_t n_slots, int32_t max_slots, int32_t depth_limit, int32_t *tiles, int32_t *faces,
                             uint8_t *coarse_source, int32_t *workspace, int32_t *counts, void *stream)
{
    using namespace am;
    if (!positions || !children || !boundaries || !tiles || !faces || !coarse_source || !workspace || !counts)
        return set_error(AM_ERR_INVALID, "am_tile_lists: null argument");
    if (n_slots < 0 || max_slots < n_slots || depth_limit < 0 || depth_limit > 29)
        return set_error(AM_ERR_INVALID, "am_tile_lists: n_slots %d, max_slots %d, depth_limit %d", n_slots,
                         max_slots, depth_limit);
    if (max_slots == 0) return AM_OK;
    ListArgs a{positions, children, boundaries, n_slots, max_slots, depth_limit, tiles, faces, coarse_source,
               workspace, counts};
    const int32_t n_blocks = (max_slots + LIST_THREADS - 1) / LIST_THREADS;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_list_counts, dim3(n_blocks), dim3(LIST_THREADS), 0, st, a);
    hipLaunchKernelGGL(k_list_offsets, dim3(1), dim3(64), 0, st, a, n_blocks);
    hipLaunchKernelGGL(k_list_scatter, dim3(n_blocks), dim3(LIST_THREADS), 0, st, a);
    return check_launch("k_list");
}
