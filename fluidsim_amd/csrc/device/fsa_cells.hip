// fsa_cells.hip -- cell hash and the cell-list build.
//
// The reference keeps per-cell singly linked lists and rebuilds them each
// step by walking the previous lists of the 27 neighbour cells with one
// thread per cell (src/pair.cu:72-207): a latency-bound pointer chase.  Here
// the cell list is a counting sort of the particles by cell id:
//   1. hash+count  one thread per particle: cell id (findPositionCell,
//                  src/pair.cu:8-24) and an atomic arrival rank in its cell
//   2. scan        exclusive prefix sum of the per-cell counts -> start[]
//   3. place       the particle's slot and .dat id at start[c] + rank
//   4. rank+gather inside each cell the final rank is the number of cell
//                  members with a smaller .dat id, which makes the order
//                  deterministic (ascending .dat id) whatever order the
//                  atomics came in; the state is gathered into the other
//                  buffer at start[c] + rank.
// All passes are streaming over HBM except the small per-cell rank loop,
// which reads the cell's ids from L1/L2.
#include "fsa_internal.h"
#include "fsa_math.h"

namespace fsa {

namespace {

__global__ __launch_bounds__(kBlock) void k_fill_i32(int32_t *__restrict__ p, int32_t v, int n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_hash_count(DevParams P, int n, const float4 *__restrict__ pos,
                                                       const int32_t *__restrict__ prev_cell,
                                                       int32_t *__restrict__ key, int32_t *__restrict__ rank,
                                                       int32_t *__restrict__ count, unsigned long long *jumps,
                                                       int count_jumps) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const float4 x = pos[s];
    const int c = find_cell(P, x.x, x.y, x.z);
    key[s] = c;
    // Particles arrive sorted by their previous cell, so the lanes of a wave
    // share few cells: one atomic per distinct cell and wave, the lanes of a
    // cell take consecutive ranks (the order is fixed later by the rank pass).
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned long long todo = __ballot(1);
    int r = 0;
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const int lc = __shfl(c, leader, 64);
        const unsigned long long same = __ballot(c == lc);
        int base = 0;
        if (lane == leader) base = atomicAdd(&count[lc], __popcll(same));
        base = __shfl(base, leader, 64);
        if (c == lc) r = base + __popcll(same & below);
        todo &= ~same;
    }
    rank[s] = r;
    if (count_jumps) {
        // more than one cell in any direction since the last build: the
        // reference's list rebuild would have lost this particle
        const int o = prev_cell[s];
        const int nyz = P.cn[1] * P.cn[2];
        const int ax = c / nyz, bx = o / nyz;
        const int ay = (c - ax * nyz) / P.cn[2], by = (o - bx * nyz) / P.cn[2];
        const int az = c % P.cn[2], bz = o % P.cn[2];
        if (abs(ax - bx) > 1 || abs(ay - by) > 1 || abs(az - bz) > 1) atomicAdd(jumps, 1ull);
    }
}

// inclusive scan over the 64 lanes of a wavefront
__device__ __forceinline__ int wave_inclusive_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// exclusive scan of the per-thread totals of a 256-thread workgroup
__device__ __forceinline__ int block_exclusive_scan(int v, int *wave_tot, int &block_total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int inc = wave_inclusive_scan(v);
    if (lane == 63) wave_tot[wave] = inc;
    __syncthreads();
    int before = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) before += (w < wave) ? wave_tot[w] : 0;
    block_total = wave_tot[0] + wave_tot[1] + wave_tot[2] + wave_tot[3];
    return before + inc - v;
}

// pass 1: per-tile inclusive scan into start[1 + i], tile totals
__global__ __launch_bounds__(kBlock) void k_scan_tiles(const int32_t *__restrict__ count, int32_t *__restrict__ start,
                                                       int32_t *__restrict__ tile_sums, int n) {
    __shared__ int wave_tot[kBlock / 64];
    const int base = blockIdx.x * kScanTile + threadIdx.x * 4;
    // count is padded to a multiple of kScanTile with zeros
    const int4 v = *reinterpret_cast<const int4 *>(count + base);
    const int t0 = v.x, t1 = t0 + v.y, t2 = t1 + v.z, t3 = t2 + v.w;
    int total;
    const int ex = block_exclusive_scan(t3, wave_tot, total);
    const int vals[4] = {ex + t0, ex + t1, ex + t2, ex + t3};
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (base + k < n) start[1 + base + k] = vals[k];
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// pass 2: one workgroup scans the tile totals (exclusive, in place)
__global__ __launch_bounds__(kBlock) void k_scan_tile_sums(int32_t *__restrict__ tile_sums, int ntiles) {
    __shared__ int wave_tot[kBlock / 64];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int b = 0; b < ntiles; b += kBlock) {
        const int i = b + threadIdx.x;
        const int v = i < ntiles ? tile_sums[i] : 0;
        int total;
        const int ex = block_exclusive_scan(v, wave_tot, total);
        const int c = carry;
        if (i < ntiles) tile_sums[i] = c + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + total;
        __syncthreads();
    }
}

// pass 3: add tile offsets; start[0] = 0
__global__ __launch_bounds__(kBlock) void k_scan_add(int32_t *__restrict__ start,
                                                     const int32_t *__restrict__ tile_sums, int n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i == 0) start[0] = 0;
    if (i < n) start[1 + i] += tile_sums[i / kScanTile];
}

__global__ __launch_bounds__(kBlock) void k_place(int n, const int32_t *__restrict__ key,
                                                  const int32_t *__restrict__ rank,
                                                  const int32_t *__restrict__ start, const int32_t *__restrict__ id,
                                                  int32_t *__restrict__ tmp_slot, int32_t *__restrict__ tmp_id) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const int t = start[key[s]] + rank[s];
    tmp_slot[t] = s;
    tmp_id[t] = id[s];
}

template <bool HAS_REF>
__global__ __launch_bounds__(kBlock) void k_rank_gather(int n, const int32_t *__restrict__ key,
                                                        const int32_t *__restrict__ start,
                                                        const int32_t *__restrict__ tmp_slot,
                                                        const int32_t *__restrict__ tmp_id,
                                                        const float4 *__restrict__ pos_in,
                                                        const float4 *__restrict__ vel_in,
                                                        const float4 *__restrict__ ref_in, float4 *__restrict__ pos_out,
                                                        float4 *__restrict__ vel_out, float4 *__restrict__ ref_out,
                                                        int32_t *__restrict__ id_out, int32_t *__restrict__ cell_out) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const int s = tmp_slot[t];
    const int my = tmp_id[t];
    const int c = key[s];
    const int b = start[c], e = start[c + 1];
    int r = 0;
    for (int u = b; u < e; ++u) r += tmp_id[u] < my ? 1 : 0;
    const int d = b + r;
    pos_out[d] = pos_in[s];
    vel_out[d] = vel_in[s];
    if (HAS_REF) ref_out[d] = ref_in[s];
    id_out[d] = my;
    cell_out[d] = c;
}

inline int grid_for(int n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

hipError_t launch_fill_i32(int32_t *p, int32_t v, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n)), dim3(kBlock), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_hash_count(const DevParams &P, KindBuffers &k, unsigned long long *jumps, bool count_jumps,
                             hipStream_t s) {
    const int padded = (P.ncell + kScanTile - 1) / kScanTile * kScanTile;
    hipError_t e = launch_fill_i32(k.count, 0, padded, s);
    if (e != hipSuccess || k.n == 0) return e;
    hipLaunchKernelGGL(k_hash_count, dim3(grid_for(k.n)), dim3(kBlock), 0, s, P, k.n, k.pos[k.st], k.cell[k.ord],
                       k.key, k.rank, k.count, jumps, count_jumps ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_scan(const int32_t *count, int32_t *start, int ncell, ScanScratch &scr, hipStream_t s) {
    const int ntiles = (ncell + kScanTile - 1) / kScanTile;
    if (ntiles > scr.ntiles_cap) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(kBlock), 0, s, count, start, scr.tile_sums, ncell);
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(kBlock), 0, s, scr.tile_sums, ntiles);
    hipLaunchKernelGGL(k_scan_add, dim3(grid_for(ncell)), dim3(kBlock), 0, s, start, scr.tile_sums, ncell);
    return hipGetLastError();
}

hipError_t launch_sort_gather(const DevParams &P, KindBuffers &k, bool has_ref, hipStream_t s) {
    (void)P;
    if (k.n == 0) return hipSuccess;
    const int g = grid_for(k.n);
    hipLaunchKernelGGL(k_place, dim3(g), dim3(kBlock), 0, s, k.n, k.key, k.rank, k.start, k.id[k.ord], k.tmp_slot,
                       k.tmp_id);
    const int so = k.st, sn = 1 - k.st, oo = k.ord, on = 1 - k.ord;
    if (has_ref)
        hipLaunchKernelGGL(k_rank_gather<true>, dim3(g), dim3(kBlock), 0, s, k.n, k.key, k.start, k.tmp_slot,
                           k.tmp_id, k.pos[so], k.vel[so], k.ref[oo], k.pos[sn], k.vel[sn], k.ref[on], k.id[on],
                           k.cell[on]);
    else
        hipLaunchKernelGGL(k_rank_gather<false>, dim3(g), dim3(kBlock), 0, s, k.n, k.key, k.start, k.tmp_slot,
                           k.tmp_id, k.pos[so], k.vel[so], (const float4 *)nullptr, k.pos[sn], k.vel[sn],
                           (float4 *)nullptr, k.id[on], k.cell[on]);
    k.st = sn;
    k.ord = on;
    return hipGetLastError();
}

}  // namespace fsa
