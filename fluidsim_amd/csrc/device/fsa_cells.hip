// fsa_cells.hip -- cell hash and the cell-list build.
//
// The reference keeps per-cell singly linked lists and rebuilds them each
// step by walking the previous lists of the 27 neighbour cells with one
// thread per cell (src/pair.cu:72-207): a latency-bound pointer chase.  Here
// the cell list is a counting sort of the particles by key, the cell id
// (findPositionCell, src/pair.cu:8-24) times kSubZ plus the eighth-cell z
// index inside the cell:
//   1. fill+hash   zero the per-key counts; one thread per particle: key
//                  and an arrival rank among its key (one atomic per run of
//                  equal keys in a wavefront)
//   2. scan        exclusive prefix sum of the per-key counts -> fstart[];
//                  the per-cell starts are every kSubZ-th of them
//   3. place       the particle's slot and .dat id at fstart[k] + rank
//   4. gather      inside each key the final rank is the number of members
//                  with a smaller .dat id, which makes the order
//                  deterministic (ascending .dat id within a key) whatever
//                  order the atomics came in; the state is gathered into the
//                  other buffer at fstart[k] + rank.
// All passes stream over HBM except the small per-cell rank loop, which
// reads the cell's ids from L1/L2.
#include "fsa_internal.h"
#include "fsa_math.h"

namespace fsa {

namespace {

__global__ __launch_bounds__(kBlock) void k_fill_i32(int32_t *__restrict__ p, int32_t v, int n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_fill_f4(float4 *__restrict__ p, float4 v, int n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

struct OpFillCountsImpl {
    static __device__ void run(const CellJob &j, int blk, int) {
        const int i = blk * kBlock + threadIdx.x;
        if (i < padded_cells(j.nk)) j.count[i] = 0;
    }
};

// per cell: kReachNear if one of its 27 cells holds a particle of the kind,
// kReachFar if one of the 125 cells within two cells of it does
__global__ __launch_bounds__(kBlock) void k_reach(DevParams P, const int32_t *__restrict__ start,
                                                  uint8_t *__restrict__ reach) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= P.ncell) return;
    int b[9], e[9];
    column_ranges(P, c, start, b, e);
    bool near = false;
#pragma unroll
    for (int k = 0; k < 9; ++k) near |= b[k] < e[k];
    bool far = near;
    const int3 ci = cell_coords(P, c);
    const int zlo = max(ci.z - 2, 0), zhi1 = min(ci.z + 2, P.cn[2] - 1) + 1;
    for (int ox = -2; ox <= 2 && !far; ++ox)
        for (int oy = -2; oy <= 2 && !far; ++oy) {
            const int gx = ci.x + ox, gy = ci.y + oy;
            if ((unsigned)gx >= (unsigned)P.cn[0] || (unsigned)gy >= (unsigned)P.cn[1]) continue;
            const int base = (gx * P.cn[1] + gy) * P.cn[2];
            far = start[base + zhi1] > start[base + zlo];
        }
    reach[c] = (uint8_t)((near ? kReachNear : 0) | (far ? kReachFar : 0));
}

struct OpHashImpl {
    static __device__ void run(const CellJob &j, int blk, int) {
        const int s = blk * kBlock + threadIdx.x;
        const bool valid = s < j.n;
        const DevParams &P = j.P;
        // every lane takes part in the wave operations; lanes past the end
        // (the tail of the last wave) carry key -1 and do no memory access
        int key = -1;
        if (valid) {
            const float4 x = j.pos_in[s];
            key = cell_key(P, x.x, x.y, x.z, j.nsub);
        }
        // Particles arrive sorted by their previous key, so equal keys come
        // in runs of adjacent lanes.  The first lane of each run (a lane
        // whose left neighbour has another key) takes the run's ranks with
        // one atomic; the others add their distance to it.  Equal keys in
        // separate runs take separate rank blocks, so ranks stay unique and
        // counts exact whatever the order (the gather pass fixes the order).
        const int lane = threadIdx.x & 63;
        const int left = __shfl_up(key, 1, 64);
        const unsigned long long heads = __ballot(lane == 0 || key != left);
        const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
        const int head = 63 - __clzll((long long)(heads & upto));
        const unsigned long long later = heads & ~upto;
        const int next = later ? __ffsll((long long)later) - 1 : 64;
        int base = 0;
        if (valid && lane == head) base = atomicAdd(&j.count[key], next - lane);
        base = __shfl(base, head, 64);
        if (!valid) return;
        const int c = key / j.nsub;
        j.key[s] = key;
        j.rank[s] = base + (lane - head);
        if (j.count_jumps) {
            // more than one cell in any direction since the last build: the
            // reference's list rebuild would have lost this particle
            const int o = j.prev_cell[s];
            const int nyz = P.cn[1] * P.cn[2];
            const int ax = c / nyz, bx = o / nyz;
            const int ay = (c - ax * nyz) / P.cn[2], by = (o - bx * nyz) / P.cn[2];
            const int az = c % P.cn[2], bz = o % P.cn[2];
            if (abs(ax - bx) > 1 || abs(ay - by) > 1 || abs(az - bz) > 1) atomicAdd(j.jumps, 1ull);
        }
    }
};

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

// exclusive scan of the per-thread totals of an NT-thread workgroup
// (wave_tot: NT / 64 ints of LDS)
template <int NT = kBlock>
__device__ __forceinline__ int block_exclusive_scan(int v, int *wave_tot, int &block_total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int inc = wave_inclusive_scan(v);
    if (lane == 63) wave_tot[wave] = inc;
    __syncthreads();
    int before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) {
        const int t = wave_tot[w];
        before += (w < wave) ? t : 0;
        all += t;
    }
    block_total = all;
    return before + inc - v;
}

constexpr int kSumsThreads = 1024;  // the one workgroup that scans the tile totals

constexpr int kPerThread = kScanTile / kBlock;  // counts per thread of a scan tile
static_assert(kPerThread % 4 == 0, "scan tiles are read as int4");

// the kPerThread counts of this thread in tile `blk` (count is padded to a
// multiple of kScanTile with zeros)
__device__ __forceinline__ void load_counts(const CellJob &j, int blk, int (&c)[kPerThread]) {
    const int4 *src = reinterpret_cast<const int4 *>(j.count + (size_t)blk * kScanTile + threadIdx.x * kPerThread);
#pragma unroll
    for (int q = 0; q < kPerThread / 4; ++q) {
        const int4 v = src[q];
        c[4 * q] = v.x;
        c[4 * q + 1] = v.y;
        c[4 * q + 2] = v.z;
        c[4 * q + 3] = v.w;
    }
}

// pass 1: the total of each scan tile
struct OpScanTilesImpl {
    static __device__ void run(const CellJob &j, int blk, int) {
        __shared__ int wave_tot[kBlock / 64];
        // a batched grid rounds each scene's block range up: whole blocks past
        // the last tile leave before the first barrier
        if (blk >= scan_tiles(j.nk)) return;
        int c[kPerThread];
        load_counts(j, blk, c);
        int t = 0;
#pragma unroll
        for (int k = 0; k < kPerThread; ++k) t += c[k];
        int total;
        (void)block_exclusive_scan(t, wave_tot, total);
        if (threadIdx.x == 0) j.tile_sums[blk] = total;
    }
};

// pass 2: one workgroup of kSumsThreads scans the tile totals (exclusive, in place)
struct OpScanSumsImpl {
    static __device__ void run(const CellJob &j, int blk, int) {
        __shared__ int wave_tot[kSumsThreads / 64];
        __shared__ int carry;
        if (blk != 0) return;  // padding blocks of a batched grid
        const int ntiles = scan_tiles(j.nk);
        if (threadIdx.x == 0) carry = 0;
        __syncthreads();
        for (int b = 0; b < ntiles; b += kSumsThreads) {
            const int i = b + threadIdx.x;
            const int v = i < ntiles ? j.tile_sums[i] : 0;
            int total;
            const int ex = block_exclusive_scan<kSumsThreads>(v, wave_tot, total);
            const int c = carry;
            if (i < ntiles) j.tile_sums[i] = c + ex;
            __syncthreads();
            if (threadIdx.x == 0) carry = c + total;
            __syncthreads();
        }
    }
};

// pass 3: the tile's inclusive scan plus the totals of the tiles before it
// into fstart[1 + i]; fstart[0] = 0.  With several keys per cell the
// per-cell starts are every nsub-th key start (cell c begins where its first
// key does), written by the thread that finishes that key start.  The counts
// go back to zero, ready for the next build's hash.
struct OpScanAddImpl {
    static __device__ void run(const CellJob &j, int blk, int) {
        __shared__ int wave_tot[kBlock / 64];
        if (blk >= scan_tiles(j.nk)) return;
        int c[kPerThread];
        load_counts(j, blk, c);
        int t = 0;
#pragma unroll
        for (int k = 0; k < kPerThread; ++k) t += c[k];
        int total;
        int run = block_exclusive_scan(t, wave_tot, total) + j.tile_sums[blk];
        const int base = blk * kScanTile + threadIdx.x * kPerThread;
        if (blk == 0 && threadIdx.x == 0) {
            j.fstart[0] = 0;
            if (j.nsub > 1) j.start[0] = 0;
        }
#pragma unroll
        for (int k = 0; k < kPerThread; ++k) {
            const int i = base + k;
            run += c[k];
            if (i < j.nk) {
                j.fstart[1 + i] = run;
                if (j.nsub > 1 && (i + 1) % j.nsub == 0) j.start[(i + 1) / j.nsub] = run;
            }
        }
        int4 *dst = reinterpret_cast<int4 *>(j.count + (size_t)base);
#pragma unroll
        for (int q = 0; q < kPerThread / 4; ++q) dst[q] = make_int4(0, 0, 0, 0);
    }
};

struct OpPlaceImpl {
    static __device__ void run(const CellJob &j, int blk, int) {
        const int s = blk * kBlock + threadIdx.x;
        if (s >= j.n) return;
        const int t = j.fstart[j.key[s]] + j.rank[s];
        j.tmp_slot[t] = s;
        j.tmp_id[t] = j.id_in[s];
    }
};

struct OpGatherImpl {
    static __device__ void run(const CellJob &j, int blk, int) {
        const int t = blk * kBlock + threadIdx.x;
        if (t >= j.n) return;
        const int s = j.tmp_slot[t];
        const int my = j.tmp_id[t];
        // what depends on the slot alone is loaded before the key's range and
        // the rank search, so those loads wait on tmp_slot only
        const int k = j.key[s];
        const float4 x = j.pos_in[s], u = j.vel_in[s];
        const int h = j.hidx_out ? j.hidx_in[s] : 0;
        const int c = k / j.nsub;
        const int b = j.fstart[k], e = j.fstart[k + 1];
        // rank of `my` among the key's ids (t lies in [b, e)), four ids per
        // round with independent loads; positions past the range re-read its
        // last id and are not counted
        int r = 0;
        for (int v = b; v < e; v += 4) {
            const int i0 = j.tmp_id[v];
            const int i1 = j.tmp_id[min(v + 1, e - 1)];
            const int i2 = j.tmp_id[min(v + 2, e - 1)];
            const int i3 = j.tmp_id[min(v + 3, e - 1)];
            r += (int)(i0 < my) + (int)((v + 1 < e) & (i1 < my)) + (int)((v + 2 < e) & (i2 < my)) +
                 (int)((v + 3 < e) & (i3 < my));
        }
        const int d = b + r;
        j.pos_out[d] = x;
        j.vel_out[d] = u;
        if (j.has_ref) j.ref_out[d] = j.ref_in[s];
        if (j.cpos_in) {
            j.cpos_out[d] = j.cpos_in[s];
            j.cvel_out[d] = j.cvel_in[s];
        }
        if (j.pk_out) j.pk_out[d] = pack_partner(j.P, x, u, c);
        j.id_out[d] = my;
        j.cell_out[d] = c;
        if (j.hidx_out) {
            j.hidx_out[d] = h;
            j.hslot[h] = d;
        }
    }
};

template <class Op>
hipError_t one(const CellJob &j, int blocks, hipStream_t s) {
    if (blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_one<Op, CellJob>), dim3(blocks), dim3(kBlock), 0, s, j);
    return hipGetLastError();
}

}  // namespace

int blocks_cells(BatchOp op, const CellJob &j) {
    if (j.n == 0) return 0;
    switch (op) {
    case kBFillCounts: return blocks_for(padded_cells(j.nk));
    case kBHash: return blocks_for(j.n);
    case kBScanTiles: return scan_tiles(j.nk);
    case kBScanSums: return 1;
    case kBScanAdd: return scan_tiles(j.nk);
    case kBPlace: return blocks_for(j.n);
    case kBGather: return blocks_for(j.n);
    default: return 0;
    }
}

hipError_t launch_many_cells(BatchOp op, const CellJob *jobs, const int *prefix, int njobs, int nblocks,
                             hipStream_t s) {
    if (nblocks <= 0 || njobs <= 0) return hipSuccess;
    const dim3 g(nblocks), b(kBlock);
    switch (op) {
    case kBFillCounts: hipLaunchKernelGGL((k_many<OpFillCountsImpl, CellJob>), g, b, 0, s, jobs, prefix, njobs); break;
    case kBHash: hipLaunchKernelGGL((k_many<OpHashImpl, CellJob>), g, b, 0, s, jobs, prefix, njobs); break;
    case kBScanTiles: hipLaunchKernelGGL((k_many<OpScanTilesImpl, CellJob>), g, b, 0, s, jobs, prefix, njobs); break;
    case kBScanSums:
        hipLaunchKernelGGL((k_many<OpScanSumsImpl, CellJob, kSumsThreads>), g, dim3(kSumsThreads), 0, s, jobs, prefix,
                           njobs);
        break;
    case kBScanAdd: hipLaunchKernelGGL((k_many<OpScanAddImpl, CellJob>), g, b, 0, s, jobs, prefix, njobs); break;
    case kBPlace: hipLaunchKernelGGL((k_many<OpPlaceImpl, CellJob>), g, b, 0, s, jobs, prefix, njobs); break;
    case kBGather: hipLaunchKernelGGL((k_many<OpGatherImpl, CellJob>), g, b, 0, s, jobs, prefix, njobs); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

CellJob make_cell_job(const DevParams &P, const KindBuffers &k, unsigned long long *jumps, bool count_jumps) {
    CellJob j{};
    j.P = P;
    j.n = k.n;
    j.nsub = k.nsub;
    j.nk = P.ncell * k.nsub;
    j.has_ref = k.ref[0] != nullptr ? 1 : 0;
    j.count_jumps = count_jumps ? 1 : 0;
    const int so = k.st, sn = 1 - k.st, oo = k.ord, on = 1 - k.ord;
    j.pos_in = k.pos[so];
    j.vel_in = k.vel[so];
    j.cpos_in = k.cpos[so];
    j.cvel_in = k.cvel[so];
    j.cpos_out = k.cpos[sn];
    j.cvel_out = k.cvel[sn];
    j.pk_out = k.pk;
    j.ref_in = k.ref[oo];
    j.id_in = k.id[oo];
    j.prev_cell = k.cell[oo];
    j.pos_out = k.pos[sn];
    j.vel_out = k.vel[sn];
    j.ref_out = k.ref[on];
    j.id_out = k.id[on];
    j.cell_out = k.cell[on];
    j.key = k.key;
    j.rank = k.rank;
    j.count = k.count;
    j.start = k.next_start();  // the build's starts become current when it flips the order
    j.fstart = k.next_fstart();
    j.tmp_slot = k.tmp_slot;
    j.tmp_id = k.tmp_id;
    j.tile_sums = k.tile_sums;
    j.jumps = jumps;
    if (k.hslot) {
        j.hidx_in = k.hidx[oo];
        j.hidx_out = k.hidx[on];
        j.hslot = k.hslot;
    }
    return j;
}

hipError_t launch_fill_i32(int32_t *p, int32_t v, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(n)), dim3(kBlock), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_fill_f4(float4 *p, float4 v, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_f4, dim3(blocks_for(n)), dim3(kBlock), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_reach(const DevParams &P, const int32_t *start, uint8_t *reach, hipStream_t s) {
    if (P.ncell <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reach, dim3(blocks_for(P.ncell)), dim3(kBlock), 0, s, P, start, reach);
    return hipGetLastError();
}

hipError_t launch_hash_one(const CellJob &j, hipStream_t s) {
    hipError_t e;
    if ((e = one<OpFillCountsImpl>(j, blocks_cells(kBFillCounts, j), s)) != hipSuccess) return e;
    return one<OpHashImpl>(j, blocks_cells(kBHash, j), s);
}

hipError_t launch_sort_one(const CellJob &j, hipStream_t s) {
    hipError_t e;
    if ((e = one<OpScanTilesImpl>(j, blocks_cells(kBScanTiles, j), s)) != hipSuccess) return e;
    hipLaunchKernelGGL((k_one<OpScanSumsImpl, CellJob, kSumsThreads>), dim3(1), dim3(kSumsThreads), 0, s, j);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = one<OpScanAddImpl>(j, blocks_cells(kBScanAdd, j), s)) != hipSuccess) return e;
    if ((e = one<OpPlaceImpl>(j, blocks_cells(kBPlace, j), s)) != hipSuccess) return e;
    return one<OpGatherImpl>(j, blocks_cells(kBGather, j), s);
}

// the counts are zero between builds (allocation, then every sort's scan),
// so a whole build starts with the hash
hipError_t launch_cells_one(const CellJob &j, hipStream_t s) {
    hipError_t e = one<OpHashImpl>(j, blocks_cells(kBHash, j), s);
    return e != hipSuccess ? e : launch_sort_one(j, s);
}

}  // namespace fsa
