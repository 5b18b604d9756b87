// fsa_pairs.hip -- fluid neighbour lists with a skin (Verlet lists).
//
// The reference finds every fluid particle's partners each step by walking
// the 27 cells around it (findFluidPair, src/fluid.cu:169-232).  Most of
// what that walk tests lies beyond the cut-off R (about 480 candidate tests
// for 80 partners at h = 1.3 dp).  A list of the candidates within
// (1 + skin) R, built from the cell lists, holds every partner until some
// particle has moved more than skin R / 2 since the build: two particles
// then close in by at most skin R, so a pair inside R now was inside
// (1 + skin) R then.  The fluid pass still tests every listed candidate
// against R exactly, so it finds the walk's partner set each step (tested
// against the oracle's all-pairs scan); only the order in which it sums them
// is the list's, the walk order of the build.
//
// A build runs after every fluid sort, on the device, when a decision pass
// finds that one is due (so the host never waits, and recorded HIP graphs
// stay valid):
//   1. decide  one workgroup: the largest move (the fluid pass's waves
//              take their maxima into kMoveWords words, which the decision
//              reads and clears) or a build forced by the host (upload).  In a fast flow one splash particle
//              can call for a build every few steps, more often than a
//              build pays for: lists that served fewer than kListMinAge
//              steps switch the scene to walking its cells for a stretch of
//              steps (kListRetry, doubled each time the lists fail again,
//              up to kListRetryMax), after which it builds and tries again.
//   2. build   the sorted state becomes the home order the fluid pass runs
//              over (home index = sorted slot), with the build positions;
//              per particle, the candidates within the list reach over the
//              5x5 columns of cells around its own (the reach exceeds one
//              cell), each column's z-window in slot order, as the byte
//              offsets of their home slots; and the wall particles within
//              the reach, as the byte offsets of their slots (the walls
//              never move, so they stay candidates as long as the fluid's).
// The build pass leaves at once when the current lists serve or the step
// walks.  A particle with more
// candidates than a list holds is marked unlisted and walks its cells in the
// fluid pass instead.
#include <algorithm>

#include "fsa_internal.h"
#include "fsa_math.h"

namespace fsa {

namespace {

constexpr int kDecideThreads = 1024;
constexpr int kBuildBatch = 8;     // candidate loads in flight in the build
static_assert(kPosPad >= kBuildBatch, "candidate batches may read past a range");

struct OpListDecide {
    static __device__ void run(const ListJob &j, int blk, int) {
        __shared__ uint32_t wmax[kDecideThreads / 64];
        if (blk != 0) return;
        // squared moves are float bits of nonnegative values (or NaN), so
        // their order is the unsigned order and a NaN exceeds every bound
        static_assert(kMoveWords == kDecideThreads, "one move word per thread");
        uint32_t m = j.L.move[threadIdx.x];
        j.L.move[threadIdx.x] = 0u;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, d, 64));
        if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x != 0) return;
        uint32_t all = 0;
#pragma unroll
        for (int w = 0; w < kDecideThreads / 64; ++w) all = max(all, wmax[w]);
        uint32_t *f = j.L.flag;
        const uint32_t age = f[kFlagAge], walk_left = f[kFlagWalkLeft];
        const bool moved = all > __float_as_uint(j.P.vthr2);
        uint32_t mode;
        if (f[kFlagForce] != 0u) {
            mode = kModeBuild;
        } else if (walk_left > 0u) {
            // walking: the lists are tried again (built afresh) when the
            // stretch is over
            f[kFlagWalkLeft] = walk_left - 1u;
            mode = walk_left == 1u ? kModeBuild : kModeWalk;
        } else if (moved && age < kListMinAge) {
            // the lists lasted too few steps to pay for their builds
            const uint32_t stretch = f[kFlagRetry] < kListRetry ? kListRetry : f[kFlagRetry];
            f[kFlagWalkLeft] = stretch - 1u;
            f[kFlagRetry] = min(2u * stretch, kListRetryMax);
            mode = kModeWalk;
        } else {
            if (moved) f[kFlagRetry] = kListRetry;  // these lists paid for themselves
            mode = moved ? kModeBuild : kModeList;
        }
        f[kFlagMode] = mode;
        f[kFlagForce] = 0u;
        if (mode == kModeBuild) {
            f[kFlagBuilds] += 1u;
            f[kFlagUnlisted] = 0u;
            f[kFlagEntries] = 0u;
            f[kFlagWallEntries] = 0u;
            f[kFlagAge] = 0u;
        } else if (mode == kModeList) {
            f[kFlagAge] = age + 1u;
        } else {
            f[kFlagWalks] += 1u;
        }
    }
};

// The z-window of column (ox, oy) of the 5x5 columns around the particle's
// cell for partners within a squared distance reach2: z_window (fsa_math.h)
// over columns up to two cells away, whose x (y) distance is the distance to
// the far face of the column between, plus one cell per column crossed.
__device__ __forceinline__ void wide_window(const DevParams &P, const Columns &cols, const int32_t *__restrict__ fine,
                                            int ox, int oy, int zlo, int zhi1, const Window &w, float reach2,
                                            int &b, int &e) {
    const float ex = ox < 0 ? w.fx + (float)(-ox - 1) : (ox > 0 ? (1.0f - w.fx) + (float)(ox - 1) : 0.0f);
    const float ey = oy < 0 ? w.fy + (float)(-oy - 1) : (oy > 0 ? (1.0f - w.fy) + (float)(oy - 1) : 0.0f);
    const float gx = fmaxf(ex - w.mx, 0.0f) * P.ival[0], gy = fmaxf(ey - w.my, 0.0f) * P.ival[1];
    const float r2 = reach2 - gx * gx - gy * gy;
    const int gxi = cols.cx + ox, gyi = cols.cy + oy;
    const bool ok = (unsigned)gxi < (unsigned)cols.nx && (unsigned)gyi < (unsigned)cols.ny && r2 > 0.0f;
    const float rzs = __builtin_amdgcn_sqrtf(fmaxf(r2, 0.0f)) * P.inv[2] * (float)kSubZ + w.mz;
    const int lo = max((int)floorf(w.tzs - rzs), zlo * kSubZ);
    const int hi = max(min((int)floorf(w.tzs + rzs) + 1, zhi1 * kSubZ), lo);
    const uint32_t k0 = ok ? (uint32_t)((cols.cbase + ox * cols.ny * cols.nz + oy * cols.nz) * kSubZ) : 0u;
    const int bb = fine[k0 + (ok ? (uint32_t)lo : 0u)], ee = fine[k0 + (ok ? (uint32_t)hi : 0u)];
    b = ok ? bb : 0;
    e = ok ? ee : 0;
}

// A hit is written straight to its row.  (Staging each lane's list in LDS
// and writing it out row by row, 64 lanes to consecutive addresses, kept the
// stores coalesced, but its LDS left room for fewer waves to hide the walk's
// load latency: 3.5 ms per build at 4 M particles against 2.0 ms.)
constexpr int kBuildThreads = 256;
constexpr int kBuildGrid = 16384;  // workgroups of one scene's build (grid-stride past 4 M particles)

// The candidates of one kind (positions pos, per-key starts fine) within
// the list reach of xi, in the walk order of the 5x5 columns, written to
// rows 0.. of ent as slot << shift -- the byte offsets of their records:
// 32-byte home records for the fluid, 16-byte positions for the walls --
// while fewer than `rows`; returns how many there are.
__device__ __forceinline__ int list_kind(const DevParams &P, const Columns &cols, const Window &w, int zlo, int zhi1,
                                         bool valid, const float4 &xi, const float4 *__restrict__ pos,
                                         const int32_t *__restrict__ fine, uint32_t *__restrict__ ent, size_t nf,
                                         int i, int rows, int shift) {
    int cnt = 0;
#pragma unroll 1
    for (int ox = -2; ox <= 2; ++ox) {
        // the row's 5 windows first, their loads in flight together
        int cb[5], cend[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            wide_window(P, cols, fine, ox, c - 2, zlo, zhi1, w, P.vreach2, cb[c], cend[c]);
            if (!valid) cend[c] = cb[c];
        }
#pragma unroll 1
        for (int col = 0; col < 5; ++col) {
            const int b = cb[col], e = cend[col];
            if (__ballot(b < e) == 0) continue;
            // the next batch is loaded while this one is tested (loads at or
            // past the window end stay in the padded array)
            float4 q[kBuildBatch];
            {
                const float4 *src = pos + min(b, e);
#pragma unroll
                for (int k = 0; k < kBuildBatch; ++k) q[k] = src[k];
            }
#pragma unroll 1
            for (int s0 = b; __ballot(s0 < e) != 0; s0 += kBuildBatch) {
                const int ss = min(s0, e);
                const int lim = e - ss;
                float4 qn[kBuildBatch];
                const float4 *nxt = pos + min(s0 + kBuildBatch, e);
#pragma unroll
                for (int k = 0; k < kBuildBatch; ++k) qn[k] = nxt[k];
#pragma unroll
                for (int k = 0; k < kBuildBatch; ++k) {
                    const float d2 = len2(q[k].x - xi.x, q[k].y - xi.y, q[k].z - xi.z);
                    if (k < lim && d2 < P.vreach2) {
                        if (cnt < rows) ent[(size_t)cnt * nf + i] = (uint32_t)(ss + k) << shift;
                        ++cnt;
                    }
                }
#pragma unroll
                for (int k = 0; k < kBuildBatch; ++k) q[k] = qn[k];
            }
        }
    }
    return cnt;
}

// wave sum of v (every lane of the wave active)
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += (uint32_t)__shfl_xor((int)v, d, 64);
    return v;
}

struct OpListBuild {
    static __device__ void run(const ListJob &j, int blk, int nblk) {
        const uint32_t mode = j.L.flag[kFlagMode];
        if (mode != kModeBuild) return;  // the current lists serve, or the fluid pass walks
        const DevParams &P = j.P;
        const int lane = threadIdx.x;
        const size_t nf = (size_t)P.nf;
        uint32_t ce = 0, cw = 0;  // listed fluid and wall entries of this thread's particles
        for (int base = blk * kBuildThreads; base < P.nf; base += nblk * kBuildThreads) {
            const int i = base + lane;
            const bool valid = i < P.nf;
            // every lane takes part in the wave-uniform loops; lanes past the
            // end have empty windows and write nothing
            const float4 xi = j.spos[valid ? i : 0];
            if (valid) {
                // the new home order (the candidates below are read from the
                // sorted state, the same values)
                j.hrec[2 * (size_t)i] = xi;
                j.hrec[2 * (size_t)i + 1] = j.svel[i];
                j.hidx[i] = i;
                j.hslot[i] = i;
                j.L.x0[i] = xi;
            }
            const int cell = find_cell(P, xi.x, xi.y, xi.z);
            const Columns cols(P, cell);
            const Window w(P, cols, xi.x, xi.y, xi.z);
            const int cz = cell - cols.cbase;
            const int zlo = max(cz - 2, 0), zhi1 = min(cz + 2, cols.nz - 1) + 1;
            // the fluid's candidates, then the walls' (the walls never move:
            // their candidates stay valid as long as the fluid's do); one
            // loop, so that the walk is inlined once
            bool unlisted = false;
#pragma unroll 1
            for (int kind = 0; kind < (j.L.went ? 2 : 1); ++kind) {
                const bool wall = kind == 1;
                // most waves have no wall within two cells (the list reach)
                if (wall && __ballot(valid && (j.vreach[cell] & kReachFar) != 0) == 0) {
                    if (valid) j.L.wmeta[i] = 0;
                    continue;
                }
                const int rows = wall ? P.vwrows : P.vrows;
                const int cnt = list_kind(P, cols, w, zlo, zhi1, valid, xi, wall ? j.vpos : j.spos,
                                          wall ? j.vfine : j.ffine, wall ? j.L.went : j.L.ent, nf, i, rows,
                                          wall ? 4 : 5);
                const bool listed = valid && cnt <= rows;
                unlisted |= valid && !listed;
                if (valid) (wall ? j.L.wmeta : j.L.meta)[i] = listed ? cnt : -1;
                (wall ? cw : ce) += listed ? (uint32_t)cnt : 0u;
            }
            if (unlisted) atomicAdd(&j.L.flag[kFlagUnlisted], 1u);
        }
        // the wave's listed entries, one atomic each (every lane is still active)
        ce = wave_sum(ce);
        cw = wave_sum(cw);
        if ((lane & 63) == 0) {
            if (ce > 0u) atomicAdd(&j.L.flag[kFlagEntries], ce);
            if (cw > 0u) atomicAdd(&j.L.flag[kFlagWallEntries], cw);
        }
    }
};

template <class Op, int THREADS = kBlock>
hipError_t one(const ListJob &j, int blocks, hipStream_t s) {
    if (blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_one<Op, ListJob, THREADS>), dim3(blocks), dim3(THREADS), 0, s, j);
    return hipGetLastError();
}

}  // namespace

int blocks_list(ListOp op, const ListJob &j) {
    if (!j.P.vlist || j.P.nf == 0) return 0;
    switch (op) {
    case kLDecide: return 1;
    case kLBuild: return std::min((j.P.nf + kBuildThreads - 1) / kBuildThreads, kBuildGrid);
    default: return 0;
    }
}

hipError_t launch_list_one(ListOp op, const ListJob &j, hipStream_t s) {
    const int nb = blocks_list(op, j);
    switch (op) {
    case kLDecide: return one<OpListDecide, kDecideThreads>(j, nb, s);
    case kLBuild: return one<OpListBuild, kBuildThreads>(j, nb, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_many_list(ListOp op, const ListJob *jobs, const int *prefix, int njobs, int nblocks,
                            hipStream_t s) {
    if (nblocks <= 0 || njobs <= 0) return hipSuccess;
    const dim3 g(nblocks);
    switch (op) {
    case kLDecide:
        hipLaunchKernelGGL((k_many<OpListDecide, ListJob, kDecideThreads>), g, dim3(kDecideThreads), 0, s, jobs,
                           prefix, njobs);
        break;
    case kLBuild:
        hipLaunchKernelGGL((k_many<OpListBuild, ListJob, kBuildThreads>), g, dim3(kBuildThreads), 0, s, jobs, prefix,
                           njobs);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_list_build(const ListJob &j, hipStream_t s) {
    hipError_t e;
    for (ListOp op : {kLDecide, kLBuild})
        if ((e = launch_list_one(op, j, s)) != hipSuccess) return e;
    return hipSuccess;
}

}  // namespace fsa
