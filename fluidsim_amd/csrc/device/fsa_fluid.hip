// fsa_fluid.hip -- the fluid pass: Riemann-SPH continuity and momentum sums
// over fluid, solid and virtual neighbours (findFluidPair and the three pair
// functions, src/fluid.cu:26-232) fused with the explicit update
// (fluidUpdate, src/fluid.cu:234-269).
//
// One thread per fluid particle in sorted order, so the 64 lanes of a
// wavefront are ~64 spatially adjacent particles whose neighbour columns
// overlap and stay in L1/L2.  Each column of the 27-cell neighbourhood is a
// contiguous slot range per kind.
//
// Only ~1 candidate in 6 lies inside the cut-off, and in SIMT a wave pays
// for the pair math whenever any lane takes that branch.  So each kind's
// partners are done in two phases: the candidate test appends the in-range
// slots to a lane-private list in LDS (entry k of lane t at [k][t], so lanes
// hit distinct banks), then a dense loop evaluates the pair terms of the
// list.  Pairs are summed in exactly the candidate order.  The list holds all
// of a particle's partners of one kind (~80 at h = 1.3 dp), so the pair loop
// runs once per kind with every lane busy: a lane's partner count varies
// little across a wave, while its count at any point of the candidate walk
// varies a lot (the walk visits the columns in a fixed order and the lanes'
// particles sit at different places in their cells).  Entries are 16 bits:
// the offset inside the run (12 bits, high) and a run-base slot (4 bits,
// low; the run's first slot is kept in LDS beside the list), so that the
// high part is the offset in bytes of 16-byte records; a run longer than
// 4096 slots continues in spare base slots.
//
// Pair arithmetic is FP32 with the hardware reciprocal, reciprocal square
// root and exp2, and the algebraically reduced forms noted inline; the
// results agree with the IEEE-division form of the reference to a few ulp per
// term (see the parity tests for the bound held).
//
// The kernel reads the state buffer and writes the updated state into the
// other buffer (Jacobi, one writer per particle), which removes the
// round-trip of du_dt and drho_dt through HBM.
#include "fsa_internal.h"
#include "fsa_math.h"

namespace fsa {

namespace {

enum PairKind { kFluidFluid = 0, kFluidSolid = 1, kFluidVirt = 2 };


struct Self {
    float x, y, z, rho, ux, uy, uz, p;
    float inv_rho, two_rho;
    float m2_inv_rho;    // -2 / rho_i
    float vis5;          // 5 mu (0 for an inviscid fluid)
    float half_inv_rho;  // 1 / (rho_i + rho_j) when rho_j = rho_i (solid partners)
    // packed form: own slot and cell, position relative to the cell centre in cells
    int i, cell;
    float fx, fy, fz;
};

// The sums leave out the particle's own factors, applied once at the end:
// drho/dt = 2 rho_i drho, a = -2/rho_i (ax, ay, az).
struct Acc {
    float drho, ax, ay, az;
};

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float exp2_hw(float x) { return __builtin_amdgcn_exp2f(x); }

// One Riemann pair term (fluid.cu:26-167) for a partner inside the cut-off.
//   p*  = (rho_i p_j + rho_j p_i)/(rho_i+rho_j) + [du>=0] rho_i rho_j du/(rho_i+rho_j) * D
//         D = min(3 du, c0) (fluid, wall) or c0 (solid); walls clamp p* >= 0
//   v*  = (rho_i u_i + rho_j u_j)/(rho_i+rho_j) + (p_i-p_j)/(rho_i+rho_j)/c0 e
//   a  += -2 p*/rho_i dx k + [viscid] 10 mu ((u_j-u_i).dx)/(|dx|^2+0.01h^2)/rho_i dx k
//   rho' += 2 rho_i (u_i - v*).dx k
// evaluated with A = rho_j/(rho_i+rho_j) (1/2 for a solid partner, whose
// rho_j is rho_i): p* = p_j + A (p_i - p_j) + rho_i A max(du, 0) D, and
// (u_i - v*).dx = A (u_i-u_j).dx - (p_i-p_j)/(rho_i+rho_j)/c0 |dx|; the
// acceleration term is -2/rho_i [p* + 5 mu (u_i-u_j).dx/(|dx|^2+0.01h^2)] dx k.
// k = kd exp(-|dx|^2/h^2) = exp2(log2(kd) - |dx|^2 log2(e)/h^2).  A padding
// slot of the pair loop holds the sentinel partner (or comes with d2 =
// kPadD2 in the packed form): its k is exactly 0 and every other factor
// finite, so its terms are exact zeros.  The self pair (d2 = 0) takes
// 1/|dx| = rsq(FLT_MIN), finite, so its du and |dx| are 0 as well.
// The term's factors: a += (dx, dy, dz) coef, rho' += kc dr.
struct Term {
    float coef, kc, dr;
};

template <int KIND>
__device__ __forceinline__ Term pair_factors(const DevParams &P, const Self &me, float dx, float dy, float dz,
                                             float d2, float ujx, float ujy, float ujz, float rhoj, float pj) {
    const float kc = exp2_hw(fmaf(-d2, P.kexp2, P.log2kd));
    const float r = rsq(fmaxf(d2, 1.17549435e-38f));  // 1/|dx|, e = dx r
    const float inv_rsum = KIND == kFluidSolid ? me.half_inv_rho : rcp(me.rho + rhoj);
    const float A = KIND == kFluidSolid ? 0.5f : rhoj * inv_rsum;
    const float dpij = me.p - pj;
    const float udx = (me.ux - ujx) * dx + (me.uy - ujy) * dy + (me.uz - ujz) * dz;  // (u_i-u_j).dx
    const float du = fmaxf(udx * r, 0.0f);  // the dissipation acts on approaching pairs only
    const float diss = KIND == kFluidSolid ? P.cs : fminf(3.0f * du, P.cs);
    float ps = fmaf(me.rho * A * du, diss, fmaf(A, dpij, pj));
    if (KIND == kFluidVirt) ps = fmaxf(ps, 0.0f);
    float coef = fmaf(me.vis5 * udx, rcp(d2 + P.eps_vis), ps);
    coef *= kc;
    const float dpc = dpij * inv_rsum * P.cs_inv;
    return {coef, kc, A * udx - dpc * (d2 * r)};
}

template <int KIND>
__device__ __forceinline__ void pair_term(const DevParams &P, const Self &me, float dx, float dy, float dz, float d2,
                                          float ujx, float ujy, float ujz, float rhoj, float pj, Acc &acc) {
    const Term t = pair_factors<KIND>(P, me, dx, dy, dz, d2, ujx, ujy, ujz, rhoj, pj);
    acc.ax = fmaf(dx, t.coef, acc.ax);
    acc.ay = fmaf(dy, t.coef, acc.ay);
    acc.az = fmaf(dz, t.coef, acc.az);
    acc.drho = fmaf(t.kc, t.dr, acc.drho);
}

// A listed fluid or wall candidate: the pair term if it lies inside the
// cut-off, else nothing -- its terms are selected away, so that a partner
// outside the cut-off leaves the sums exactly as the walk, which never
// evaluates it, leaves them (a non-finite one included).
template <int KIND>
__device__ __forceinline__ void listed_pair(const DevParams &P, const Self &me, const float4 &pj, const float4 &vj,
                                            Acc &acc, int &hits) {
    const float dx = pj.x - me.x, dy = pj.y - me.y, dz = pj.z - me.z;
    const float d2 = len2(dx, dy, dz);
    const bool hit = d2 < P.R2;
    const Term t = pair_factors<KIND>(P, me, dx, dy, dz, d2, vj.x, vj.y, vj.z, pj.w, vj.w);
    const float ax = fmaf(dx, t.coef, acc.ax), ay = fmaf(dy, t.coef, acc.ay), az = fmaf(dz, t.coef, acc.az);
    const float dr = fmaf(t.kc, t.dr, acc.drho);
    acc.ax = hit ? ax : acc.ax;
    acc.ay = hit ? ay : acc.ay;
    acc.az = hit ? az : acc.az;
    acc.drho = hit ? dr : acc.drho;
    hits += hit ? 1 : 0;
}

// squared distance given to padding slots of the packed form: exp2 of its
// (negated, scaled) value is 0, its square root and reciprocal square root
// stay finite
constexpr float kPadD2 = 1.0e30f;

template <int KIND>
__device__ __forceinline__ void one_pair(const DevParams &P, const Self &me, const float4 &pj, const float4 &vj,
                                         const float4 &n, Acc &acc) {
    const float dx = pj.x - me.x, dy = pj.y - me.y, dz = pj.z - me.z;
    const float d2 = len2(dx, dy, dz);
    if (KIND == kFluidSolid) {
        // mirrored solid velocity (fluid.cu:79-83); rho_j = rho_i, p_j = p_s
        const float m = 2.0f * ((me.ux - vj.x) * n.x + (me.uy - vj.y) * n.y + (me.uz - vj.z) * n.z);
        pair_term<KIND>(P, me, dx, dy, dz, d2, me.ux - m * n.x, me.uy - m * n.y, me.uz - m * n.z, me.rho, n.w,
                        acc);
    } else {
        pair_term<KIND>(P, me, dx, dy, dz, d2, vj.x, vj.y, vj.z, pj.w, vj.w, acc);
    }
}

// A fluid partner from its FP16 record (FSA_FORM_PACKED).  The partner lies
// in one of the 27 cells around the reader's, so the cell-id difference
// (dix nyz + diy nz + diz, each d in {-1, 0, 1}) decodes by comparisons once
// shifted to nonnegative digits (unambiguous with >= 3 cells per axis).  The
// offset between the two particles is then (d + r_j - f_i) cells.  The
// reader's own slot gives dx = 0 exactly, as in the FP32 form.
__device__ __forceinline__ void one_pair_packed(const DevParams &P, const Self &me, const uint4 &rec, int cj, int j,
                                                bool valid, Acc &acc) {
    const int nz = P.cn[2], nyz = P.cn[1] * nz;
    int t = cj - me.cell + nyz + nz + 1;
    const int ax = (t >= nyz ? 1 : 0) + (t >= 2 * nyz ? 1 : 0);
    t -= ax * nyz;
    const int ay = (t >= nz ? 1 : 0) + (t >= 2 * nz ? 1 : 0);
    t -= ay * nz;
    const float2 rxy = unpack_h2(rec.x), rzd = unpack_h2(rec.y), uxy = unpack_h2(rec.z), uzp = unpack_h2(rec.w);
    float dx = ((float)(ax - 1) + (rxy.x - me.fx)) * P.ival[0];
    float dy = ((float)(ay - 1) + (rxy.y - me.fy)) * P.ival[1];
    float dz = ((float)(t - 1) + (rzd.x - me.fz)) * P.ival[2];
    if (j == me.i) dx = dy = dz = 0.0f;
    const float d2 = valid ? len2(dx, dy, dz) : kPadD2;
    const float rhoj = P.rho0 + P.rho0 * rzd.y;
    pair_term<kFluidFluid>(P, me, dx, dy, dz, d2, uxy.x, uxy.y, uzp.x, rhoj, uzp.y * P.rho2p, acc);
}

// workgroup sizes of the fluid pass (one to four waves; LDS per wave bounds
// the occupancy, so the best size depends on the scene: fsa_autotune_fluid)
constexpr int kThreadChoices[] = {64, 128, 256};
// waves per SIMD the fluid kernels keep registers for: three (168 VGPRs
// each), as many as LDS admits with workgroups of up to 128 threads
constexpr int fluid_waves(int threads) { return threads > 128 ? 2 : 3; }
// LDS per lane: (kListCap + kListPad) 16-bit entries + kBaseSlots 32-bit bases = 240 B, so that ten wave64s fit a
// CU's 160 KiB (a lane more of either, and LDS allocation granules leave room for nine)
constexpr int kListCap = 89;    // 16-bit list entries per lane
constexpr int kListPad = 7;     // sentinel entries past the last one (the rounds and their prefetch read them)
constexpr int kColumns = 9;     // neighbour columns; column c's run starts in base slot c
constexpr int kBaseSlots = 12;  // run bases per lane (4-bit ids): the columns, spares for long runs, the sentinel
constexpr int kSentinelSlot = kBaseSlots - 1;  // holds the kind's particle count: the first padding slot
constexpr int kSpareSlots = kSentinelSlot - kColumns;
constexpr int kPairBatch = 4;   // partners per round of the pair loop (the next round is gathered meanwhile)
// listed candidates per round of the list pass (4, 5, 6, 7 and 8 measured:
// 6 keeps the most partner bytes in flight without running past the rounds
// the waves can hide; 0.957 ms against 1.009 at 4 per round, 4 M particles)
constexpr int kListBatch = 6;
static_assert(3 * kListBatch - 1 <= kListAhead, "the list pass reads up to 3 kListBatch - 2 rows past a list");
constexpr int kScanBatch = 8;   // candidate loads in flight in the test loop
constexpr uint16_t kSentinelEntry = kSentinelSlot;  // offset 0 of the sentinel's run
static_assert(kPosPad >= kScanBatch, "candidate batches may read kScanBatch-1 entries past a range");
static_assert(kBaseSlots <= 16 && kBaseSlots * kRunMax <= 65536, "entries are 16 bits");
static_assert(kListPad >= 2 * kPairBatch - 1, "a round and the prefetch of the next read up to 2 kPairBatch - 1 "
                                              "entries past the last one");

// The run bases in LDS are byte offsets of the run's first slot in the
// kind's float4 arrays (slot counts stay below 2^28, so they fit 32 bits);
// a partner's arrays are read at base + 32-bit byte offset.
__device__ __forceinline__ const float4 &at(const float4 *__restrict__ base, uint32_t byte_off) {
    return *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(base) + byte_off);
}
__device__ __forceinline__ uint32_t slot_bytes(int slot) { return (uint32_t)slot << 4; }

// byte offset of the slot named by list entry v
template <int T>
__device__ __forceinline__ uint32_t entry_bytes(const uint32_t *base, uint32_t v) {
    return base[(v & 15u) * T] + (v & 0xfff0u);
}

// lane-private LDS views of a T-thread workgroup: entry k at list[k * T],
// base b at base[b * T]
template <int T>
struct Lists {
    uint16_t *list;
    uint32_t *base;
};

// partner state of list entries k .. k+kPairBatch-1 (the entries past the
// last one name the sentinel partner)
template <int KIND, int T>
__device__ __forceinline__ void gather(const float4 *__restrict__ pos, const float4 *__restrict__ vel,
                                       const float4 *__restrict__ np, const Lists<T> &L, int k, float4 *pj,
                                       float4 *vj, float4 *nj) {
#pragma unroll
    for (int q = 0; q < kPairBatch; ++q) {
        const uint32_t off = entry_bytes<T>(L.base, L.list[(k + q) * T]);
        pj[q] = at(pos, off);
        vj[q] = at(vel, off);
        if (KIND == kFluidSolid) nj[q] = at(np, off);
    }
}

// FP16 records and cells of list entries k .. k+kPairBatch-1 (packed form)
template <int T>
__device__ __forceinline__ void gather_packed(const uint4 *__restrict__ pk, const int32_t *__restrict__ pcell,
                                              const Lists<T> &L, int k, int cnt, uint4 *rq, int *cq, int *jq) {
#pragma unroll
    for (int q = 0; q < kPairBatch; ++q) {
        const int j = (int)(entry_bytes<T>(L.base, L.list[min(k + q, cnt - 1) * T]) >> 4);
        rq[q] = pk[j];
        cq[q] = pcell[j];
        jq[q] = j;
    }
}

// Packed partner records of the kind's sources (null in the FP32 form)
struct PackedSrc {
    const uint4 *pk;
    const int32_t *cell;
};

// pair terms of list entries [0, cnt), in list order.  The partner state of
// the next kPairBatch entries is gathered while the current ones are
// evaluated; all kPairBatch slots of a round are evaluated, the ones past
// cnt with weight 0 (the sentinel partner; in the packed form the last entry
// again, with d2 = kPadD2).  PACK: fluid partners come from their FP16
// records.
template <int KIND, int T, bool PACK, bool COUNT>
__device__ __forceinline__ void evaluate(const DevParams &P, const Self &me, const float4 *__restrict__ pos,
                                         const float4 *__restrict__ vel, const float4 *__restrict__ np,
                                         const PackedSrc &ps, const Lists<T> &L, int cnt, Acc &acc, int &found) {
    if (COUNT) found += cnt;
    if (__ballot(cnt > 0) == 0) return;
    if constexpr (PACK && KIND == kFluidFluid) {
        uint4 rq[kPairBatch];
        int cq[kPairBatch], jq[kPairBatch];
        if (cnt > 0) gather_packed<T>(ps.pk, ps.cell, L, 0, cnt, rq, cq, jq);
        for (int k = 0; k < cnt; k += kPairBatch) {
            uint4 rn[kPairBatch];
            int cn[kPairBatch], jn[kPairBatch];
            gather_packed<T>(ps.pk, ps.cell, L, k + kPairBatch, cnt, rn, cn, jn);
#pragma unroll
            for (int q = 0; q < kPairBatch; ++q)
                one_pair_packed(P, me, rq[q], cq[q], jq[q], k + q < cnt, acc);
#pragma unroll
            for (int q = 0; q < kPairBatch; ++q) {
                rq[q] = rn[q];
                cq[q] = cn[q];
                jq[q] = jn[q];
            }
        }
        return;
    }
#pragma unroll
    for (int q = 0; q < kListPad; ++q) L.list[(cnt + q) * T] = kSentinelEntry;
    float4 pj[kPairBatch], vj[kPairBatch], nj[kPairBatch];
    gather<KIND, T>(pos, vel, np, L, 0, pj, vj, nj);
    for (int k = 0; k < cnt; k += kPairBatch) {
        float4 pn[kPairBatch], vn[kPairBatch], nn[kPairBatch];
        gather<KIND, T>(pos, vel, np, L, k + kPairBatch, pn, vn, nn);
#pragma unroll
        for (int q = 0; q < kPairBatch; ++q) one_pair<KIND>(P, me, pj[q], vj[q], nj[q], acc);
#pragma unroll
        for (int q = 0; q < kPairBatch; ++q) {
            pj[q] = pn[q];
            vj[q] = vn[q];
            nj[q] = nn[q];
        }
    }
}

// candidate test over slots [b, e) of column `col` of one kind, appending
// hits to the list.  Base slot `col` holds b (written by the kind's pass);
// a run longer than kRunMax slots continues in spare base slots.  Control
// flow is uniform across the wave: the batch loop runs while any lane still
// has candidates (a lane past its run re-tests the padded tail at the run
// end, masked out), and a full list or a full spare table is evaluated by
// all lanes together.  The next batch of candidates is loaded while the
// current one is tested.
template <int KIND, int T, bool PACK, bool COUNT>
__device__ __forceinline__ void scan(const DevParams &P, const Self &me, int col, int b, int e,
                                     const float4 *__restrict__ pos, const float4 *__restrict__ vel,
                                     const float4 *__restrict__ np, const PackedSrc &ps, const Lists<T> &L, int &cnt,
                                     int &spare, Acc &acc, int &found) {
    int sb = b;
    for (int sub = 0; __ballot(sb < e) != 0; ++sub, sb += kRunMax) {
        const bool has = sb < e;
        int slot = col;
        if (sub > 0) {
            if (__ballot(has && spare == kSpareSlots) != 0) {
                evaluate<KIND, T, PACK, COUNT>(P, me, pos, vel, np, ps, L, cnt, acc, found);
                cnt = 0;
                spare = 0;
            }
            if (has) {
                // a spare slot, or without spares the column's own one, free after the evaluation
                if (kSpareSlots > 0) slot = kColumns + spare++;
                L.base[slot * T] = slot_bytes(sb);
            }
        }
        const int se = min(sb + kRunMax, e);  // <= sb for a lane without work
        // entry = tag + 16 * (slot index of the partner), in 16 bits
        const uint32_t tag = (uint32_t)slot - ((uint32_t)sb << 4);
        // branch-free batches: loads at or past the run end stay inside the
        // padded array (se <= n, kPosPad >= kScanBatch) and are masked out of
        // the test; every candidate writes list slot cnt, which only a hit
        // keeps (cnt <= kListCap-1 here: see the evaluation below)
        float4 q[kScanBatch];
        {
            const float4 *src = pos + min(sb, se);
#pragma unroll
            for (int k = 0; k < kScanBatch; ++k) q[k] = src[k];
        }
        for (int j = sb; __ballot(j < se) != 0; j += kScanBatch) {
            const int jj = min(j, se);
            const int lim = se - jj;  // candidates k < lim of the batch lie in the run
            const uint32_t e0 = tag + ((uint32_t)jj << 4);
            float4 qn[kScanBatch];
            const float4 *nxt = pos + min(j + kScanBatch, se);
#pragma unroll
            for (int k = 0; k < kScanBatch; ++k) qn[k] = nxt[k];
            // the write position advances by one entry (T lanes) per hit: a
            // select and an add per candidate, no address arithmetic
            uint16_t *wp = L.list + cnt * T;
#pragma unroll
            for (int k = 0; k < kScanBatch; ++k) {
                const float d2 = len2(q[k].x - me.x, q[k].y - me.y, q[k].z - me.z);
                const bool hit = (k < lim) & (d2 < P.R2);
                *wp = (uint16_t)(e0 + 16u * k);
                wp += hit ? T : 0;
            }
            // (the low 32 bits of an LDS pointer are its LDS byte address)
            cnt = (int)(((uint32_t)(uintptr_t)wp - (uint32_t)(uintptr_t)L.list) / (2u * T));
#pragma unroll
            for (int k = 0; k < kScanBatch; ++k) q[k] = qn[k];
            // at most kScanBatch entries are added per round; the bases stay
            // valid across an evaluation
            if (__ballot(cnt > kListCap - kScanBatch) != 0) {
                evaluate<KIND, T, PACK, COUNT>(P, me, pos, vel, np, ps, L, cnt, acc, found);
                cnt = 0;
            }
        }
    }
}

// all partners of one kind (n particles): the 9 columns in order, then one
// evaluation; fluid and wall partners are read from each column's z-window
// only (solid partners, sorted by cell alone, from the whole column).  The
// slot ranges of all 9 columns are taken first (their loads in flight
// together); a kind with no candidate for any lane of the wave is skipped,
// and so is a column without one.  Solid and wall particles are far from
// most fluid particles.  The solid's slots are sorted by cell id, whose
// leading digit is the cell's x index, so its first and last slots bound the
// solid in x; the walls have a per-cell reach flag.  Either decides without
// loads of column ranges; then the whole-column ranges do, and only a wave
// with walls in reach takes their z-windows.
template <int KIND, int T, bool PACK, bool COUNT>
__device__ __forceinline__ void pass(const DevParams &P, const Self &me, const Columns &cols, const Window &w,
                                     int n, const int32_t *__restrict__ start,
                                     const int32_t *__restrict__ fine, const float4 *__restrict__ pos,
                                     const float4 *__restrict__ vel, const float4 *__restrict__ np,
                                     const PackedSrc &ps, const Lists<T> &L, Acc &acc, int &found,
                                     const int32_t *__restrict__ sorted_cells = nullptr,
                                     const uint8_t *__restrict__ reach = nullptr, bool active = true) {
    if (KIND == kFluidSolid && sorted_cells) {
        const int nyz = P.cn[1] * P.cn[2];
        const int lo = sorted_cells[0] / nyz, hi = sorted_cells[n - 1] / nyz;
        if (__ballot(cols.cx + 1 >= lo && cols.cx - 1 <= hi) == 0) return;
    }
    if (KIND == kFluidVirt && reach && __ballot(active && (reach[me.cell] & kReachNear) != 0) == 0) return;
    int cb[kColumns], ce[kColumns];
    if (KIND != kFluidFluid) {
        column_ranges(cols, start, cb, ce);
        bool any = false;
#pragma unroll
        for (int col = 0; col < kColumns; ++col) any |= cb[col] < ce[col];
        if (__ballot(active && any) == 0) return;
    }
    if (KIND != kFluidSolid) {
#pragma unroll
        for (int col = 0; col < kColumns; ++col) z_window(P, cols, fine, col, w, P.R2w, cb[col], ce[col]);
    }
    {
        // (lanes that are not active, e.g. listed particles, take no part)
        bool any = false;
#pragma unroll
        for (int col = 0; col < kColumns; ++col) {
            if (!active) ce[col] = cb[col];
            any |= cb[col] < ce[col];
        }
        if (__ballot(any) == 0) return;
    }
    int cnt = 0, spare = 0;
    L.base[kSentinelSlot * T] = slot_bytes(n);
#pragma unroll 1
    for (int col = 0; col < kColumns; ++col) {
        const int b = cb[col], e = ce[col];
        if (__ballot(b < e) == 0) continue;
        L.base[col * T] = slot_bytes(b);
        scan<KIND, T, PACK, COUNT>(P, me, col, b, e, pos, vel, np, ps, L, cnt, spare, acc, found);
    }
    evaluate<KIND, T, PACK, COUNT>(P, me, pos, vel, np, ps, L, cnt, acc, found);
}

// Fluid (or wall) partners from the particle's neighbour list (home
// particle i; entries are byte offsets into pos and vel, `sentinel` the
// offset of the kind's sentinel partner): every listed candidate
// is evaluated in list order (the walk order of the build) and counts only
// inside the cut-off.  About nine in ten listed fluid candidates are
// partners, so this costs less than the walk's test-then-evaluate, and each
// candidate's state is read once.  Rows past the lane's list read the
// sentinel partner.  Three rounds are in flight: the entries of the round
// after next, the partner state of the next, the current one's arithmetic.
template <int KIND, bool COUNT>
__device__ __forceinline__ void list_pass(const DevParams &P, const Self &me, int i, int cnt,
                                          const uint32_t *__restrict__ list, const float4 *__restrict__ pos,
                                          const float4 *__restrict__ vel, uint32_t sentinel, Acc &acc,
                                          int &found) {
    if (__ballot(cnt > 0) == 0) return;
    const size_t nf = (size_t)P.nf;
    const uint32_t *__restrict__ ent = list + i;
    uint32_t en[kListBatch];
    float4 pc[kListBatch], vc[kListBatch];
#pragma unroll
    for (int q = 0; q < kListBatch; ++q) {
        const uint32_t e = ent[(size_t)q * nf];
        en[q] = ent[(size_t)(kListBatch + q) * nf];
        const uint32_t off = q < cnt ? e : sentinel;
        pc[q] = at(pos, off);
        vc[q] = at(vel, off);
    }
    int hits = 0;
    for (int k = 0; __ballot(k < cnt) != 0; k += kListBatch) {
        const int kn = k + kListBatch;
        float4 pn[kListBatch], vn[kListBatch];
#pragma unroll
        for (int q = 0; q < kListBatch; ++q) {
            const uint32_t off = kn + q < cnt ? en[q] : sentinel;
            pn[q] = at(pos, off);
            vn[q] = at(vel, off);
        }
#pragma unroll
        for (int q = 0; q < kListBatch; ++q) en[q] = ent[(size_t)(kn + kListBatch + q) * nf];
#pragma unroll
        for (int q = 0; q < kListBatch; ++q) listed_pair<KIND>(P, me, pc[q], vc[q], acc, hits);
#pragma unroll
        for (int q = 0; q < kListBatch; ++q) {
            pc[q] = pn[q];
            vc[q] = vn[q];
        }
    }
    if (COUNT) found += hits;
}

// With neighbour lists in use (P.vlist, never with PACK, and the step not a
// walking one) one thread per home particle, else one per sorted slot;
// outputs in sorted order go to slot o.  Whether lists are in use is decided
// at run time, so that the fluid walk (for every particle, or for the
// unlisted ones) has one call site: a second inlined copy of it spills
// registers.  A walking step writes no home state: the next build makes
// the sorted order home again.
template <bool RATES, int T, bool PACK>
__device__ __forceinline__ void fluid_body(const DevParams &P, const FluidStepArgs &a, const Lists<T> &list,
                                           uint32_t *s_wave, int blk, int nblk) {
    const int i = xcd_tile(blk, nblk) * T + threadIdx.x;
    if (i >= P.nf) return;  // lists are lane-private: no barrier follows
    const bool LIST = !PACK && P.vlist != 0 && a.L.flag[kFlagMode] != kModeWalk;

    const float4 pi4 = LIST ? a.hrec[2 * (size_t)i] : a.fpos[i];
    const float4 vi4 = LIST ? a.hrec[2 * (size_t)i + 1] : a.fvel[i];
    Self me;
    me.x = pi4.x;
    me.y = pi4.y;
    me.z = pi4.z;
    me.rho = pi4.w;
    me.ux = vi4.x;
    me.uy = vi4.y;
    me.uz = vi4.z;
    me.p = vi4.w;
    me.inv_rho = 1.0f / me.rho;
    me.two_rho = 2.0f * me.rho;
    me.m2_inv_rho = -2.0f * me.inv_rho;
    me.vis5 = P.viscid ? 0.5f * P.visc10 : 0.0f;
    me.half_inv_rho = 0.5f * me.inv_rho;
    Acc acc{0.0f, 0.0f, 0.0f, 0.0f};
    // (the cell the sort found: the hash took it from this same position)
    const int cell = LIST ? find_cell(P, me.x, me.y, me.z) : a.fcell[i];
    me.i = i;
    me.cell = cell;
    me.fx = me.fy = me.fz = 0.0f;
    if (PACK) {
        const int3 ci = cell_coords(P, cell);
        me.fx = cell_offset(P, me.x, 0, ci.x);
        me.fy = cell_offset(P, me.y, 1, ci.y);
        me.fz = cell_offset(P, me.z, 2, ci.z);
    }

    // partners kind by kind: fluid, solid, wall (the canonical order)
    // (the cell's columns and the particle's place in its cell serve all three kinds)
    const PackedSrc fsrc{a.fpk, a.fcell}, none{nullptr, nullptr};
    const Columns cols(P, cell);
    const Window w(P, cols, me.x, me.y, me.z);
    // (debug runs count each kind's partners: the neighbour sets the walk found)
    int found[3] = {0, 0, 0};
    // listed particles test their lists; the others (all of them without
    // lists) walk their cells, over the sorted state of this step (the same
    // values as the home state)
    const int meta = LIST ? a.L.meta[i] : -1;
    // (home records: position and velocity 16 bytes apart, entries of 32-byte records)
    if (!PACK && LIST)
        list_pass<kFluidFluid, RATES>(P, me, i, meta, a.L.ent, a.hrec, a.hrec + 1, (uint32_t)P.nf << 5, acc,
                                      found[0]);
    if (__ballot(meta < 0) != 0)
        pass<kFluidFluid, T, PACK, RATES>(P, me, cols, w, P.nf, a.fstart, a.ffine, a.fpos, a.fvel, nullptr, fsrc,
                                          list, acc, found[0], nullptr, nullptr, meta < 0);
    pass<kFluidSolid, T, false, RATES>(P, me, cols, w, P.ns, a.sstart, nullptr, a.spos, a.svel, a.snp, none, list,
                                       acc, found[1], a.scell);
    // wall partners likewise (their lists are built with the fluid's)
    const int wmeta = LIST && a.L.went ? a.L.wmeta[i] : -1;
    if (!PACK && LIST && a.L.went)
        list_pass<kFluidVirt, RATES>(P, me, i, wmeta, a.L.went, a.vpos, a.vvel, slot_bytes(P.nv), acc, found[2]);
    if (__ballot(wmeta < 0) != 0)
        pass<kFluidVirt, T, false, RATES>(P, me, cols, w, P.nv, a.vstart, a.vfine, a.vpos, a.vvel, nullptr, none,
                                          list, acc, found[2], nullptr, a.vreach, wmeta < 0);
    const int o = LIST ? a.hslot[i] : i;
    if (RATES) a.counts[o] = make_int4(found[0], found[1], found[2], 0);

    // fluidUpdate (fluid.cu:234-269): NaN scrub, explicit step with the old
    // velocity for the position, density clamp, EOS, domain clamp, speed cap
    float drho = me.two_rho * acc.drho, ax = me.m2_inv_rho * acc.ax, ay = me.m2_inv_rho * acc.ay,
          az = me.m2_inv_rho * acc.az;
    if (!(drho == drho)) drho = 0.0f;
    if (!(ax == ax && ay == ay && az == az)) ax = ay = az = 0.0f;
    if (RATES) a.rates[o] = make_float4(ax, ay, az, drho);
    const float dt = P.dt;
    float x, y, z, ux, uy, uz;
    float4 cp = make_float4(0.0f, 0.0f, 0.0f, 0.0f), cv = cp;
    if (a.fcpos) {
        // compensated form: x and u carry their rounding error (cp, cv)
        cp = a.fcpos[o];
        cv = a.fcvel[o];
        const float sx = me.ux * dt, sy = me.uy * dt, sz = me.uz * dt;
        x = kahan_add(me.x, sx, cp.x);
        y = kahan_add(me.y, sy, cp.y);
        z = kahan_add(me.z, sz, cp.z);
        const float vx = (ax + P.gx) * dt, vy = (ay + P.gy) * dt, vz = (az + P.gz) * dt;
        ux = kahan_add(me.ux, vx, cv.x);
        uy = kahan_add(me.uy, vy, cv.y);
        uz = kahan_add(me.uz, vz, cv.z);
    } else {
        x = me.x + me.ux * dt;
        y = me.y + me.uy * dt;
        z = me.z + me.uz * dt;
        ux = me.ux + (ax + P.gx) * dt;
        uy = me.uy + (ay + P.gy) * dt;
        uz = me.uz + (az + P.gz) * dt;
    }
    float rho = me.rho + drho * dt;
    rho = fminf(fmaxf(rho, P.rho_min), P.rho_max);
    // Tait equation of state (densityToPressure, fluid.cu:16-24)
    const float p = fmaxf((powf(rho * P.rho0_inv, P.gamma) - 1.0f) * P.rho2p, P.p_min);
    // a clamped or rescaled value drops its carried error
    const float xc = fminf(fmaxf(x, P.dmin[0]), P.dmax[0]);
    const float yc = fminf(fmaxf(y, P.dmin[1]), P.dmax[1]);
    const float zc = fminf(fmaxf(z, P.dmin[2]), P.dmax[2]);
    if (xc != x) cp.x = 0.0f;
    if (yc != y) cp.y = 0.0f;
    if (zc != z) cp.z = 0.0f;
    const float speed = sqrtf(len2(ux, uy, uz));
    if (speed > P.max_speed) {
        const float f = P.max_speed / speed;
        ux *= f;
        uy *= f;
        uz *= f;
        cv = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    // a non-finite velocity survives every clamp above (the scrub covers the
    // pair sums only): the scene has diverged
    if (!(__builtin_isfinite(ux) && __builtin_isfinite(uy) && __builtin_isfinite(uz)))
        atomicOr(a.health, (unsigned long long)FSA_HEALTH_FLUID_NONFINITE);
    const float4 xo = make_float4(xc, yc, zc, rho), uo = make_float4(ux, uy, uz, p);
    a.fpos_out[o] = xo;
    a.fvel_out[o] = uo;
    if (a.fcpos) {
        a.fcpos_out[o] = cp;
        a.fcvel_out[o] = cv;
    }
    if (LIST) {
        a.hrec_out[2 * (size_t)i] = xo;
        a.hrec_out[2 * (size_t)i + 1] = uo;
        // the wave's largest squared move since the last list build, for the
        // next build decision (float bits of a nonnegative value or NaN:
        // unsigned order); through the wave's first lane's base slot, which
        // is free now.  Lane 0 is active whenever any lane of the wave is.
        const float4 x0 = a.L.x0[i];
        const uint32_t m2 = __float_as_uint(len2(xc - x0.x, yc - x0.y, zc - x0.z));
        const bool lead = (threadIdx.x & 63) == 0;
        if (lead) *s_wave = 0u;
        __builtin_amdgcn_wave_barrier();
        atomicMax(s_wave, m2);
        __builtin_amdgcn_wave_barrier();
        if (lead) atomicMax(&a.L.move[(i >> 6) & (kMoveWords - 1)], *s_wave);
    }
}

template <bool RATES, int T, bool PACK>
struct OpFluid {
    static __device__ void run(const FluidJob &j, int blk, int nblk) {
        __shared__ uint16_t s_list[(kListCap + kListPad) * T];
        __shared__ uint32_t s_base[kBaseSlots * T];
        const Lists<T> list{s_list + threadIdx.x, s_base + threadIdx.x};
        uint32_t *s_wave = s_base + (threadIdx.x & ~63u);  // base slot 0 of the wave's first lane
        fluid_body<RATES, T, PACK>(j.P, j.a, list, s_wave, blk, nblk);
    }
};

}  // namespace

bool fluid_threads_valid(int threads) {
    for (int t : kThreadChoices)
        if (t == threads) return true;
    return false;
}

int blocks_fluid(const FluidJob &j, int threads) { return (int)(((long long)j.P.nf + threads - 1) / threads); }

namespace {
template <int T, bool PACK>
hipError_t one_fluid(const FluidJob &j, hipStream_t s) {
    const dim3 grid(blocks_fluid(j, T));
    if (j.a.rates)
        hipLaunchKernelGGL((k_one<OpFluid<true, T, PACK>, FluidJob, T, fluid_waves(T)>), grid, dim3(T), 0, s, j);
    else
        hipLaunchKernelGGL((k_one<OpFluid<false, T, PACK>, FluidJob, T, fluid_waves(T)>), grid, dim3(T), 0, s, j);
    return hipGetLastError();
}
template <int T, bool PACK>
hipError_t many_fluid(const FluidJob *jobs, const int *prefix, int njobs, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL((k_many<OpFluid<false, T, PACK>, FluidJob, T, fluid_waves(T)>), dim3(nblocks), dim3(T), 0, s, jobs, prefix,
                       njobs);
    return hipGetLastError();
}
template <bool PACK>
hipError_t one_fluid_t(const FluidJob &j, int threads, hipStream_t s) {
    switch (threads) {
    case 64: return one_fluid<64, PACK>(j, s);
    case 128: return one_fluid<128, PACK>(j, s);
    case 256: return one_fluid<256, PACK>(j, s);
    default: return hipErrorInvalidValue;
    }
}
template <bool PACK>
hipError_t many_fluid_t(const FluidJob *jobs, const int *prefix, int njobs, int nblocks, int threads, hipStream_t s) {
    switch (threads) {
    case 64: return many_fluid<64, PACK>(jobs, prefix, njobs, nblocks, s);
    case 128: return many_fluid<128, PACK>(jobs, prefix, njobs, nblocks, s);
    case 256: return many_fluid<256, PACK>(jobs, prefix, njobs, nblocks, s);
    default: return hipErrorInvalidValue;
    }
}
}  // namespace

hipError_t launch_fluid(const DevParams &P, const FluidStepArgs &a, int threads, hipStream_t s) {
    if (P.nf == 0) return hipSuccess;
    const FluidJob j{P, a};
    return (P.form & FSA_FORM_PACKED) ? one_fluid_t<true>(j, threads, s) : one_fluid_t<false>(j, threads, s);
}

hipError_t launch_many_fluid(const FluidJob *jobs, const int *prefix, int njobs, int nblocks, int threads,
                             bool packed, hipStream_t s) {
    if (nblocks <= 0 || njobs <= 0) return hipSuccess;
    return packed ? many_fluid_t<true>(jobs, prefix, njobs, nblocks, threads, s)
                  : many_fluid_t<false>(jobs, prefix, njobs, nblocks, threads, s);
}

}  // namespace fsa
