// fsa_solid.hip -- the solid side: coupling fields (K6), total-Lagrangian
// deformation and stress (K10), forces (K11) and the sub-step integrator
// (K12); src/solid.cu:10-359.
//
// Neighbours are found in the cell lists of the step (built from
// start-of-step positions) and, for solid-solid terms, filtered by
// reference-configuration distance, exactly as the reference walks its lists
// (src/solid.cu:53-69, 195-203).
//
// A gate has only tens of thousands of particles, too few for one thread per
// particle to fill 256 CUs, and each particle walks ~500 candidates.  So a
// group of kSub = 16 lanes serves one particle: the lanes take the candidates
// of every run round-robin, and the partial sums are combined by a fixed
// butterfly, after which lane 0 of the group holds the sum and writes.
// Matrices live in 9 SoA planes; the per-particle factors of the pair force
// that do not depend on the partner (P_i L_i and det(F_i) F_i^-T) are
// computed once in the deformation kernel instead of once per pair.
#include "fsa_internal.h"
#include "fsa_math.h"

namespace fsa {

namespace {

constexpr int kSub = kSolidLanes;

struct M3 {
    float m[3][3];  // row-major: m[r][c]
};

__device__ __forceinline__ M3 mzero() {
    M3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[i][j] = 0.0f;
    return r;
}
__device__ __forceinline__ M3 mident() {
    M3 r = mzero();
    r.m[0][0] = r.m[1][1] = r.m[2][2] = 1.0f;
    return r;
}
__device__ __forceinline__ M3 mmul(const M3 &a, const M3 &b) {
    M3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
    return r;
}
__device__ __forceinline__ float mdet(const M3 &a) {
    const auto &m = a.m;
    return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
           m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}
// adjugate / determinant
__device__ __forceinline__ M3 minverse(const M3 &a) {
    const auto &m = a.m;
    const float c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    const float c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    const float c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    const float inv_det = 1.0f / (m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02);
    M3 r;
    r.m[0][0] = c00 * inv_det;
    r.m[1][0] = c01 * inv_det;
    r.m[2][0] = c02 * inv_det;
    r.m[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * inv_det;
    r.m[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * inv_det;
    r.m[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * inv_det;
    r.m[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * inv_det;
    r.m[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * inv_det;
    r.m[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * inv_det;
    return r;
}
__device__ __forceinline__ M3 load9(const float *__restrict__ base, int n, int i) {
    M3 r;
#pragma unroll
    for (int e = 0; e < 9; ++e) r.m[e / 3][e % 3] = base[(size_t)e * n + i];
    return r;
}
__device__ __forceinline__ void store9(float *__restrict__ base, int n, int i, const M3 &a) {
#pragma unroll
    for (int e = 0; e < 9; ++e) base[(size_t)e * n + i] = a.m[e / 3][e % 3];
}

// sum over the kSub lanes of a group; lane 0 of the group always combines
// in the same order, so its result is deterministic
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
    for (int d = kSub / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d, kSub);
    return v;
}

struct Lane {
    int i;    // solid slot served by this group
    int sub;  // lane within the group
};
__device__ __forceinline__ Lane lane_of(int blk) {
    const int t = blk * kBlock + threadIdx.x;
    return {t / kSub, t % kSub};
}

// ---- K6: solidCoupleUpdate (src/solid.cu:262-344) ------------------------
__device__ __forceinline__ void couple_body(const DevParams &P, const SolidArgs &a, int blk) {
    const Lane L = lane_of(blk);
    if (L.i >= P.ns) return;  // whole groups leave together
    const float4 xi4 = a.spos[L.i];
    const float xi = xi4.x, yi = xi4.y, zi = xi4.z;
    float nx = 0.0f, ny = 0.0f, nz = 0.0f, pressure = 0.0f, wsum = 0.0f;
    // fluid candidates within R_rec, recorded for the force passes of the step
    int32_t *rec = a.x.fnbr + (size_t)L.i * (kSolidListCap * kSub) + L.sub;
    int nrec = 0;
    for_each_column(P, a.scell[L.i], [&](int c0, int c1) {
        // Shepard pressure from fluid (solidPressureCalculate, solid.cu:275-291)
        for (int j = a.fstart[c0] + L.sub, e = a.fstart[c1]; j < e; j += kSub) {
            const float4 pj = a.fpos[j];
            const float dx = pj.x - xi, dy = pj.y - yi, dz = pj.z - zi;
            const float d2 = len2(dx, dy, dz);
            if (d2 < P.R2_rec) {
                if (nrec < kSolidListCap) rec[nrec * kSub] = j;
                ++nrec;
            }
            if (d2 < P.R2) {
                const float w = (expf(-(d2 * P.h_inv2)) - P.c1) * P.c2;
                wsum += w;
                pressure += a.fvel[j].w * w;
            }
        }
        // normal from solid neighbours (solidNormalCalculate, solid.cu:262-273)
        for (int j = a.sstart[c0] + L.sub, e = a.sstart[c1]; j < e; j += kSub) {
            const float4 pj = a.spos[j];
            const float dx = pj.x - xi, dy = pj.y - yi, dz = pj.z - zi;
            const float d2 = len2(dx, dy, dz);
            if (d2 < P.R2) {
                const float kc = expf(-(d2 * P.h_inv2)) * P.kd;
                nx -= dx * kc;
                ny -= dy * kc;
                nz -= dz * kc;
            }
        }
    });
    nx = group_sum(nx);
    ny = group_sum(ny);
    nz = group_sum(nz);
    pressure = group_sum(pressure);
    wsum = group_sum(wsum);
    a.x.fcnt[(size_t)L.i * kSub + L.sub] = nrec;
    if (L.sub != 0) return;
    a.x.x0[L.i] = xi4;
    const float n2 = len2(nx, ny, nz);
    if (n2 < kEpsForSum) {
        nx = ny = nz = 0.0f;
    } else {
        const float r = 1.0f / sqrtf(n2);
        nx *= r;
        ny *= r;
        nz *= r;
    }
    const float p = wsum > kEpsFloat ? fmaxf(pressure / wsum, P.p_min) : 0.0f;
    a.x.np[L.i] = make_float4(nx, ny, nz, p);
}

// ---- solid partners by reference distance ----------------------------------
// The solid-solid terms (src/solid.cu:53-69, 195-203) take the solids j in
// the 27 cells of i's cell with |X_j - X_i|^2 < R^2.  Lane `sub` of i's group
// takes the candidates sub, sub + kSub, ... of every column run.  The cells
// are fixed for the whole step, so this set and order are the same in every
// sub-step: the first deformation pass of a step walks the candidates and
// records each lane's partners (kWalkRecord), later passes replay the record
// (kReplay).  A lane with more than kSolidListCap partners walks again, so
// each lane visits the same partners in the same order on either path and
// the sums have the same bits.  (Gathering 2 or 4 recorded partners per
// round was measured slower: the extra registers cost more occupancy than
// the overlapped loads gained.)
enum PartnerMode { kWalkRecord = 0, kReplay = 1 };

// (a partner's reference position X_j is read at Xsrc[XSTRIDE * j])
template <int MODE, int XSTRIDE, class F>
__device__ __forceinline__ void solid_partners(const DevParams &P, const SolidArgs &a, int i, int sub,
                                               const float4 &Xi, const float4 *__restrict__ Xsrc, F &&body) {
    int32_t *rec = a.x.nbr + (size_t)i * (kSolidListCap * kSub) + sub;
    int32_t *rec_n = a.x.ncnt + (size_t)i * kSub + sub;
    if (MODE == kReplay && a.solid_rec) {
        const int cnt = *rec_n;
        if (cnt <= kSolidListCap) {
            // the next partner's index is loaded while this one is evaluated
            int jn = cnt > 0 ? rec[0] : 0;
            for (int k = 0; k < cnt; ++k) {
                const int j = jn;
                if (k + 1 < cnt) jn = rec[(k + 1) * kSub];
                const float4 Xj = Xsrc[XSTRIDE * (size_t)j];
                const float dX[3] = {Xj.x - Xi.x, Xj.y - Xi.y, Xj.z - Xi.z};
                body(j, dX, len2(dX[0], dX[1], dX[2]));
            }
            return;
        }
    }
    int cnt = 0;
    for_each_column(P, a.scell[i], [&](int c0, int c1) {
        for (int j = a.sstart[c0] + sub, e = a.sstart[c1]; j < e; j += kSub) {
            const float4 Xj = Xsrc[XSTRIDE * (size_t)j];
            const float dX[3] = {Xj.x - Xi.x, Xj.y - Xi.y, Xj.z - Xi.z};
            const float d2 = len2(dX[0], dX[1], dX[2]);
            if (d2 < P.R2) {
                if (MODE == kWalkRecord && cnt < kSolidListCap) rec[cnt * kSub] = j;
                ++cnt;
                body(j, dX, d2);
            }
        }
    });
    if (MODE == kWalkRecord) *rec_n = cnt;
}

// ---- K10: solidDeformationCalculate (src/solid.cu:10-79) -----------------
// The per-particle part, from the particle's sums F = sum dx dX^T k and B =
// sum dX dX^T k (or the step's L): the corrected deformation gradient, the
// stress, and what the force pass reads.
__device__ __forceinline__ void deformation_tail(const DevParams &P, const SolidArgs &a, int i, M3 F, const M3 &B,
                                                 bool reuse_L) {
    const float4 Xi = a.sref[i];
    const float4 xi = a.spos[i];
    const M3 Lc = reuse_L ? load9(a.x.L, P.ns, i) : (mdet(B) != 0.0f ? minverse(B) : mident());
    F = mmul(F, Lc);
    // St Venant-Kirchhoff (solidStressCalculation, solid.cu:27-36)
    M3 E;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float ftf = F.m[0][r] * F.m[0][c] + F.m[1][r] * F.m[1][c] + F.m[2][r] * F.m[2][c];
            E.m[r][c] = 0.5f * (ftf - (r == c ? 1.0f : 0.0f));
        }
    const float tr = E.m[0][0] + E.m[1][1] + E.m[2][2];
    M3 S;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) S.m[r][c] = 2.0f * P.mu * E.m[r][c] + (r == c ? P.lambda * tr : 0.0f);
    const M3 Pk = mmul(F, S);
    const int n = P.ns;
    store9(a.x.F, n, i, F);
    store9(a.x.P, n, i, Pk);
    if (!reuse_L) store9(a.x.L, n, i, Lc);
    {
        const M3 PL = mmul(Pk, Lc);
        float4 *rec = a.x.frec + kForceRec * (size_t)i;
        rec[0] = Xi;
        rec[1] = xi;
        rec[2] = a.svel[i];
#pragma unroll
        for (int r = 0; r < 3; ++r) rec[3 + r] = make_float4(PL.m[r][0], PL.m[r][1], PL.m[r][2], 0.0f);
    }
    // det(F) F^-T L, the particle's factor of the artificial-viscosity force
    // (whose pair term is ((det F F^-T visc) L) dX k, solid.cu:156: visc
    // and k are scalars of the pair, so the matrix product is the particle's)
    const M3 Fi = minverse(F);
    const float dF = mdet(F);
    M3 A;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) A.m[r][c] = dF * Fi.m[c][r];
    store9(a.x.A, n, i, mmul(A, Lc));
}

// The pair sums, kSub lanes per particle.  SPLIT (batched launches): the
// group's sums go to dsum (F in planes 0-8, B in 9-17) and OpDeformTail
// finishes, one thread per particle; otherwise lane 0 finishes here, while
// the group's other lanes wait.  The same arithmetic either way: the same bits.
template <int MODE, bool SPLIT>
__device__ __forceinline__ void deformation_body(const DevParams &P, const SolidArgs &a, int blk) {
    const Lane L = lane_of(blk);
    if (L.i >= P.ns) return;
    const int i = L.i;
    const float4 Xi = a.sref[i];
    const float4 xi = a.spos[i];
    // The correction matrix L = B^-1, B = sum dX dX^T k, depends on the
    // reference positions and the partner order only, both fixed within a
    // step: a replaying pass (the later sub-steps) reads the L the step's
    // first pass stored, the same bits as summing B again.
    const bool reuse_L = MODE == kReplay && a.solid_rec != 0;
    M3 F = mzero(), B = mzero();
    solid_partners<MODE, 1>(P, a, i, L.sub, Xi, a.sref, [&](int j, const float (&dX)[3], float d2) {
        const float4 xj = a.spos[j];
        const float dx[3] = {xj.x - xi.x, xj.y - xi.y, xj.z - xi.z};
        const float kc = hw_exp2(fmaf(-d2, P.kexp2, P.log2kd));  // kd exp(-|dX|^2/h^2)
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) F.m[r][c] += (dx[r] * dX[c]) * kc;
        if (!reuse_L) {
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) B.m[r][c] += (dX[r] * dX[c]) * kc;
        }
    });
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) F.m[r][c] = group_sum(F.m[r][c]);
    if (!reuse_L) {
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) B.m[r][c] = group_sum(B.m[r][c]);
    }
    if (L.sub != 0) return;
    if (SPLIT) {
        store9(a.x.dsum, P.ns, i, F);
        if (!reuse_L) store9(a.x.dsum + 9 * (size_t)P.ns, P.ns, i, B);
        return;
    }
    deformation_tail(P, a, i, F, B, reuse_L);
}

// the per-particle part of a split deformation pass, one thread per particle
template <int MODE>
__device__ __forceinline__ void deformation_tail_body(const DevParams &P, const SolidArgs &a, int blk) {
    const int i = blk * kBlock + threadIdx.x;
    if (i >= P.ns) return;
    const bool reuse_L = MODE == kReplay && a.solid_rec != 0;
    const M3 F = load9(a.x.dsum, P.ns, i);
    const M3 B = reuse_L ? mzero() : load9(a.x.dsum + 9 * (size_t)P.ns, P.ns, i);
    deformation_tail(P, a, i, F, B, reuse_L);
}

// ---- K11: solidForceCalculate (src/solid.cu:81-207) ----------------------
__device__ __forceinline__ void force_body(const DevParams &P, const SolidArgs &a, int blk) {
    const Lane L = lane_of(blk);
    if (L.i >= P.ns) return;
    const int i = L.i;
    const int n = P.ns;
    const float4 xi = a.spos[i], ui = a.svel[i], Xi = a.sref[i];
    const float4 np = a.x.np[i];
    M3 PLi;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const float4 row = a.x.frec[kForceRec * (size_t)i + 3 + r];
        PLi.m[r][0] = row.x;
        PLi.m[r][1] = row.y;
        PLi.m[r][2] = row.z;
    }
    const M3 ALi = load9(a.x.A, n, i);  // det(F_i) F_i^-T L_i
    float cx = 0.0f, cy = 0.0f, cz = 0.0f;  // couple_dudt
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    const float eps = 0.01f * P.h2;
    // fluid pressure on the solid (solidFluidCouple, solid.cu:81-117): the
    // solid side takes the fluid velocity mirrored about its normal.  Lane
    // `sub` takes the fluid candidates sub, sub + kSub, ... of every column
    // run.  The coupling pass of the step recorded each lane's candidates
    // within R_rec of the step-start position x0; while the solid is within d
    // of x0 they hold every candidate within R, in walk order, so replaying
    // them gives the walk's sums bit for bit.  Otherwise the lane walks.
    const auto fluid_term = [&](const float4 &pj, const float4 &uj, float dx, float dy, float dz, float d2) {
        const float m = 2.0f * ((uj.x - ui.x) * np.x + (uj.y - ui.y) * np.y + (uj.z - ui.z) * np.z);
        const float uix = uj.x - m * np.x, uiy = uj.y - m * np.y, uiz = uj.z - m * np.z;
        const float kc = hw_exp2(fmaf(-d2, P.kexp2, P.log2kd));
        // rho_i = rho_j: p* = (p_f + p_s)/2 + [du>=0] rho du/2 c0
        const float rho = pj.w;
        const float r = d2 != 0.0f ? hw_rsq(d2) : 0.0f;  // 1/|dx|, e = dx r
        float ps = 0.5f * (uj.w + np.w);
        const float du = ((uix - uj.x) * dx + (uiy - uj.y) * dy + (uiz - uj.z) * dz) * r;
        if (du >= 0.0f) ps += 0.5f * rho * du * P.cs;
        const float s = -2.0f * ps * P.rho_s_inv;
        cx += (s * dx) * kc;
        cy += (s * dy) * kc;
        cz += (s * dz) * kc;
    };
    bool replay = false;
    int nrec = 0;
    const int32_t *rec = a.x.fnbr + (size_t)i * (kSolidListCap * kSub) + L.sub;
    if (a.fluid_rec) {
        const float4 x0 = a.x.x0[i];
        nrec = a.x.fcnt[(size_t)i * kSub + L.sub];
        replay = nrec <= kSolidListCap && len2(xi.x - x0.x, xi.y - x0.y, xi.z - x0.z) <= P.disp2_max;
    }
    if (replay) {
        for (int k = 0; k < nrec; ++k) {
            const int j = rec[k * kSub];
            const float4 pj = a.fpos[j];
            const float dx = pj.x - xi.x, dy = pj.y - xi.y, dz = pj.z - xi.z;
            const float d2 = len2(dx, dy, dz);
            if (d2 < P.R2) fluid_term(pj, a.fvel[j], dx, dy, dz, d2);
        }
    } else {
        for_each_column(P, a.scell[i], [&](int c0, int c1) {
            for (int j = a.fstart[c0] + L.sub, e = a.fstart[c1]; j < e; j += kSub) {
                const float4 pj = a.fpos[j];
                const float dx = pj.x - xi.x, dy = pj.y - xi.y, dz = pj.z - xi.z;
                const float d2 = len2(dx, dy, dz);
                if (d2 < P.R2) fluid_term(pj, a.fvel[j], dx, dy, dz, d2);
            }
        });
    }
    // stress divergence + artificial viscosity (solidSolidForce, solid.cu:119-157):
    //   sum_j (P_i L_i + P_j L_j) dX k / rho_s + (det F_i F_i^-T L_i) dX (visc k)
    // with the particle's own matrices taken out of the sums: each lane sums
    // P_j L_j dX ks, dX ks and dX visc k over its partners, then applies
    // P_i L_i and det F_i F_i^-T L_i to its two vector sums once.
    float sum_ks[3] = {0.0f, 0.0f, 0.0f}, sum_vk[3] = {0.0f, 0.0f, 0.0f};
    solid_partners<kReplay, kForceRec>(P, a, i, L.sub, Xi, a.x.frec, [&](int j, const float (&dX)[3], float d2) {
        const float kc = hw_exp2(fmaf(-d2, P.kexp2, P.log2kd));
        const float ks = kc * P.rho_s_inv;
        float v[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const float4 row = a.x.frec[kForceRec * (size_t)j + 3 + r];
            v[r] = (row.x * dX[0] + row.y * dX[1] + row.z * dX[2]) * ks;
        }
        ax += v[0];
        ay += v[1];
        az += v[2];
        const float4 xj = a.x.frec[kForceRec * (size_t)j + 1], uj = a.x.frec[kForceRec * (size_t)j + 2];
        const float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
        const float dudx = (uj.x - ui.x) * dx + (uj.y - ui.y) * dy + (uj.z - ui.z) * dz;
        const float nu = P.h * dudx * hw_rcp(len2(dx, dy, dz) + eps);
        float visc = P.alpha * nu * P.cs_s;
        visc -= P.beta * nu * nu;
        const float vk = visc * kc;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            sum_ks[c] += dX[c] * ks;
            sum_vk[c] += dX[c] * vk;
        }
    });
    {
        float v[3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
            v[r] = (PLi.m[r][0] * sum_ks[0] + PLi.m[r][1] * sum_ks[1] + PLi.m[r][2] * sum_ks[2]) +
                   (ALi.m[r][0] * sum_vk[0] + ALi.m[r][1] * sum_vk[1] + ALi.m[r][2] * sum_vk[2]);
        ax += v[0];
        ay += v[1];
        az += v[2];
    }
    cx = group_sum(cx);
    cy = group_sum(cy);
    cz = group_sum(cz);
    ax = group_sum(ax);
    ay = group_sum(ay);
    az = group_sum(az);
    if (L.sub != 0) return;
    a.x.couple[i] = make_float4(cx, cy, cz, 0.0f);
    a.x.acc[i] = make_float4(ax + cx, ay + cy, az + cz, 0.0f);
}

// ---- K12: solidUpdate (src/solid.cu:209-239) -----------------------------
__device__ __forceinline__ void update_body(const DevParams &P, const SolidArgs &a, int blk) {
    const int i = blk * kBlock + threadIdx.x;
    if (i >= P.ns) return;
    const int type = __float_as_int(a.sref[i].w);
    const float dt = P.dt_s;
    if (type == FSA_FIXED_SOLID) {
        a.svel[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        a.x.acc[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    } else if (type == FSA_UNCONSTRAINED_SOLID) {
        const float4 x = a.spos[i], u = a.svel[i], ac = a.x.acc[i];
        float px, py, pz;
        float4 un;
        if (a.scpos) {
            // compensated form: x and u carry their rounding error
            float4 cp = a.scpos[i], cv = a.scvel[i];
            const float sx = u.x * dt, sy = u.y * dt, sz = u.z * dt;
            px = kahan_add(x.x, sx, cp.x);
            py = kahan_add(x.y, sy, cp.y);
            pz = kahan_add(x.z, sz, cp.z);
            const float vx = (ac.x + P.gx) * dt, vy = (ac.y + P.gy) * dt, vz = (ac.z + P.gz) * dt;
            un = make_float4(kahan_add(u.x, vx, cv.x), kahan_add(u.y, vy, cv.y), kahan_add(u.z, vz, cv.z), 0.0f);
            const float qx = fminf(fmaxf(px, P.dmin[0]), P.dmax[0]);
            const float qy = fminf(fmaxf(py, P.dmin[1]), P.dmax[1]);
            const float qz = fminf(fmaxf(pz, P.dmin[2]), P.dmax[2]);
            if (qx != px) cp.x = 0.0f;
            if (qy != py) cp.y = 0.0f;
            if (qz != pz) cp.z = 0.0f;
            px = qx;
            py = qy;
            pz = qz;
            a.scpos[i] = cp;
            a.scvel[i] = cv;
        } else {
            px = x.x + u.x * dt;
            py = x.y + u.y * dt;
            pz = x.z + u.z * dt;
            un = make_float4(u.x + (ac.x + P.gx) * dt, u.y + (ac.y + P.gy) * dt, u.z + (ac.z + P.gz) * dt, 0.0f);
            px = fminf(fmaxf(px, P.dmin[0]), P.dmax[0]);
            py = fminf(fmaxf(py, P.dmin[1]), P.dmax[1]);
            pz = fminf(fmaxf(pz, P.dmin[2]), P.dmax[2]);
        }
        if (!(__builtin_isfinite(un.x) && __builtin_isfinite(un.y) && __builtin_isfinite(un.z)))
            atomicOr(a.health, (unsigned long long)FSA_HEALTH_SOLID_NONFINITE);
        a.spos[i] = make_float4(px, py, pz, 0.0f);
        a.svel[i] = un;
    }
}

struct OpCouple {
    static __device__ void run(const SolidJob &j, int blk, int) { couple_body(j.P, j.a, blk); }
};
template <int MODE, bool SPLIT = false>
struct OpDeform {
    static __device__ void run(const SolidJob &j, int blk, int) { deformation_body<MODE, SPLIT>(j.P, j.a, blk); }
};
template <int MODE>
struct OpDeformTail {
    static __device__ void run(const SolidJob &j, int blk, int) { deformation_tail_body<MODE>(j.P, j.a, blk); }
};
struct OpForce {
    static __device__ void run(const SolidJob &j, int blk, int) { force_body(j.P, j.a, blk); }
};
struct OpSolidUpdate {
    static __device__ void run(const SolidJob &j, int blk, int) { update_body(j.P, j.a, blk); }
};

template <class Op>
hipError_t one(const DevParams &P, const SolidArgs &a, int blocks, hipStream_t s) {
    if (blocks <= 0) return hipSuccess;
    const SolidJob j{P, a};
    hipLaunchKernelGGL((k_one<Op, SolidJob>), dim3(blocks), dim3(kBlock), 0, s, j);
    return hipGetLastError();
}

}  // namespace

int blocks_solid(BatchOp op, const SolidJob &j) {
    if (j.P.ns == 0) return 0;
    if (op == kBSolidUpdate || op == kBDeformTailRecord || op == kBDeformTail) return blocks_for(j.P.ns);
    return blocks_for((long long)j.P.ns * kSub);  // couple, deformation, force: kSub lanes per particle
}

hipError_t launch_couple(const DevParams &P, const SolidArgs &a, hipStream_t s) {
    return one<OpCouple>(P, a, blocks_solid(kBCouple, SolidJob{P, a}), s);
}
hipError_t launch_solid_deformation(const DevParams &P, const SolidArgs &a, bool record, hipStream_t s) {
    const int blocks = blocks_solid(kBDeform, SolidJob{P, a});
    return record ? one<OpDeform<kWalkRecord>>(P, a, blocks, s) : one<OpDeform<kReplay>>(P, a, blocks, s);
}
hipError_t launch_solid_force(const DevParams &P, const SolidArgs &a, hipStream_t s) {
    return one<OpForce>(P, a, blocks_solid(kBForce, SolidJob{P, a}), s);
}
hipError_t launch_solid_update(const DevParams &P, const SolidArgs &a, hipStream_t s) {
    return one<OpSolidUpdate>(P, a, blocks_solid(kBSolidUpdate, SolidJob{P, a}), s);
}

hipError_t launch_many_solid(BatchOp op, const SolidJob *jobs, const int *prefix, int njobs, int nblocks,
                             hipStream_t s) {
    if (nblocks <= 0 || njobs <= 0) return hipSuccess;
    const dim3 g(nblocks), b(kBlock);
    switch (op) {
    case kBCouple: hipLaunchKernelGGL((k_many<OpCouple, SolidJob>), g, b, 0, s, jobs, prefix, njobs); break;
    // batched deformation passes are split: pair sums, then the per-particle part
    case kBDeformRecord:
        hipLaunchKernelGGL((k_many<OpDeform<kWalkRecord, true>, SolidJob>), g, b, 0, s, jobs, prefix, njobs);
        break;
    case kBDeform:
        hipLaunchKernelGGL((k_many<OpDeform<kReplay, true>, SolidJob>), g, b, 0, s, jobs, prefix, njobs);
        break;
    case kBDeformTailRecord:
        hipLaunchKernelGGL((k_many<OpDeformTail<kWalkRecord>, SolidJob>), g, b, 0, s, jobs, prefix, njobs);
        break;
    case kBDeformTail:
        hipLaunchKernelGGL((k_many<OpDeformTail<kReplay>, SolidJob>), g, b, 0, s, jobs, prefix, njobs);
        break;
    case kBForce: hipLaunchKernelGGL((k_many<OpForce, SolidJob>), g, b, 0, s, jobs, prefix, njobs); break;
    case kBSolidUpdate: hipLaunchKernelGGL((k_many<OpSolidUpdate, SolidJob>), g, b, 0, s, jobs, prefix, njobs); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace fsa
