// fsa_virt.hip -- virtual wall particles (getVirtData, src/virt.cu:9-75).
//
// Per wall particle: the nearest fluid particle over the 27 neighbour cells
// (strict <, first in traversal order wins, no cut-off) gives density,
// hydrostatically corrected pressure and the velocity that is mirrored about
// the wall normal; with no fluid in reach the wall keeps rho0, p = 0, u = 0.
// The normal is the normalised negative kernel-gradient sum over the other
// wall particles in reach.  Wall particles never move, so the normal is the
// same every step: it is computed on the first update and read back after,
// which gives bit-identical results to recomputing it.
#include "fsa_internal.h"
#include "fsa_math.h"

namespace fsa {

namespace {

// A wall particle with no fluid in its 27 cells gets rho0, p = 0 and u = 0,
// the same values on every update; once they are written (the w of its
// normal record is then 1) it is left alone until fluid comes into reach.
template <bool NORMALS>
__device__ __forceinline__ void virt_body(const DevParams &P, const VirtArgs &a, int blk) {
    const int i = blk * kBlock + threadIdx.x;
    if (i >= P.nv) return;
    const int cell = a.vcell[i];
    int cb[9], ce[9];
    column_ranges(P, cell, a.fstart, cb, ce);
    bool any = false;
#pragma unroll
    for (int col = 0; col < 9; ++col) any |= cb[col] < ce[col];
    const float4 n4 = NORMALS ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : a.vnrm[i];
    if (!NORMALS && !any && n4.w != 0.0f) return;

    const float4 xi4 = a.vpos_in[i];
    const float xi = xi4.x, yi = xi4.y, zi = xi4.z;
    float ux = 0.0f, uy = 0.0f, uz = 0.0f;
    float nx = 0.0f, ny = 0.0f, nz = 0.0f;
    float pressure = 0.0f, density = P.rho0, nearest = kLargeFloat;
    int best = -1;

    // nearest fluid particle over the 9 columns in order: strict <, so the
    // first in traversal order wins ties; kScan loads in flight, reads past
    // the run end stay in the padded array and are masked.  Most walls have
    // no fluid in reach: their column ranges come back empty in one round.
    // Once a candidate is found, a later column is read only in its z-window
    // within the distance found so far (the fluid is sorted by eighth-cell z
    // inside a cell): the slots left out are farther, and a candidate that is
    // not strictly nearer never replaces the one found, so the result is the
    // whole columns' walk.
    {
        const Columns cols(P, cell);
        const Window w(P, cols, xi, yi, zi);
        constexpr int kScan = 8;
        static_assert(kPosPad >= kScan, "padded position arrays");
#pragma unroll 1
        for (int col = 0; any && col < 9; ++col) {
            int b = cb[col], e = ce[col];
            if (b < e && nearest < kLargeFloat) z_window(P, cols, a.ffine, col, w, nearest * (1.0f + 1.0e-5f), b, e);
            for (int j = b; j < e; j += kScan) {
                float4 q[kScan];
                const float4 *src = a.fpos + j;
#pragma unroll
                for (int k = 0; k < kScan; ++k) q[k] = src[k];
#pragma unroll
                for (int k = 0; k < kScan; ++k) {
                    const float d2 = len2(q[k].x - xi, q[k].y - yi, q[k].z - zi);
                    const bool better = (j + k < e) & (d2 < nearest);
                    nearest = better ? d2 : nearest;
                    best = better ? j + k : best;
                }
            }
        }
    }
    if (NORMALS) {
        // normal from the other wall particles in reach (first update only)
        for_each_column(P, cell, [&](int c0, int c1) {
            for (int j = a.vstart[c0], e = a.vstart[c1]; j < e; ++j) {
                const float4 pj = a.vpos_in[j];
                const float dx = pj.x - xi, dy = pj.y - yi, dz = pj.z - zi;
                const float d2 = len2(dx, dy, dz);
                if (d2 < P.R2 && j != i) {
                    const float kc = expf(-(d2 * P.h_inv2)) * P.kd;
                    nx += (-dx) * kc;
                    ny += (-dy) * kc;
                    nz += (-dz) * kc;
                }
            }
        });
    }

    if (best >= 0) {
        const float4 pj = a.fpos[best];
        const float4 vj = a.fvel[best];
        const float bdx = pj.x - xi, bdy = pj.y - yi, bdz = pj.z - zi;
        density = pj.w;
        pressure = vj.w - density * (P.gx * bdx + P.gy * bdy + P.gz * bdz);
        ux = vj.x;
        uy = vj.y;
        uz = vj.z;
    }
    if (NORMALS) {
        const float n2 = len2(nx, ny, nz);
        if (n2 < kEpsForSum) {
            nx = ny = nz = 0.0f;
        } else {
            const float r = 1.0f / sqrtf(n2);
            nx *= r;
            ny *= r;
            nz *= r;
        }
        a.vnrm[i] = make_float4(nx, ny, nz, any ? 0.0f : 1.0f);
    } else {
        nx = n4.x;
        ny = n4.y;
        nz = n4.z;
        if ((n4.w != 0.0f) == any) reinterpret_cast<float *>(a.vnrm.get() + i)[3] = any ? 0.0f : 1.0f;
    }
    const float m = 2.0f * (ux * nx + uy * ny + uz * nz);
    a.vpos[i] = make_float4(xi, yi, zi, density);
    a.vvel[i] = make_float4(ux - m * nx, uy - m * ny, uz - m * nz, pressure);
}

struct OpVirt {
    static __device__ void run(const VirtJob &j, int blk, int) {
        if (j.a.compute_normals) virt_body<true>(j.P, j.a, blk);
        else virt_body<false>(j.P, j.a, blk);
    }
};

}  // namespace

int blocks_virt(const VirtJob &j) { return blocks_for(j.P.nv); }

hipError_t launch_virt(const DevParams &P, const VirtArgs &a, hipStream_t s) {
    if (P.nv == 0) return hipSuccess;
    const VirtJob j{P, a};
    hipLaunchKernelGGL((k_one<OpVirt, VirtJob>), dim3(blocks_virt(j)), dim3(kBlock), 0, s, j);
    return hipGetLastError();
}

hipError_t launch_many_virt(const VirtJob *jobs, const int *prefix, int njobs, int nblocks, hipStream_t s) {
    if (nblocks <= 0 || njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_many<OpVirt, VirtJob>), dim3(nblocks), dim3(kBlock), 0, s, jobs, prefix, njobs);
    return hipGetLastError();
}

}  // namespace fsa
