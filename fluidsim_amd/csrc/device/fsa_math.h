// fsa_math.h -- device helpers shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "fsa_internal.h"

namespace fsa {

// include/global.h:6-8
constexpr float kEpsForSum = 0.1f;
constexpr float kEpsFloat = 1.0e-6f;
constexpr float kLargeFloat = 1.0e6f;

// hardware reciprocal, reciprocal square root and exp2 (about 1 ulp; the
// IEEE forms are several instructions each)
__device__ __forceinline__ float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float hw_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float hw_exp2(float x) { return __builtin_amdgcn_exp2f(x); }

// squared length fma(z,z,fma(y,y,x*x)): the one form used for every
// glm::length2 of the reference, so candidate tests are bit-reproducible
__device__ __forceinline__ float len2(float x, float y, float z) {
    return __builtin_fmaf(z, z, __builtin_fmaf(y, y, x * x));
}

// findPositionCell, src/pair.cu:8-24; indices are clamped into the grid
// (in-grid positions are unaffected)
__device__ __forceinline__ int find_cell(const DevParams &P, float x, float y, float z) {
    int ix = (int)((x - P.cmin[0]) * P.inv[0]);
    int iy = (int)((y - P.cmin[1]) * P.inv[1]);
    int iz = (int)((z - P.cmin[2]) * P.inv[2]);
    ix = min(max(ix, 0), P.cn[0] - 1);
    iy = min(max(iy, 0), P.cn[1] - 1);
    iz = min(max(iz, 0), P.cn[2] - 1);
    return (ix * P.cn[1] + iy) * P.cn[2] + iz;
}

// sort key of a position with nsub keys per cell: cell * nsub + the
// 1/nsub-cell z index inside the cell, (int)(t nsub) - nsub (int)t with t
// the z coordinate in cells (nsub t is exact for nsub a power of 2), clamped like
// the cell indices
__device__ __forceinline__ int cell_key(const DevParams &P, float x, float y, float z, int nsub) {
    const int c = find_cell(P, x, y, z);
    const float tz = (z - P.cmin[2]) * P.inv[2];
    const int iz = min(max((int)tz, 0), P.cn[2] - 1);
    const int sub = min(max((int)(tz * (float)nsub) - nsub * iz, 0), nsub - 1);
    return c * nsub + sub;
}

// cell coordinates (ix, iy, iz) of cell id c = (ix ny + iy) nz + iz
__device__ __forceinline__ int3 cell_coords(const DevParams &P, int c) {
    const int nz = P.cn[2], nyz = P.cn[1] * nz;
    const int ix = c / nyz;
    const int rem = c - ix * nyz;
    const int iy = rem / nz;
    return make_int3(ix, iy, rem - iy * nz);
}

// ---- precision forms -----------------------------------------------------

// v += d with compensated (Kahan) summation: c carries the rounding error of
// the previous additions (the increment is rounded in its own statement, so
// no multiply-add is fused into it)
__device__ __forceinline__ float kahan_add(float v, float d, float &c) {
    const float y = d - c;
    const float t = v + y;
    c = (t - v) - y;
    return t;
}

typedef _Float16 fsa_half2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ unsigned pack_h2(float a, float b) {
    const fsa_half2 h = {(_Float16)a, (_Float16)b};  // round to nearest even
    return __builtin_bit_cast(unsigned, h);
}
__device__ __forceinline__ float2 unpack_h2(unsigned v) {
    const fsa_half2 h = __builtin_bit_cast(fsa_half2, v);
    return make_float2((float)h.x, (float)h.y);
}

// position on axis a in cells, relative to the centre of cell index i
// (in [-0.5, 0.5) for the particle's own cell)
__device__ __forceinline__ float cell_offset(const DevParams &P, float x, int a, int i) {
    return (x - P.cmin[a]) * P.inv[a] - ((float)i + 0.5f);
}

// FP16 partner record of a fluid particle in cell c:
//   (rx, ry) (rz, rho / rho0 - 1) (ux, uy) (uz, p / B)
// with r the position relative to the cell centre in cells
__device__ __forceinline__ uint4 pack_partner(const DevParams &P, float4 x, float4 u, int c) {
    const int3 ci = cell_coords(P, c);
    return make_uint4(pack_h2(cell_offset(P, x.x, 0, ci.x), cell_offset(P, x.y, 1, ci.y)),
                      pack_h2(cell_offset(P, x.z, 2, ci.z), x.w * P.rho0_inv - 1.0f),
                      pack_h2(u.x, u.y), pack_h2(u.z, u.w * P.B_inv));
}

// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD
// b % 8), and each XCD has its own L2.  Renumbering so that XCD x walks one
// contiguous range of tiles keeps neighbouring cell columns in one L2.
// A bijection of [0, nb).
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    constexpr int kXcd = 8;
    const int x = b % kXcd, q = nb / kXcd, r = nb % kXcd;
    return x * q + min(x, r) + b / kXcd;
}

// The 27 neighbour cells of cell c as 9 columns (ox, oy ascending), each
// the contiguous z-run cz-1..cz+1 clipped to the grid: `body(c0, c1)` gets
// the first cell and one past the last cell of each column.
template <class F>
__device__ __forceinline__ void for_each_column(const DevParams &P, int c, F &&body) {
    const int ny = P.cn[1], nz = P.cn[2];
    const int nyz = ny * nz;
    const int cx = c / nyz;
    const int rem = c - cx * nyz;
    const int cy = rem / nz;
    const int cz = rem - cy * nz;
    const int zlo = max(cz - 1, 0), zhi = min(cz + 1, nz - 1);
#pragma unroll 1
    for (int ox = -1; ox <= 1; ++ox) {
        const int gx = cx + ox;
        if ((unsigned)gx >= (unsigned)P.cn[0]) continue;
#pragma unroll 1
        for (int oy = -1; oy <= 1; ++oy) {
            const int gy = cy + oy;
            if ((unsigned)gy >= (unsigned)ny) continue;
            const int base = (gx * ny + gy) * nz;
            body(base + zlo, base + zhi + 1);
        }
    }
}

// The 9 columns of cell c (the order of for_each_column) as cell ranges
// [c0, c1) of one z-run each; columns outside the grid are the empty range
// [0, 0).  Column (ox, oy) starts at the particle's own column start
// (c - cz) plus ox ny nz + oy nz, wave-uniform offsets: one add per column.
struct Columns {
    int cx, cy, zlo, zhi1, ny, nz, nx, cbase;
    __device__ __forceinline__ Columns(const DevParams &P, int c) {
        ny = P.cn[1];
        nz = P.cn[2];
        nx = P.cn[0];
        const int nyz = ny * nz;
        cx = c / nyz;
        const int rem = c - cx * nyz;
        cy = rem / nz;
        const int cz = rem - cy * nz;
        cbase = c - cz;
        zlo = max(cz - 1, 0);
        zhi1 = min(cz + 1, nz - 1) + 1;
    }
    __device__ __forceinline__ bool inside(int col) const {
        const int gx = cx + col / 3 - 1, gy = cy + col % 3 - 1;
        return (unsigned)gx < (unsigned)nx && (unsigned)gy < (unsigned)ny;
    }
    __device__ __forceinline__ int base(int col) const { return cbase + (col / 3 - 1) * (ny * nz) + (col % 3 - 1) * nz; }
    __device__ __forceinline__ int c0(int col) const { return inside(col) ? base(col) + zlo : 0; }
    __device__ __forceinline__ int c1(int col) const { return inside(col) ? base(col) + zhi1 : 0; }
};

// The slot ranges [b[k], e[k]) of the 9 columns (the order of
// for_each_column; empty outside the grid) in the starts of one kind.  All
// 18 loads are issued together, where a column-by-column walk waits for
// each column's pair in turn.
__device__ __forceinline__ void column_ranges(const Columns &cols, const int32_t *__restrict__ start, int (&b)[9],
                                              int (&e)[9]) {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const bool in = cols.inside(k);
        const uint32_t base = (uint32_t)cols.base(k);
        b[k] = start[in ? base + (uint32_t)cols.zlo : 0u];
        e[k] = start[in ? base + (uint32_t)cols.zhi1 : 0u];
    }
}
__device__ __forceinline__ void column_ranges(const DevParams &P, int c, const int32_t *__restrict__ start,
                                              int (&b)[9], int (&e)[9]) {
    column_ranges(Columns(P, c), start, b, e);
}

// Where a particle sits, for its z-windows: fx, fy its position in its
// cell in cells, tzs its z in sub-cells (1/kSubZ cell), and the margins that
// cover the rounding of cell coordinates.  A coordinate t = (x - cell_min) /
// interval carries a relative error below 2^-23 (two roundings), on the
// particle's side and on the partner's, so the margins grow with t: 1e-4
// cell + 3e-7 t across x and y, 1e-3 sub-cell + 3e-7 t in z.
struct Window {
    float fx, fy, tzs, mx, my, mz;
    __device__ __forceinline__ Window(const DevParams &P, const Columns &cols, float x, float y, float z) {
        const float tx = (x - P.cmin[0]) * P.inv[0], ty = (y - P.cmin[1]) * P.inv[1];
        fx = tx - (float)cols.cx;
        fy = ty - (float)cols.cy;
        tzs = (z - P.cmin[2]) * P.inv[2] * (float)kSubZ;
        mx = fmaf(tx, 3.0e-7f, 1.0e-4f);
        my = fmaf(ty, 3.0e-7f, 1.0e-4f);
        mz = fmaf(tzs, 3.0e-7f, 1.0e-3f);
    }
};

// The z-window of column `col` for partners within a squared distance
// reach2 (with its own rounding allowance): inside a cell the slots of the
// fluid and the walls are sorted by their sub-cell z index, so a column's
// slots with z in a range are the contiguous key range of that z-range.
// The column is at least (gx, gy) away from the particle across x and y
// (its distance to the cell faces the column lies beyond, less the margin),
// so partners within reach lie within rz = sqrt(reach2 - gx^2 - gy^2) in z;
// the window is the sub-cells of [z - rz, z + rz], widened by the z margin.
// The caller still tests every slot of the window exactly, so what it finds
// is the whole column's, in the same order.  The square root is the
// hardware one (within 1 ulp; the z margin is ~10^4 ulp of rz).  The code is
// branch-free: a column outside the grid or out of reach reads key 0 twice
// and yields the empty range, so that a caller's unrolled 9 windows stay
// straight-line code.
__device__ __forceinline__ void z_window(const DevParams &P, const Columns &cols, const int32_t *__restrict__ fine,
                                         int col, const Window &w, float reach2, int &b, int &e) {
    const int ox = col / 3 - 1, oy = col % 3 - 1;
    const float ex = ox < 0 ? w.fx : (ox > 0 ? 1.0f - w.fx : 0.0f);
    const float ey = oy < 0 ? w.fy : (oy > 0 ? 1.0f - w.fy : 0.0f);
    const float gx = fmaxf(ex - w.mx, 0.0f) * P.ival[0], gy = fmaxf(ey - w.my, 0.0f) * P.ival[1];
    const float r2 = reach2 - gx * gx - gy * gy;
    const bool ok = cols.inside(col) && r2 > 0.0f;
    const float rzs = __builtin_amdgcn_sqrtf(fmaxf(r2, 0.0f)) * P.inv[2] * (float)kSubZ + w.mz;
    const int lo = max((int)floorf(w.tzs - rzs), cols.zlo * kSubZ);
    const int hi = max(min((int)floorf(w.tzs + rzs) + 1, cols.zhi1 * kSubZ), lo);
    const uint32_t k0 = ok ? (uint32_t)(cols.base(col) * kSubZ) : 0u;
    const int bb = fine[k0 + (ok ? (uint32_t)lo : 0u)], ee = fine[k0 + (ok ? (uint32_t)hi : 0u)];
    b = ok ? bb : 0;
    e = ok ? ee : 0;
}

}  // namespace fsa
