// fsa_observe.hip -- run-level observables: FP64 conservation sums, surge
// front, gate-tip displacement, and Shepard probe pressure.
//
// The sums are deterministic: a fixed grid of kSumBlocks workgroups
// accumulates grid-stride partial sums in FP64, a tree in LDS combines each
// workgroup, and one workgroup adds the partials in index order.
#include "fsa_internal.h"
#include "fsa_math.h"

namespace fsa {

namespace {

constexpr int kSumBlocks = 256;

}  // namespace

__global__ void k_fill_u64(unsigned long long *p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

namespace {

template <int N>
__device__ __forceinline__ void block_reduce(double (&v)[N], double (*sh)[kBlock]) {
#pragma unroll
    for (int k = 0; k < N; ++k) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
#pragma unroll
            for (int k = 0; k < N; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = sh[k][0];
}

// partial sums: [0] rho, [1..3] rho u, [4] rho |u|^2 / 2, [5] rho (-g.x),
// [7..9] solid u, [10..12] tip (x - X), [13] tip count; maxima: [6] x,
// [14] fluid |u|^2, [15] solid |u|^2 (all in FP64)
__device__ __forceinline__ bool is_max_slot(int k) { return k == 6 || k >= 14; }

__global__ __launch_bounds__(kBlock) void k_sums(DevParams P, ObserveArgs a, double *__restrict__ partial) {
    __shared__ double sh[kObserveSums][kBlock];
    double v[kObserveSums];
#pragma unroll
    for (int k = 0; k < kObserveSums; ++k) v[k] = 0.0;
    double front = -1.0e300, fu2 = 0.0, su2 = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < P.nf; i += kSumBlocks * kBlock) {
        const float4 x = a.fpos[i], u = a.fvel[i];
        const double rho = x.w;
        v[0] += rho;
        v[1] += rho * u.x;
        v[2] += rho * u.y;
        v[3] += rho * u.z;
        v[4] += 0.5 * rho * ((double)u.x * u.x + (double)u.y * u.y + (double)u.z * u.z);
        v[5] -= rho * ((double)P.gx * x.x + (double)P.gy * x.y + (double)P.gz * x.z);
        front = fmax(front, (double)x.x);
        fu2 = fmax(fu2, (double)u.x * u.x + (double)u.y * u.y + (double)u.z * u.z);
    }
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < P.ns; i += kSumBlocks * kBlock) {
        const float4 x = a.spos[i], u = a.svel[i], X = a.sref[i];
        v[7] += u.x;
        v[8] += u.y;
        v[9] += u.z;
        su2 = fmax(su2, (double)u.x * u.x + (double)u.y * u.y + (double)u.z * u.z);
        const int type = __float_as_int(X.w);
        if (type == FSA_UNCONSTRAINED_SOLID && X.y >= a.tip_y) {
            // compensated form: the position is x - c (c the carried rounding error)
            const float4 c = a.scpos ? a.scpos[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            v[10] += ((double)x.x - c.x) - X.x;
            v[11] += ((double)x.y - c.y) - X.y;
            v[12] += ((double)x.z - c.z) - X.z;
            v[13] += 1.0;
        }
    }
    // the maxima are combined separately from the sums
    double mx[3] = {front, fu2, su2};
    double bmax[3];
    for (int m = 0; m < 3; ++m) {
        sh[0][threadIdx.x] = mx[m];
        __syncthreads();
        for (int s = kBlock / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) sh[0][threadIdx.x] = fmax(sh[0][threadIdx.x], sh[0][threadIdx.x + s]);
            __syncthreads();
        }
        bmax[m] = sh[0][0];
        __syncthreads();
    }
    block_reduce(v, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < kObserveSums; ++k) partial[(size_t)blockIdx.x * kObserveSums + k] = v[k];
        partial[(size_t)blockIdx.x * kObserveSums + 6] = bmax[0];
        partial[(size_t)blockIdx.x * kObserveSums + 14] = bmax[1];
        partial[(size_t)blockIdx.x * kObserveSums + 15] = bmax[2];
    }
}

__global__ __launch_bounds__(kBlock) void k_sums_final(const double *__restrict__ partial, double *__restrict__ out) {
    if (threadIdx.x >= kObserveSums) return;
    const int k = threadIdx.x;
    double acc = k == 6 ? -1.0e300 : 0.0;
    for (int b = 0; b < kSumBlocks; ++b) {
        const double p = partial[(size_t)b * kObserveSums + k];
        acc = is_max_slot(k) ? fmax(acc, p) : acc + p;
    }
    out[k] = acc;
}

// Shepard pressure at probe points (the solidPressureCalculate weights);
// the points are packed (x, y, z) triples
__global__ __launch_bounds__(kBlock) void k_probe(DevParams P, ObserveArgs a, const float *__restrict__ pts, int n,
                                                  float *__restrict__ out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float3 xi = make_float3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
    float pressure = 0.0f, wsum = 0.0f;
    for_each_column(P, find_cell(P, xi.x, xi.y, xi.z), [&](int c0, int c1) {
        for (int j = a.fstart[c0], e = a.fstart[c1]; j < e; ++j) {
            const float4 pj = a.fpos[j];
            const float d2 = len2(pj.x - xi.x, pj.y - xi.y, pj.z - xi.z);
            if (d2 < P.R2) {
                const float w = (expf(-(d2 * P.h_inv2)) - P.c1) * P.c2;
                wsum += w;
                pressure += a.fvel[j].w * w;
            }
        }
    });
    out[i] = wsum > kEpsFloat ? pressure / wsum : 0.0f;
}

}  // namespace

int observe_partials() { return kSumBlocks * kObserveSums; }

hipError_t launch_observe(const DevParams &P, const ObserveArgs &a, double *partial, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_sums, dim3(kSumBlocks), dim3(kBlock), 0, s, P, a, partial);
    hipLaunchKernelGGL(k_sums_final, dim3(1), dim3(kBlock), 0, s, partial, out);
    return hipGetLastError();
}

hipError_t launch_probe(const DevParams &P, const ObserveArgs &a, const float *pts, int n, float *out,
                        hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_probe, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, P, a, pts, n, out);
    return hipGetLastError();
}

}  // namespace fsa

namespace fsa {

namespace {

// candidates and cut-off hits of the fluid pass, per partner kind
__global__ __launch_bounds__(kBlock) void k_census(DevParams P, CensusArgs a, unsigned long long *__restrict__ out) {
    __shared__ unsigned long long sh[6];
    if (threadIdx.x < 6) sh[threadIdx.x] = 0ull;
    __syncthreads();
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < P.nf) {
        const float4 xi = a.fpos[i];
        unsigned long long cand[3] = {0, 0, 0}, hit[3] = {0, 0, 0};
        const int32_t *starts[3] = {a.fstart, a.sstart, a.vstart};
        const float4 *pos[3] = {a.fpos, a.spos, a.vpos};
        for_each_column(P, a.fcell[i], [&](int c0, int c1) {
            for (int k = 0; k < 3; ++k) {
                const int b = starts[k][c0], e = starts[k][c1];
                cand[k] += (unsigned long long)(e - b);
                for (int j = b; j < e; ++j) {
                    const float4 pj = pos[k][j];
                    hit[k] += len2(pj.x - xi.x, pj.y - xi.y, pj.z - xi.z) < P.R2 ? 1ull : 0ull;
                }
            }
        });
        for (int k = 0; k < 3; ++k) {
            atomicAdd(&sh[k], cand[k]);
            atomicAdd(&sh[3 + k], hit[k]);
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) atomicAdd(&out[threadIdx.x], sh[threadIdx.x]);
}

}  // namespace

hipError_t launch_census(const DevParams &P, const CensusArgs &a, unsigned long long *out, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, out, 6);
    if (P.nf > 0)
        hipLaunchKernelGGL(k_census, dim3((P.nf + kBlock - 1) / kBlock), dim3(kBlock), 0, s, P, a, out);
    return hipGetLastError();
}

}  // namespace fsa
