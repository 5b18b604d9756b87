// fsa_fluid.hip -- the fluid pass: Riemann-SPH continuity and momentum sums
// over fluid, solid and virtual neighbours (findFluidPair and the three pair
// functions, src/fluid.cu:26-232) fused with the explicit update
// (fluidUpdate, src/fluid.cu:234-269).
//
// One thread per fluid particle in sorted order, so the 64 lanes of a
// wavefront are ~64 spatially adjacent particles whose neighbour columns
// overlap and stay in L1/L2.  Each column of the 27-cell neighbourhood is a
// contiguous slot range per kind, read as float4 (x,y,z,rho) for the
// candidate test and float4 (u,p) only for partners inside the cut-off.
// The kernel reads the state buffer and writes the updated state into the
// other buffer (Jacobi, one writer per particle), which removes the
// round-trip of du_dt and drho_dt through HBM.
#include "fsa_internal.h"
#include "fsa_math.h"

namespace fsa {

namespace {

enum PairKind { kFluidFluid = 0, kFluidSolid = 1, kFluidVirt = 2 };

// One Riemann pair term.  Partner state (uj, rhoj, pj) is already mirrored
// for solids.  The solid partner dissipates with c0 instead of
// min(3 du, c0) (fluid.cu:105-106); the wall partner clamps p* at 0
// (fluid.cu:154).
template <int KIND>
__device__ __forceinline__ void riemann_pair(const DevParams &P, float dx, float dy, float dz, float d2, float uix,
                                             float uiy, float uiz, float rhoi, float pi, float ujx, float ujy,
                                             float ujz, float rhoj, float pj, float &drho, float &ax, float &ay,
                                             float &az) {
    const float q2 = d2 * P.h_inv2;
    const float kc = expf(-q2) * P.kd;
    float ex = 0.0f, ey = 0.0f, ez = 0.0f;
    if (q2 != 0.0f) {
        const float r = 1.0f / sqrtf(d2);
        ex = dx * r;
        ey = dy * r;
        ez = dz * r;
    }
    const float rsum = rhoi + rhoj;
    float ps = (rhoi * pj + rhoj * pi) / rsum;
    const float du = (uix - ujx) * ex + (uiy - ujy) * ey + (uiz - ujz) * ez;
    if (du >= 0.0f) {
        const float diss = KIND == kFluidSolid ? P.cs : fminf(3.0f * du, P.cs);
        ps += rhoi * rhoj * du / rsum * diss;
    }
    if (KIND == kFluidVirt) ps = fmaxf(ps, 0.0f);
    const float dp = (pi - pj) / rsum / P.cs;
    const float vsx = (rhoi * uix + rhoj * ujx) / rsum + dp * ex;
    const float vsy = (rhoi * uiy + rhoj * ujy) / rsum + dp * ey;
    const float vsz = (rhoi * uiz + rhoj * ujz) / rsum + dp * ez;
    const float s = -2.0f * ps / rhoi;
    ax += (s * dx) * kc;
    ay += (s * dy) * kc;
    az += (s * dz) * kc;
    drho += 2.0f * rhoi * ((uix - vsx) * dx + (uiy - vsy) * dy + (uiz - vsz) * dz) * kc;
    if (P.viscid) {
        float vis = (ujx - uix) * dx + (ujy - uiy) * dy + (ujz - uiz) * dz;
        vis /= d2 + 0.01f * P.h2;
        const float t = 10.0f * P.visc * vis / rhoi;
        ax += (t * dx) * kc;
        ay += (t * dy) * kc;
        az += (t * dz) * kc;
    }
}

// Tait equation of state, densityToPressure (src/fluid.cu:16-24)
__device__ __forceinline__ float density_to_pressure(const DevParams &P, float rho) {
    const float p = (powf(rho * P.rho0_inv, P.gamma) - 1.0f) * P.rho2p;
    return fmaxf(p, P.p_min);
}

template <bool RATES>
__global__ __launch_bounds__(kBlock) void k_fluid(DevParams P, FluidStepArgs a) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= P.nf) return;
    const float4 pi4 = a.fpos[i];
    const float4 vi4 = a.fvel[i];
    const float xi = pi4.x, yi = pi4.y, zi = pi4.z, rhoi = pi4.w;
    const float uix = vi4.x, uiy = vi4.y, uiz = vi4.z, pi = vi4.w;
    const float R2 = P.R2;
    float drho = 0.0f, ax = 0.0f, ay = 0.0f, az = 0.0f;

    for_each_column(P, a.fcell[i], [&](int c0, int c1) {
        // fluid partners (fluidFluidCalculate, fluid.cu:26-69)
        for (int j = a.fstart[c0], e = a.fstart[c1]; j < e; ++j) {
            const float4 pj = a.fpos[j];
            const float dx = pj.x - xi, dy = pj.y - yi, dz = pj.z - zi;
            const float d2 = len2(dx, dy, dz);
            if (d2 < R2) {
                const float4 vj = a.fvel[j];
                riemann_pair<kFluidFluid>(P, dx, dy, dz, d2, uix, uiy, uiz, rhoi, pi, vj.x, vj.y, vj.z, pj.w, vj.w,
                                          drho, ax, ay, az);
            }
        }
        // solid partners (fluidSolidCalculate, fluid.cu:71-119): the partner
        // velocity is u_i mirrored about the solid normal, rho_j = rho_i,
        // p_j = the solid's Shepard pressure
        for (int j = a.sstart[c0], e = a.sstart[c1]; j < e; ++j) {
            const float4 pj = a.spos[j];
            const float dx = pj.x - xi, dy = pj.y - yi, dz = pj.z - zi;
            const float d2 = len2(dx, dy, dz);
            if (d2 < R2) {
                const float4 us = a.svel[j];
                const float4 np = a.snp[j];
                const float m = 2.0f * ((uix - us.x) * np.x + (uiy - us.y) * np.y + (uiz - us.z) * np.z);
                riemann_pair<kFluidSolid>(P, dx, dy, dz, d2, uix, uiy, uiz, rhoi, pi, uix - m * np.x,
                                          uiy - m * np.y, uiz - m * np.z, rhoi, np.w, drho, ax, ay, az);
            }
        }
        // wall partners (fluidVirtCalculate, fluid.cu:121-167)
        for (int j = a.vstart[c0], e = a.vstart[c1]; j < e; ++j) {
            const float4 pj = a.vpos[j];
            const float dx = pj.x - xi, dy = pj.y - yi, dz = pj.z - zi;
            const float d2 = len2(dx, dy, dz);
            if (d2 < R2) {
                const float4 vj = a.vvel[j];
                riemann_pair<kFluidVirt>(P, dx, dy, dz, d2, uix, uiy, uiz, rhoi, pi, vj.x, vj.y, vj.z, pj.w, vj.w,
                                         drho, ax, ay, az);
            }
        }
    });

    // fluidUpdate (fluid.cu:234-269): NaN scrub, explicit step with the old
    // velocity for the position, density clamp, EOS, domain clamp, speed cap
    if (!(drho == drho)) drho = 0.0f;
    if (!(ax == ax && ay == ay && az == az)) ax = ay = az = 0.0f;
    if (RATES) a.rates[i] = make_float4(ax, ay, az, drho);
    const float dt = P.dt;
    float x = xi + uix * dt, y = yi + uiy * dt, z = zi + uiz * dt;
    float ux = uix + (ax + P.gx) * dt, uy = uiy + (ay + P.gy) * dt, uz = uiz + (az + P.gz) * dt;
    float rho = rhoi + drho * dt;
    rho = fminf(fmaxf(rho, P.rho_min), P.rho_max);
    const float p = density_to_pressure(P, rho);
    x = fminf(fmaxf(x, P.dmin[0]), P.dmax[0]);
    y = fminf(fmaxf(y, P.dmin[1]), P.dmax[1]);
    z = fminf(fmaxf(z, P.dmin[2]), P.dmax[2]);
    const float speed = sqrtf(len2(ux, uy, uz));
    const float max_speed = 0.3f * P.cs;
    if (speed > max_speed) {
        const float f = max_speed / speed;
        ux *= f;
        uy *= f;
        uz *= f;
    }
    a.fpos_out[i] = make_float4(x, y, z, rho);
    a.fvel_out[i] = make_float4(ux, uy, uz, p);
}

}  // namespace

hipError_t launch_fluid(const DevParams &P, const FluidStepArgs &a, hipStream_t s) {
    if (P.nf == 0) return hipSuccess;
    const dim3 grid((P.nf + kBlock - 1) / kBlock);
    if (a.rates)
        hipLaunchKernelGGL(k_fluid<true>, grid, dim3(kBlock), 0, s, P, a);
    else
        hipLaunchKernelGGL(k_fluid<false>, grid, dim3(kBlock), 0, s, P, a);
    return hipGetLastError();
}

}  // namespace fsa
