// fsa_internal.h -- device-side data layout and launch interfaces of
// libfluidsim_amd.so.  Not part of the public ABI.
//
// HBM layout (all per-kind arrays are in *sorted order*: particles grouped by
// cell id, ascending .dat index inside a cell):
//   fluid  pos  float4 (x, y, z, rho)   vel float4 (ux, uy, uz, p)
//   solid  pos  float4 (x, y, z, -)     vel float4 (ux, uy, uz, -)
//          ref  float4 (X, Y, Z, type)  np  float4 (nx, ny, nz, p)
//          F, P, L, PL = P*L, A = det(F) F^-T : 9 SoA planes of ns floats
//          (plane e = row-major element (e/3, e%3))
//   virt   pos  float4 (x, y, z, rho)   vel float4 (ux, uy, uz, p)
//          nrm  float4 (nx, ny, nz, -)
//   cells  start[kind][ncell+1]: slots of cell c are [start[c], start[c+1])
// Fluid and solid state is double-buffered: the cell sort gathers into the
// other buffer, and the fused fluid pair+update kernel reads one buffer and
// writes the other, so every kernel is a Jacobi sweep with one writer per
// output element.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "fluidsim_amd.h"

namespace fsa {

constexpr int kBlock = 256;      // 4 wave64 per workgroup
constexpr int kScanTile = 1024;  // elements per scan workgroup (256 threads x 4)
// position arrays carry kPosPad zeroed entries past the last particle, so a
// batch of candidate loads may run past the end of a slot range (the
// results past the range are masked out of every test)
constexpr int kPosPad = 16;

// the kernels' view of fsa_params
struct DevParams {
    int nf, ns, nv;
    float gx, gy, gz;
    float cs, rho0, rho0_inv, gamma, visc, rho2p, p_min, rho_min, rho_max;
    int viscid;
    float rho_s, lambda, mu, cs_s, alpha, beta;
    float dmin[3], dmax[3], cmin[3], inv[3];
    int cn[3];
    int ncell;
    float dt, dt_s;
    int sub;
    float h, h2, h_inv2, R2, c1, c2, kd;
    // host-folded constants of the hot loops
    float kexp2;      // h^-2 * log2(e): exp(-q^2) = exp2(-d2 * kexp2)
    float cs_inv;     // 1 / c0
    float eps_vis;    // 0.01 h^2
    float visc10;     // 10 * mu
    float max_speed;  // 0.3 c0
};

DevParams make_dev_params(const fsa_params &p);

// buffers of one particle kind
struct KindBuffers {
    int n = 0;
    int ord = 0;  // which id/cell (and solid ref) buffer holds the current order
    int st = 0;   // which pos/vel buffer holds the current state
    float4 *pos[2] = {nullptr, nullptr};
    float4 *vel[2] = {nullptr, nullptr};
    float4 *ref[2] = {nullptr, nullptr};  // solid only
    int32_t *id[2] = {nullptr, nullptr};
    int32_t *cell[2] = {nullptr, nullptr};
    int32_t *start = nullptr;  // ncell + 1
    int32_t *count = nullptr;  // ncell padded to kScanTile
    int32_t *key = nullptr;    // cell of each slot of the current order
    int32_t *rank = nullptr;   // arrival slot inside the cell (atomic)
    int32_t *tmp_slot = nullptr;
    int32_t *tmp_id = nullptr;
};

struct SolidExtra {
    float4 *np = nullptr;   // normal + coupling pressure
    float4 *acc = nullptr;  // du_dt (w unused)
    float4 *couple = nullptr;
    float *F = nullptr, *P = nullptr, *L = nullptr, *PL = nullptr, *A = nullptr;  // 9 * ns each
};

struct VirtExtra {
    float4 *nrm = nullptr;
    bool normals_ready = false;
};

struct ScanScratch {
    int32_t *tile_sums = nullptr;
    int ntiles_cap = 0;
};

// ---- launchers (fsa_cells.hip) -------------------------------------------
hipError_t launch_fill_i32(int32_t *p, int32_t v, int n, hipStream_t s);
// cell id of every slot, arrival ranks, per-cell counts, jump diagnostic
hipError_t launch_hash_count(const DevParams &P, KindBuffers &k, unsigned long long *jumps, bool count_jumps,
                             hipStream_t s);
// exclusive scan count[0..ncell) -> start[0..ncell]
hipError_t launch_scan(const int32_t *count, int32_t *start, int ncell, ScanScratch &scr, hipStream_t s);
// place + rank + gather into the other buffers; flips ord/st
hipError_t launch_sort_gather(const DevParams &P, KindBuffers &k, bool has_ref, hipStream_t s);

// ---- launchers (fsa_fluid.hip) -------------------------------------------
struct FluidStepArgs {
    const float4 *fpos, *fvel;
    const int32_t *fcell;
    const int32_t *fstart, *sstart, *vstart;
    const float4 *spos, *svel, *snp;
    const float4 *vpos, *vvel;
    float4 *fpos_out, *fvel_out;
    float4 *rates;  // (ax, ay, az, drho) or null
};
hipError_t launch_fluid(const DevParams &P, const FluidStepArgs &a, hipStream_t s);

// ---- launchers (fsa_virt.hip) --------------------------------------------
struct VirtArgs {
    const float4 *vpos_in;
    float4 *vpos, *vvel, *vnrm;
    const int32_t *vcell;
    const int32_t *fstart, *vstart;
    const float4 *fpos, *fvel;
    bool compute_normals;
};
hipError_t launch_virt(const DevParams &P, const VirtArgs &a, hipStream_t s);

// ---- launchers (fsa_solid.hip) -------------------------------------------
struct SolidArgs {
    float4 *spos, *svel;
    const float4 *sref;
    const int32_t *scell;
    const int32_t *fstart, *sstart;
    const float4 *fpos, *fvel;
    SolidExtra x;  // device pointers, passed by value
};
hipError_t launch_couple(const DevParams &P, const SolidArgs &a, hipStream_t s);
hipError_t launch_solid_deformation(const DevParams &P, const SolidArgs &a, hipStream_t s);
hipError_t launch_solid_force(const DevParams &P, const SolidArgs &a, hipStream_t s);
hipError_t launch_solid_update(const DevParams &P, const SolidArgs &a, hipStream_t s);

// ---- launchers (fsa_observe.hip) -------------------------------------------
constexpr int kObserveSums = 14;
struct ObserveArgs {
    const float4 *fpos, *fvel;
    const int32_t *fstart;
    const float4 *spos, *svel, *sref;
    float tip_y;  // reference y at and above which an unconstrained solid is in the tip row
};
struct CensusArgs {
    const float4 *fpos, *spos, *vpos;
    const int32_t *fcell, *fstart, *sstart, *vstart;
};
__global__ void k_fill_u64(unsigned long long *p, int n);
hipError_t launch_census(const DevParams &P, const CensusArgs &a, unsigned long long *out, hipStream_t s);
int observe_partials();  // doubles of scratch the sums need
hipError_t launch_observe(const DevParams &P, const ObserveArgs &a, double *partial, double *out, hipStream_t s);
hipError_t launch_probe(const DevParams &P, const ObserveArgs &a, const float4 *pts, int n, float *out,
                        hipStream_t s);

}  // namespace fsa
