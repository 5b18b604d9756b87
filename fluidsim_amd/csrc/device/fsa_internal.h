// fsa_internal.h -- device-side data layout and launch interfaces of
// libfluidsim_amd.so.  Not part of the public ABI.
//
// HBM layout (all per-kind arrays are in *sorted order*: particles grouped by
// cell id, ascending .dat index inside a cell):
//   fluid  pos  float4 (x, y, z, rho)   vel float4 (ux, uy, uz, p)
//   solid  pos  float4 (x, y, z, -)     vel float4 (ux, uy, uz, -)
//          ref  float4 (X, Y, Z, type)  np  float4 (nx, ny, nz, p)
//          F, P, L, A = det(F) F^-T : 9 SoA planes of ns floats
//          (plane e = row-major element (e/3, e%3)); frec: per particle
//          the force pass's partner record, 6 float4 (X, x, u, P L rows)
//   virt   pos  float4 (x, y, z, rho)   vel float4 (ux, uy, uz, p)
//          nrm  float4 (nx, ny, nz, -)
//   cells  start[kind][ncell+1]: slots of cell c are [start[c], start[c+1])
// Fluid and solid state is double-buffered: the cell sort gathers into the
// other buffer, and the fused fluid pair+update kernel reads one buffer and
// writes the other, so every kernel is a Jacobi sweep with one writer per
// output element.
//
// Every kernel is written as an operation `Op::run(job, block, nblocks)`
// over one scene's job descriptor.  k_one<Op> launches it for one scene
// (the job travels as the kernel argument); k_many<Op> launches it for many
// scenes at once: the grid is the concatenation of the scenes' block ranges,
// each block finds its scene in a prefix table and reads the scene's job
// from device memory.  Both run the same code, so a batched step gives the
// same bits as stepping every scene on its own.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <type_traits>

#include "fluidsim_amd.h"

namespace fsa {

constexpr int kBlock = 256;      // 4 wave64 per workgroup
constexpr int kScanTile = 1024;  // elements per scan workgroup (256 threads x 4)
// Position, velocity and solid-normal arrays carry kPosPad entries past the
// last particle, so a batch of candidate loads may run past the end of a slot
// range (the results past the range are masked out of every test).  Each of
// those entries holds a sentinel partner: position (kFarPos, kFarPos,
// kFarPos), density w = rho0, zero velocity, pressure and normal.  A pair
// term with it is finite and its kernel weight is exactly 0, so the fluid
// pass pads its partner rounds with the sentinel instead of masking.
constexpr int kPosPad = 16;
constexpr float kFarPos = 1.0e15f;
// slot byte offsets are 32-bit (slot * 16 B)
constexpr int kMaxKindParticles = (1 << 28) - kPosPad;
// and so are the neighbour lists' entries (home record * 32 B): a larger
// fluid walks its cells every step
constexpr int kMaxListedFluid = (1 << 27) - kPosPad;
constexpr int kSolidLanes = 16;  // lanes serving one solid particle
constexpr int kForceRec = 6;     // float4s of a solid particle's force-pass record
// 16-bit partner entries of the fluid pass (and of the neighbour lists):
// offset inside a run of slots (kRunBits, high), run-base id (4 bits, low)
constexpr int kRunBits = 12;
constexpr int kRunMax = 1 << kRunBits;
// Inside a cell, fluid and wall particles are sorted by an eighth-cell z
// index (then by .dat index): their sort key is cell * kSubZ + sub.  The
// per-cell start array every pass uses is taken from the per-key one, which
// the fluid pass also reads to test only the z-range of a column its cut-off
// reaches.  Solid particles keep one key per cell (cell, then .dat index):
// their cell build runs every step, and the per-key arrays would cost more
// than the few fluid particles near the gate save.
constexpr int kSubZ = 8;
// solid-solid partners recorded per lane in the first sub-step of a step
// (entries per lane; a lane with more partners walks its candidates instead)
constexpr int kSolidListCap = 256 / kSolidLanes;

// the kernels' view of fsa_params
// A pointer to global memory, as held by the job descriptors.  A batched
// launch reads its scenes' descriptors from device memory, so the compiler
// could not tell where the pointers they hold point: it would address them
// with flat instructions, which take no scalar base and whose waits cover LDS
// traffic too (measured: the batched fluid pass ran 40 % slower per particle
// than the same scenes launched one by one).  Typed as address space 1 in
// device code, every access through them is a global one.  The layout is a
// plain pointer on both sides.
#if defined(__HIP_DEVICE_COMPILE__)
#define FSA_GLOBAL_AS __attribute__((address_space(1)))
#else
#define FSA_GLOBAL_AS
#endif
template <class T>
struct gptr {
    FSA_GLOBAL_AS T *p;
    gptr() = default;
    __host__ __device__ gptr(T *q) : p((FSA_GLOBAL_AS T *)q) {}
    __host__ __device__ gptr(std::nullptr_t) : p(nullptr) {}
    template <class U, class = typename std::enable_if<std::is_convertible<U *, T *>::value &&
                                                       !std::is_same<U, T>::value>::type>
    __host__ __device__ gptr(const gptr<U> &o) : p((FSA_GLOBAL_AS T *)o.p) {}
    __host__ __device__ operator T *() const { return (T *)p; }
    __host__ __device__ T *get() const { return (T *)p; }
};
static_assert(sizeof(gptr<float4>) == sizeof(void *), "a job pointer is a plain pointer");

struct DevParams {
    int nf, ns, nv;
    float gx, gy, gz;
    float cs, rho0, rho0_inv, gamma, visc, rho2p, p_min, rho_min, rho_max;
    int viscid;
    float rho_s, lambda, mu, cs_s, alpha, beta;
    float rho_s_inv;  // 1 / rho_s
    float dmin[3], dmax[3], cmin[3], inv[3];
    int cn[3];
    int ncell;
    float dt, dt_s;
    int sub;
    float h, h2, h_inv2, R2, c1, c2, kd;
    // host-folded constants of the hot loops
    float kexp2;      // h^-2 * log2(e): exp(-q^2) = exp2(-d2 * kexp2)
    float log2kd;     // log2(kd): kd exp(-q^2) = exp2(log2kd - d2 * kexp2)
    float cs_inv;     // 1 / c0
    float eps_vis;    // 0.01 h^2
    float visc10;     // 10 * mu
    float max_speed;  // 0.3 c0
    // fluid partners of the solid force pass are recorded by the coupling
    // pass within R_rec = R + d of the step-start solid position; they cover
    // every fluid partner of a solid that has moved at most d since
    float R2w;        // R^2 (1 + 1e-5): the cut-off the fluid pass's z-windows are drawn with
    float R2_rec;     // R_rec^2 (with a rounding margin)
    float disp2_max;  // d^2
    // precision form (FSA_FORM_* bits) and its constants
    int form;
    float ival[3];  // cell interval (the packed positions are in cells)
    float B_inv;    // 1 / B, B = rho0 c0^2 / gamma (packed pressure is p / B)
    // fluid neighbour lists (vlist = 0: the fluid pass walks the cells)
    int vlist;
    int vrows;      // fluid candidates one list holds
    int vwrows;     // wall candidates one list holds
    float vreach2;  // squared list reach ((1 + skin) R)^2, widened by a rounding margin
    float vthr2;    // squared move since the last build that calls for a new one, (skin R / 2)^2 less a margin
};

DevParams make_dev_params(const fsa_params &p);
// R_rec of the coupling records, for fluid speeds up to vmax (see fsa_capi.hip)
void set_record_reach(DevParams &d, const fsa_params &p, double vmax);

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
    int nsub = 1;  // sort keys per cell: kSubZ (fluid) or 1
    // per-cell starts, ncell + 1: slots of cell c are [start[c], start[c+1]);
    // per-key starts, ncell * nsub + 1: slots of key k are [fstart[k],
    // fstart[k+1]) (the per-cell ones if nsub is 1).  Indexed like id/cell
    // by `ord`: a build writes the other pair, so that passes still reading
    // the current order may run beside the next build.  Kinds whose passes
    // never overlap a build keep one pair (both entries alike).
    int32_t *start_buf[2] = {nullptr, nullptr};
    int32_t *fstart_buf[2] = {nullptr, nullptr};
    int32_t *cur_start() const { return start_buf[ord]; }
    int32_t *cur_fstart() const { return fstart_buf[ord]; }
    int32_t *next_start() const { return start_buf[1 - ord]; }
    int32_t *next_fstart() const { return fstart_buf[1 - ord]; }
    int32_t *count = nullptr;   // ncell * nsub padded to kScanTile; zero between builds
    int32_t *key = nullptr;    // cell of each slot of the current order
    int32_t *rank = nullptr;   // arrival slot inside the cell (atomic)
    int32_t *tmp_slot = nullptr;
    int32_t *tmp_id = nullptr;
    int32_t *tile_sums = nullptr;  // one per scan tile
    // compensated form: the rounding error carried by pos.xyz and vel.xyz,
    // double-buffered and sorted with the state
    float4 *cpos[2] = {nullptr, nullptr};
    float4 *cvel[2] = {nullptr, nullptr};
    // packed form (fluid): the FP16 partner records of the current order
    uint4 *pk = nullptr;
    // neighbour lists (fluid): the fluid pass reads and writes the state in
    // its home order -- the sorted order of the last list build -- in hrec
    // (double-buffered by hst), and also writes it into the sorted buffers,
    // which every other pass reads.  A home record is 32 bytes, position
    // (x, y, z, rho) at hrec[2 h] and velocity (u, p) at hrec[2 h + 1], so
    // that a listed candidate's two loads fall in one cache line; kPosPad
    // padding records hold the sentinel partner.  hidx holds the home index
    // of each sorted slot (indexed by ord, moved by every sort), hslot the
    // sorted slot of each home index (written by every sort).
    float4 *hrec[2] = {nullptr, nullptr};
    int hst = 0;
    int32_t *hidx[2] = {nullptr, nullptr};
    int32_t *hslot = nullptr;
};

// Fluid neighbour lists.  A list holds the candidates within the list reach
// in the walk order of its build (the 25 columns of the 5x5 cells around the
// particle's, x-major, each column's z-window in slot order), each as the
// byte offset of its home slot in the home state arrays.
constexpr int kListAhead = 18;     // rows past the longest list the fluid pass may read (and discard)
constexpr int kMoveWords = 1024;   // words the fluid pass's wavefronts spread their largest moves over
constexpr int kListRowsMax = 256;  // candidates a list holds at most
enum ListFlag {
    kFlagMode = 0,      // this step's ListMode (set by the decision, read by the build and fluid passes)
    kFlagForce = 1,     // build at the next decision whatever the moves (set by the host)
    kFlagBuilds = 2,    // builds since the context's last fluid upload
    kFlagUnlisted = 3,  // particles the last build could not list
    kFlagEntries = 4,   // candidates the last build listed (all particles)
    kFlagAge = 5,       // steps the current lists have served
    kFlagWalkLeft = 6,  // steps left to walk before the lists are tried again
    kFlagWalks = 7,     // steps walked since the context's last fluid upload
    kFlagRetry = 8,     // length of the next walking stretch
    kFlagWallEntries = 9,  // wall candidates the last build listed (all particles)
    kFlagWords = 12
};
// What the fluid pass of a step does with lists: use the current ones,
// build new ones first, or walk every particle's cells over the sorted order.
enum ListMode { kModeList = 0, kModeBuild = 1, kModeWalk = 2 };
// A list that has to be rebuilt after fewer than kListMinAge steps did not
// pay for its build (a build costs about as much as the list saves in that
// many steps at 4 M particles): the scene then walks its cells for
// kListRetry steps before it builds again, twice as long after each failure
// in a row, at most kListRetryMax.
constexpr uint32_t kListMinAge = 8;
constexpr uint32_t kListRetry = 64;
constexpr uint32_t kListRetryMax = 1024;
struct NeighbourLists {
    gptr<uint32_t> ent = nullptr;  // entry k of home particle i at ent[k * nf + i]; vrows + kListAhead rows
    gptr<int32_t> meta = nullptr;  // entries, or -1: not listed (the particle walks its cells)
    // wall candidates (byte offsets of wall slots), the same layout with
    // vwrows + kListAhead rows; wmeta -1: the particle walks the wall cells
    gptr<uint32_t> went = nullptr;
    gptr<int32_t> wmeta = nullptr;
    gptr<float4> x0 = nullptr;     // position at the last build
    // the largest squared move since the last build (float bits), in
    // kMoveWords words: each wavefront of the fluid pass takes the maximum
    // into one of them, the decision reduces them and clears them
    gptr<uint32_t> move = nullptr;
    gptr<uint32_t> flag = nullptr;  // ListFlag words
};

struct SolidExtra {
    gptr<float4> np = nullptr;   // normal + coupling pressure
    gptr<float4> acc = nullptr;  // du_dt (w unused)
    gptr<float4> couple = nullptr;
    gptr<float> F = nullptr, P = nullptr, L = nullptr, A = nullptr;  // 9 * ns each
    // batched deformation passes: each particle's pair sums F and B (18 planes of ns)
    gptr<float> dsum = nullptr;
    // what the force pass reads of a partner, written by the deformation
    // pass of the same sub-step: frec[kForceRec * i + ...] = reference
    // position X (w: type), position x, velocity u, then P L by rows (w
    // unused) -- 96 contiguous bytes instead of six arrays
    gptr<float4> frec = nullptr;
    // reference-configuration partners of the step: lane l of particle i holds
    // its hits in walk order at nbr[(i * kSolidListCap + k) * kSolidLanes + l],
    // their number (uncapped) at ncnt[i * kSolidLanes + l]
    gptr<int32_t> nbr = nullptr;
    gptr<int32_t> ncnt = nullptr;
    // fluid candidates within R_rec of the step-start position x0, recorded
    // per lane by the coupling pass in the same layout
    gptr<int32_t> fnbr = nullptr;
    gptr<int32_t> fcnt = nullptr;
    gptr<float4> x0 = nullptr;
};

struct VirtExtra {
    float4 *nrm = nullptr;
    bool normals_ready = false;
    // per cell: kReachNear if a wall particle lies in one of its 27 cells,
    // kReachFar if one lies within two cells (the walls never move: set once
    // at their upload)
    uint8_t *reach = nullptr;
};

__host__ __device__ inline int blocks_for(long long n) { return (int)((n + kBlock - 1) / kBlock); }
__host__ __device__ inline int scan_tiles(int ncell) { return (ncell + kScanTile - 1) / kScanTile; }
__host__ __device__ inline int padded_cells(int ncell) { return scan_tiles(ncell) * kScanTile; }

// ---- generic wrappers -------------------------------------------------------

// WAVES: the register allocation leaves room for at least this many waves
// per SIMD
template <class Op, class Job, int THREADS = kBlock, int WAVES = 1>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_one(Job job) {
    Op::run(job, (int)blockIdx.x, (int)gridDim.x);
}

// prefix[s] is the first block of job s, prefix[n] the grid size
template <class Op, class Job, int THREADS = kBlock, int WAVES = 1>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_many(
    const Job *__restrict__ jobs, const int *__restrict__ prefix, int n) {
    const int b = (int)blockIdx.x;
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= b) lo = mid;
        else hi = mid;
    }
    const Job job = jobs[lo];  // a private copy: fields stay in registers across the body's stores
    Op::run(job, b - prefix[lo], prefix[lo + 1] - prefix[lo]);
}

// ---- cell build (fsa_cells.hip) ---------------------------------------------
// One kind of one scene.  Grids: Fill/ScanAdd over cells, Hash/Place/Gather
// over particles, ScanTiles over scan tiles, ScanSums a single block.
struct CellJob {
    DevParams P;
    int n;
    int nsub;  // keys per cell: kSubZ or 1
    int nk;    // keys: ncell * nsub
    int has_ref;
    int count_jumps;
    gptr<const float4> pos_in, vel_in, ref_in;
    gptr<const float4> cpos_in, cvel_in;  // compensation terms (null unless compensated)
    gptr<const int32_t> id_in, prev_cell;
    gptr<float4> pos_out, vel_out, ref_out;
    gptr<float4> cpos_out, cvel_out;
    gptr<uint4> pk_out;  // FP16 partner records of the sorted state (null unless packed)
    gptr<int32_t> id_out, cell_out;
    // neighbour lists (else null): home index of each slot in and out, and
    // the sorted slot of each home index
    gptr<const int32_t> hidx_in;
    gptr<int32_t> hidx_out, hslot;
    gptr<int32_t> key, rank, count, start, fstart, tmp_slot, tmp_id, tile_sums;
    gptr<unsigned long long> jumps;
};
// the job of kind k reading the current buffers and writing the other ones
CellJob make_cell_job(const DevParams &P, const KindBuffers &k, unsigned long long *jumps, bool count_jumps);
// the whole build (hash, scan, place, gather) for one job; the counts must
// be zero on entry, as every sort leaves them
hipError_t launch_cells_one(const CellJob &j, hipStream_t s);
hipError_t launch_hash_one(const CellJob &j, hipStream_t s);     // fill (zero the counts) + hash
hipError_t launch_sort_one(const CellJob &j, hipStream_t s);     // scan + place + gather
hipError_t launch_fill_i32(int32_t *p, int32_t v, int n, hipStream_t s);
hipError_t launch_fill_f4(float4 *p, float4 v, int n, hipStream_t s);
// per cell: whether a particle of the kind with starts `start` lies in one of
// its 27 cells (kReachNear) or within two cells of it (kReachFar)
constexpr uint8_t kReachNear = 1, kReachFar = 2;
hipError_t launch_reach(const DevParams &P, const int32_t *start, uint8_t *reach, hipStream_t s);

// ---- neighbour lists (fsa_pairs.hip) -------------------------------------------
// The two passes of a list build, run after every fluid sort: the decision
// (one workgroup: a build when the host forced one or some particle has
// moved more than half the skin), then, only if it says so, the new home
// order (the sorted state copied into the home buffers the next fluid pass
// reads, home index = sorted slot) and the lists.
struct ListJob {
    DevParams P;
    gptr<const float4> spos, svel;  // sorted state of the current order
    gptr<float4> hrec;              // home records the next fluid pass reads
    gptr<int32_t> hidx;              // home index of each sorted slot of the current order
    gptr<int32_t> hslot;             // sorted slot of each home index
    gptr<const int32_t> ffine;       // per-key starts of the current order
    gptr<const float4> vpos;         // wall positions (sorted once, at their upload)
    gptr<const int32_t> vfine;       // per-key starts of the walls
    gptr<const uint8_t> vreach;      // per cell: walls near (kReachNear) or within two cells (kReachFar)
    NeighbourLists L;
};
enum ListOp { kLDecide, kLBuild };
hipError_t launch_list_one(ListOp op, const ListJob &j, hipStream_t s);
hipError_t launch_many_list(ListOp op, const ListJob *jobs, const int *prefix, int njobs, int nblocks, hipStream_t s);
int blocks_list(ListOp op, const ListJob &j);
// both passes for one scene
hipError_t launch_list_build(const ListJob &j, hipStream_t s);

// ---- fluid pass (fsa_fluid.hip) -----------------------------------------------
struct FluidStepArgs {
    gptr<const float4> fpos, fvel;
    gptr<const float4> fcpos, fcvel;     // compensation terms in (null unless compensated)
    gptr<float4> fcpos_out, fcvel_out;   // and out
    gptr<const uint4> fpk;                // FP16 partner records (null unless packed)
    gptr<const int32_t> fcell;
    gptr<const int32_t> fstart, sstart, vstart;
    gptr<const int32_t> ffine, vfine;    // per-key starts of the fluid and the walls (kSubZ keys per cell)
    gptr<const float4> spos, svel, snp;
    gptr<const float4> vpos, vvel;
    gptr<float4> fpos_out, fvel_out;
    gptr<float4> rates;  // (ax, ay, az, drho) or null
    gptr<int4> counts;   // with rates: the partners found per kind (fluid, solid, wall, 0)
    gptr<const int32_t> scell;   // the solid's cells in slot order (null without solid particles)
    gptr<const uint8_t> vreach;  // per cell: walls within its 27 cells (kReachNear) or two cells (null without walls)
    gptr<unsigned long long> health;  // FSA_HEALTH_* bits of the scene (an updated velocity not finite)
    // neighbour lists (P.vlist): one thread per home particle, which reads
    // its state from the home records hrec, writes the update to hrec_out and
    // to the sorted buffers at hslot (fpos_out, fvel_out, the compensation
    // terms, rates and counts, all in sorted order)
    gptr<const float4> hrec;
    gptr<float4> hrec_out;
    gptr<const int32_t> hslot;
    NeighbourLists L;
};
struct FluidJob {
    DevParams P;
    FluidStepArgs a;
};
// workgroup sizes of the fluid pass: 64, 128 or 256 threads (identical results)
constexpr int kFluidThreadsDefault = 64;
bool fluid_threads_valid(int threads);
hipError_t launch_fluid(const DevParams &P, const FluidStepArgs &a, int threads, hipStream_t s);

// ---- walls (fsa_virt.hip) ------------------------------------------------------
struct VirtArgs {
    gptr<const float4> vpos_in;
    gptr<float4> vpos, vvel, vnrm;
    gptr<const int32_t> vcell;
    gptr<const int32_t> fstart, vstart;
    gptr<const int32_t> ffine;  // per-key starts of the fluid (kSubZ keys per cell)
    gptr<const float4> fpos, fvel;
    int compute_normals;
};
struct VirtJob {
    DevParams P;
    VirtArgs a;
};
hipError_t launch_virt(const DevParams &P, const VirtArgs &a, hipStream_t s);

// ---- solid (fsa_solid.hip) -------------------------------------------------------
struct SolidArgs {
    gptr<float4> spos, svel;
    gptr<float4> scpos, scvel;  // compensation terms (null unless compensated)
    gptr<const float4> sref;
    gptr<const int32_t> scell;
    gptr<const int32_t> fstart, sstart;
    gptr<const float4> fpos, fvel;
    SolidExtra x;  // device pointers, passed by value
    int fluid_rec;  // the coupling pass of this step (this order) recorded the fluid partners
    int solid_rec;  // later passes replay the solid partners the first deformation pass recorded
    gptr<unsigned long long> health;  // FSA_HEALTH_* bits of the scene
};
struct SolidJob {
    DevParams P;
    SolidArgs a;
};
hipError_t launch_couple(const DevParams &P, const SolidArgs &a, hipStream_t s);
// record = true in the first sub-step of a step: walks the candidates and
// records every lane's partners; later passes replay the records
hipError_t launch_solid_deformation(const DevParams &P, const SolidArgs &a, bool record, hipStream_t s);
hipError_t launch_solid_force(const DevParams &P, const SolidArgs &a, hipStream_t s);
hipError_t launch_solid_update(const DevParams &P, const SolidArgs &a, hipStream_t s);

// ---- batched launches ----------------------------------------------------------
// Each launches one grid over all jobs of a table: `jobs` and `prefix` are
// device arrays, prefix[s] the first block of job s and prefix[njobs] = nblocks.
// blocks_* give the grid size of one job (0 when it has no work); the host
// rounds every job's range up to whole XCD rounds (multiples of 8 blocks) so
// a block's XCD is the same in the batched grid as in a single-scene grid.
enum BatchOp {
    kBFillCounts, kBHash, kBScanTiles, kBScanSums, kBScanAdd, kBPlace, kBGather,  // CellJob
    kBCouple, kBDeformRecord, kBDeform, kBDeformTailRecord, kBDeformTail,         // SolidJob
    kBForce, kBSolidUpdate
};
hipError_t launch_many_cells(BatchOp op, const CellJob *jobs, const int *prefix, int njobs, int nblocks, hipStream_t s);
int blocks_cells(BatchOp op, const CellJob &j);
// packed: the jobs' scenes use FSA_FORM_PACKED (all of them or none)
hipError_t launch_many_fluid(const FluidJob *jobs, const int *prefix, int njobs, int nblocks, int threads,
                             bool packed, hipStream_t s);
int blocks_fluid(const FluidJob &j, int threads);
hipError_t launch_many_virt(const VirtJob *jobs, const int *prefix, int njobs, int nblocks, hipStream_t s);
int blocks_virt(const VirtJob &j);
hipError_t launch_many_solid(BatchOp op, const SolidJob *jobs, const int *prefix, int njobs, int nblocks,
                             hipStream_t s);
int blocks_solid(BatchOp op, const SolidJob &j);

// ---- observables (fsa_observe.hip) ------------------------------------------------
constexpr int kObserveSums = 16;
struct ObserveArgs {
    const float4 *fpos, *fvel;
    const int32_t *fstart;
    const float4 *spos, *svel, *sref;
    const float4 *scpos;  // carried rounding error of the solid positions (compensated form) or null
    float tip_y;  // reference y at and above which an unconstrained solid is in the tip row
};
struct CensusArgs {
    const float4 *fpos, *spos, *vpos;
    const int32_t *fcell, *fstart, *sstart, *vstart;
};
hipError_t launch_census(const DevParams &P, const CensusArgs &a, unsigned long long *out, hipStream_t s);
int observe_partials();  // doubles of scratch the sums need
hipError_t launch_observe(const DevParams &P, const ObserveArgs &a, double *partial, double *out, hipStream_t s);
// pts: n packed (x, y, z) triples
hipError_t launch_probe(const DevParams &P, const ObserveArgs &a, const float *pts, int n, float *out,
                        hipStream_t s);

}  // namespace fsa
