// fsa_capi.hip -- the C ABI of fluidsim_amd.h: per-scene context, buffers,
// uploads/downloads, and the step stages in the reference's order
// (src/calculation.cu:23-33).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "fluidsim_amd.h"
#include "fsa_context.h"
#include "fsa_internal.h"

using namespace fsa;

namespace {
thread_local std::string g_err;
}  // namespace

int fsa::set_err(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr) FSA_HIP_TRY(expr)

// ---- the staging area of fsa_probe_pressure_noalloc ----------------------------
// Points (x, y, z) of up to FSA_PROBE_NOALLOC_MAX probes, then their
// pressures: 4 floats per probe, 1 MiB, one copy on every device the library
// runs on, held until the process ends.  In host code `scratch` names a
// host-side stand-in, not the device array: host code uses only the device
// address hipGetSymbolAddress returns for the current device, kept per device
// below and per context in fsa_sim::probe_scratch.
static __device__ __align__(256) float scratch[4 * FSA_PROBE_NOALLOC_MAX];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_addr[kMaxDevices] = {};

// the device address of `scratch` on `device`, which must be current; taken
// once per device (at the first fsa_create on it)
int scratch_address(int device, float **out) {
    if (device < 0 || device >= kMaxDevices) return set_err(FSA_ERR_NODEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_err(FSA_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_addr[device] = static_cast<float *>(p);
    }
    *out = g_scratch_addr[device];
    return FSA_OK;
}
}  // namespace

namespace fsa {
// The coupling pass records the fluid candidates of a solid particle within
// R_rec of its step-start position x0, from the fluid positions before the
// fluid update; the force passes replay them against the updated positions
// while the solid is within d = 0.04 R of x0.  A fluid partner within R of the
// solid then lay within R + d + f of x0 before the update, where f bounds the
// update's move |u| dt: |u| is at most the speed cap 0.3 c0 after any update
// (src/fluid.cu:265-268) and at most the largest uploaded speed before the
// first one.  The clamp to the domain only shortens the move; the
// compensated form adds at most an ulp of the coordinate, which the absolute
// margin (8 ulp of the largest cell-grid coordinate) covers.
void set_record_reach(DevParams &d, const fsa_params &p, double vmax) {
    const double R = std::sqrt((double)p.impact_distance_square), disp = 0.04 * R;
    double xmax = 0.0;
    for (int a = 0; a < 3; ++a)
        xmax = std::max({xmax, std::fabs((double)p.cell_min[a]),
                         std::fabs((double)p.cell_min[a] + (double)p.interval[a] * p.cell_number[a])});
    const double f = vmax * (double)p.dt * (1.0 + 1.0e-4) + 8.0 * xmax * 0x1p-23;
    d.R2_rec = (float)((R + disp + f) * (R + disp + f) * (1.0 + 1.0e-5));
}

DevParams make_dev_params(const fsa_params &p) {
    DevParams d{};
    d.nf = p.n_fluid;
    d.ns = p.n_solid;
    d.nv = p.n_virt;
    d.gx = p.gravity[0];
    d.gy = p.gravity[1];
    d.gz = p.gravity[2];
    d.cs = p.fluid_sound_speed;
    d.rho0 = p.fluid_reference_density;
    d.rho0_inv = p.fluid_reference_density_inverse;
    d.gamma = p.fluid_gamma;
    d.visc = p.fluid_viscosity;
    d.rho2p = p.fluid_coefficient_rho2p;
    d.p_min = p.fluid_min_pressure;
    d.rho_min = p.fluid_min_density;
    d.rho_max = p.fluid_max_density;
    d.viscid = p.fluid_is_viscid ? 1 : 0;
    d.rho_s = p.solid_reference_density;
    d.rho_s_inv = p.solid_reference_density != 0.0f ? 1.0f / p.solid_reference_density : 0.0f;
    d.lambda = p.solid_lambda;
    d.mu = p.solid_nu;
    d.cs_s = p.solid_sound_speed;
    d.alpha = p.solid_artificial_viscosity[0];
    d.beta = p.solid_artificial_viscosity[1];
    for (int a = 0; a < 3; ++a) {
        d.dmin[a] = p.domain_min[a];
        d.dmax[a] = p.domain_max[a];
        d.cmin[a] = p.cell_min[a];
        d.inv[a] = p.interval_inv[a];
        d.cn[a] = p.cell_number[a];
    }
    d.ncell = p.cell_number_total;
    d.dt = p.dt;
    d.sub = p.solid_sub_step;
    // solidUpdate, src/solid.cu:219
    d.dt_s = p.solid_sub_step > 0 ? p.dt / (float)p.solid_sub_step : p.dt;
    d.h = p.smoothing_length;
    d.h2 = p.smoothing_length_square;
    d.h_inv2 = p.smoothing_length_inverse_square;
    d.R2 = p.impact_distance_square;
    d.c1 = p.kernel_coefficient_1;
    d.c2 = p.kernel_coefficient_2;
    d.kd = p.kernel_differential;
    d.kexp2 = (float)((double)p.smoothing_length_inverse_square * 1.4426950408889634);
    d.log2kd = p.kernel_differential > 0.0f ? (float)std::log2((double)p.kernel_differential) : -1.0e30f;
    d.cs_inv = 1.0f / p.fluid_sound_speed;
    d.eps_vis = 0.01f * p.smoothing_length_square;
    d.visc10 = 10.0f * p.fluid_viscosity;
    d.max_speed = 0.3f * p.fluid_sound_speed;
    {
        const double R = std::sqrt((double)p.impact_distance_square), disp = 0.04 * R;
        d.disp2_max = (float)(disp * disp);
        d.R2w = (float)(R * R * (1.0 + 1.0e-5));
    }
    set_record_reach(d, p, d.max_speed);
    d.form = p.precision_form;
    for (int a = 0; a < 3; ++a) d.ival[a] = p.interval[a];
    d.B_inv = p.fluid_coefficient_rho2p != 0.0f ? 1.0f / p.fluid_coefficient_rho2p : 0.0f;
    // neighbour lists (fsa_pairs.hip); not with the packed form, whose
    // partner records are written in sorted order by the sort
    const double skin = p.neighbour_skin;
    d.vlist = (skin > 0.0 && !(p.precision_form & FSA_FORM_PACKED) && p.n_fluid > 0 &&
               p.n_fluid <= kMaxListedFluid) ? 1 : 0;
    if (d.vlist) {
        const double R = std::sqrt((double)p.impact_distance_square), reach = (1.0 + skin) * R;
        // FP32 positions and squared distances carry relative errors of a few
        // 1e-7: the margins keep every pair inside R now inside the reach of
        // the build, with the moves bounded by the threshold
        d.vreach2 = (float)(reach * reach * (1.0 + 2.0e-5));
        d.vthr2 = (float)(0.25 * skin * skin * R * R * (1.0 - 2.0e-5));
        // list rows: 1.5 x the lattice sites within the reach (room for a
        // compressed or clustered neighbourhood; longer lists walk instead)
        const double dp = p.particle_diameter > 0.0f ? (double)p.particle_diameter : R / 2.6;
        const double sites = 4.0 / 3.0 * 3.141592653589793 * std::pow(reach / dp, 3.0);
        d.vrows = std::min(std::max((int)std::ceil(1.5 * sites / 8.0) * 8, 32), kListRowsMax);
        // walls: a fluid particle at a flat wall sees about a third of the
        // sites within the reach in wall layers, at a corner two thirds;
        // corner particles with more wall candidates walk the wall cells
        d.vwrows = p.n_virt > 0 ? std::max(d.vrows / 2 / 8 * 8, 32) : 0;
    }
    return d;
}

ListJob list_job(const fsa_sim *s, const KindBuffers &f) {
    ListJob j{};
    j.P = s->P;
    j.spos = f.pos[f.st];
    j.svel = f.vel[f.st];
    j.hrec = f.hrec[f.hst];
    j.hidx = f.hidx[f.ord];
    j.hslot = f.hslot;
    j.ffine = f.cur_fstart();
    j.vpos = s->virt.pos[0];
    j.vfine = s->virt.cur_fstart();
    j.vreach = s->P.nv ? s->vx.reach : nullptr;
    j.L = s->lists;
    return j;
}
}  // namespace fsa

namespace {

int check_params(const fsa_params &p) {
    if (p.n_fluid < 0 || p.n_solid < 0 || p.n_virt < 0) return set_err(FSA_ERR_INVALID, "negative particle count");
    if (p.n_fluid > kMaxKindParticles || p.n_solid > kMaxKindParticles || p.n_virt > kMaxKindParticles)
        return set_err(FSA_ERR_INVALID, "more than 268,435,440 particles of one kind");
    for (int a = 0; a < 3; ++a)
        if (p.cell_number[a] <= 0) return set_err(FSA_ERR_INVALID, "cell_number must be positive on every axis");
    const int64_t total = (int64_t)p.cell_number[0] * p.cell_number[1] * p.cell_number[2];
    if (total != p.cell_number_total) return set_err(FSA_ERR_INVALID, "cell_number_total != product of cell_number");
    if (total * kSubZ >= (int64_t)INT32_MAX - 2 * kScanTile)
        return set_err(FSA_ERR_INVALID, "cell grid too large for int32 sort keys");
    if (!(p.dt > 0.0f)) return set_err(FSA_ERR_INVALID, "dt must be positive");
    if (p.solid_sub_step < 1) return set_err(FSA_ERR_INVALID, "solid_sub_step must be >= 1");
    if (p.precision_form < 0 || p.precision_form > (FSA_FORM_PACKED | FSA_FORM_COMPENSATED))
        return set_err(FSA_ERR_INVALID, "precision_form must be a combination of FSA_FORM_PACKED and FSA_FORM_COMPENSATED");
    // a packed partner's cell is decoded from the cell-id difference, which is
    // unambiguous with 3 or more cells per axis
    if ((p.precision_form & FSA_FORM_PACKED) &&
        (p.cell_number[0] < 3 || p.cell_number[1] < 3 || p.cell_number[2] < 3))
        return set_err(FSA_ERR_INVALID, "the packed form needs at least 3 cells along every axis");
    if (!(p.impact_distance_square > 0.0f)) return set_err(FSA_ERR_INVALID, "impact_distance_square must be positive");
    // the list build reads the 5x5 cells around a particle's: a reach of at
    // most two cut-offs (two cells)
    if (!(p.neighbour_skin >= 0.0f && p.neighbour_skin <= 1.0f))
        return set_err(FSA_ERR_INVALID, "neighbour_skin must be in [0, 1]");
    // the 27-cell search needs cells at least one cut-off wide
    for (int a = 0; a < 3; ++a) {
        const float w = p.interval[a];
        if (!(w * w * 1.0001f >= p.impact_distance_square))
            return set_err(FSA_ERR_INVALID, "cell interval smaller than the kernel cut-off");
    }
    return FSA_OK;
}

int alloc_kind(fsa_sim *s, KindBuffers &k, int n, bool with_ref, bool double_state, int nsub, bool compensated = false,
               bool packed = false, bool double_starts = false) {
    k.n = n;
    k.nsub = nsub;
    const int ncell = s->P.ncell;
    const int nk = ncell * nsub;
    const size_t padded = (size_t)(nk + kScanTile - 1) / kScanTile * kScanTile;
    int rc;
    for (int b = 0; b < (double_starts ? 2 : 1); ++b) {
        if ((rc = s->alloc(&k.start_buf[b], (size_t)ncell + 1))) return rc;
        if (nsub > 1) {
            if ((rc = s->alloc(&k.fstart_buf[b], (size_t)nk + 1))) return rc;
        } else {
            k.fstart_buf[b] = k.start_buf[b];  // one key per cell: the per-key starts are the per-cell ones
        }
    }
    if (!double_starts) {
        k.start_buf[1] = k.start_buf[0];
        k.fstart_buf[1] = k.fstart_buf[0];
    }
    if ((rc = s->alloc(&k.count, padded))) return rc;
    if (!s->dry) HIP_TRY(hipMemsetAsync(k.count, 0, sizeof(int32_t) * padded, s->stream));  // zero between builds
    if (n == 0) return FSA_OK;
    for (int b = 0; b < 2; ++b) {
        if (b == 0 || double_state) {
            if ((rc = s->alloc(&k.pos[b], (size_t)n + kPosPad))) return rc;
            if ((rc = s->alloc(&k.vel[b], (size_t)n + kPosPad))) return rc;
            if (!s->dry) {
                // the padding entries hold the sentinel partner
                HIP_TRY(hipMemsetAsync(k.pos[b], 0, sizeof(float4) * (size_t)n, s->stream));
                HIP_TRY(launch_fill_f4(k.pos[b] + n, make_float4(kFarPos, kFarPos, kFarPos, s->P.rho0), kPosPad,
                                       s->stream));
                HIP_TRY(hipMemsetAsync(k.vel[b], 0, sizeof(float4) * ((size_t)n + kPosPad), s->stream));
            }
            if (compensated) {
                if ((rc = s->alloc(&k.cpos[b], n)) || (rc = s->alloc(&k.cvel[b], n))) return rc;
                // the upload's state carries no rounding error yet
                if (!s->dry) {
                    HIP_TRY(hipMemsetAsync(k.cpos[b], 0, sizeof(float4) * (size_t)n, s->stream));
                    HIP_TRY(hipMemsetAsync(k.cvel[b], 0, sizeof(float4) * (size_t)n, s->stream));
                }
            }
        }
        if ((rc = s->alloc(&k.id[b], n))) return rc;
        if ((rc = s->alloc(&k.cell[b], n))) return rc;
        if (with_ref && (rc = s->alloc(&k.ref[b], n))) return rc;
    }
    if (packed && (rc = s->alloc(&k.pk, n))) return rc;
    if ((rc = s->alloc(&k.key, n))) return rc;
    if ((rc = s->alloc(&k.rank, n))) return rc;
    if ((rc = s->alloc(&k.tmp_slot, n))) return rc;
    if ((rc = s->alloc(&k.tmp_id, n))) return rc;
    if ((rc = s->alloc(&k.tile_sums, (size_t)scan_tiles(nk)))) return rc;
    return FSA_OK;
}

// the home state and the neighbour lists of the fluid (fsa_pairs.hip)
int alloc_lists(fsa_sim *s) {
    KindBuffers &f = s->fluid;
    NeighbourLists &L = s->lists;
    const size_t n = (size_t)s->P.nf;
    int rc;
    for (int b = 0; b < 2; ++b) {
        if ((rc = s->alloc(&f.hrec[b], 2 * (n + kPosPad))) || (rc = s->alloc(&f.hidx[b], n))) return rc;
        if (!s->dry) {
            // the padding records hold the sentinel partner
            std::vector<float4> pad(2 * kPosPad, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
            for (int k = 0; k < kPosPad; ++k) pad[2 * k] = make_float4(kFarPos, kFarPos, kFarPos, s->P.rho0);
            HIP_TRY(hipMemcpyAsync(f.hrec[b] + 2 * n, pad.data(), sizeof(float4) * pad.size(), hipMemcpyHostToDevice,
                                   s->stream));
            HIP_TRY(hipStreamSynchronize(s->stream));
        }
    }
    if ((rc = s->alloc(&f.hslot, n)) || (rc = s->alloc(&L.ent, (size_t)(s->P.vrows + kListAhead) * n)) ||
        (rc = s->alloc(&L.meta, n)) ||
        (s->P.vwrows > 0 && ((rc = s->alloc(&L.went, (size_t)(s->P.vwrows + kListAhead) * n)) ||
                             (rc = s->alloc(&L.wmeta, n)))) ||
        (rc = s->alloc(&L.x0, n)) || (rc = s->alloc(&L.move, (size_t)kMoveWords)) ||
        (rc = s->alloc(&L.flag, (size_t)kFlagWords)))
        return rc;
    if (!s->dry) {
        HIP_TRY(hipMemsetAsync(L.move, 0, sizeof(uint32_t) * kMoveWords, s->stream));
        HIP_TRY(hipMemsetAsync(L.flag, 0, sizeof(uint32_t) * kFlagWords, s->stream));
    }
    return FSA_OK;
}

// every device buffer of a context (dry: only count the bytes)
int alloc_all(fsa_sim *s) {
    int rc;
    const bool comp = (s->P.form & FSA_FORM_COMPENSATED) != 0, packed = (s->P.form & FSA_FORM_PACKED) != 0;
    // the next step's fluid cell build may run beside the solid sub-steps,
    // which read the fluid cells of the current order
    if ((rc = alloc_kind(s, s->fluid, s->P.nf, false, true, kSubZ, comp, packed, true))) return rc;
    if (s->P.vlist && (rc = alloc_lists(s))) return rc;
    if ((rc = alloc_kind(s, s->solid, s->P.ns, true, true, 1, comp, false))) return rc;
    // walls never move: their one sort (at upload) takes the fluid's keys, so
    // that the fluid pass reads wall partners through z-windows as well
    if ((rc = alloc_kind(s, s->virt, s->P.nv, false, false, kSubZ))) return rc;
    const int ns = s->P.ns;
    if ((rc = s->alloc(&s->sx.np, (size_t)ns + kPosPad)) || (rc = s->alloc(&s->sx.acc, ns)) ||
        (rc = s->alloc(&s->sx.couple, ns)) ||
        (rc = s->alloc(&s->sx.F, 9 * (size_t)ns)) || (rc = s->alloc(&s->sx.P, 9 * (size_t)ns)) ||
        (rc = s->alloc(&s->sx.L, 9 * (size_t)ns)) || (rc = s->alloc(&s->sx.frec, kForceRec * (size_t)ns)) ||
        (rc = s->alloc(&s->sx.A, 9 * (size_t)ns)) || (rc = s->alloc(&s->sx.dsum, 18 * (size_t)ns)) ||
        (rc = s->alloc(&s->sx.nbr, (size_t)ns * kSolidListCap * kSolidLanes)) ||
        (rc = s->alloc(&s->sx.ncnt, (size_t)ns * kSolidLanes)) ||
        (rc = s->alloc(&s->sx.fnbr, (size_t)ns * kSolidListCap * kSolidLanes)) ||
        (rc = s->alloc(&s->sx.fcnt, (size_t)ns * kSolidLanes)) || (rc = s->alloc(&s->sx.x0, ns)))
        return rc;
    // the sentinel solid partner's normal and pressure are zero
    if (!s->dry && ns > 0) HIP_TRY(hipMemsetAsync(s->sx.np, 0, sizeof(float4) * ((size_t)ns + kPosPad), s->stream));
    if ((rc = s->alloc(&s->vx.nrm, s->P.nv)) || (rc = s->alloc(&s->virt_x, s->P.nv))) return rc;
    if (s->P.nv > 0 && (rc = s->alloc(&s->vx.reach, (size_t)s->P.ncell))) return rc;
    if ((rc = s->alloc(&s->jumps, 2))) return rc;
    s->health = s->jumps ? s->jumps + 1 : nullptr;
    if ((rc = s->alloc(&s->obs_partial, (size_t)observe_partials())) || (rc = s->alloc(&s->obs_out, kObserveSums)))
        return rc;
    return FSA_OK;
}

// start arrays of kinds without particles are all zero (every cell empty)
int zero_starts(fsa_sim *s, KindBuffers &k) {
    for (int b = 0; b < 2; ++b) {
        if (b == 1 && k.start_buf[1] == k.start_buf[0]) break;
        HIP_TRY(launch_fill_i32(k.start_buf[b], 0, s->P.ncell + 1, s->stream));
        if (k.fstart_buf[b] != k.start_buf[b])
            HIP_TRY(launch_fill_i32(k.fstart_buf[b], 0, s->P.ncell * k.nsub + 1, s->stream));
    }
    return FSA_OK;
}

// host-side grid check of uploaded positions (findPositionCell must land
// inside the grid; the reference indexes out of bounds otherwise)
int check_in_grid(const fsa_sim *s, const float *x, int n, const char *what) {
    const DevParams &P = s->P;
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            const float v = x[3 * i + a];
            if (!(v >= P.cmin[a]))
                return set_err(FSA_ERR_INVALID, std::string(what) + " particle " + std::to_string(i) +
                                                    " lies outside the cell grid");
            const int c = (int)((v - P.cmin[a]) * P.inv[a]);
            if (c < 0 || c >= P.cn[a])
                return set_err(FSA_ERR_INVALID, std::string(what) + " particle " + std::to_string(i) +
                                                    " lies outside the cell grid");
        }
    return FSA_OK;
}

// zero the compensation terms of buffer 0 (compensated form)
int clear_compensation(fsa_sim *s, KindBuffers &k) {
    if (!k.cpos[0] || k.n == 0) return FSA_OK;
    HIP_TRY(hipMemsetAsync(k.cpos[0], 0, sizeof(float4) * (size_t)k.n, s->stream));
    HIP_TRY(hipMemsetAsync(k.cvel[0], 0, sizeof(float4) * (size_t)k.n, s->stream));
    return FSA_OK;
}

// the sort gathered into the other buffers: they now hold the state
void flip(KindBuffers &k) {
    k.st ^= 1;
    k.ord ^= 1;
}

// full cell build of one kind: hash + count, scan, place + rank + gather
int build_cells(fsa_sim *s, KindBuffers &k, bool count_jumps) {
    if (k.n == 0) return FSA_OK;
    int rc = drop_pending_hash(s);
    if (rc) return rc;
    HIP_TRY(launch_cells_one(make_cell_job(s->P, k, s->jumps, count_jumps), s->stream));
    flip(k);
    return FSA_OK;
}

template <class T> int download_sorted(fsa_sim *s, const T *dev, int n, std::vector<T> &host) {
    host.resize(n);
    if (n == 0) return FSA_OK;
    HIP_TRY(hipMemcpyAsync(host.data(), dev, sizeof(T) * (size_t)n, hipMemcpyDeviceToHost, s->stream));
    return FSA_OK;
}

int sorted_ids(fsa_sim *s, const KindBuffers &k, std::vector<int32_t> &ids) {
    int rc = download_sorted(s, k.id[k.ord], k.n, ids);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(s->stream));
    return FSA_OK;
}

// fold every recorded sample of one stage into the totals (waits for them)
int stage_collect(fsa_sim *s, int st) {
    const int ring = fsa_sim::kRing;
    while (s->ev_pending[st] > 0) {
        const int slot = (s->ev_head[st] - s->ev_pending[st] + ring) % ring;
        HIP_TRY(hipEventSynchronize(s->ev_end[st][slot]));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, s->ev_begin[st][slot], s->ev_end[st][slot]));
        s->stage_ms[st] += ms;
        s->stage_n[st] += 1;
        s->ev_pending[st] -= 1;
    }
    return FSA_OK;
}
// a stage's bracket: both events on the stream its work runs on
int stage_open(fsa_sim *s, int st, hipStream_t on = nullptr) {
    if (!((s->timing >> st) & 1u)) return FSA_OK;
    if (s->ev_pending[st] >= fsa_sim::kRing) {
        int rc = stage_collect(s, st);
        if (rc) return rc;
    }
    HIP_TRY(hipEventRecord(s->ev_begin[st][s->ev_head[st]], on ? on : s->stream));
    s->ev_open[st] = true;
    return FSA_OK;
}
void stage_end(fsa_sim *s, int st, hipStream_t on = nullptr) {
    if (!((s->timing >> st) & 1u) || !s->ev_open[st]) return;
    if (hipEventRecord(s->ev_end[st][s->ev_head[st]], on ? on : s->stream) != hipSuccess) return;
    s->ev_head[st] = (s->ev_head[st] + 1) % fsa_sim::kRing;
    s->ev_pending[st] += 1;
    s->ev_open[st] = false;
}

#define CHECK_SIM(s)                                                                   \
    do {                                                                               \
        if (!(s)) return set_err(FSA_ERR_INVALID, "null context");                     \
    } while (0)

// The side stream and its fork/join events.  fsa_create makes them right
// after the context's stream: streams share the device's few hardware queues
// in creation order, and a side stream made much later (after the streams of
// many other contexts) can land on its context stream's queue, where the two
// run one after the other.
int ensure_side(fsa_sim *s) {
    if (s->side) return FSA_OK;
    HIP_TRY(hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming));
    return FSA_OK;
}

}  // namespace

// A stage-call hash (fsa_particle_cell_update) not yet consumed by a sort
// leaves counts behind and ranks of the old order: zero the counts and drop
// the hash (the next fill_cell then needs a new one).
int fsa::drop_pending_hash(fsa_sim *s) {
    if (!s->hashed) return FSA_OK;
    for (KindBuffers *k : {&s->fluid, &s->solid})
        if (k->n) HIP_TRY(launch_fill_i32(k->count, 0, padded_cells(s->P.ncell * k->nsub), s->stream));
    s->hashed = false;
    return FSA_OK;
}

int fsa::require_uploaded(fsa_sim *s) {
    const int n[3] = {s->P.nf, s->P.ns, s->P.nv};
    for (int k = 0; k < 3; ++k)
        if (n[k] > 0 && !s->uploaded[k])
            return set_err(FSA_ERR_STATE, "stage called before every particle kind was uploaded");
    return FSA_OK;
}

SolidArgs fsa::solid_args(const fsa_sim *s, const KindBuffers &f, const KindBuffers &so) {
    SolidArgs a{};
    a.spos = so.pos[so.st];
    a.svel = so.vel[so.st];
    a.scpos = so.cpos[so.st];
    a.scvel = so.cvel[so.st];
    a.sref = so.ref[so.ord];
    a.scell = so.cell[so.ord];
    a.fstart = f.cur_start();
    a.sstart = so.cur_start();
    a.fpos = f.pos[f.st];
    a.fvel = f.vel[f.st];
    a.x = s->sx;
    a.health = s->health;
    return a;
}

VirtArgs fsa::virt_args(const fsa_sim *s, const KindBuffers &f, int compute_normals) {
    VirtArgs a{};
    a.vpos_in = s->virt_x;
    a.vpos = s->virt.pos[0];
    a.vvel = s->virt.vel[0];
    a.vnrm = s->vx.nrm;
    a.vcell = s->virt.cell[s->virt.ord];
    a.fstart = f.cur_start();
    a.vstart = s->virt.cur_start();
    a.ffine = f.cur_fstart();
    a.fpos = f.pos[f.st];
    a.fvel = f.vel[f.st];
    a.compute_normals = compute_normals;
    return a;
}

FluidStepArgs fsa::fluid_args(const fsa_sim *s, const KindBuffers &f, const KindBuffers &so) {
    FluidStepArgs a{};
    if (s->P.vlist) {
        a.hrec = f.hrec[f.hst];
        a.hrec_out = f.hrec[1 - f.hst];
        a.hslot = f.hslot;
        a.L = s->lists;
    }
    a.fpos = f.pos[f.st];
    a.fvel = f.vel[f.st];
    a.fcpos = f.cpos[f.st];
    a.fcvel = f.cvel[f.st];
    a.fcpos_out = f.cpos[1 - f.st];
    a.fcvel_out = f.cvel[1 - f.st];
    a.fpk = f.pk;
    a.fcell = f.cell[f.ord];
    a.fstart = f.cur_start();
    a.ffine = f.cur_fstart();
    a.sstart = so.cur_start();
    a.vstart = s->virt.cur_start();
    a.vfine = s->virt.cur_fstart();
    a.spos = so.pos[so.st];
    a.svel = so.vel[so.st];
    a.snp = s->sx.np;
    a.vpos = s->virt.pos[0];
    a.vvel = s->virt.vel[0];
    a.fpos_out = f.pos[1 - f.st];
    a.fvel_out = f.vel[1 - f.st];
    a.rates = nullptr;
    a.scell = so.n ? so.cell[so.ord] : nullptr;
    a.vreach = s->P.nv ? s->vx.reach : nullptr;
    a.health = s->health;
    return a;
}

extern "C" {

int fsa_abi_version(void) { return FSA_ABI_VERSION; }

const char *fsa_last_error(void) { return g_err.c_str(); }

int fsa_device_count(int *count) {
    if (!count) return set_err(FSA_ERR_INVALID, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    int good = 0;
    for (int d = 0; d < n; ++d) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, d) == hipSuccess && std::strncmp(prop.gcnArchName, "gfx950", 6) == 0)
            ++good;
    }
    *count = good;
    return FSA_OK;
}

namespace {
// unique per context state over the process lifetime (see fsa_sim::serial)
uint64_t next_serial() {
    static std::atomic<uint64_t> counter{1};
    return counter.fetch_add(1);
}
}  // namespace

int fsa_create(const fsa_params *params, void *stream, fsa_sim **out) {
    if (!params || !out) return set_err(FSA_ERR_INVALID, "null argument");
    *out = nullptr;
    int rc = check_params(*params);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_err(FSA_ERR_NODEVICE, "no HIP device visible");
    fsa_sim *s = new (std::nothrow) fsa_sim;
    if (!s) return set_err(FSA_ERR_NOMEM, "host allocation failed");
    s->params = *params;
    s->P = make_dev_params(*params);
    s->serial = next_serial();
    {
        hipDeviceProp_t prop;
        if (hipGetDevice(&s->device) != hipSuccess || hipGetDeviceProperties(&prop, s->device) != hipSuccess ||
            std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
            delete s;
            return set_err(FSA_ERR_NODEVICE, "current device is not gfx950");
        }
    }
    if (stream) {
        s->stream = static_cast<hipStream_t>(stream);
    } else {
        if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
            delete s;
            return set_err(FSA_ERR_HIP, "hipStreamCreate failed");
        }
        s->own_stream = true;
    }
    auto fail = [&](int code) {
        fsa_destroy(s);
        return code;
    };
    if ((rc = ensure_side(s))) return fail(rc);
    if (!s->dry && (rc = scratch_address(s->device, &s->probe_scratch))) return fail(rc);
    if ((rc = alloc_all(s))) return fail(rc);
    if (hipMemsetAsync(s->jumps, 0, 2 * sizeof(unsigned long long), s->stream) != hipSuccess)
        return fail(set_err(FSA_ERR_HIP, "hipMemsetAsync failed"));
    if ((rc = zero_starts(s, s->fluid)) || (rc = zero_starts(s, s->solid)) || (rc = zero_starts(s, s->virt)))
        return fail(rc);
    if (hipStreamSynchronize(s->stream) != hipSuccess) return fail(set_err(FSA_ERR_HIP, "stream sync failed"));
    *out = s;
    return FSA_OK;
}

int fsa_footprint(const fsa_params *params, int64_t *resident, int64_t *upload_peak) {
    if (!params || !resident) return set_err(FSA_ERR_INVALID, "null argument");
    int rc = check_params(*params);
    if (rc) return rc;
    fsa_sim s;
    s.dry = true;
    s.params = *params;
    s.P = make_dev_params(*params);
    if ((rc = alloc_all(&s))) return rc;
    *resident = s.bytes;
    // fsa_upload_virt sorts the walls through two temporary float4 arrays
    if (upload_peak) *upload_peak = s.bytes + 2 * (int64_t)sizeof(float4) * s.P.nv;
    return FSA_OK;
}

int fsa_device_memory(int64_t *free_bytes, int64_t *total_bytes) {
    if (!free_bytes || !total_bytes) return set_err(FSA_ERR_INVALID, "null argument");
    size_t f = 0, t = 0;
    HIP_TRY(hipMemGetInfo(&f, &t));
    *free_bytes = (int64_t)f;
    *total_bytes = (int64_t)t;
    return FSA_OK;
}

int fsa_destroy(fsa_sim *s) {
    if (!s) return FSA_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (void *p : s->allocations) (void)hipFree(p);
    for (int st = 0; st < FSA_STAGE_COUNT; ++st)
        for (int r = 0; r < fsa_sim::kRing; ++r) {
            if (s->ev_begin[st][r]) (void)hipEventDestroy(s->ev_begin[st][r]);
            if (s->ev_end[st][r]) (void)hipEventDestroy(s->ev_end[st][r]);
        }
    if (s->side) {
        (void)hipStreamSynchronize(s->side);
        (void)hipStreamDestroy(s->side);
    }
    if (s->ev_fork) (void)hipEventDestroy(s->ev_fork);
    if (s->ev_join) (void)hipEventDestroy(s->ev_join);
    if (s->own_stream && s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return FSA_OK;
}

int fsa_upload_fluid(fsa_sim *s, const float *x, const float *u, const float *rho, const float *p) {
    CHECK_SIM(s);
    const int n = s->P.nf;
    if (n > 0 && (!x || !u || !rho || !p)) return set_err(FSA_ERR_INVALID, "null fluid array");
    int rc = check_in_grid(s, x, n, "fluid");
    if (rc) return rc;
    std::vector<float4> pos(n), vel(n);
    std::vector<int32_t> id(n);
    double vmax = s->P.max_speed;
    for (int i = 0; i < n; ++i) {
        pos[i] = make_float4(x[3 * i], x[3 * i + 1], x[3 * i + 2], rho[i]);
        vel[i] = make_float4(u[3 * i], u[3 * i + 1], u[3 * i + 2], p[i]);
        id[i] = i;
        const double v2 = (double)u[3 * i] * u[3 * i] + (double)u[3 * i + 1] * u[3 * i + 1] +
                          (double)u[3 * i + 2] * u[3 * i + 2];
        if (std::isfinite(v2) && v2 > vmax * vmax) vmax = std::sqrt(v2);
    }
    // the first fluid update moves a particle by its uploaded speed, which may
    // exceed the cap: the coupling records reach that far (until the next
    // upload).  A non-finite velocity is a diverged state (fsa_get_health),
    // not a bounded move.
    set_record_reach(s->P, s->params, vmax);
    // a batch keys its recorded launches on the serial: a re-uploaded scene is a new one
    s->serial = next_serial();
    KindBuffers &k = s->fluid;
    s->couple_fresh = false;  // the coupling records name slots of the old order
    HIP_TRY(hipMemsetAsync(s->health, 0, sizeof(unsigned long long), s->stream));  // a new state
    if (n > 0) {
        k.st = 0;
        k.ord = 0;
        HIP_TRY(hipMemcpyAsync(k.pos[0], pos.data(), sizeof(float4) * n, hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(k.vel[0], vel.data(), sizeof(float4) * n, hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(k.id[0], id.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s->stream));
        if ((rc = clear_compensation(s, k))) return rc;  // an uploaded state carries no rounding error
        if (s->P.vlist) {
            // any home order will do until the build below makes the sorted one home
            k.hst = 0;
            HIP_TRY(hipMemcpyAsync(k.hidx[0], id.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s->stream));
        }
        if ((rc = build_cells(s, k, false))) return rc;
        if (s->P.vlist) {
            // the counts start again and a build is due whatever the moves say
            HIP_TRY(launch_fill_i32(reinterpret_cast<int32_t *>(s->lists.flag.get()), 0, kFlagWords, s->stream));
            HIP_TRY(launch_fill_i32(reinterpret_cast<int32_t *>(s->lists.flag.get()) + kFlagForce, 1, 1, s->stream));
            HIP_TRY(launch_list_build(list_job(s, k), s->stream));
        }
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->uploaded[FSA_FLUID] = true;
    return FSA_OK;
}

int fsa_upload_solid(fsa_sim *s, const float *x, const float *u, const float *X, const int32_t *type) {
    CHECK_SIM(s);
    const int n = s->P.ns;
    if (n > 0 && (!x || !u || !X || !type)) return set_err(FSA_ERR_INVALID, "null solid array");
    int rc = check_in_grid(s, x, n, "solid");
    if (rc) return rc;
    std::vector<float4> pos(n), vel(n), ref(n);
    std::vector<int32_t> id(n);
    for (int i = 0; i < n; ++i) {
        if (type[i] != FSA_UNCONSTRAINED_SOLID && type[i] != FSA_FIXED_SOLID)
            return set_err(FSA_ERR_INVALID, "solid type must be 0 (unconstrained) or 1 (fixed)");
        pos[i] = make_float4(x[3 * i], x[3 * i + 1], x[3 * i + 2], 0.0f);
        vel[i] = make_float4(u[3 * i], u[3 * i + 1], u[3 * i + 2], 0.0f);
        float tbits;
        std::memcpy(&tbits, &type[i], sizeof tbits);
        ref[i] = make_float4(X[3 * i], X[3 * i + 1], X[3 * i + 2], tbits);
        id[i] = i;
    }
    KindBuffers &k = s->solid;
    s->couple_fresh = false;  // the coupling records name slots of the old order
    HIP_TRY(hipMemsetAsync(s->health, 0, sizeof(unsigned long long), s->stream));  // a new state
    if (n > 0) {
        k.st = 0;
        k.ord = 0;
        HIP_TRY(hipMemcpyAsync(k.pos[0], pos.data(), sizeof(float4) * n, hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(k.vel[0], vel.data(), sizeof(float4) * n, hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(k.ref[0], ref.data(), sizeof(float4) * n, hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(k.id[0], id.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s->stream));
        if ((rc = clear_compensation(s, k))) return rc;  // an uploaded state carries no rounding error
        if ((rc = build_cells(s, k, false))) return rc;
        // Solid_Host_List::addList defaults (include/host_data.cuh:92-99):
        // F = I, P = 0, L = I, n = 0, p = 0, du_dt = 0
        std::vector<float> eye(9 * (size_t)n, 0.0f), zero9(9 * (size_t)n, 0.0f);
        for (int e : {0, 4, 8})
            for (int i = 0; i < n; ++i) eye[(size_t)e * n + i] = 1.0f;
        std::vector<float4> z4(n, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        HIP_TRY(hipMemcpyAsync(s->sx.F, eye.data(), 36 * (size_t)n, hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(s->sx.L, eye.data(), 36 * (size_t)n, hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(s->sx.P, zero9.data(), 36 * (size_t)n, hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(s->sx.np, z4.data(), 16 * (size_t)n, hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(s->sx.acc, z4.data(), 16 * (size_t)n, hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(s->sx.couple, z4.data(), 16 * (size_t)n, hipMemcpyHostToDevice, s->stream));
        // gate tip row: unconstrained solids within dp/2 of the largest reference y
        float ymax = -3.0e38f;
        for (int i = 0; i < n; ++i)
            if (type[i] == FSA_UNCONSTRAINED_SOLID) ymax = std::max(ymax, X[3 * i + 1]);
        s->tip_y = ymax > -3.0e38f ? ymax - 0.5f * s->params.particle_diameter : 3.0e38f;
        HIP_TRY(hipStreamSynchronize(s->stream));
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->uploaded[FSA_SOLID] = true;
    return FSA_OK;
}

int fsa_upload_virt(fsa_sim *s, const float *x, const int32_t *type) {
    CHECK_SIM(s);
    const int n = s->P.nv;
    if (n > 0 && (!x || !type)) return set_err(FSA_ERR_INVALID, "null virt array");
    int rc = check_in_grid(s, x, n, "virtual");
    if (rc) return rc;
    KindBuffers &k = s->virt;
    if (n > 0) {
        // Virt_Host_List::addList defaults: u = 0, p = 0, rho = rho0
        std::vector<float4> pos(n), vel(n, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        std::vector<int32_t> id(n);
        for (int i = 0; i < n; ++i) {
            pos[i] = make_float4(x[3 * i], x[3 * i + 1], x[3 * i + 2], s->params.fluid_reference_density);
            id[i] = i;
        }
        // walls are static: sort once, with pos/vel scratch in the second
        // buffers of the fluid-sized temporaries
        float4 *pos_tmp = nullptr, *vel_tmp = nullptr;
        HIP_TRY(hipMalloc(&pos_tmp, sizeof(float4) * n));
        if (hipMalloc(&vel_tmp, sizeof(float4) * n) != hipSuccess) {
            (void)hipFree(pos_tmp);
            return set_err(FSA_ERR_NOMEM, "hipMalloc failed for wall scratch");
        }
        k.pos[1] = pos_tmp;
        k.vel[1] = vel_tmp;
        k.st = 1;
        k.ord = 0;
        int err = FSA_OK;
        if (hipMemcpyAsync(pos_tmp, pos.data(), sizeof(float4) * n, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
            hipMemcpyAsync(vel_tmp, vel.data(), sizeof(float4) * n, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
            hipMemcpyAsync(k.id[0], id.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s->stream) != hipSuccess)
            err = set_err(FSA_ERR_HIP, "wall upload copy failed");
        if (!err) err = build_cells(s, k, false);  // gathers into pos[0]/vel[0]
        if (!err && launch_reach(s->P, k.cur_start(), s->vx.reach, s->stream) != hipSuccess)
            err = set_err(FSA_ERR_HIP, "wall reach launch failed");
        if (!err && hipMemcpyAsync(s->virt_x, k.pos[0], sizeof(float4) * n, hipMemcpyDeviceToDevice, s->stream) !=
                        hipSuccess)
            err = set_err(FSA_ERR_HIP, "wall position copy failed");
        if (hipStreamSynchronize(s->stream) != hipSuccess && !err) err = set_err(FSA_ERR_HIP, "wall upload failed");
        (void)hipFree(pos_tmp);
        (void)hipFree(vel_tmp);
        k.pos[1] = nullptr;
        k.vel[1] = nullptr;
        if (err) return err;
        (void)type;  // FIXED_VIRT only; MOVING_VIRT is declared but unused by the reference
        s->vx.normals_ready = false;
        // the fluid's neighbour lists hold wall candidates: the next step builds them again
        if (s->P.vlist)
            HIP_TRY(launch_fill_i32(reinterpret_cast<int32_t *>(s->lists.flag.get()) + kFlagForce, 1, 1, s->stream));
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->uploaded[FSA_VIRT] = true;
    return FSA_OK;
}

int fsa_download_fluid(fsa_sim *s, float *x, float *u, float *rho, float *p) {
    CHECK_SIM(s);
    const KindBuffers &k = s->fluid;
    std::vector<int32_t> ids;
    std::vector<float4> pos, vel;
    int rc;
    if ((rc = download_sorted(s, k.pos[k.st], k.n, pos)) || (rc = download_sorted(s, k.vel[k.st], k.n, vel)) ||
        (rc = sorted_ids(s, k, ids)))
        return rc;
    for (int t = 0; t < k.n; ++t) {
        const int i = ids[t];
        if (x) { x[3 * i] = pos[t].x; x[3 * i + 1] = pos[t].y; x[3 * i + 2] = pos[t].z; }
        if (u) { u[3 * i] = vel[t].x; u[3 * i + 1] = vel[t].y; u[3 * i + 2] = vel[t].z; }
        if (rho) rho[i] = pos[t].w;
        if (p) p[i] = vel[t].w;
    }
    return FSA_OK;
}

int fsa_download_fluid_rates(fsa_sim *s, float *drho, float *a) {
    CHECK_SIM(s);
    if (!s->rates) return set_err(FSA_ERR_STATE, "pair sums are recorded only with FSA_DEBUG_PAIR_SUMS");
    const KindBuffers &k = s->fluid;
    std::vector<int32_t> ids;
    std::vector<float4> r;
    int rc;
    if ((rc = download_sorted(s, (const float4 *)s->rates, k.n, r)) || (rc = sorted_ids(s, k, ids))) return rc;
    for (int t = 0; t < k.n; ++t) {
        const int i = ids[t];
        if (drho) drho[i] = r[t].w;
        if (a) { a[3 * i] = r[t].x; a[3 * i + 1] = r[t].y; a[3 * i + 2] = r[t].z; }
    }
    return FSA_OK;
}

int fsa_download_pair_counts(fsa_sim *s, int32_t *counts) {
    CHECK_SIM(s);
    if (!s->pair_counts) return set_err(FSA_ERR_STATE, "pair counts are recorded only with FSA_DEBUG_PAIR_SUMS");
    if (!counts) return set_err(FSA_ERR_INVALID, "null output");
    const KindBuffers &k = s->fluid;
    std::vector<int32_t> ids;
    std::vector<float4> r;
    int rc;
    if ((rc = download_sorted(s, reinterpret_cast<const float4 *>(s->pair_counts), k.n, r)) ||
        (rc = sorted_ids(s, k, ids)))
        return rc;
    for (int t = 0; t < k.n; ++t) {
        const int i = ids[t];
        int4 c;
        std::memcpy(&c, &r[t], sizeof c);
        counts[3 * i] = c.x;
        counts[3 * i + 1] = c.y;
        counts[3 * i + 2] = c.z;
    }
    return FSA_OK;
}

int fsa_download_solid(fsa_sim *s, float *x, float *u, float *p, float *nrm, float *du_dt, float *couple, float *F,
                       float *Pk, float *L) {
    CHECK_SIM(s);
    const KindBuffers &k = s->solid;
    const int n = k.n;
    std::vector<int32_t> ids;
    std::vector<float4> pos, vel, np, acc, cpl;
    std::vector<float> mF, mP, mL;
    int rc;
    if ((rc = download_sorted(s, k.pos[k.st], n, pos)) || (rc = download_sorted(s, k.vel[k.st], n, vel)) ||
        (rc = download_sorted(s, (const float4 *)s->sx.np, n, np)) ||
        (rc = download_sorted(s, (const float4 *)s->sx.acc, n, acc)) ||
        (rc = download_sorted(s, (const float4 *)s->sx.couple, n, cpl)) ||
        (rc = download_sorted(s, (const float *)s->sx.F, 9 * n, mF)) ||
        (rc = download_sorted(s, (const float *)s->sx.P, 9 * n, mP)) ||
        (rc = download_sorted(s, (const float *)s->sx.L, 9 * n, mL)) || (rc = sorted_ids(s, k, ids)))
        return rc;
    auto put3 = [](float *o, int i, const float4 &v) {
        if (o) { o[3 * i] = v.x; o[3 * i + 1] = v.y; o[3 * i + 2] = v.z; }
    };
    // SoA row-major planes -> glm column-major 9 floats per particle
    auto put9 = [n](float *o, int i, int t, const std::vector<float> &m) {
        if (!o) return;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) o[9 * i + 3 * c + r] = m[(size_t)(3 * r + c) * n + t];
    };
    for (int t = 0; t < n; ++t) {
        const int i = ids[t];
        put3(x, i, pos[t]);
        put3(u, i, vel[t]);
        if (p) p[i] = np[t].w;
        put3(nrm, i, np[t]);
        put3(du_dt, i, acc[t]);
        put3(couple, i, cpl[t]);
        put9(F, i, t, mF);
        put9(Pk, i, t, mP);
        put9(L, i, t, mL);
    }
    return FSA_OK;
}

int fsa_download_virt(fsa_sim *s, float *x, float *nrm, float *u, float *rho, float *p) {
    CHECK_SIM(s);
    const KindBuffers &k = s->virt;
    const int n = k.n;
    std::vector<int32_t> ids;
    std::vector<float4> pos, vel, nn;
    int rc;
    if ((rc = download_sorted(s, k.pos[0], n, pos)) || (rc = download_sorted(s, k.vel[0], n, vel)) ||
        (rc = download_sorted(s, (const float4 *)s->vx.nrm, n, nn)) || (rc = sorted_ids(s, k, ids)))
        return rc;
    for (int t = 0; t < n; ++t) {
        const int i = ids[t];
        if (x) { x[3 * i] = pos[t].x; x[3 * i + 1] = pos[t].y; x[3 * i + 2] = pos[t].z; }
        if (nrm) {
            const bool ready = s->vx.normals_ready;
            nrm[3 * i] = ready ? nn[t].x : 0.0f;
            nrm[3 * i + 1] = ready ? nn[t].y : 0.0f;
            nrm[3 * i + 2] = ready ? nn[t].z : 0.0f;
        }
        if (u) { u[3 * i] = vel[t].x; u[3 * i + 1] = vel[t].y; u[3 * i + 2] = vel[t].z; }
        if (rho) rho[i] = pos[t].w;
        if (p) p[i] = vel[t].w;
    }
    return FSA_OK;
}

// ---- stages ----------------------------------------------------------------

int fsa_particle_cell_update(fsa_sim *s) {
    CHECK_SIM(s);
    int rc = require_uploaded(s);
    if (rc) return rc;
    if ((rc = stage_open(s, FSA_STAGE_CELL))) return rc;
    // wall cell ids never change (the walls do not move): only fluid and solid
    if (s->fluid.n) HIP_TRY(launch_hash_one(make_cell_job(s->P, s->fluid, s->jumps, true), s->stream));
    if (s->solid.n) HIP_TRY(launch_hash_one(make_cell_job(s->P, s->solid, s->jumps, true), s->stream));
    s->hashed = true;
    return FSA_OK;
}

int fsa_fill_cell(fsa_sim *s) {
    CHECK_SIM(s);
    int rc = require_uploaded(s);
    if (rc) return rc;
    // the sort consumes the ranks of the last hash, which describe the
    // current order only until the next sort
    if (!s->hashed) return set_err(FSA_ERR_STATE, "fill_cell needs particle_cell_update first");
    s->hashed = false;
    s->couple_fresh = false;  // the coupling records describe the old order
    for (KindBuffers *k : {&s->fluid, &s->solid}) {
        if (!k->n) continue;
        HIP_TRY(launch_sort_one(make_cell_job(s->P, *k, s->jumps, false), s->stream));
        flip(*k);
    }
    if (s->P.vlist) HIP_TRY(launch_list_build(list_job(s, s->fluid), s->stream));
    stage_end(s, FSA_STAGE_CELL);
    return FSA_OK;
}

int fsa_couple_calculation(fsa_sim *s) {
    CHECK_SIM(s);
    int rc = require_uploaded(s);
    if (rc) return rc;
    // DeviceCalculation::coupleCalculation, src/solid.cu:346-359
    if (s->P.nf == 0 || s->P.ns == 0) return FSA_OK;
    if ((rc = stage_open(s, FSA_STAGE_COUPLE))) return rc;
    HIP_TRY(launch_couple(s->P, solid_args(s), s->stream));
    s->couple_fresh = true;
    stage_end(s, FSA_STAGE_COUPLE);
    return FSA_OK;
}

int fsa_virt_update(fsa_sim *s) {
    CHECK_SIM(s);
    int rc = require_uploaded(s);
    if (rc) return rc;
    if (s->P.nv == 0) return FSA_OK;
    if ((rc = stage_open(s, FSA_STAGE_VIRT))) return rc;
    const VirtArgs a = virt_args(s, s->vx.normals_ready ? 0 : 1);
    HIP_TRY(launch_virt(s->P, a, s->stream));
    s->vx.normals_ready = true;
    stage_end(s, FSA_STAGE_VIRT);
    return FSA_OK;
}

int fsa_fluid_calculation(fsa_sim *s) {
    CHECK_SIM(s);
    int rc = require_uploaded(s);
    if (rc) return rc;
    if (s->P.nf == 0) return FSA_OK;
    if ((rc = stage_open(s, FSA_STAGE_FLUID))) return rc;
    KindBuffers &f = s->fluid;
    FluidStepArgs a = fluid_args(s);
    a.rates = (s->debug & FSA_DEBUG_PAIR_SUMS) ? s->rates : nullptr;
    a.counts = (s->debug & FSA_DEBUG_PAIR_SUMS) ? s->pair_counts : nullptr;
    HIP_TRY(launch_fluid(s->P, a, s->fluid_threads, s->stream));
    f.st = 1 - f.st;
    if (s->P.vlist) f.hst = 1 - f.hst;
    stage_end(s, FSA_STAGE_FLUID);
    return FSA_OK;
}

// the solid sub-steps of solidCalculation (src/solid.cu:241-260)
static int solid_sub_steps(fsa_sim *s) {
    SolidArgs a = solid_args(s);
    const bool walk = (s->debug & FSA_DEBUG_WALK) != 0;
    a.fluid_rec = s->couple_fresh && !walk ? 1 : 0;
    a.solid_rec = walk ? 0 : 1;
    for (int sub = 0; sub < s->P.sub; ++sub) {
        HIP_TRY(launch_solid_deformation(s->P, a, sub == 0 || walk, s->stream));
        HIP_TRY(launch_solid_force(s->P, a, s->stream));
        HIP_TRY(launch_solid_update(s->P, a, s->stream));
    }
    return FSA_OK;
}

int fsa_solid_calculation(fsa_sim *s) {
    CHECK_SIM(s);
    int rc = require_uploaded(s);
    if (rc) return rc;
    if (s->P.ns == 0) return FSA_OK;
    if ((rc = stage_open(s, FSA_STAGE_SOLID))) return rc;
    if ((rc = solid_sub_steps(s))) return rc;
    stage_end(s, FSA_STAGE_SOLID);
    return FSA_OK;
}


// Steps in the reference's stage order (src/calculation.cu:23-33), with the
// head of step i+1 moved beside the tail of step i: the fluid cell build and
// the wall pass of the next step need only the fluid state the fluid pass
// has just written, so they run on the side stream while the solid
// sub-steps (which read that state and the current fluid cells: the build
// writes the other state buffer and the other start arrays) and the solid
// cell build run on the context's stream.  The two join before the next
// coupling pass.  Every kernel sees the same inputs as in the sequential
// order, so the results are the same bits.  Stage timing keeps the overlap:
// the side stream's fluid cell build and wall pass are timed there, and the
// solid kind's cell build is timed with the solid stage it follows.
static int step_overlapped(fsa_sim *s, int n_steps) {
    int rc;
    if ((rc = require_uploaded(s)) || (rc = ensure_side(s))) return rc;
    if ((rc = fsa_particle_cell_update(s)) || (rc = fsa_fill_cell(s))) return rc;
    bool walls_done = false;
    for (int i = 0; i < n_steps; ++i) {
        if ((rc = fsa_couple_calculation(s))) return rc;
        if (!walls_done && (rc = fsa_virt_update(s))) return rc;
        if ((rc = fsa_fluid_calculation(s))) return rc;
        if (i == n_steps - 1) return fsa_solid_calculation(s);
        HIP_TRY(hipEventRecord(s->ev_fork, s->stream));
        HIP_TRY(hipStreamWaitEvent(s->side, s->ev_fork, 0));
        // context stream: this step's solid sub-steps (their arguments name
        // the current fluid buffers), then the next solid cell build
        if (s->P.ns) {
            if ((rc = stage_open(s, FSA_STAGE_SOLID)) || (rc = solid_sub_steps(s))) return rc;
            HIP_TRY(launch_cells_one(make_cell_job(s->P, s->solid, s->jumps, true), s->stream));
            flip(s->solid);
            stage_end(s, FSA_STAGE_SOLID);
        }
        // side stream: the next fluid cell build, then the next wall pass
        if (s->fluid.n) {
            if ((rc = stage_open(s, FSA_STAGE_CELL, s->side))) return rc;
            HIP_TRY(launch_cells_one(make_cell_job(s->P, s->fluid, s->jumps, true), s->side));
            flip(s->fluid);
            if (s->P.vlist) HIP_TRY(launch_list_build(list_job(s, s->fluid), s->side));
            stage_end(s, FSA_STAGE_CELL, s->side);
        }
        if (s->P.nv) {
            if ((rc = stage_open(s, FSA_STAGE_VIRT, s->side))) return rc;
            HIP_TRY(launch_virt(s->P, virt_args(s, s->vx.normals_ready ? 0 : 1), s->side));
            s->vx.normals_ready = true;
            stage_end(s, FSA_STAGE_VIRT, s->side);
        }
        HIP_TRY(hipEventRecord(s->ev_join, s->side));
        HIP_TRY(hipStreamWaitEvent(s->stream, s->ev_join, 0));
        s->couple_fresh = false;  // the coupling records describe the old order
        walls_done = true;
    }
    return FSA_OK;
}

int fsa_step(fsa_sim *s, int n_steps) {
    CHECK_SIM(s);
    if (n_steps < 0) return set_err(FSA_ERR_INVALID, "negative step count");
    if (n_steps > 1) return step_overlapped(s, n_steps);
    for (int i = 0; i < n_steps; ++i) {
        int rc;
        if ((rc = fsa_particle_cell_update(s)) || (rc = fsa_fill_cell(s)) || (rc = fsa_couple_calculation(s)) ||
            (rc = fsa_virt_update(s)) || (rc = fsa_fluid_calculation(s)) || (rc = fsa_solid_calculation(s)))
            return rc;
    }
    return FSA_OK;
}

int fsa_synchronize(fsa_sim *s) {
    CHECK_SIM(s);
    HIP_TRY(hipStreamSynchronize(s->stream));
    return FSA_OK;
}

// ---- introspection ---------------------------------------------------------

static KindBuffers *kind_of(fsa_sim *s, int kind) {
    switch (kind) {
    case FSA_FLUID: return &s->fluid;
    case FSA_SOLID: return &s->solid;
    case FSA_VIRT: return &s->virt;
    default: return nullptr;
    }
}

int fsa_get_cell_ids(fsa_sim *s, int kind, int32_t *out) {
    CHECK_SIM(s);
    KindBuffers *k = kind_of(s, kind);
    if (!k || !out) return set_err(FSA_ERR_INVALID, "bad kind or null output");
    std::vector<int32_t> ids, cells;
    int rc;
    if ((rc = download_sorted(s, (const int32_t *)k->cell[k->ord], k->n, cells)) || (rc = sorted_ids(s, *k, ids)))
        return rc;
    for (int t = 0; t < k->n; ++t) out[ids[t]] = cells[t];
    return FSA_OK;
}

int fsa_get_cell_start(fsa_sim *s, int kind, int32_t *out) {
    CHECK_SIM(s);
    KindBuffers *k = kind_of(s, kind);
    if (!k || !out) return set_err(FSA_ERR_INVALID, "bad kind or null output");
    HIP_TRY(hipMemcpyAsync(out, k->cur_start(), sizeof(int32_t) * ((size_t)s->P.ncell + 1), hipMemcpyDeviceToHost,
                           s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return FSA_OK;
}

int fsa_get_sorted_ids(fsa_sim *s, int kind, int32_t *out) {
    CHECK_SIM(s);
    KindBuffers *k = kind_of(s, kind);
    if (!k || !out) return set_err(FSA_ERR_INVALID, "bad kind or null output");
    std::vector<int32_t> ids;
    int rc = sorted_ids(s, *k, ids);
    if (rc) return rc;
    std::copy(ids.begin(), ids.end(), out);
    return FSA_OK;
}

int fsa_get_jump_count(fsa_sim *s, int64_t *count) {
    CHECK_SIM(s);
    if (!count) return set_err(FSA_ERR_INVALID, "null output");
    unsigned long long v = 0;
    HIP_TRY(hipMemcpyAsync(&v, s->jumps, sizeof v, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    *count = (int64_t)v;
    return FSA_OK;
}

int fsa_get_list_stats(fsa_sim *s, int64_t *builds, int64_t *unlisted, int64_t *entries, int64_t *walks) {
    CHECK_SIM(s);
    uint32_t v[kFlagWords] = {};
    if (s->P.vlist) {
        HIP_TRY(hipMemcpyAsync(v, s->lists.flag, sizeof v, hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
    }
    if (builds) *builds = (int64_t)v[kFlagBuilds];
    if (unlisted) *unlisted = (int64_t)v[kFlagUnlisted];
    if (entries) *entries = (int64_t)v[kFlagEntries];
    if (walks) *walks = (int64_t)v[kFlagWalks];
    return FSA_OK;
}

int fsa_get_health(fsa_sim *s, unsigned *flags) {
    CHECK_SIM(s);
    if (!flags) return set_err(FSA_ERR_INVALID, "null output");
    unsigned long long v = 0;
    HIP_TRY(hipMemcpyAsync(&v, s->health, sizeof v, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    *flags = (unsigned)v;
    return FSA_OK;
}

int fsa_set_debug(fsa_sim *s, unsigned flags) {
    CHECK_SIM(s);
    if ((flags & FSA_DEBUG_PAIR_SUMS) && !s->rates && s->P.nf > 0) {
        int rc = s->alloc(&s->rates, s->P.nf);
        if (rc || (rc = s->alloc(&s->pair_counts, s->P.nf))) return rc;
    }
    s->debug = flags;
    return FSA_OK;
}

int fsa_set_timing_stages(fsa_sim *s, unsigned stage_mask) {
    CHECK_SIM(s);
    if (stage_mask >> FSA_STAGE_COUNT) return set_err(FSA_ERR_INVALID, "unknown stage in the timing mask");
    for (int st = 0; st < FSA_STAGE_COUNT; ++st) {
        int rc = stage_collect(s, st);
        if (rc) return rc;
        s->ev_open[st] = false;
        if (!((stage_mask >> st) & 1u)) continue;
        for (int r = 0; r < fsa_sim::kRing; ++r) {
            if (!s->ev_begin[st][r]) HIP_TRY(hipEventCreate(&s->ev_begin[st][r]));
            if (!s->ev_end[st][r]) HIP_TRY(hipEventCreate(&s->ev_end[st][r]));
        }
    }
    s->timing = stage_mask;
    return FSA_OK;
}

int fsa_set_timing(fsa_sim *s, int enabled) {
    return fsa_set_timing_stages(s, enabled ? (1u << FSA_STAGE_COUNT) - 1u : 0u);
}

int fsa_get_stage_times(fsa_sim *s, double *ms, int64_t *samples) {
    CHECK_SIM(s);
    for (int st = 0; st < FSA_STAGE_COUNT; ++st) {
        int rc = stage_collect(s, st);
        if (rc) return rc;
        if (ms) ms[st] = s->stage_ms[st];
        if (samples) samples[st] = s->stage_n[st];
    }
    return FSA_OK;
}

int fsa_reset_stage_times(fsa_sim *s) {
    CHECK_SIM(s);
    for (int st = 0; st < FSA_STAGE_COUNT; ++st) {
        int rc = stage_collect(s, st);
        if (rc) return rc;
        s->stage_ms[st] = 0.0;
        s->stage_n[st] = 0;
    }
    return FSA_OK;
}

static ObserveArgs observe_args(fsa_sim *s) {
    ObserveArgs a{};
    a.fpos = s->fluid.pos[s->fluid.st];
    a.fvel = s->fluid.vel[s->fluid.st];
    a.fstart = s->fluid.cur_start();
    a.spos = s->solid.pos[s->solid.st];
    a.svel = s->solid.vel[s->solid.st];
    a.sref = s->solid.ref[s->solid.ord];
    a.scpos = s->solid.cpos[s->solid.st];
    a.tip_y = s->tip_y;
    return a;
}

int fsa_observe(fsa_sim *s, fsa_observables *out) {
    CHECK_SIM(s);
    if (!out) return set_err(FSA_ERR_INVALID, "null output");
    int rc = require_uploaded(s);
    if (rc) return rc;
    HIP_TRY(launch_observe(s->P, observe_args(s), s->obs_partial, s->obs_out, s->stream));
    double v[kObserveSums];
    HIP_TRY(hipMemcpyAsync(v, s->obs_out, sizeof v, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    const double dp = s->params.particle_diameter;
    const double V = dp * dp * dp;
    const double ms = (double)s->params.solid_reference_density * V;
    *out = fsa_observables{};
    out->fluid_mass = v[0] * V;
    for (int a = 0; a < 3; ++a) {
        out->fluid_momentum[a] = v[1 + a] * V;
        out->solid_momentum[a] = v[7 + a] * ms;
    }
    out->fluid_kinetic_energy = v[4] * V;
    out->fluid_potential_energy = v[5] * V;
    out->front_x = s->P.nf > 0 ? v[6] : 0.0;
    out->gate_tip_count = (int64_t)v[13];
    for (int a = 0; a < 3; ++a) out->gate_tip_displacement[a] = v[13] > 0 ? v[10 + a] / v[13] : 0.0;
    // CFL bounds from the FP64 speed maxima
    out->max_fluid_speed = std::sqrt(v[14]);
    out->max_solid_speed = std::sqrt(v[15]);
    const double h = s->params.smoothing_length;
    out->cfl_dt = 0.25 * h / ((double)s->params.fluid_sound_speed + out->max_fluid_speed);
    out->solid_cfl_dt = s->P.ns > 0 ? 0.3 * h / ((double)s->params.solid_sound_speed + out->max_solid_speed) : 0.0;
    return FSA_OK;
}

int fsa_pair_census(fsa_sim *s, int64_t counts[6]) {
    CHECK_SIM(s);
    if (!counts) return set_err(FSA_ERR_INVALID, "null output");
    int rc = require_uploaded(s);
    if (rc) return rc;
    if (!s->census) {
        if ((rc = s->alloc(&s->census, 6))) return rc;
    }
    CensusArgs a{};
    a.fpos = s->fluid.pos[s->fluid.st];
    a.spos = s->solid.pos[s->solid.st];
    a.vpos = s->virt.pos[0];
    a.fcell = s->fluid.cell[s->fluid.ord];
    a.fstart = s->fluid.cur_start();
    a.sstart = s->solid.cur_start();
    a.vstart = s->virt.cur_start();
    HIP_TRY(launch_census(s->P, a, s->census, s->stream));
    unsigned long long v[6];
    HIP_TRY(hipMemcpyAsync(v, s->census, sizeof v, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (int k = 0; k < 6; ++k) counts[k] = (int64_t)v[k];
    return FSA_OK;
}

int fsa_probe_pressure(fsa_sim *s, const float *points, int n, float *pressure) {
    CHECK_SIM(s);
    if (n < 0 || (n > 0 && (!points || !pressure))) return set_err(FSA_ERR_INVALID, "bad probe arguments");
    int rc = require_uploaded(s);
    if (rc) return rc;
    if (n == 0) return FSA_OK;
    if ((rc = check_in_grid(s, points, n, "probe"))) return rc;
    if (n > s->probe_cap) {
        float *p = nullptr, *o = nullptr;
        if ((rc = s->alloc(&p, 3 * (size_t)n)) || (rc = s->alloc(&o, (size_t)n))) return rc;
        s->probe_pts = p;  // the old buffers stay owned by the context until fsa_destroy
        s->probe_out = o;
        s->probe_cap = n;
    }
    HIP_TRY(hipMemcpyAsync(s->probe_pts, points, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(launch_probe(s->P, observe_args(s), s->probe_pts, n, s->probe_out, s->stream));
    HIP_TRY(hipMemcpyAsync(pressure, s->probe_out, sizeof(float) * n, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return FSA_OK;
}

int fsa_probe_pressure_noalloc(fsa_sim *s, const float *points, int n, float *pressure) {
    CHECK_SIM(s);
    if (n < 0 || (n > 0 && (!points || !pressure))) return set_err(FSA_ERR_INVALID, "bad probe arguments");
    if (n > FSA_PROBE_NOALLOC_MAX)
        return set_err(FSA_ERR_INVALID, "fsa_probe_pressure_noalloc takes at most " +
                                            std::to_string(FSA_PROBE_NOALLOC_MAX) + " points per call");
    int rc = require_uploaded(s);
    if (rc) return rc;
    if (n == 0) return FSA_OK;
    if ((rc = check_in_grid(s, points, n, "probe"))) return rc;
    // the call writes every point it reads and every pressure it returns:
    // nothing in the staging area needs clearing
    float *pts = s->probe_scratch;
    float *out = s->probe_scratch + 3 * FSA_PROBE_NOALLOC_MAX;
    HIP_TRY(hipMemcpyAsync(pts, points, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(launch_probe(s->P, observe_args(s), pts, n, out, s->stream));
    HIP_TRY(hipMemcpyAsync(pressure, out, sizeof(float) * n, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return FSA_OK;
}

int fsa_set_fluid_threads(fsa_sim *s, int threads) {
    CHECK_SIM(s);
    if (!fluid_threads_valid(threads)) return set_err(FSA_ERR_INVALID, "fluid workgroup size must be 64, 128 or 256");
    s->fluid_threads = threads;
    return FSA_OK;
}

int fsa_get_fluid_threads(fsa_sim *s, int *threads) {
    CHECK_SIM(s);
    if (!threads) return set_err(FSA_ERR_INVALID, "null output");
    *threads = s->fluid_threads;
    return FSA_OK;
}

int fsa_autotune_fluid(fsa_sim *s, int repeats, int *threads) {
    CHECK_SIM(s);
    if (repeats < 1) return set_err(FSA_ERR_INVALID, "repeats must be >= 1");
    int rc = require_uploaded(s);
    if (rc) return rc;
    if (s->P.nf > 0) {
        // the pass reads the current state buffer and writes the other one,
        // which the next step overwrites: timing it leaves the scene as it was
        const FluidStepArgs a = fluid_args(s);
        hipEvent_t t0 = nullptr, t1 = nullptr;
        HIP_TRY(hipEventCreate(&t0));
        if (hipEventCreate(&t1) != hipSuccess) {
            (void)hipEventDestroy(t0);
            return set_err(FSA_ERR_HIP, "hipEventCreate failed");
        }
        float best = 3.0e38f;
        int best_t = s->fluid_threads;
        hipError_t e = hipSuccess;
        for (int t : {64, 128, 256}) {
            if ((e = launch_fluid(s->P, a, t, s->stream)) != hipSuccess) break;  // warm-up
            if ((e = hipEventRecord(t0, s->stream)) != hipSuccess) break;
            for (int r = 0; r < repeats && e == hipSuccess; ++r) e = launch_fluid(s->P, a, t, s->stream);
            if (e != hipSuccess || (e = hipEventRecord(t1, s->stream)) != hipSuccess ||
                (e = hipEventSynchronize(t1)) != hipSuccess)
                break;
            float ms = 0.0f;
            if ((e = hipEventElapsedTime(&ms, t0, t1)) != hipSuccess) break;
            if (ms < best) {
                best = ms;
                best_t = t;
            }
        }
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        HIP_TRY(e);
        s->fluid_threads = best_t;
    }
    if (threads) *threads = s->fluid_threads;
    return FSA_OK;
}

int fsa_memory_used(fsa_sim *s, int64_t *bytes) {
    CHECK_SIM(s);
    if (!bytes) return set_err(FSA_ERR_INVALID, "null output");
    *bytes = s->bytes;
    return FSA_OK;
}

}  // extern "C"
