// fsa_context.h -- the per-scene context behind the opaque fsa_sim handle and
// the helpers the C ABI files share.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "fluidsim_amd.h"
#include "fsa_internal.h"

namespace fsa {

// records the message for fsa_last_error() and returns `code`
int set_err(int code, const std::string &msg);

}  // namespace fsa

#define FSA_HIP_TRY(expr)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return ::fsa::set_err(e_ == hipErrorOutOfMemory ? FSA_ERR_NOMEM : FSA_ERR_HIP,         \
                                  std::string(#expr) + ": " + hipGetErrorString(e_));              \
    } while (0)

struct fsa_sim {
    fsa_params params{};
    fsa::DevParams P{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = 0;
    // fsa_step runs the next step's fluid cell build and wall pass here,
    // beside the solid sub-steps on `stream` (forked from and joined back
    // into it); created on the first step that uses it
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;

    fsa::KindBuffers fluid, solid, virt;
    fsa::NeighbourLists lists;  // fluid neighbour lists (P.vlist)
    fsa::SolidExtra sx;
    fsa::VirtExtra vx;
    float4 *virt_x = nullptr;  // constant wall positions (x, y, z, 0)
    float4 *rates = nullptr;   // fluid (ax, ay, az, drho) when debugging
    int4 *pair_counts = nullptr;  // and the partners each found per kind
    unsigned long long *jumps = nullptr;   // [0]: particles that moved more than a cell in a step
    unsigned long long *health = nullptr;  // = jumps + 1: FSA_HEALTH_* bits since the last upload
    // observables
    double *obs_partial = nullptr, *obs_out = nullptr;
    float tip_y = 3.0e38f;  // reference y of the gate tip row (none: above everything)
    unsigned long long *census = nullptr;
    float *probe_pts = nullptr;  // (x, y, z) per point
    float *probe_out = nullptr;
    int probe_cap = 0;
    // this device's staging area of fsa_probe_pressure_noalloc (a device
    // address of the library's static array, shared by every context on the device)
    float *probe_scratch = nullptr;

    int fluid_threads = fsa::kFluidThreadsDefault;  // workgroup size of the fluid pass

    bool uploaded[3] = {false, false, false};
    bool hashed = false;  // cell ids/ranks of the current order are fresh
    // the coupling pass ran on the current order, so the solid force pass may
    // replay its fluid-partner records
    bool couple_fresh = false;
    unsigned debug = 0;
    // unique per context over the process lifetime: a batch keys its recorded
    // launches on it (a new context may reuse a freed one's address)
    uint64_t serial = 0;

    // stage timing: a ring of event pairs per stage, read back lazily so the
    // host never waits on the GPU while it enqueues steps
    static constexpr int kRing = 256;
    unsigned timing = 0;  // stages timed: bit 1u << FSA_STAGE_x
    hipEvent_t ev_begin[FSA_STAGE_COUNT][kRing] = {};
    hipEvent_t ev_end[FSA_STAGE_COUNT][kRing] = {};
    int ev_head[FSA_STAGE_COUNT] = {};     // next ring slot
    int ev_pending[FSA_STAGE_COUNT] = {};  // recorded, not yet read back
    bool ev_open[FSA_STAGE_COUNT] = {};
    double stage_ms[FSA_STAGE_COUNT] = {};
    int64_t stage_n[FSA_STAGE_COUNT] = {};

    std::vector<void *> allocations;
    int64_t bytes = 0;
    bool dry = false;  // only count the bytes (fsa_footprint)

    template <class T> int alloc(fsa::gptr<T> *p, size_t count) {
        T *q = nullptr;
        const int rc = alloc(&q, count);
        *p = q;
        return rc;
    }
    template <class T> int alloc(T **p, size_t count) {
        *p = nullptr;
        if (count == 0) return FSA_OK;
        if (dry) {
            bytes += (int64_t)(count * sizeof(T));
            return FSA_OK;
        }
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, count * sizeof(T));
        if (e != hipSuccess)
            return fsa::set_err(FSA_ERR_NOMEM, std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
                                                   " bytes: " + hipGetErrorString(e));
        allocations.push_back(q);
        bytes += (int64_t)(count * sizeof(T));
        *p = static_cast<T *>(q);
        return FSA_OK;
    }
};

namespace fsa {

// FSA_ERR_STATE unless every kind with particles has been uploaded
int require_uploaded(fsa_sim *s);
// clears the counts a stage-call hash left without a sort (see fsa_capi.hip)
int drop_pending_hash(fsa_sim *s);

// stage arguments over given fluid/solid buffer indices (a batch plans a
// step ahead of the context's own indices) ...
SolidArgs solid_args(const fsa_sim *s, const KindBuffers &f, const KindBuffers &so);
VirtArgs virt_args(const fsa_sim *s, const KindBuffers &f, int compute_normals);
FluidStepArgs fluid_args(const fsa_sim *s, const KindBuffers &f, const KindBuffers &so);  // rates off
// the list-build job of the fluid buffers f (after a fluid sort)
ListJob list_job(const fsa_sim *s, const KindBuffers &f);
// ... and over the context's current ones
inline SolidArgs solid_args(const fsa_sim *s) { return solid_args(s, s->fluid, s->solid); }
inline VirtArgs virt_args(const fsa_sim *s, int compute_normals) { return virt_args(s, s->fluid, compute_normals); }
inline FluidStepArgs fluid_args(const fsa_sim *s) { return fluid_args(s, s->fluid, s->solid); }

}  // namespace fsa
