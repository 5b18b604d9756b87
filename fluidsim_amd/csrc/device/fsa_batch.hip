// fsa_batch.hip -- many scenes advanced by one launch per stage.
//
// The reference runs one scene per process (process-global parameter and
// pointer blocks, src/sim.cu:9-10, src/memory_copy.cu:5-6) and its step is a
// fixed sequence of stages (src/calculation.cu:23-33).  A batch keeps that
// sequence but launches every stage once for all of its scenes: the grid of
// a launch is the concatenation of the scenes' grids, each scene's range
// rounded up to whole XCD rounds (8 blocks), and a block finds its scene by a
// binary search in the launch's prefix table (k_many).  Each scene runs the
// same operation code as when it is stepped alone, so results are
// bit-identical to fsa_step on each scene.
//
// Scene-dependent launch arguments change with the double-buffer indices,
// which repeat every two steps (the fluid order buffer and the solid state
// and order buffers flip once per step; the fluid state buffer flips twice).
// A plan therefore holds the job and prefix tables of two consecutive steps
// in device memory, each recorded as one hipGraph, and a step is one graph
// launch.  The plan is keyed on the scene list (context serials) and on
// every scene's buffer indices; any mismatch -- another scene list, or a
// scene stepped alone in between -- rebuilds it.
//
// Solid sub-steps: the scenes' solid_sub_step counts differ (a stiffer gate
// needs more).  Sub-step k launches the three solid stages over the scenes
// with more than k sub-steps; the others get empty ranges.
//
// Concurrency: the scenes are split over a few chains (fsa_batch_set_streams,
// 2 by default), each with its own stream and its own recorded steps.  A step
// launches every chain's graph; the chains are independent, so while one
// runs its latency-bound stages (the solid sub-steps launch a few hundred
// workgroups each, the cell build a string of short passes) the other's
// fluid pass fills the CUs.  The split is deterministic in the scene list
// (heaviest first to the least loaded chain), so a list stepped again finds
// its chains' plans.
//
// Launch shape: the fluid pass runs in workgroups of 64, 128 or 256 threads
// (identical results).  With FSA_BATCH_AUTOTUNE every recording first times
// the batched fluid pass at each size on the scenes' current state -- the
// pass reads the current state buffer and writes the other one, which the
// next step overwrites, so the scenes do not advance -- and keeps the
// fastest; otherwise it takes the first listed scene's size.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "fluidsim_amd.h"
#include "fsa_context.h"
#include "fsa_internal.h"

using namespace fsa;

#define HIP_TRY(expr) FSA_HIP_TRY(expr)

namespace {

constexpr int kXcdRound = 8;  // blocks per round of the XCD dispatcher
constexpr int kMaxChains = 8;  // concurrent launch chains of a batch
constexpr size_t kAlign = 256;

inline int round_xcd(int blocks) { return (blocks + kXcdRound - 1) / kXcdRound * kXcdRound; }
inline size_t align_up(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

// indices of the double buffers of one scene at the start of a step
struct Phase {
    int fst, ford, fhst, sst, sord;
    bool operator==(const Phase &o) const {
        return fst == o.fst && ford == o.ford && fhst == o.fhst && sst == o.sst && sord == o.sord;
    }
};
Phase phase_of(const fsa_sim *s) {
    return {s->fluid.st, s->fluid.ord, s->fluid.hst, s->solid.st, s->solid.ord};
}
// the indices one step later: the sort flips state and order of every kind
// with particles, the fluid pass flips the fluid state back (and the home
// state of a scene with neighbour lists)
Phase next_phase(const fsa_sim *s, Phase p) {
    if (s->P.nf > 0) p.ford ^= 1;
    if (s->P.vlist) p.fhst ^= 1;
    if (s->P.ns > 0) {
        p.sst ^= 1;
        p.sord ^= 1;
    }
    return p;
}
void apply_phase(fsa_sim *s, const Phase &p) {
    s->fluid.st = p.fst;
    s->fluid.ord = p.ford;
    s->fluid.hst = p.fhst;
    s->solid.st = p.sst;
    s->solid.ord = p.sord;
}

// one launch of a step: which op, which job table, its prefix table
struct Launch {
    enum Family { kCells, kList, kCouple, kVirt, kFluid, kSolid } family;
    BatchOp op;
    ListOp lop;  // kList
    int njobs;
    size_t prefix_off;  // byte offset of the prefix table in the blob
    int nblocks;
};

// the launches and tables of one step
struct StepTables {
    size_t cells_off = 0, list_off = 0, couple_off = 0, virt_off = 0, fluid_off = 0, solid_off = 0;
    int ncells = 0, nscenes = 0;
    std::vector<Launch> launches;
};

// One chain of launches: the scenes it steps, its stream, its plan.
struct Chain {
    hipStream_t stream = nullptr;  // owned; default priority
    unsigned flags = 0;

    // the plan
    std::vector<uint64_t> serials;
    std::vector<Phase> phase0;  // buffer indices at the start of step 0 of the plan
    StepTables tables[2];
    char *dev_blob = nullptr;
    size_t blob_cap = 0;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    bool planned = false;
    int fluid_threads = 0;  // workgroup size of the recorded fluid launches
    bool packed = false;    // the scenes use the packed precision form (all or none)
    int64_t plans_built = 0;
    // the autotuner's last choice and the chain's fluid total when it was
    // made: a plan rebuild re-times only when the total has moved by more
    // than kRetuneFraction (continuous batching rebuilds on every admit and
    // finish); its device tables persist between timings
    int tuned_threads = 0;
    int64_t tuned_fluid = 0;
    char *tune_blob = nullptr;
    size_t tune_cap = 0;
    int64_t tunes = 0;
    double tune_ms = 0.0;  // host wall time spent timing
};

constexpr double kRetuneFraction = 0.25;

}  // namespace

struct fsa_batch {
    int device = 0;
    hipStream_t stream = nullptr;  // the batch's stream: work is ordered after it and it waits for the chains
    bool own_stream = false;
    unsigned flags = 0;
    hipEvent_t ev_sync = nullptr;
    int nchains = 3;               // concurrent launch chains the scenes are split over
    std::vector<Chain *> chains;   // created as needed
    std::vector<hipEvent_t> ev_chain;
    int64_t steps_run = 0;
};

namespace {

void drop_graphs(Chain *b) {
    for (int p = 0; p < 2; ++p) {
        if (b->exec[p]) (void)hipGraphExecDestroy(b->exec[p]);
        if (b->graph[p]) (void)hipGraphDestroy(b->graph[p]);
        b->exec[p] = nullptr;
        b->graph[p] = nullptr;
    }
}

// host image of one step's tables, appended to `blob`
template <class T> size_t put(std::vector<char> &blob, const T *data, size_t count) {
    const size_t off = align_up(blob.size());
    blob.resize(off + sizeof(T) * count);
    if (count) std::memcpy(blob.data() + off, data, sizeof(T) * count);
    return off;
}

void add_launch(std::vector<char> &blob, StepTables &t, Launch::Family fam, BatchOp op,
                const std::vector<int> &blocks, ListOp lop = kLDecide) {
    std::vector<int> prefix(blocks.size() + 1, 0);
    for (size_t i = 0; i < blocks.size(); ++i) prefix[i + 1] = prefix[i] + round_xcd(blocks[i]);
    if (prefix.back() == 0) return;  // no scene has work in this stage
    Launch l{};
    l.family = fam;
    l.op = op;
    l.lop = lop;
    l.njobs = (int)blocks.size();
    l.nblocks = prefix.back();
    l.prefix_off = put(blob, prefix.data(), prefix.size());
    t.launches.push_back(l);
}

// tables of one step for scenes whose buffer indices are `ph`
void build_step(const std::vector<fsa_sim *> &sims, const std::vector<Phase> &ph, int fluid_threads,
                std::vector<char> &blob, StepTables &t) {
    const int n = (int)sims.size();
    std::vector<CellJob> cells;
    std::vector<ListJob> lists(n);
    std::vector<VirtJob> virt(n);
    std::vector<FluidJob> fluid(n);
    std::vector<SolidJob> couple(n), solid(n);
    int smax = 0;
    for (int i = 0; i < n; ++i) {
        const fsa_sim *s = sims[i];
        KindBuffers f = s->fluid, so = s->solid;
        f.st = ph[i].fst;
        f.ord = ph[i].ford;
        f.hst = ph[i].fhst;
        so.st = ph[i].sst;
        so.ord = ph[i].sord;
        cells.push_back(make_cell_job(s->P, f, s->jumps, true));
        cells.push_back(make_cell_job(s->P, so, s->jumps, true));
        // after the sort
        if (f.n) {
            f.st ^= 1;
            f.ord ^= 1;
        }
        if (so.n) {
            so.st ^= 1;
            so.ord ^= 1;
        }
        lists[i] = list_job(s, f);
        couple[i] = SolidJob{s->P, solid_args(s, f, so)};
        virt[i] = VirtJob{s->P, virt_args(s, f, 0)};
        fluid[i] = FluidJob{s->P, fluid_args(s, f, so)};
        // the solid sub-steps see the fluid after its pass (src/calculation.cu:31-32)
        if (f.n) f.st ^= 1;
        solid[i] = SolidJob{s->P, solid_args(s, f, so)};
        solid[i].a.fluid_rec = (s->P.nf > 0 && s->P.ns > 0) ? 1 : 0;  // the couple launch of this step records
        solid[i].a.solid_rec = 1;
        smax = std::max(smax, s->P.ns > 0 ? s->P.sub : 0);
    }
    t = StepTables{};
    t.nscenes = n;
    t.ncells = (int)cells.size();
    t.cells_off = put(blob, cells.data(), cells.size());
    t.list_off = put(blob, lists.data(), lists.size());
    t.couple_off = put(blob, couple.data(), couple.size());
    t.virt_off = put(blob, virt.data(), virt.size());
    t.fluid_off = put(blob, fluid.data(), fluid.size());
    t.solid_off = put(blob, solid.data(), solid.size());

    std::vector<int> blocks;
    // (the counts are zero between builds: no fill)
    for (BatchOp op : {kBHash, kBScanTiles, kBScanSums, kBScanAdd, kBPlace, kBGather}) {
        blocks.clear();
        for (const CellJob &j : cells) blocks.push_back(blocks_cells(op, j));
        add_launch(blob, t, Launch::kCells, op, blocks);
    }
    // neighbour lists: the build decision, then (if due) the new home order and the lists
    for (ListOp op : {kLDecide, kLBuild}) {
        blocks.clear();
        for (const ListJob &j : lists) blocks.push_back(blocks_list(op, j));
        add_launch(blob, t, Launch::kList, kBFillCounts, blocks, op);
    }
    // couple (src/solid.cu:346-359: only with both fluid and solid)
    blocks.clear();
    for (const SolidJob &j : couple) blocks.push_back(j.P.nf > 0 ? blocks_solid(kBCouple, j) : 0);
    add_launch(blob, t, Launch::kCouple, kBCouple, blocks);
    blocks.clear();
    for (const VirtJob &j : virt) blocks.push_back(blocks_virt(j));
    add_launch(blob, t, Launch::kVirt, kBFillCounts, blocks);
    blocks.clear();
    for (const FluidJob &j : fluid) blocks.push_back(blocks_fluid(j, fluid_threads));
    add_launch(blob, t, Launch::kFluid, kBFillCounts, blocks);
    for (int sub = 0; sub < smax; ++sub)
        for (BatchOp op : {sub == 0 ? kBDeformRecord : kBDeform, sub == 0 ? kBDeformTailRecord : kBDeformTail,
                           kBForce, kBSolidUpdate}) {
            blocks.clear();
            for (const SolidJob &j : solid) blocks.push_back(sub < j.P.sub ? blocks_solid(op, j) : 0);
            add_launch(blob, t, Launch::kSolid, op, blocks);
        }
}

hipError_t enqueue_step(Chain *b, const StepTables &t) {
    const char *base = b->dev_blob;
    for (const Launch &l : t.launches) {
        const int *prefix = reinterpret_cast<const int *>(base + l.prefix_off);
        hipError_t e = hipSuccess;
        switch (l.family) {
        case Launch::kCells:
            e = launch_many_cells(l.op, reinterpret_cast<const CellJob *>(base + t.cells_off), prefix, l.njobs,
                                  l.nblocks, b->stream);
            break;
        case Launch::kList:
            e = launch_many_list(l.lop, reinterpret_cast<const ListJob *>(base + t.list_off), prefix, l.njobs,
                                 l.nblocks, b->stream);
            break;
        case Launch::kCouple:
            e = launch_many_solid(l.op, reinterpret_cast<const SolidJob *>(base + t.couple_off), prefix, l.njobs,
                                  l.nblocks, b->stream);
            break;
        case Launch::kVirt:
            e = launch_many_virt(reinterpret_cast<const VirtJob *>(base + t.virt_off), prefix, l.njobs, l.nblocks,
                                 b->stream);
            break;
        case Launch::kFluid:
            e = launch_many_fluid(reinterpret_cast<const FluidJob *>(base + t.fluid_off), prefix, l.njobs,
                                  l.nblocks, b->fluid_threads, b->packed, b->stream);
            break;
        case Launch::kSolid:
            e = launch_many_solid(l.op, reinterpret_cast<const SolidJob *>(base + t.solid_off), prefix, l.njobs,
                                  l.nblocks, b->stream);
            break;
        }
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// the fluid workgroup size for these scenes, timed on their current state
// (see the file comment): each size's best of three interleaved timings,
// and the smallest size unless a larger one is faster by more than 2 %.
// Timed alone, 64- and 128-thread launches come within the noise of each
// other; run beside the other launch chain, the smaller workgroups fill the
// gaps better (measured 6.71-6.77 against 6.80-6.82 ms per 64-scene step).
// 0 on failure with the error recorded.
int tune_fluid(Chain *b, const std::vector<fsa_sim *> &sims) {
    const int n = (int)sims.size();
    std::vector<FluidJob> jobs(n);
    for (int i = 0; i < n; ++i) jobs[i] = FluidJob{sims[i]->P, fluid_args(sims[i])};
    static constexpr int kChoices[3] = {64, 128, 256};
    std::vector<char> blob;
    const size_t joff = put(blob, jobs.data(), jobs.size());
    size_t poff[3];
    int nblk[3];
    for (int c = 0; c < 3; ++c) {
        std::vector<int> prefix(n + 1, 0);
        for (int i = 0; i < n; ++i) prefix[i + 1] = prefix[i] + round_xcd(blocks_fluid(jobs[i], kChoices[c]));
        nblk[c] = prefix.back();
        poff[c] = put(blob, prefix.data(), prefix.size());
    }
    if (nblk[0] == 0) return kFluidThreadsDefault;  // no fluid anywhere
    hipEvent_t t0 = nullptr, t1 = nullptr;
    auto fail = [&](hipError_t e) {
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
        set_err(e == hipErrorOutOfMemory ? FSA_ERR_NOMEM : FSA_ERR_HIP,
                std::string("batch fluid autotune: ") + hipGetErrorString(e));
        return 0;
    };
    hipError_t e;
    if (blob.size() > b->tune_cap) {
        if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) return fail(e);  // the old tables may be read
        if (b->tune_blob) (void)hipFree(b->tune_blob);
        b->tune_blob = nullptr;
        b->tune_cap = 0;
        const size_t cap = std::max(blob.size(), 2 * b->tune_cap);
        if ((e = hipMalloc(&b->tune_blob, cap)) != hipSuccess) return fail(e);
        b->tune_cap = cap;
    }
    char *dev = b->tune_blob;
    if ((e = hipMemcpyAsync(dev, blob.data(), blob.size(), hipMemcpyHostToDevice, b->stream)) != hipSuccess ||
        (e = hipEventCreate(&t0)) != hipSuccess || (e = hipEventCreate(&t1)) != hipSuccess)
        return fail(e);
    float ms_of[3] = {3.0e38f, 3.0e38f, 3.0e38f};
    const FluidJob *dj = reinterpret_cast<const FluidJob *>(dev + joff);
    for (int round = 0; round < 3; ++round)
        for (int c = 0; c < 3; ++c) {
            const int *dp = reinterpret_cast<const int *>(dev + poff[c]);
            const int t = kChoices[c];
            if ((e = launch_many_fluid(dj, dp, n, nblk[c], t, b->packed, b->stream)) != hipSuccess ||  // warm-up
                (e = hipEventRecord(t0, b->stream)) != hipSuccess)
                return fail(e);
            for (int r = 0; r < 2; ++r)
                if ((e = launch_many_fluid(dj, dp, n, nblk[c], t, b->packed, b->stream)) != hipSuccess)
                    return fail(e);
            float ms = 0.0f;
            if ((e = hipEventRecord(t1, b->stream)) != hipSuccess || (e = hipEventSynchronize(t1)) != hipSuccess ||
                (e = hipEventElapsedTime(&ms, t0, t1)) != hipSuccess)
                return fail(e);
            ms_of[c] = std::min(ms_of[c], ms);
        }
    int best_c = 0;
    for (int c = 1; c < 3; ++c)
        if (ms_of[c] < 0.98f * ms_of[best_c]) best_c = c;
    const int best_t = kChoices[best_c];
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return best_t;
}

int build_plan(Chain *b, const std::vector<fsa_sim *> &sims) {
    drop_graphs(b);
    b->planned = false;
    const int n = (int)sims.size();
    // one fluid launch serves every scene: the packed form is a kernel variant
    b->packed = (sims[0]->P.form & FSA_FORM_PACKED) != 0;
    for (const fsa_sim *s : sims)
        if (((s->P.form & FSA_FORM_PACKED) != 0) != b->packed)
            return set_err(FSA_ERR_INVALID, "the scenes of one batch must all use the packed form or none");
    if (b->flags & FSA_BATCH_AUTOTUNE) {
        int64_t total = 0;
        for (const fsa_sim *s : sims) total += s->P.nf;
        if (b->tuned_threads == 0 || std::abs((double)(total - b->tuned_fluid)) > kRetuneFraction * b->tuned_fluid) {
            const auto w0 = std::chrono::steady_clock::now();
            const int t = tune_fluid(b, sims);
            if (t == 0) return FSA_ERR_HIP;
            b->tune_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
            b->tunes += 1;
            b->tuned_threads = t;
            b->tuned_fluid = total;
        }
        b->fluid_threads = b->tuned_threads;
    } else {
        b->fluid_threads = sims[0]->fluid_threads;
    }
    std::vector<Phase> ph(n), ph1(n);
    for (int i = 0; i < n; ++i) {
        ph[i] = phase_of(sims[i]);
        ph1[i] = next_phase(sims[i], ph[i]);
    }
    std::vector<char> blob;
    build_step(sims, ph, b->fluid_threads, blob, b->tables[0]);
    build_step(sims, ph1, b->fluid_threads, blob, b->tables[1]);
    if (blob.size() > b->blob_cap) {
        HIP_TRY(hipStreamSynchronize(b->stream));  // the old tables may still be read
        if (b->dev_blob) (void)hipFree(b->dev_blob);
        b->dev_blob = nullptr;
        b->blob_cap = 0;
        const size_t cap = std::max(blob.size(), 2 * b->blob_cap);
        HIP_TRY(hipMalloc(&b->dev_blob, cap));
        b->blob_cap = cap;
    }
    HIP_TRY(hipMemcpyAsync(b->dev_blob, blob.data(), blob.size(), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));  // `blob` is pageable host memory and goes out of scope
    if (!(b->flags & FSA_BATCH_NO_GRAPH)) {
        for (int p = 0; p < 2; ++p) {
            HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            const hipError_t e = enqueue_step(b, b->tables[p]);
            hipGraph_t g = nullptr;
            const hipError_t e2 = hipStreamEndCapture(b->stream, &g);
            if (e != hipSuccess || e2 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return set_err(FSA_ERR_HIP, std::string("batch step capture: ") +
                                                hipGetErrorString(e != hipSuccess ? e : e2));
            }
            b->graph[p] = g;
            HIP_TRY(hipGraphInstantiate(&b->exec[p], g, nullptr, nullptr, 0));
        }
    }
    b->serials.resize(n);
    for (int i = 0; i < n; ++i) b->serials[i] = sims[i]->serial;
    b->phase0 = ph;
    b->planned = true;
    b->plans_built += 1;
    return FSA_OK;
}

// which recorded step matches the scenes' current buffer indices (-1: none)
int plan_phase(const Chain *b, const std::vector<fsa_sim *> &sims) {
    if (!b->planned || sims.size() != b->serials.size()) return -1;
    if (!(b->flags & FSA_BATCH_AUTOTUNE) && sims[0]->fluid_threads != b->fluid_threads) return -1;
    for (size_t i = 0; i < sims.size(); ++i)
        if (sims[i]->serial != b->serials[i]) return -1;
    for (int p = 0; p < 2; ++p) {
        bool all = true;
        for (size_t i = 0; i < sims.size() && all; ++i) {
            const Phase want = p == 0 ? b->phase0[i] : next_phase(sims[i], b->phase0[i]);
            all = phase_of(sims[i]) == want;
        }
        if (all) return p;
    }
    return -1;
}

// a chain of its own, with its own stream (default priority)
Chain *new_chain(unsigned flags) {
    Chain *c = new (std::nothrow) Chain;
    if (!c) return nullptr;
    c->flags = flags;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return nullptr;
    }
    return c;
}

void free_chain(Chain *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);
    if (c->dev_blob) (void)hipFree(c->dev_blob);
    if (c->tune_blob) (void)hipFree(c->tune_blob);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

// Split the scenes over `nch` chains: heaviest first, each to the chain with
// the least work so far (work ~ fluid particles + 4 x solid particles x
// sub-steps).  Deterministic in the scene list, so a list stepped again
// finds its chains' plans.
std::vector<std::vector<fsa_sim *>> split(const std::vector<fsa_sim *> &sims, int nch) {
    std::vector<int> order(sims.size());
    std::vector<double> work(sims.size());
    for (size_t i = 0; i < sims.size(); ++i) {
        order[i] = (int)i;
        work[i] = (double)sims[i]->P.nf + 4.0 * sims[i]->P.ns * sims[i]->P.sub + 1.0;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return work[a] > work[c]; });
    std::vector<std::vector<fsa_sim *>> part(nch);
    std::vector<double> load(nch, 0.0);
    for (int i : order) {
        const int c = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        part[c].push_back(sims[i]);
        load[c] += work[i];
    }
    // each chain keeps its scenes in list order
    for (auto &v : part)
        std::stable_sort(v.begin(), v.end(), [&](fsa_sim *a, fsa_sim *c) {
            return std::find(sims.begin(), sims.end(), a) < std::find(sims.begin(), sims.end(), c);
        });
    return part;
}

}  // namespace

extern "C" {

int fsa_batch_create(void *stream, unsigned flags, fsa_batch **out) {
    if (!out) return set_err(FSA_ERR_INVALID, "null argument");
    *out = nullptr;
    if (flags & ~(unsigned)(FSA_BATCH_NO_GRAPH | FSA_BATCH_AUTOTUNE))
        return set_err(FSA_ERR_INVALID, "unknown batch flag");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_err(FSA_ERR_NODEVICE, "no HIP device visible");
    fsa_batch *b = new (std::nothrow) fsa_batch;
    if (!b) return set_err(FSA_ERR_NOMEM, "host allocation failed");
    b->flags = flags;
    (void)hipGetDevice(&b->device);
    if (stream) {
        b->stream = static_cast<hipStream_t>(stream);
    } else {
        if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
            delete b;
            return set_err(FSA_ERR_HIP, "hipStreamCreate failed");
        }
        b->own_stream = true;
    }
    if (hipEventCreateWithFlags(&b->ev_sync, hipEventDisableTiming) != hipSuccess) {
        fsa_batch_destroy(b);
        return set_err(FSA_ERR_HIP, "hipEventCreate failed");
    }
    *out = b;
    return FSA_OK;
}

int fsa_batch_destroy(fsa_batch *b) {
    if (!b) return FSA_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (Chain *c : b->chains) free_chain(c);
    for (hipEvent_t e : b->ev_chain) (void)hipEventDestroy(e);
    if (b->ev_sync) (void)hipEventDestroy(b->ev_sync);
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return FSA_OK;
}

int fsa_batch_set_streams(fsa_batch *b, int n) {
    if (!b) return set_err(FSA_ERR_INVALID, "null batch");
    if (n < 1 || n > kMaxChains) return set_err(FSA_ERR_INVALID, "stream count must be 1..8");
    b->nchains = n;
    return FSA_OK;
}

int fsa_batch_step(fsa_batch *b, fsa_sim *const *scenes, int n_scenes, int n_steps) {
    if (!b) return set_err(FSA_ERR_INVALID, "null batch");
    if (n_scenes < 0 || n_steps < 0) return set_err(FSA_ERR_INVALID, "negative scene or step count");
    if (n_scenes > 0 && !scenes) return set_err(FSA_ERR_INVALID, "null scene list");
    if (n_scenes == 0 || n_steps == 0) return FSA_OK;
    std::vector<fsa_sim *> sims(scenes, scenes + n_scenes);
    for (int i = 0; i < n_scenes; ++i) {
        fsa_sim *s = sims[i];
        if (!s) return set_err(FSA_ERR_INVALID, "null scene in the list");
        if (s->device != b->device) return set_err(FSA_ERR_INVALID, "scene lives on another device than the batch");
        int rc = require_uploaded(s);
        if (rc) return rc;
        for (int j = 0; j < i; ++j)
            if (sims[j] == s) return set_err(FSA_ERR_INVALID, "scene listed twice");
    }
    // one batch launch serves every scene of a chain: the packed form is a
    // kernel variant, so the whole list must agree on it
    const bool packed = (sims[0]->P.form & FSA_FORM_PACKED) != 0;
    for (const fsa_sim *s : sims)
        if (((s->P.form & FSA_FORM_PACKED) != 0) != packed)
            return set_err(FSA_ERR_INVALID, "the scenes of one batch must all use the packed form or none");
    const int nch = std::min(b->nchains, n_scenes);
    while ((int)b->chains.size() < nch) {
        Chain *c = new_chain(b->flags);
        hipEvent_t e = nullptr;
        if (!c || hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            free_chain(c);
            return set_err(FSA_ERR_HIP, "batch chain creation failed");
        }
        b->chains.push_back(c);
        b->ev_chain.push_back(e);
    }
    // a scene hashed by a stage call but not sorted has counts to clear
    for (fsa_sim *s : sims) {
        int rc = drop_pending_hash(s);
        if (rc) return rc;
    }
    // the chains start after the work already queued on the scenes' streams
    // and on the batch stream
    for (fsa_sim *s : sims) {
        if (s->stream == b->stream) continue;
        HIP_TRY(hipEventRecord(b->ev_sync, s->stream));
        HIP_TRY(hipStreamWaitEvent(b->stream, b->ev_sync, 0));
    }
    HIP_TRY(hipEventRecord(b->ev_sync, b->stream));
    const std::vector<std::vector<fsa_sim *>> part = split(sims, nch);
    std::vector<int> phase(nch, 0);
    for (int c = 0; c < nch; ++c) {
        Chain *ch = b->chains[c];
        HIP_TRY(hipStreamWaitEvent(ch->stream, b->ev_sync, 0));
        phase[c] = plan_phase(ch, part[c]);
        if (phase[c] < 0) {
            int rc = build_plan(ch, part[c]);
            if (rc) return rc;
            phase[c] = 0;
        }
        // wall normals are computed once, on a scene's first update (the
        // walls do not move); the step's own wall pass then reads them
        for (fsa_sim *s : part[c]) {
            if (s->P.nv == 0 || s->vx.normals_ready) continue;
            HIP_TRY(launch_virt(s->P, virt_args(s, 1), ch->stream));
            s->vx.normals_ready = true;
        }
    }
    for (int k = 0; k < n_steps; ++k) {
        for (int c = 0; c < nch; ++c) {
            Chain *ch = b->chains[c];
            if (ch->exec[phase[c]]) {
                HIP_TRY(hipGraphLaunch(ch->exec[phase[c]], ch->stream));
            } else {
                HIP_TRY(enqueue_step(ch, ch->tables[phase[c]]));
            }
            phase[c] ^= 1;
        }
        for (fsa_sim *s : sims) {
            apply_phase(s, next_phase(s, phase_of(s)));
            s->hashed = false;
            s->couple_fresh = s->P.nf > 0 && s->P.ns > 0;
        }
        b->steps_run += 1;
    }
    // the batch stream, then the scenes' streams, wait for the chains
    for (int c = 0; c < nch; ++c) {
        HIP_TRY(hipEventRecord(b->ev_chain[c], b->chains[c]->stream));
        HIP_TRY(hipStreamWaitEvent(b->stream, b->ev_chain[c], 0));
    }
    HIP_TRY(hipEventRecord(b->ev_sync, b->stream));
    for (fsa_sim *s : sims) {
        if (s->stream == b->stream) continue;
        HIP_TRY(hipStreamWaitEvent(s->stream, b->ev_sync, 0));
    }
    return FSA_OK;
}

int fsa_batch_synchronize(fsa_batch *b) {
    if (!b) return set_err(FSA_ERR_INVALID, "null batch");
    HIP_TRY(hipStreamSynchronize(b->stream));
    return FSA_OK;
}

int fsa_batch_stats(fsa_batch *b, int64_t *plans_built, int64_t *steps_run, int *fluid_threads) {
    if (!b) return set_err(FSA_ERR_INVALID, "null batch");
    int64_t plans = 0;
    for (const Chain *c : b->chains) plans += c->plans_built;
    if (plans_built) *plans_built = plans;
    if (steps_run) *steps_run = b->steps_run;
    if (fluid_threads) *fluid_threads = b->chains.empty() ? 0 : b->chains[0]->fluid_threads;
    return FSA_OK;
}

}  // extern "C"

int fsa_batch_tuning(fsa_batch *b, int64_t *tunes, double *tune_ms) {
    if (!b) return set_err(FSA_ERR_INVALID, "null batch");
    int64_t n = 0;
    double ms = 0.0;
    for (const Chain *c : b->chains) {
        n += c->tunes;
        ms += c->tune_ms;
    }
    if (tunes) *tunes = n;
    if (tune_ms) *tune_ms = ms;
    return FSA_OK;
}
