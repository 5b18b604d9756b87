/*
 * fluidsim_amd.h -- C ABI of the MI355X-native FluidSim step pipeline.
 *
 * The reference has no plugin API: its step pipeline is a set of static C++
 * entry points over process-global state (include/calculation.cuh:9-21,
 * src/calculation.cu:14-35) fed from `__constant__` copies of `Parameter`
 * and `Data_Pointer` (src/memory_copy.cu:5-12).  This header replaces that
 * interface with an opaque per-scene context, the same stage granularity,
 * plain pointers and sizes, and integer status codes.
 *
 * Conventions
 *  - Every function returns FSA_OK (0) or a non-zero FSA_ERR_* code; the
 *    message of the last failure on the calling thread is fsa_last_error().
 *    Nothing in the library calls exit() (the reference's CHECK macro did,
 *    include/cuda_macro.cuh:19-33).
 *  - Host arrays are caller-owned and in *reference order* (the order of the
 *    .dat files); vectors are packed xyz triples (3*N floats), 3x3 matrices
 *    are 9 floats per particle in glm column-major order (m[col][row]).
 *  - Device arrays are owned by the context.  Particles are kept sorted by
 *    cell inside the context; downloads undo the permutation.
 *  - One context is not thread-safe; distinct contexts on distinct streams
 *    are independent.
 */
#ifndef FLUIDSIM_AMD_H
#define FLUIDSIM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FSA_ABI_VERSION 3

enum fsa_status {
    FSA_OK = 0,
    FSA_ERR_INVALID = 1, /* bad argument, size or particle outside the cell grid */
    FSA_ERR_HIP = 2,     /* HIP runtime failure (message holds hipGetErrorString) */
    FSA_ERR_NOMEM = 3,   /* device allocation failed */
    FSA_ERR_STATE = 4,   /* call out of order (e.g. step before upload) */
    FSA_ERR_NODEVICE = 5 /* no usable gfx950 device */
};

/* particle kinds, same numbering as the reference ParticleType (include/types.h:4-9) */
enum fsa_kind { FSA_FLUID = 0, FSA_SOLID = 1, FSA_VIRT = 2 };

/* solid types, same numbering as SolidType (include/types.h:11-15) */
enum fsa_solid_type { FSA_UNCONSTRAINED_SOLID = 0, FSA_FIXED_SOLID = 1 };

/*
 * Derived parameter block.  Field for field this is the part of the
 * reference `Parameter` (include/parameter.h:17-49) that the device code
 * reads through `dev_par`; the derivations are those of
 * FluidProperties/SolidProperties::initiate (src/properties.cpp:47-66),
 * Physics::readParameter (src/physics.cpp:9-19), Kernel::initiate
 * (src/kernel.cpp:25-60), Domain::initiate (src/domain.cpp:28-47) and
 * SimTime::readParameter (src/sim_time.cpp:8-24).  The host layer
 * (Parameter::initiate) fills it from config/input.ini.
 */
typedef struct fsa_params {
    int32_t n_fluid, n_solid, n_virt;

    /* Physics */
    float gravity[3];

    /* FluidProperties */
    float fluid_sound_speed;
    float fluid_reference_density;
    float fluid_reference_density_inverse;
    float fluid_gamma;
    float fluid_gamma_inv;
    float fluid_viscosity; /* dynamic viscosity */
    float fluid_coefficient_p2rho;
    float fluid_coefficient_rho2p;
    float fluid_min_pressure;
    float fluid_min_density;
    float fluid_max_density;
    int32_t fluid_is_viscid;

    /* SolidProperties */
    float solid_reference_density;
    float solid_youngs_modulus;
    float solid_poisson_ratio;
    float solid_lambda;
    float solid_nu; /* shear modulus (the reference calls it nu) */
    float solid_bulk_modulus;
    float solid_sound_speed;
    float solid_artificial_viscosity[2]; /* alpha, beta */

    /* Domain */
    float domain_min[3], domain_max[3];
    float cell_min[3], cell_max[3];
    float interval[3], interval_inv[3];
    int32_t cell_number[3];
    int32_t cell_number_total;

    /* SimTime */
    float dt;
    int32_t istart, iend;
    int32_t solid_sub_step;
    int32_t result_interval;
    int32_t sort_every_step;

    /* Kernel */
    float smoothing_length;
    float particle_diameter;
    float impact_length_hsml_ratio;
    float smoothing_length_square;
    float smoothing_length_inverse;
    float smoothing_length_inverse_square;
    float impact_distance_square;
    float kernel_coefficient_1;
    float kernel_coefficient_2;
    float kernel_differential;
    int32_t thread_num; /* read for key compatibility and not used: the fluid pass runs
                           workgroups of 64, 128 or 256 threads (64 by default; see
                           fsa_set_fluid_threads, fsa_autotune_fluid), the other passes 256 */

    /* Precision form: FSA_FORM_* bits, from the optional ini key
     * `precision_form` (no counterpart in the reference; 0 = its plain FP32
     * arithmetic). */
    int32_t precision_form;

    /* Neighbour lists: the fluid pass tests each fluid particle's fluid
     * partners from a list of the candidates within (1 + neighbour_skin) R
     * (R the cut-off), built from the cell lists and rebuilt on the device
     * once some particle has moved more than neighbour_skin R / 2 since the
     * last build; the exact cut-off test still runs every step, so every
     * step finds the same partner set as a walk of the 27 cells.  From the
     * optional ini key `neighbour_skin` (no counterpart in the reference;
     * 0 = walk the cells every step).  Not used with FSA_FORM_PACKED, nor
     * beyond 2^27 - 16 fluid particles (32-bit list entries). */
    float neighbour_skin;
} fsa_params;

/* The fluid pass gathers its fluid partners' attributes from 16-byte FP16
 * records written by the cell build: position relative to the partner's cell
 * centre (in cells), density as rho/rho0 - 1, velocity, pressure / B.  Pair
 * sums stay FP32, and the cut-off test still reads FP32 positions, so the
 * neighbour sets are the plain form's.  Needs at least 3 cells on every axis. */
#define FSA_FORM_PACKED 1
/* Compensated (Kahan) FP32 updates of fluid and solid positions and
 * velocities: each carries its rounding error into the next step. */
#define FSA_FORM_COMPENSATED 2

typedef struct fsa_sim fsa_sim;

/* ---- library ---------------------------------------------------------- */

/* ABI version of the loaded library (FSA_ABI_VERSION at build time). */
int fsa_abi_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char *fsa_last_error(void);

/* Number of visible gfx950 devices (0 when none). */
int fsa_device_count(int *count);

/* ---- context lifetime (replaces Sim::initiate + MemoryCopy::copyConstant,
 *      src/sim.cu:304-321, src/memory_copy.cu:8-12) --------------------- */

/* Allocate a context and its device buffers on the current device.
 * `stream` is a hipStream_t or NULL for a stream owned by the context. */
int fsa_create(const fsa_params *params, void *stream, fsa_sim **out);
int fsa_destroy(fsa_sim *sim);

/* Upload particles from host arrays in .dat order (Sim::readParticleData +
 * Sim::copyData, src/sim.cu:40-151, 304-314).  Initial per-particle state is
 * that of the *_Host_List::addList defaults (include/host_data.cuh:20-163).
 * Every particle must lie inside the cell grid, else FSA_ERR_INVALID. */
int fsa_upload_fluid(fsa_sim *sim, const float *position, const float *velocity,
                     const float *density, const float *pressure);
int fsa_upload_solid(fsa_sim *sim, const float *position, const float *velocity,
                     const float *reference_position, const int32_t *solid_type);
int fsa_upload_virt(fsa_sim *sim, const float *position, const int32_t *virt_type);

/* Download state in .dat order; any pointer may be NULL to skip a field.
 * (Replaces the thrust host_vector copies of src/output.cu.) */
int fsa_download_fluid(fsa_sim *sim, float *position, float *velocity,
                       float *density, float *pressure);
int fsa_download_solid(fsa_sim *sim, float *position, float *velocity,
                       float *pressure, float *normal, float *du_dt, float *couple_dudt,
                       float *deformation, float *stress_1st_piola, float *correction);
int fsa_download_virt(fsa_sim *sim, float *position, float *normal, float *velocity,
                      float *density, float *pressure);

/* ---- step stages (DeviceCalculation, include/calculation.cuh:12-20) ----
 * Stages are enqueued on the context stream and return without waiting. */

/* K1-K3: cell id of every fluid/solid particle (src/pair.cu:8-70). */
int fsa_particle_cell_update(fsa_sim *sim);
/* Cell lists: counting sort of particles by (cell id, particle id) and
 * per-cell start offsets; replaces the linked-list rebuild of
 * src/pair.cu:72-207.  Every cell holds the same particle set as the
 * reference lists; the order inside a cell is ascending .dat index, and a
 * particle that moved more than one cell is kept instead of dropped. */
int fsa_fill_cell(fsa_sim *sim);
/* K6: solid normals and Shepard pressure (src/solid.cu:262-359). */
int fsa_couple_calculation(fsa_sim *sim);
/* K7: virtual wall particle state (src/virt.cu:9-87). */
int fsa_virt_update(fsa_sim *sim);
/* K8+K9 fused: Riemann-SPH pair sums and explicit update (src/fluid.cu:26-284). */
int fsa_fluid_calculation(fsa_sim *sim);
/* K10-K12 x solid_sub_step (src/solid.cu:10-260). */
int fsa_solid_calculation(fsa_sim *sim);
/* The whole step loop body of src/calculation.cu:23-33, n_steps times.  With
 * n_steps > 1 the next step's fluid cell build and wall pass run on an
 * internal second stream beside the solid sub-steps (forked from and joined
 * back into the context's stream, so the call is still ordered on it); the
 * results are those of the stage calls made one by one, bit for bit. */
int fsa_step(fsa_sim *sim, int n_steps);
/* Wait for all work enqueued on the context stream. */
int fsa_synchronize(fsa_sim *sim);

/* ---- introspection for parity checks and diagnostics ------------------ */

/* Cell id of every particle of `kind`, in .dat order. */
int fsa_get_cell_ids(fsa_sim *sim, int kind, int32_t *cell_id);
/* Per-cell start offsets into the sorted order (cell_number_total+1 ints). */
int fsa_get_cell_start(fsa_sim *sim, int kind, int32_t *cell_start);
/* .dat index of the particle at each sorted slot (N ints). */
int fsa_get_sorted_ids(fsa_sim *sim, int kind, int32_t *sorted_id);
/* Fluid pair sums of the last fluid stage (drho_dt N floats, du_dt 3N),
 * in .dat order.  Only recorded while FSA_DEBUG_PAIR_SUMS is set. */
int fsa_download_fluid_rates(fsa_sim *sim, float *drho_dt, float *du_dt);
/* Partners the last fluid stage found per fluid particle and kind (3N ints:
 * fluid, solid, wall; the particle itself counts as a fluid partner), in
 * .dat order: the neighbour sets of findFluidPair (src/fluid.cu:188-227).
 * Only recorded while FSA_DEBUG_PAIR_SUMS is set. */
int fsa_download_pair_counts(fsa_sim *sim, int32_t *counts);
/* Number of particles that moved more than one cell in one step since the
 * context was created (the reference silently dropped such particles from
 * its lists, src/pair.cu:87-150; this build keeps them). */
int fsa_get_jump_count(fsa_sim *sim, int64_t *count);
/* Neighbour lists (fsa_params.neighbour_skin > 0): how many times the
 * lists were built since the last fluid upload (the upload builds them
 * once); how many fluid particles the last build could not list (more
 * candidates than a list holds: those walk their cells every step until the
 * next build); the candidates the last build listed over all particles;
 * and the steps in which every particle walked its cells instead, because
 * the lists did not last long enough to pay for their builds (a fast flow:
 * after a stretch of such steps -- 64, doubled each time the lists fail
 * again, at most 1024 -- the lists are built and tried again).  Any pointer
 * may be NULL.  All zero without lists.  Reading syncs the context's
 * stream. */
int fsa_get_list_stats(fsa_sim *sim, int64_t *builds, int64_t *unlisted, int64_t *entries, int64_t *walks);
/* Divergence guard (no counterpart in the reference, which only zeroes NaN
 * pair sums, src/fluid.cu:246-249): the fluid and solid updates set a bit
 * when a velocity they write is not finite, which no later clamp repairs
 * (a NaN position clamps to the domain box, a NaN velocity stays).  Bits
 * accumulate from the last upload of the kind on; reading syncs the
 * context's stream.  A batch leaves each scene's bits in its own context, so
 * a caller finds a diverged scene of a batch by asking each. */
#define FSA_HEALTH_FLUID_NONFINITE 1u
#define FSA_HEALTH_SOLID_NONFINITE 2u
int fsa_get_health(fsa_sim *sim, unsigned *flags);

#define FSA_DEBUG_PAIR_SUMS 1u
/* The solid sub-steps walk their candidate cells in every pass instead of
 * replaying the partners recorded earlier in the step (same sums, bit for
 * bit; the records only save the walk). */
#define FSA_DEBUG_WALK 2u
int fsa_set_debug(fsa_sim *sim, unsigned flags);

/* ---- run-level observables ----------------------------------------------
 * The reference has no probes (src/probe.cu is commented out) and no
 * conservation diagnostics; these are computed on the device from the
 * current state.  Sums are accumulated in FP64.  V = particle_diameter^3. */
typedef struct fsa_observables {
    double fluid_mass;                  /* sum rho V */
    double fluid_momentum[3];           /* sum rho V u */
    double fluid_kinetic_energy;        /* sum rho V |u|^2 / 2 */
    double fluid_potential_energy;      /* sum rho V (-g . x) */
    double front_x;                     /* largest fluid x (the surge front along the flow axis) */
    double solid_momentum[3];           /* sum rho_s V u over solids */
    double gate_tip_displacement[3];    /* mean x - X over the gate tip row (in the compensated
                                           form x includes its carried rounding error) */
    int64_t gate_tip_count;             /* solids in the tip row: unconstrained solids whose
                                           reference y is within dp/2 of the largest one */
    double max_fluid_speed;             /* max |u| over the fluid (FP64 reduction) */
    double max_solid_speed;             /* max |u| over the solid */
    double cfl_dt;                      /* 0.25 h / (c0 + max fluid |u|): the acoustic CFL bound of
                                           the time step (the dam-break scenes run at 0.25 h / c0) */
    double solid_cfl_dt;                /* 0.3 h / (c_s + max solid |u|): the bound of one solid sub-step;
                                           0 without solids */
} fsa_observables;

int fsa_observe(fsa_sim *sim, fsa_observables *out);

/* Shepard-interpolated fluid pressure at n points (x,y,z triples), with the
 * weights of solidPressureCalculate (src/solid.cu:275-291):
 * p = sum p_j w_j / sum w_j, w = (exp(-q^2) - c1) c2 over fluid within K h;
 * 0 where the sum of weights is <= 1e-6.  Points must lie in the cell grid. */
int fsa_probe_pressure(fsa_sim *sim, const float *points, int n, float *pressure);
/* fsa_probe_pressure without allocating device memory.  No counterpart in
 * the reference: this is for callers that cannot allocate (code between
 * steps of a serving loop or a harness that forbids device allocation);
 * fsa_probe_pressure allocates staging buffers when n grows.  The points and
 * their pressures are staged in a static area of the library, one per
 * device (1 MiB, held until the process ends), so n is at most
 * FSA_PROBE_NOALLOC_MAX (else FSA_ERR_INVALID), and calls that overlap on
 * one GPU -- from any context on it, on any thread -- race on that area.
 * Results are bit-identical to fsa_probe_pressure. */
#define FSA_PROBE_NOALLOC_MAX 65536
int fsa_probe_pressure_noalloc(fsa_sim *sim, const float *points, int n, float *pressure);

/* Work census of the fluid pass on the current state: counts[k] = candidate
 * tests and counts[3+k] = partners inside the cut-off, for partner kind k
 * (fluid, solid, virt), summed over all fluid particles (self included).
 * Used to price the pass in flops; the fluid pass itself is not run. */
int fsa_pair_census(fsa_sim *sim, int64_t counts[6]);

/* ---- timing ------------------------------------------------------------ */

enum fsa_stage {
    FSA_STAGE_CELL = 0,   /* particleCellUpdate + fillCell */
    FSA_STAGE_COUPLE = 1, /* coupleCalculation */
    FSA_STAGE_VIRT = 2,   /* virtUpdate */
    FSA_STAGE_FLUID = 3,  /* fluidCalculation (the fused pair+update kernel) */
    FSA_STAGE_SOLID = 4,  /* solidCalculation (all sub-steps) */
    FSA_STAGE_COUNT = 5
};

/* When enabled, HIP events bracket every stage on the context stream. */
int fsa_set_timing(fsa_sim *sim, int enabled);
/* Timing for the stages in stage_mask only (bit 1u << FSA_STAGE_x; 0 = off).
 * Every timed stage adds two event records per step to its stream, and
 * each costs the stream a few microseconds; timing one stage keeps a timed
 * run's step close to an untimed one.  No counterpart in the reference
 * (a measurement aid, like fsa_set_timing). */
int fsa_set_timing_stages(fsa_sim *sim, unsigned stage_mask);
/* Sum of event-measured stage times (ms) and number of samples per stage
 * since the last reset; synchronizes the stream. */
int fsa_get_stage_times(fsa_sim *sim, double *ms, int64_t *samples);
int fsa_reset_stage_times(fsa_sim *sim);

/* Device bytes held by the context. */
int fsa_memory_used(fsa_sim *sim, int64_t *bytes);

/* ---- launch shape of the fluid pass --------------------------------------
 * The reference launches every kernel with `thread_num` threads per block
 * (src/fluid.cu:276-282).  Here the fluid pass (K8+K9)
 * runs in workgroups of 64, 128 or 256 threads; all three give identical
 * results, and the fastest depends on the scene (LDS per wave bounds the
 * occupancy).  The default is 64. */
int fsa_set_fluid_threads(fsa_sim *sim, int threads);
int fsa_get_fluid_threads(fsa_sim *sim, int *threads);
/* Time the fluid pass at each workgroup size on the current state (one
 * warm-up and `repeats` timed launches each), keep the fastest and return it
 * in *threads (may be NULL).  The scene does not advance: the pass reads the
 * current state buffer and writes the other one, which the next step
 * overwrites.  Synchronizes the context stream. */
int fsa_autotune_fluid(fsa_sim *sim, int repeats, int *threads);

/* Device bytes fsa_create would allocate for these parameters (= what
 * fsa_memory_used reports after creation, before optional debug, census and
 * probe buffers), and the peak while fsa_upload_virt runs (it sorts the walls
 * through temporary arrays).  Allocates nothing and needs no device: a
 * scheduler admits scenes by this figure.  The reference sizes its device
 * arrays from the .dat counts the same way (Sim::copyData, src/sim.cu:304-308;
 * the list initiate functions, include/particle.cuh:73, 155, 218). */
int fsa_footprint(const fsa_params *params, int64_t *resident, int64_t *upload_peak);

/* Free and total bytes of the current device (hipMemGetInfo). */
int fsa_device_memory(int64_t *free_bytes, int64_t *total_bytes);

/* ---- batches: many scenes per launch ------------------------------------
 *
 * The reference runs one scene per process: its parameter and pointer
 * blocks are process-global (src/sim.cu:9-10, src/memory_copy.cu:5-6) and its
 * step loop is DeviceCalculation::calculation (src/calculation.cu:23-33).  A
 * batch advances any number of independent contexts through that same stage
 * sequence with one launch per stage for all of them, recorded as a HIP graph
 * per step.  Every context's result is bit-identical to fsa_step on it alone;
 * contexts may differ in every parameter, particle count and solid_sub_step.
 *
 * A batch is an executor, not an owner: the caller names the contexts on
 * every fsa_batch_step call and may change the list between calls
 * (continuous batching).  The recorded launches are reused while the list
 * and the contexts' buffers stay the same and rebuilt otherwise (a new list,
 * or a context stepped on its own in between).  All contexts must live on
 * the batch's device.  The batch stream first waits for work queued on each
 * context's stream, and each context's stream then waits for the batch, so
 * stage calls and downloads on a context stay ordered after the batch step.
 * Per-stage timing and FSA_DEBUG_PAIR_SUMS are not recorded in a batch.
 *
 * The contexts are split over a few launch chains (3 by default), each with
 * its own stream and recording, so that one chain's latency-bound stages
 * (solid sub-steps, cell build) overlap another chain's fluid pass.  The
 * split is deterministic in the list; results do not depend on it. */
typedef struct fsa_batch fsa_batch;

#define FSA_BATCH_NO_GRAPH 1u /* launch every stage directly instead of replaying a HIP graph */
#define FSA_BATCH_AUTOTUNE 2u /* time the batched fluid pass at each workgroup size when the
                                 launches are recorded, and keep the fastest; a chain re-times
                                 only when its fluid particle total has moved by more than a
                                 quarter since it last did */

/* stream: hipStream_t to run on, or NULL for a batch-owned stream. */
int fsa_batch_create(void *stream, unsigned flags, fsa_batch **out);
int fsa_batch_destroy(fsa_batch *batch);
/* n_steps full steps of every listed context (no context listed twice). */
int fsa_batch_step(fsa_batch *batch, fsa_sim *const *scenes, int n_scenes, int n_steps);
/* Number of concurrent launch chains (streams) the contexts are split over,
 * 1..8 (default 3); takes effect on the next fsa_batch_step.  The batch
 * scheme has no counterpart in the reference (one scene per process). */
int fsa_batch_set_streams(fsa_batch *batch, int n);
int fsa_batch_synchronize(fsa_batch *batch);
/* How often the launches were (re)recorded and steps run since creation, and
 * the fluid workgroup size of the current recording (0 before the first).
 * Without FSA_BATCH_AUTOTUNE a recording uses the first listed scene's size
 * (fsa_set_fluid_threads). */
int fsa_batch_stats(fsa_batch *batch, int64_t *plans_built, int64_t *steps_run, int *fluid_threads);
/* FSA_BATCH_AUTOTUNE: how many timings the chains ran and the host wall time
 * they took, in ms (no counterpart in the reference). */
int fsa_batch_tuning(fsa_batch *batch, int64_t *tunes, double *tune_ms);

#ifdef __cplusplus
}
#endif

#endif /* FLUIDSIM_AMD_H */
