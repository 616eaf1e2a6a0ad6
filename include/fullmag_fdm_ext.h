/*
 * fullmag_fdm_ext.h -- entry points of the MI355X FDM engine that have no
 * counterpart in the reference ABI (native/include/fullmag_fdm.h).  They are
 * additive: a caller that only knows fullmag_fdm.h never needs them.
 */
#ifndef FULLMAG_FDM_EXT_H
#define FULLMAG_FDM_EXT_H

#include "fullmag_fdm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ABI revision of this library (bumped when an entry point is added). */
#define FULLMAG_FDM_EXT_VERSION 9
int fullmag_fdm_ext_version(void);

/* Time-dependent Zeeman waveforms: H_ext(t) = H0 + f(t) * H1 on the cells
 * selected by field_mask (all magnetic cells when NULL). */
#define FULLMAG_FDM_WAVE_NONE 0
#define FULLMAG_FDM_WAVE_RAMP 1  /* f = clamp((t - t0)/(t1 - t0), 0, 1)            */
#define FULLMAG_FDM_WAVE_SINC 2  /* f = sin(x)/x, x = 2 pi frequency (t - t0)       */
#define FULLMAG_FDM_WAVE_SINE 3  /* f = sin(2 pi frequency t + phase)               */

/*
 * Plan extension (no counterpart in the reference ABI).  struct_size must be
 * sizeof(fullmag_fdm_plan_ext); fields added later go at the end.  Every
 * pointer is optional (NULL = off) and only read during create.
 */
typedef struct {
    uint32_t      struct_size;
    const double *ms_field;               /* f64[cell_count]: per-cell Ms (A/m)          */
    const double *anisotropy_axis_field;  /* f64[3*cell_count] AoS: per-cell uniaxial axis */
    uint32_t      field_waveform;         /* FULLMAG_FDM_WAVE_*                          */
    double        field_amplitude_am[3];  /* H1 (A/m)                                    */
    double        field_t0;               /* s */
    double        field_t1;               /* s (ramp end)                                */
    double        field_frequency;        /* Hz */
    double        field_phase;            /* rad */
    const uint8_t *field_mask;            /* u8[cell_count]                              */
    double        adaptive_rtol;          /* as fullmag_fdm_backend_set_adaptive_ext     */
    double        adaptive_growth_limit;
    double        adaptive_shrink_limit;
    /* version 5: thermal field of the plan's temperature (the CPU engine's
     * FdmProblem::thermal_dt / thermal_seed, fdm_problem.rs:27-35).  A struct
     * of the version-4 size (ending at adaptive_shrink_limit) is accepted. */
    double        thermal_dt;             /* s; 0 = the dt of each step attempt          */
    uint64_t      thermal_seed;           /* counter-generator seed; 0 = 42              */
    /* version 6: periodic boundaries of the CPU engine (FdmBoundaryPolicy,
     * fdm_types.rs:94-132; demag by truncated image sums,
     * fdm_fft.rs:140-229,611-717).  Bit 0/1/2: x/y/z periodic; images per
     * periodic axis (0 = the CPU engine's default 10).  Smaller structs
     * (versions 4 and 5) are accepted. */
    uint32_t      periodic_axes;
    uint32_t      periodic_images[3];
    /* version 7: multilayer stack (crates/fullmag-engine/src/multilayer.rs,
     * shifted kernels crates/fullmag-fdm-demag/src/shifted_kernel.rs): the
     * grid's nz planes (2..8) are separate single-cell layers whose centres
     * lie at layer_z[k] (m); they couple only through demag, each plane keeps
     * its own exchange.  NULL = uniform grid.  Smaller structs accepted. */
    const double *layer_z;
} fullmag_fdm_plan_ext;

/* create() with the extension struct (ext may be NULL: same as create()). */
fullmag_fdm_backend *fullmag_fdm_backend_create_ext(const fullmag_fdm_plan_desc *plan,
                                                    const fullmag_fdm_plan_ext *ext);

/*
 * Device-resident runner loop: repeat accepted steps until time t_end or
 * max_steps accepted steps, whichever comes first, with the CPU runner's
 * semantics (crates/fullmag-runner/src/cpu_reference.rs:622-639):
 * dt_step = min(dt, t_end - t); for DP45 the next dt is the controller's
 * suggestion and every trial step is clamped to [dt_min, dt_max].  Step
 * acceptance and dt control run on the device; the host only launches and
 * checks a completion flag.  out_stats receives the last accepted step,
 * *out_steps the number of accepted steps taken (either may be NULL).
 */
int fullmag_fdm_backend_advance(fullmag_fdm_backend *handle, double t_end, uint64_t max_steps,
                                double dt_seconds, fullmag_fdm_step_stats *out_stats,
                                uint64_t *out_steps);

/*
 * Mixed-tolerance step control of the CPU engine's AdaptiveStepConfig
 * (crates/fullmag-engine/src/fdm_types.rs:182-209): per-cell error divided
 * by (atol + rtol*|m0|) and compared with 1 when rtol > 0; step growth and
 * shrink factors limited to growth_limit (<= 0: unlimited) and shrink_limit.
 */
int fullmag_fdm_backend_set_adaptive_ext(fullmag_fdm_backend *handle, double rtol,
                                         double growth_limit, double shrink_limit);

/* Counters since create: accepted steps, rejected trial steps, time. */
int fullmag_fdm_backend_get_counters(fullmag_fdm_backend *handle, uint64_t *accepted,
                                     uint64_t *rejected, double *time_seconds);

/*
 * Kernel timing with HIP events on the engine's stream.  Runs `reps`
 * launches of each kernel of one RHS evaluation (stage 2 of the configured
 * method) and of the step controller, and writes the mean duration in
 * microseconds of kernel kind i to out_us[i]:
 *   0 x-forward (stage combine + R2C; stage combine alone without demag),
 *   1 y pass (thin film fused / 3-D forward), 2 z pass (3-D),
 *   3 y inverse (3-D), 4 x-inverse + local pass, 5 controller,
 *   6 fused x-inverse + local pass of stage 2 and x-forward of stage 3
 *   (the launch that joins consecutive stages inside a trial step).
 * Kinds that do not exist for the grid are reported as 0.  The state is
 * left consistent (a following step recomputes every stage).  n_out must be
 * at least 7.
 */
int fullmag_fdm_backend_time_kernels(fullmag_fdm_backend *handle, int reps, double *out_us,
                                     int n_out);

/*
 * Kernel timing inside whole trial steps (no reference counterpart): runs
 * `trials` trial steps of the configured method (after one untimed) as
 * separate launches in the order of the captured trial step, an event
 * after each, so that every kernel reads what the kernel before it wrote,
 * as in the stepping loop; out_us[i] is the mean time between the events
 * around the first launch of kind i in a trial step (kinds as
 * fullmag_fdm_backend_time_kernels, at stage 2; kind 4, the final x pass,
 * at the last stage with the error norm; 0 for kinds a trial step does not
 * have), less the events' own cost per launch (the same number of trial
 * steps with events only around each step measures it).  With n_out > 7,
 * out_us[7] is the time of a whole trial step without the inner events and
 * out_us[8] (n_out > 8) that per-launch event cost.  The control block is restored
 * after every trial step, so the state does not change.  Demag plans
 * without layer stacks; n_out at least 7.
 */
int fullmag_fdm_backend_time_trial_kernels(fullmag_fdm_backend *handle, int trials, double *out_us,
                                           int n_out);

/*
 * Tile tuner (no reference counterpart): times the x-pass launches of a
 * trial step for each x tile (rows per work-group 1, 2, 4; register
 * four-step or LDS Stockham FFT) at `reps` launches each and, on 3-D grids,
 * the y-line and z passes for each of their tiles (128 or 256 threads per
 * y-line group; all, half or a quarter of the ky lines per z group), keeps
 * the fastest for this handle and records it for later creates of the same
 * precision, grid and integrator in this process.  Every tile gives
 * bit-identical fields.  The state is not changed.  Outputs
 * (optional): the chosen rows per work-group and 1 if the four-step FFT is
 * used (both 0 without demag).
 */
int fullmag_fdm_backend_autotune(fullmag_fdm_backend *handle, int reps, int *out_rows_per_group,
                                 int *out_four_step_x);

/*
 * Seeded random unit vectors of the planner's RandomSeeded initial state
 * (crates/fullmag-plan/src/util.rs:118-147, xorshift64 13/7/17, components
 * uniform in [-1,1], rejection of near-zero norms, then normalised).  Host
 * code, no device needed.  out: 3*count doubles AoS.  Seed 0 is rejected
 * (the generator would never leave its rejection loop).
 */
int fullmag_fdm_random_unit_vectors(uint64_t seed, uint64_t count, double *out);

/*
 * Adjoint sensitivities of a fixed-step run (no reference counterpart; the
 * reference has no differentiable path).  The handle's plan must use a
 * fixed-step integrator (RK4 or Heun) and uniform exchange (no region
 * table).  Objective over the states m_n after steps n = first..steps:
 *   J = sum_{n=first}^{steps} sum_i w_i m_{c,i}(t_n)^2 / (steps - first + 1)
 * with c = objective_component and w = objective_weight.
 *
 * adjoint_forward runs `steps` steps of size dt from the current state and
 * time, keeping the state every checkpoint_interval steps (0: ceil(sqrt(
 * steps))), and returns J; the trailing segments that fit the record budget
 * (at least the last one) run in recording form (their stage states and
 * fields are kept).  adjoint_backward then sweeps the steps in reverse (the
 * recorded segments from the forward's record, every earlier segment
 * recomputed once from its checkpoint with its stage states and fields
 * recorded; when the budget holds the whole run nothing is recomputed) and
 * returns dJ/dMs_i for every cell
 * (grad_ms, n values, 0 on inactive cells) and dJ/d field_amplitude_am of
 * the plan's waveform (grad_field_amplitude, 3 values).  Either output may
 * be NULL.  After either call the handle holds the final state m_steps at
 * t0 + steps*dt.  Device memory: 3N state values per checkpoint, 7 (nstage-1) N
 * per recorded step (stage states, fields and pre-normalisation norms), and
 * 27N FP64 cotangents.  FP32 plans with
 * demag convolve the field cotangent with an FP64 copy of the demag tensor
 * and transforms (an FP64 spectrum of the grid, created on the first
 * adjoint run) unless flags has FULLMAG_FDM_ADJOINT_BETA_STATE_PRECISION:
 * the far field of a localised cotangent is a small difference of FFT
 * terms that FP32 resolves only to ~1e-7 of its near field.
 */
#define FULLMAG_FDM_ADJOINT_BETA_STATE_PRECISION 1u
typedef struct fullmag_fdm_adjoint_desc {
    uint64_t struct_size;             /* sizeof(fullmag_fdm_adjoint_desc) */
    uint64_t steps;                   /* fixed steps of the run (>= 1) */
    double dt;                        /* step size [s] */
    uint64_t objective_first_step;    /* first step whose state enters J (1..steps) */
    uint32_t objective_component;     /* 0 x, 1 y, 2 z */
    uint32_t flags;                   /* FULLMAG_FDM_ADJOINT_* bits (0: defaults) */
    const double *objective_weight;   /* n per-cell weights */
    uint64_t objective_weight_len;
    uint64_t checkpoint_interval;     /* 0 = ceil(sqrt(steps)) */
    /* device memory for the forward's record of the trailing segments
       (0: half the free device memory); a budget below one segment's
       record keeps only the last segment.  Absent (struct_size without
       it): 0. */
    uint64_t record_budget_bytes;
} fullmag_fdm_adjoint_desc;

int fullmag_fdm_adjoint_forward(fullmag_fdm_backend *handle, const fullmag_fdm_adjoint_desc *desc,
                                double *out_objective);
int fullmag_fdm_adjoint_backward(fullmag_fdm_backend *handle, double *grad_ms, double *grad_field_amplitude);

/*
 * Batched sweep of independent systems (no reference counterpart; for
 * parameter sweeps of many small systems that cannot fill the device one at
 * a time).  Each plan becomes a system with its own state, material, step
 * control and adaptive dt.  Systems whose launch geometry agrees (grid
 * shape, precision, integrator, demag and exchange form, tuned tile) form a
 * group: every kernel of a trial step covers the whole group (one row of
 * work-groups per system), one captured graph per trial step, and the
 * groups run concurrently on their own streams.
 *
 * create: plans[i] with optional exts[i] (exts or its entries may be NULL);
 * returns non-NULL with last_error set on failure, like backend_create.
 * Each group's x tile is chosen by timing the group's own launches (as
 * fullmag_fdm_backend_autotune does for one handle) and recorded for later
 * creates of the same shape.
 * advance: integrates every system to t_end (at most max_steps accepted
 * steps each) with its own dt control starting from dt0, as
 * fullmag_fdm_backend_advance does for one system; out_wall_seconds[i]
 * (optional) receives the wall time after which system i was seen
 * finished (resolution: one polling chunk), out_steps[i] (optional) its
 * accepted steps.  system(i): the handle of system i, owned by the batch
 * and valid until batch_destroy, for field copies, snapshots and stats
 * (not for stepping while a batch advance runs).
 */
typedef struct fullmag_fdm_batch fullmag_fdm_batch;

fullmag_fdm_batch *fullmag_fdm_batch_create(const fullmag_fdm_plan_desc *const *plans,
                                            const fullmag_fdm_plan_ext *const *exts, uint64_t count);
int fullmag_fdm_batch_groups(const fullmag_fdm_batch *batch, uint64_t *out_groups);
int fullmag_fdm_batch_advance(fullmag_fdm_batch *batch, double t_end, uint64_t max_steps, double dt0,
                              double *out_wall_seconds, uint64_t *out_steps);
fullmag_fdm_backend *fullmag_fdm_batch_system(fullmag_fdm_batch *batch, uint64_t index);
const char *fullmag_fdm_batch_last_error(const fullmag_fdm_batch *batch);
void fullmag_fdm_batch_destroy(fullmag_fdm_batch *batch);

/*
 * No counterpart in the reference.  fullmag_fdm_backend_copy_field_preview_f64
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving or preview loops, harnesses that forbid hipMalloc):
 * the bins are staged in a fixed device array of 2 MiB per GPU
 * (FULLMAG_FDM_PREVIEW_NOALLOC_MAX_CELLS preview cells), which the library
 * holds until the process ends; the call allocates nothing, on the device or
 * otherwise.  Larger previews return FULLMAG_FDM_ERR_INVALID with
 * last_error set.  All calls on one GPU share that array: calls that
 * overlap on one GPU (several threads, or handles on the same device) race
 * and must be serialised by the caller.  Arguments, result and errors are
 * otherwise those of copy_field_preview_f64.
 */
#define FULLMAG_FDM_PREVIEW_NOALLOC_MAX_CELLS 87381u
int fullmag_fdm_backend_copy_field_preview_f64_noalloc(fullmag_fdm_backend *handle, fullmag_fdm_observable observable,
                                                       uint32_t preview_nx, uint32_t preview_ny, uint32_t preview_nz,
                                                       uint32_t z_origin, uint32_t z_stride, double *out_xyz,
                                                       uint64_t out_len);

#ifdef __cplusplus
}
#endif

#endif /* FULLMAG_FDM_EXT_H */
