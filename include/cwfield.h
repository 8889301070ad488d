/*
 * cwfield: C ABI of the MI355X (gfx950) diffusion field solve.
 *
 * This is the boundary a thin FFI crate binds so that cellweave's
 * ScalarField (crates/engine/src/pde.rs:30-53) can delegate its steady-state
 * solve to the GPU without changing any Rust signature. A cw_field handle
 * owns the field's state on the device (values, held, uptake, production),
 * the ping-pong buffers, the reduction scratch and one HIP stream; the host
 * fills it by bulk upload or by the device-side gathers from a cw_lattice
 * (the cell-id lattice of crates/engine/src/cpm.rs:198).
 *
 * Plain C types only: pointers, sizes, doubles. Every entry point returns a
 * cw_status; on failure cw_last_error() gives a thread-local message.
 * Distinct handles may be used from distinct threads at once; one handle is
 * used by one thread at a time (ScalarField::relax takes &mut self).
 *
 * Layout of every per-voxel array: flat index i = (z*H + y)*W + x, x fastest
 * (pde.rs:27-29, pde.rs:184-188, cpm.rs:194-198).
 */
#ifndef CWFIELD_H
#define CWFIELD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cw_field cw_field;
typedef struct cw_lattice cw_lattice;

/* Error codes. CW_ERR_CONFIG maps onto CellweaveError::Config and
 * CW_ERR_NOT_CONVERGED onto CellweaveError::Simulation
 * (crates/core/src/error.rs:7-14). */
typedef enum cw_status {
    CW_OK = 0,
    CW_ERR_CONFIG = 1,        /* bad dimensions, non-finite or negative D/decay */
    CW_ERR_NOT_CONVERGED = 2, /* max_sweeps ran out (pde.rs:285-287) */
    CW_ERR_NONFINITE = 3,     /* a residual or value turned NaN/inf */
    CW_ERR_DEVICE = 4,        /* HIP runtime error, or no gfx950 device */
    CW_ERR_ARG = 5            /* null pointer, size mismatch, bad option */
} cw_status;

/* Stopping rule. */
typedef enum cw_criterion {
    /* The reference rule (pde.rs:279-284): stop after the first sweep in
     * which no voxel moved by `tolerance` or more. Exact sweep count. */
    CW_CRIT_MAX_CHANGE = 0,
    /* New rule: stop when ||rt||_2 / ||b_eff||_2 <= tolerance, where rt is
     * the projected (natural) residual of the floored problem, tested in
     * FP64 on the state before a pass. */
    CW_CRIT_REL_RESIDUAL = 1
} cw_criterion;

typedef enum cw_precision {
    CW_PREC_FP64 = 0,        /* plain FP64 red-black projected SOR */
    CW_PREC_FP32_REFINE = 1  /* FP32 smoothing of the defect equation with
                                FP64 residual, correction and test */
} cw_precision;

typedef struct cw_relax_opts {
    double omega;             /* 0: the reference rule (pde.rs:267-277); < 0: each field's
                                 2/(1+sin(pi/L)), L its longest free-span edge */
    int32_t criterion;        /* cw_criterion */
    double tolerance;         /* absolute max change, or relative residual */
    uint64_t max_sweeps;      /* hard limit; reaching it is CW_ERR_NOT_CONVERGED */
    int32_t precision;        /* cw_precision */
    int32_t sweeps_per_pass;  /* temporal blocking depth s (full sweeps per
                                 HBM pass); 0 = library choice */
    int32_t check_every;      /* residual test every k passes; 0 = library choice */
    double inner_tolerance;   /* FP32_REFINE: reduction asked of one segment; 0 = 2e-2 */
    uint64_t max_outer;       /* FP32_REFINE: outer iteration cap; 0 = default */
    int32_t tile_shape;       /* sweep kernel shape index (-1 = from
                                 sweeps_per_pass); cw_autotune ranks them */
    int32_t use_graph;        /* 1: replay steady chunks of passes from a
                                 captured hipGraph (one launch per chunk) */
} cw_relax_opts;

typedef struct cw_relax_stats {
    uint64_t sweeps;          /* smoothing sweeps run (FP32_REFINE: inner sweeps summed) */
    uint64_t passes;          /* sweep kernel launches that did work */
    uint64_t outer_iters;     /* FP32_REFINE defect-correction steps */
    uint64_t span_voxels;     /* voxels in the swept box (pde.rs:305-320) */
    double omega;             /* factor used */
    double rel_residual;      /* FP64 relative residual at exit (NaN if not computed) */
    double seconds;           /* wall time of the call */
    int32_t status;           /* cw_status of this field */
    int32_t sweeps_per_pass;  /* s actually used */
    int32_t tile_shape;       /* sweep tile shape index used (volumes: the volume table) */
    int32_t reserved;         /* 0 */
} cw_relax_stats;

/* Fill opts with the reference drop-in: reference omega, max-change
 * criterion, FP64. */
void cw_relax_opts_default(cw_relax_opts *opts);

/* Thread-local description of the last failure on this thread. */
const char *cw_last_error(void);

/* Library version string and the device name it runs on. */
const char *cw_version(void);

/* Replaces ScalarField::with_depth / ScalarField::new (pde.rs:82-131):
 * same validation, all values `initial`, nothing held, no sinks. */
cw_status cw_field_create(uint32_t width, uint32_t height, uint32_t depth, double diff,
                          double decay, double initial, cw_field **out);
/* Drop. Null is accepted. */
cw_status cw_field_destroy(cw_field *field);

/* ScalarField::with_source (pde.rs:135-138). */
cw_status cw_field_set_source(cw_field *field, double source);

/* Bulk upload of the host state before a solve: values (f64), held (one byte
 * per voxel, nonzero = held, pde.rs:45-46), uptake and production (f64,
 * pde.rs:47-50). Any pointer may be NULL to keep what the device holds.
 * Replaces the per-voxel setters (pde.rs:146-220, 468-640) crossing FFI: the
 * Rust side marks arrays dirty and uploads them here in one call. */
cw_status cw_field_upload(cw_field *field, const double *values, const uint8_t *held,
                          const double *uptake, const double *production);

/* ScalarField::fill (pde.rs:212-215): every voxel of the device values set
 * to `value`, held ones included, with no host traffic. */
cw_status cw_field_fill(cw_field *field, double value);

/* Copy the device values back to the host (n doubles). */
cw_status cw_field_download(const cw_field *field, double *values);

/* ScalarField::relax (pde.rs:245-288) with its exact semantics: the free-span
 * box, the reference omega, red then black half sweeps, stop on the first
 * sweep whose largest change is below `tolerance`, Ok(1) when nothing is free,
 * CW_ERR_NOT_CONVERGED after max_sweeps. *sweeps gets Ok's count. */
cw_status cw_field_relax(cw_field *field, double tolerance, uint64_t max_sweeps,
                         uint64_t *sweeps);

/* cw_field_relax that allocates nothing (no counterpart in the reference; it
 * is for callers that cannot allocate device memory: tight serving loops and
 * harnesses that forbid hipMalloc). Same semantics, kernels and bits as
 * cw_field_relax, with eager launches instead of graph replays. The launch
 * plan (problems, control blocks, tile tables, residual partials) lives in a
 * static 4 MiB device array, one per GPU, kept until the process ends; a
 * field whose plan needs more (about 9e4 elementwise tiles of 128 x 4 voxels,
 * e.g. beyond 400^3) gets CW_ERR_ARG before any sweep. So does a field that
 * keeps production and uptake apart (a source, or voxels with both) before an
 * allocating call has created its buffer for that. Calls on one GPU share the
 * array: two calls that overlap on one GPU race. The verdicts are read on the
 * host, so it cannot be captured into a graph. */
cw_status cw_field_relax_noalloc(cw_field *field, double tolerance, uint64_t max_sweeps,
                                 uint64_t *sweeps);

/* relax with every knob: omega, criterion, precision, blocking depth. */
cw_status cw_field_relax_ex(cw_field *field, const cw_relax_opts *opts,
                            cw_relax_stats *stats);

/* FP64 relative residual (see CW_CRIT_REL_RESIDUAL) of the device values. */
cw_status cw_field_residual(cw_field *field, double *rel_residual);

/* Batched solve (new; the campaign of src/campaign.rs:196-204 run together).
 * Five or more fields: one launch sequence, every field swept by the same
 * kernel launches. Two to four fields on distinct streams: concurrent solves,
 * one internal thread and the field's own stream each. Either way each field
 * stops on its own criterion. Under CW_CRIT_MAX_CHANGE each field gets the
 * bits it gets when solved alone (an exact stop replays a launch's tail);
 * under CW_CRIT_REL_RESIDUAL and in FP32 refinement it gets the bits it gets
 * alone at the same tile_shape: with tile_shape -1 the library picks the
 * shape from the tiles of the whole batch, and the shape sets the sweeps
 * between residual tests and the segment lengths. stats (may be NULL) has
 * `count` entries. Returns CW_ERR_NOT_CONVERGED if
 * any field did not converge. The fields must not be in use elsewhere during
 * the call. */
cw_status cw_batch_relax(cw_field *const *fields, uint32_t count, const cw_relax_opts *opts,
                         cw_relax_stats *stats);

/* The campaign solved on the device (new; the histories that
 * src/campaign.rs:196-204 runs on CPU threads, solved together): planes only,
 * relative-residual rule, FP64 or FP32 refinement. Field i belongs to group
 * group[i] < ngroups (a size class), and each group is swept with its own
 * plane tile shape shapes[g] (-1 or shapes NULL: the library's choice; FP64
 * takes 8 or 9, FP32 15 or 18 to 27), as cw_autotune ranks them. Every field
 * keeps its own state on the device: one kernel between two steps moves each
 * field to its next phase (FP64: sweeps with a residual test every
 * check_every steps; FP32: FP64 residual, a smoothing segment whose length it
 * picks from the field's residual history, FP64 correction) and lists the
 * tiles that work next, and a persistent kernel works them. One captured
 * graph holds every group's chain of steps; the host launches it as often as
 * the model problem predicts and reads the fields' states once (once more per
 * extra launch if a field is still running). FP64 gives the bits and sweeps
 * of cw_batch_relax at the same shape and check_every. A segment never runs
 * past max_sweeps. stats has `count` entries. */
cw_status cw_campaign_relax(cw_field *const *fields, uint32_t count, const uint32_t *group,
                            const int32_t *shapes, uint32_t ngroups, const cw_relax_opts *opts,
                            cw_relax_stats *stats);

/* Result of cw_field_adjoint. */
typedef struct cw_adjoint_result {
    double loss;          /* 0.5 * sum over all voxels of (c - target)^2 */
    double dl_dd;         /* dL/dD, D in px^2 per unit time (divide by dx^2 for um^2/s) */
    double dl_dg;         /* dL/d(value of every held voxel) */
    uint64_t sweeps;      /* adjoint SOR sweeps */
    double rel_residual;  /* ||(c - target) - A lambda|| / ||c - target|| at exit */
    double seconds;
    int32_t status;
    int32_t pad_;
} cw_adjoint_result;

/* Adjoint gradients of the steady state (new: cellweave fits parameters by
 * hand). The device values must be the solved field (after a relax). Solves
 * A lambda = (c - target) on the free voxels in FP64 to opts->tolerance
 * (relative residual; opts->omega, opts->max_sweeps apply) and returns
 *   dL/dD = sum_i lambda_i sum_{real nbrs j} (c_j - c_i),
 *   dL/dg = D sum_i lambda_i #held_nbrs(i) + sum_held (c_h - target_h),
 *   dl_du[k] = dL/d(uptake rate of cell k) = -sum of lambda over the free
 *              voxels of cell k (grid ids, x fastest) whose uptake is > 0,
 * for k < ncells (dl_du and grid may be NULL). CW_ERR_ARG if a free voxel
 * sits on the floor (the gradient of the floored problem is not linear). */
cw_status cw_field_adjoint(cw_field *field, const double *target, const uint32_t *grid,
                           uint32_t ncells, const cw_relax_opts *opts, double *dl_du,
                           cw_adjoint_result *result);

/* Autotuner (no reference counterpart): times `launches` back-to-back launches
 * of every sweep-kernel shape that fits these fields and this precision
 * (2D or 3D) and writes up to `cap` shapes ranked by microseconds per
 * sweep. The fields' values are swept by the call. */
cw_status cw_autotune(cw_field *const *fields, uint32_t count, const cw_relax_opts *opts,
                      uint32_t launches, int32_t *shapes, double *us_per_sweep, uint32_t cap,
                      uint32_t *n_out);

/* Measurement helper (no reference counterpart): builds the same launch plan
 * cw_batch_relax would for `opts` and times `passes` back-to-back launches of
 * its sweep kernel with HIP events on the launching stream. *ms_per_pass gets
 * the average; with two to four fields, whose launches run concurrently, the
 * longest per-field average. The fields' values are swept (not solved). */
cw_status cw_batch_time_passes(cw_field *const *fields, uint32_t count,
                               const cw_relax_opts *opts, uint32_t passes,
                               double *ms_per_pass);

/* ---- the field's state on the device: lattice gathers, holds, step ---------
 * Each call below is synchronous, like the Rust function it replaces, and
 * marks the coefficients the next solve derives from the state as stale. */

/* A device copy of CpmLattice::grid (cpm.rs:198): u32 cell ids, 0 = medium,
 * 0xFFFFFFFF = obstacle (cpm.rs:137). Upload after every lattice change; the
 * gathers below read it on the device instead of one FFI call per voxel. */
cw_status cw_lattice_create(uint32_t width, uint32_t height, uint32_t depth, cw_lattice **out);
cw_status cw_lattice_destroy(cw_lattice *lattice);
cw_status cw_lattice_upload(cw_lattice *lattice, const uint32_t *ids);
/* The voxels a lattice step changed (new; no reference counterpart, it
 * replaces re-uploading cpm.rs:198's whole grid after a step's accepted
 * copies): voxel voxels[j] (flat index (z*H + y)*W + x, pde.rs:27-29) takes
 * id ids[j], in list order, so a voxel listed twice keeps its last id.
 * CW_ERR_ARG for an index outside the lattice (nothing is changed then). */
cw_status cw_lattice_set(cw_lattice *lattice, const uint64_t *voxels, const uint32_t *ids,
                         uint64_t count);

/* coupling.rs:26-39 apply_uptake: clear the sinks, then every voxel of a
 * cell id < ncells with consumes[id] != 0 (coupling.rs:15-20, a living kind)
 * takes up max(rate, 0). */
cw_status cw_field_apply_uptake(cw_field *field, cw_lattice *lattice, const uint8_t *consumes,
                                uint32_t ncells, double rate);
/* apply_uptake with one rate per cell (the per-cell uptake vector the adjoint
 * differentiates; no single counterpart in the reference): a voxel of cell id
 * takes up rates[id] where that is positive. consumes may be NULL (all). */
cw_status cw_field_apply_uptake_per_cell(cw_field *field, cw_lattice *lattice,
                                         const uint8_t *consumes, const double *rates,
                                         uint32_t ncells);
/* coupling.rs:47-53 apply_background_uptake: every medium voxel (id 0) takes
 * up max(rate, 0); other voxels keep their sink. */
cw_status cw_field_apply_background_uptake(cw_field *field, cw_lattice *lattice, double rate);
/* coupling.rs:61-73 apply_production: clear the sources, then every voxel of
 * a cell id < ncells secretes per_cell[id] where that is positive. */
cw_status cw_field_apply_production(cw_field *field, cw_lattice *lattice, const double *per_cell,
                                    uint32_t ncells);

/* coupling.rs:80-82 hold_obstacles_at: hold every obstacle voxel at value. */
cw_status cw_field_hold_obstacles_at(cw_field *field, cw_lattice *lattice, double value);
/* pde.rs:491-498 hold_where with the predicate as a byte mask (nonzero =
 * hold): held, and the value set. */
cw_status cw_field_hold_where(cw_field *field, const uint8_t *mask, double value);
/* pde.rs:550-631 hold_bath_at with is_open as a byte mask (nonzero = open):
 * the bath floods from the outer layer of the box around the closed voxels
 * (one voxel wider) through the faces of open voxels not already held;
 * everything outside that box and everything reached is held, and then every
 * held voxel takes `value`. */
cw_status cw_field_hold_bath_at(cw_field *field, const uint8_t *open, double value);
/* coupling.rs:156-158 hold_medium_bath: hold_bath_at with open = medium (id 0). */
cw_status cw_field_hold_medium_bath(cw_field *field, cw_lattice *lattice, double value);
/* pde.rs:468-470 release_all. */
cw_status cw_field_release_all(cw_field *field);

/* pde.rs:460-463 free_count: voxels not held. */
cw_status cw_field_free_count(cw_field *field, uint64_t *count);
/* pde.rs:415-456 inflow: the flux D (c_j - c_i) over every face between a
 * free voxel i and a held voxel j. Summed in a fixed tree (the same bits on
 * every run); the reference's sequential order differs by rounding only. */
cw_status cw_field_inflow(cw_field *field, double *flux);
/* DiffusionField::step (pde.rs:705-728, laplacian 651-667, check_step
 * 673-701): one explicit Euler step with no-flux walls, floored at zero;
 * CW_ERR_CONFIG for a step the scheme cannot take. */
cw_status cw_field_step(cw_field *field, double dt);
/* coupling.rs:164-182 mean_per_cell: out[k] for k < ncells is the mean value
 * over the voxels of cell k (0 for the medium and for a cell with none),
 * added in increasing voxel order as the reference adds them. */
cw_status cw_field_mean_per_cell(cw_field *field, cw_lattice *lattice, uint32_t ncells,
                                 double *out);
/* Host copies (n entries each; any pointer may be NULL) of the held mask,
 * uptake and production, for the Rust getters (pde.rs:184-204). */
cw_status cw_field_download_state(cw_field *field, uint8_t *held, double *uptake,
                                  double *production);

#ifdef __cplusplus
}
#endif
#endif
