/*
 * CPU FP64 oracle for the cellweave diffusion field solve.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the GPU path is
 * compared against. Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it; the product library (libcwfield.so) never links
 * or calls it, and nothing here is shipped as the solver.
 *
 * It restates, in C, the arithmetic and control flow of the reference's
 * ScalarField (crates/engine/src/pde.rs) and the lattice coupling
 * (crates/engine/src/coupling.rs), keeping the expression order of every
 * floating-point operation so that an FP64 result can be compared bit for bit.
 * Built with -O2 -ffp-contract=off and no fast-math.
 *
 * Functions that the reference does not have (the relative residual, the
 * relax-to-residual driver, the per-cell uptake gather, the adjoint) are new
 * definitions; their pinning tests are listed beside each declaration.
 */
#ifndef CW_PDE_ORACLE_H
#define CW_PDE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORF_OK = 0, ORF_ERR_CONFIG = 1, ORF_ERR_NOT_SETTLED = 2 };

/* State of one field: pde.rs:30-53. held is one byte per voxel. */
typedef struct orf_field {
    uint32_t w, h, d;
    size_t n;
    double *values, *next, *uptake, *production;
    uint8_t *held;
    double diff, decay, source;
    size_t *frontier;
    size_t frontier_len, frontier_cap;
} orf_field;

/* pde.rs:98-131 (with_depth): validation then allocation. */
int orf_create(uint32_t w, uint32_t h, uint32_t d, double diff, double decay,
               double initial, orf_field **out);
void orf_destroy(orf_field *f);

double *orf_values(orf_field *f);
double *orf_next(orf_field *f);
double *orf_uptake(orf_field *f);
double *orf_production(orf_field *f);
uint8_t *orf_held(orf_field *f);
size_t orf_len(const orf_field *f);

/* pde.rs:135-138 */
void orf_set_source(orf_field *f, double source);
/* pde.rs:146-152, 159-165, 170-182: clamp with max(0), out of range ignored. */
void orf_set_uptake(orf_field *f, uint32_t x, uint32_t y, double rate);
void orf_set_production(orf_field *f, uint32_t x, uint32_t y, double rate);
void orf_set_uptake_at(orf_field *f, size_t i, double rate);
void orf_set_production_at(orf_field *f, size_t i, double rate);
/* pde.rs:206-220 */
void orf_clear_production(orf_field *f);
void orf_clear_uptake(orf_field *f);
void orf_fill(orf_field *f, double value);

/* pde.rs:305-320. span = {x0,x1,y0,y1,z0,z1}; returns 0 when nothing is free. */
int orf_free_span(const orf_field *f, uint32_t span[6]);

/* pde.rs:267-277: the relaxation factor relax() picks for this span. */
double orf_reference_omega(const orf_field *f, const uint32_t span[6]);

/* pde.rs:329-405. threads <= 1 runs serially; otherwise OpenMP over rows in
 * chunks of 4 (ROWS_PER_TASK, pde.rs:63-69). Returns the largest change. */
double orf_half_sweep(orf_field *f, double omega, uint32_t colour,
                      const uint32_t span[6], int threads);
/* pde.rs:297-301 */
double orf_sor_sweep(orf_field *f, double omega, const uint32_t span[6], int threads);

/* pde.rs:245-288. omega is NaN for the reference rule. *done gets the sweep
 * count (also on failure, the number run). Returns ORF_ERR_NOT_SETTLED after
 * max_sweeps. */
int orf_relax(orf_field *f, double tolerance, size_t max_sweeps, double omega,
              int threads, size_t *done);

/* pde.rs:415-456: flux from held into free pixels, summed in fixed order. */
double orf_inflow(const orf_field *f);
/* pde.rs:460-463 */
size_t orf_free_count(const orf_field *f);
/* pde.rs:468-470 */
void orf_release_all(orf_field *f);
/* pde.rs:491-498: hold every voxel whose mask byte is set. */
void orf_hold_where(orf_field *f, double value, const uint8_t *mask);
/* pde.rs:550-640: bath flood from the box around the closed voxels. */
void orf_hold_bath_at(orf_field *f, double value, const uint8_t *is_open);

/* pde.rs:673-701 (check_step) and 705-728 (step). */
int orf_step(orf_field *f, double dt);

/* coupling.rs:26-39. consumes[id] is 1 for a living kind (coupling.rs:15-20);
 * ids at or past ncells, 0 and 0xFFFFFFFF are skipped. */
void orf_apply_uptake(orf_field *f, const uint32_t *grid, const uint8_t *consumes,
                      size_t ncells, double rate);
/* New: the per-cell form of apply_uptake, shaped like apply_production
 * (coupling.rs:61-73): rate per_cell[id], set when the cell consumes. */
void orf_apply_uptake_per_cell(orf_field *f, const uint32_t *grid,
                               const uint8_t *consumes, const double *per_cell,
                               size_t ncells);
/* coupling.rs:47-53 */
void orf_apply_background_uptake(orf_field *f, const uint32_t *grid, double rate);
/* coupling.rs:61-73 */
void orf_apply_production(orf_field *f, const uint32_t *grid, const double *per_cell,
                          size_t ncells);
/* coupling.rs:164-182. out has ncells entries (cell_count + 1). */
void orf_mean_per_cell(const orf_field *f, const uint32_t *grid, size_t ncells,
                       double *out);

/* New criterion (not in the reference). For a free voxel
 *   r_i = (source + production_i - uptake_i) + D * sum_{real nbrs} c_j
 *         - (real_i * D + decay) * c_i,
 * projected to the natural residual rt_i = A_ii*(max(0, c_i + r_i/A_ii) - c_i),
 * and divided by ||b_eff||_2 with b_eff_i = b_i + D * sum_{held nbrs} c_j.
 * Pinned by the dense assembled-matrix tests in tests/test_oracle_math.py.
 * Returns the relative norm; *r2 and *b2 (if not NULL) get the squared norms. */
double orf_rel_residual(const orf_field *f, double *r2, double *b2);
/* Residual vector (projected) into out (n entries, zero on held voxels). */
void orf_residual_vector(const orf_field *f, double *out);

/* New: relax until the relative residual of the field is <= rel_tol, testing
 * it before every sweep. *done gets the sweeps run. */
int orf_relax_residual(orf_field *f, double rel_tol, size_t max_sweeps,
                       double omega, int threads, size_t *done);

/* New: FP64 adjoint of the steady state with loss 0.5*sum_all (c - target)^2.
 * Solves A_FF lambda = (c - target)_F by SOR to a relative residual rel_tol,
 * then the closed-form gradients (valid when no free voxel is floored):
 *   dL/dD = sum_i lambda_i sum_{real nbrs j} (c_j - c_i)        (D in px^2/s)
 *   dL/du_k = -sum_{free i in cell k} lambda_i                  (grid may be NULL)
 *   dL/dg = D sum_i lambda_i #held_nbrs(i) + sum_held (c_h - t_h)
 * lambda (n entries) may be NULL. Returns ORF_ERR_NOT_SETTLED if the adjoint
 * solve does not reach rel_tol. */
int orf_adjoint(const orf_field *f, const double *target, double omega,
                double rel_tol, size_t max_sweeps, const uint32_t *grid,
                size_t ncells, double *lambda, double *dl_dd, double *dl_dg,
                double *dl_du, size_t *sweeps);

#ifdef __cplusplus
}
#endif
#endif
