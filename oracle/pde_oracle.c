/*
 * CPU FP64 oracle for the cellweave diffusion field solve: a C restatement of
 * crates/engine/src/pde.rs and crates/engine/src/coupling.rs.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
 * Every function cites the reference lines it follows. Compile with
 * -O2 -ffp-contract=off (no fast-math): each floating-point expression is
 * written in the reference's evaluation order so FP64 results match bit for
 * bit.
 *
 * One deliberate normalisation: the reference floors with Rust's
 * f64::max(x, 0.0) (pde.rs:395, pde.rs:724), which returns 0.0 for NaN and
 * leaves the sign of a zero result unspecified. Here the floor is
 * (x > 0.0 ? x : 0.0): the same number in every case, with +0.0 for a zero.
 * Results are therefore equal to the reference's numerically, and the GPU path
 * uses the same form so that bit comparisons are well defined.
 */
#include "pde_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OBSTACLE_ID 0xFFFFFFFFu

static inline double floor0(double x) { return x > 0.0 ? x : 0.0; }
static inline double clamp_rate(double r) { return r > 0.0 ? r : 0.0; } /* rate.max(0.0) */

/* pde.rs:98-131 */
int orf_create(uint32_t w, uint32_t h, uint32_t d, double diff, double decay,
               double initial, orf_field **out) {
    *out = NULL;
    if (w == 0 || h == 0 || d == 0) return ORF_ERR_CONFIG;
    if (!isfinite(diff) || !isfinite(decay) || diff < 0.0 || decay < 0.0)
        return ORF_ERR_CONFIG;
    orf_field *f = (orf_field *)calloc(1, sizeof(orf_field));
    if (!f) return ORF_ERR_CONFIG;
    f->w = w; f->h = h; f->d = d;
    f->n = (size_t)w * h * d;
    f->values = (double *)malloc(f->n * sizeof(double));
    f->next = (double *)malloc(f->n * sizeof(double));
    f->uptake = (double *)calloc(f->n, sizeof(double));
    f->production = (double *)calloc(f->n, sizeof(double));
    f->held = (uint8_t *)calloc(f->n, 1);
    if (!f->values || !f->next || !f->uptake || !f->production || !f->held) {
        orf_destroy(f);
        return ORF_ERR_CONFIG;
    }
    for (size_t i = 0; i < f->n; i++) { f->values[i] = initial; f->next[i] = initial; }
    f->diff = diff; f->decay = decay; f->source = 0.0;
    *out = f;
    return ORF_OK;
}

void orf_destroy(orf_field *f) {
    if (!f) return;
    free(f->values); free(f->next); free(f->uptake); free(f->production);
    free(f->held); free(f->frontier); free(f);
}

double *orf_values(orf_field *f) { return f->values; }
double *orf_next(orf_field *f) { return f->next; }
double *orf_uptake(orf_field *f) { return f->uptake; }
double *orf_production(orf_field *f) { return f->production; }
uint8_t *orf_held(orf_field *f) { return f->held; }
size_t orf_len(const orf_field *f) { return f->n; }

void orf_set_source(orf_field *f, double source) { f->source = source; } /* pde.rs:135-138 */

/* pde.rs:146-152 */
void orf_set_uptake(orf_field *f, uint32_t x, uint32_t y, double rate) {
    if (x >= f->w || y >= f->h) return;
    f->uptake[(size_t)y * f->w + x] = clamp_rate(rate);
}
/* pde.rs:159-165 */
void orf_set_production(orf_field *f, uint32_t x, uint32_t y, double rate) {
    if (x >= f->w || y >= f->h) return;
    f->production[(size_t)y * f->w + x] = clamp_rate(rate);
}
/* pde.rs:170-174 */
void orf_set_uptake_at(orf_field *f, size_t i, double rate) {
    if (i < f->n) f->uptake[i] = clamp_rate(rate);
}
/* pde.rs:178-182 */
void orf_set_production_at(orf_field *f, size_t i, double rate) {
    if (i < f->n) f->production[i] = clamp_rate(rate);
}
/* pde.rs:206-220 */
void orf_clear_production(orf_field *f) { memset(f->production, 0, f->n * sizeof(double)); }
void orf_clear_uptake(orf_field *f) { memset(f->uptake, 0, f->n * sizeof(double)); }
void orf_fill(orf_field *f, double value) {
    for (size_t i = 0; i < f->n; i++) { f->values[i] = value; f->next[i] = value; }
}

/* pde.rs:305-320: x and y take min/max, z runs from the first free voxel's
 * slice to the last one's. */
int orf_free_span(const orf_field *f, uint32_t span[6]) {
    size_t w = f->w, h = f->h;
    int found = 0;
    for (size_t i = 0; i < f->n; i++) {
        if (f->held[i]) continue;
        uint32_t x = (uint32_t)(i % w), y = (uint32_t)((i / w) % h), z = (uint32_t)(i / (w * h));
        if (!found) {
            span[0] = span[1] = x; span[2] = span[3] = y; span[4] = span[5] = z;
            found = 1;
        }
        if (x < span[0]) span[0] = x;
        if (x > span[1]) span[1] = x;
        if (y < span[2]) span[2] = y;
        if (y > span[3]) span[3] = y;
        span[5] = z;
    }
    return found;
}

/* pde.rs:267-277 */
double orf_reference_omega(const orf_field *f, const uint32_t span[6]) {
    const double pi = 3.14159265358979323846;
    if (f->d > 1) {
        uint32_t longest = 0;
        for (int a = 0; a < 3; a++) {
            uint32_t side = span[2 * a + 1] - span[2 * a] + 1;
            if (side > longest) longest = side;
        }
        if (longest < 2) longest = 2;
        return 2.0 / (1.0 + sin(pi / (double)longest));
    }
    double n = (double)(f->w > f->h ? f->w : f->h);
    double om = 2.0 / (1.0 + sin(pi / n));
    return om < 1.95 ? om : 1.95;
}

/* One row of half_sweep (pde.rs:346-399). */
static double sweep_row(orf_field *f, double omega, uint32_t colour, const uint32_t span[6],
                        size_t stacked) {
    const size_t w = f->w, h = f->h, slice = w * h;
    const uint32_t y = (uint32_t)(stacked % h), z = (uint32_t)(stacked / h);
    const double d = f->diff, decay = f->decay, source = f->source;
    const double *values = f->values;
    double *next = f->next;
    double moved = 0.0;
    if (y < span[2] || y > span[3]) return moved;
    for (uint32_t x = span[0]; x <= span[1]; x++) {
        size_t i = stacked * w + x;
        double before = values[i];
        if ((x + y + z) % 2 != colour || f->held[i]) {
            next[i] = before;
            continue;
        }
        double sum = 0.0, real = 0.0;
        if (x > 0) { sum += values[i - 1]; real += 1.0; }
        if (x + 1 < f->w) { sum += values[i + 1]; real += 1.0; }
        if (y > 0) { sum += values[i - w]; real += 1.0; }
        if (y + 1 < f->h) { sum += values[i + w]; real += 1.0; }
        if (z > 0) { sum += values[i - slice]; real += 1.0; }
        if (z + 1 < f->d) { sum += values[i + slice]; real += 1.0; }
        double balanced = (d * sum + source + f->production[i] - f->uptake[i]) / (real * d + decay);
        double after = floor0((1.0 - omega) * before + omega * balanced);
        next[i] = after;
        double m = fabs(after - before);
        if (m > moved) moved = m;
    }
    return moved;
}

/* pde.rs:329-405 */
double orf_half_sweep(orf_field *f, double omega, uint32_t colour, const uint32_t span[6],
                      int threads) {
    const size_t h = f->h;
    const size_t first = (size_t)span[4] * h + span[2];
    const size_t last = (size_t)span[5] * h + span[3];
    double moved = 0.0;
    if (threads > 1) {
#ifdef _OPENMP
        long nrows = (long)(last - first + 1);
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads) reduction(max : moved)
        for (long k = 0; k < nrows; k++) {
            double m = sweep_row(f, omega, colour, span, first + (size_t)k);
            if (m > moved) moved = m;
        }
#else
        for (size_t s = first; s <= last; s++) {
            double m = sweep_row(f, omega, colour, span, s);
            if (m > moved) moved = m;
        }
#endif
    } else {
        for (size_t s = first; s <= last; s++) {
            double m = sweep_row(f, omega, colour, span, s);
            if (m > moved) moved = m;
        }
    }
    double *t = f->values; f->values = f->next; f->next = t; /* pde.rs:403 */
    return moved;
}

/* pde.rs:297-301 */
double orf_sor_sweep(orf_field *f, double omega, const uint32_t span[6], int threads) {
    double red = orf_half_sweep(f, omega, 0, span, threads);
    double black = orf_half_sweep(f, omega, 1, span, threads);
    return red > black ? red : black;
}

/* pde.rs:245-288 */
int orf_relax(orf_field *f, double tolerance, size_t max_sweeps, double omega, int threads,
              size_t *done) {
    uint32_t span[6];
    if (!orf_free_span(f, span)) { *done = 1; return ORF_OK; }
    if (isnan(omega)) omega = orf_reference_omega(f, span);
    memcpy(f->next, f->values, f->n * sizeof(double));
    for (size_t k = 1; k <= max_sweeps; k++) {
        double moved = orf_sor_sweep(f, omega, span, threads);
        if (moved < tolerance) { *done = k; return ORF_OK; }
    }
    *done = max_sweeps;
    return ORF_ERR_NOT_SETTLED;
}

/* pde.rs:415-456 */
double orf_inflow(const orf_field *f) {
    uint32_t span[6];
    if (!orf_free_span(f, span)) return 0.0;
    const size_t w = f->w, slice = w * f->h;
    double flux = 0.0;
    for (uint32_t z = span[4]; z <= span[5]; z++)
        for (uint32_t y = span[2]; y <= span[3]; y++)
            for (uint32_t x = span[0]; x <= span[1]; x++) {
                size_t i = (size_t)z * slice + (size_t)y * w + x;
                if (f->held[i]) continue;
#define ACROSS(j) do { if (f->held[j]) flux += f->diff * (f->values[j] - f->values[i]); } while (0)
                if (x > 0) ACROSS(i - 1);
                if (x + 1 < f->w) ACROSS(i + 1);
                if (y > 0) ACROSS(i - w);
                if (y + 1 < f->h) ACROSS(i + w);
                if (z > 0) ACROSS(i - slice);
                if (z + 1 < f->d) ACROSS(i + slice);
#undef ACROSS
            }
    return flux;
}

/* pde.rs:460-463 */
size_t orf_free_count(const orf_field *f) {
    size_t c = 0;
    for (size_t i = 0; i < f->n; i++) c += !f->held[i];
    return c;
}

/* pde.rs:468-470 */
void orf_release_all(orf_field *f) { memset(f->held, 0, f->n); }

/* pde.rs:491-498 */
void orf_hold_where(orf_field *f, double value, const uint8_t *mask) {
    for (size_t i = 0; i < f->n; i++)
        if (mask[i]) { f->held[i] = 1; f->values[i] = value; }
}

/* pde.rs:634-640 */
static void reach(orf_field *f, size_t i, const uint8_t *is_open) {
    if (f->held[i] || !is_open[i]) return;
    f->held[i] = 1;
    if (f->frontier_len == f->frontier_cap) {
        size_t cap = f->frontier_cap ? 2 * f->frontier_cap : 1024;
        f->frontier = (size_t *)realloc(f->frontier, cap * sizeof(size_t));
        f->frontier_cap = cap;
    }
    f->frontier[f->frontier_len++] = i;
}

/* pde.rs:550-631 */
void orf_hold_bath_at(orf_field *f, double value, const uint8_t *is_open) {
    f->frontier_len = 0;
    const size_t w = f->w, h = f->h, d = f->d, slice = w * h;
    size_t s[3][2];
    int closed = 0;
    for (size_t i = 0; i < f->n; i++) {
        if (is_open[i]) continue;
        size_t p[3] = {i % w, (i / w) % h, i / slice};
        if (!closed) {
            for (int a = 0; a < 3; a++) s[a][0] = s[a][1] = p[a];
            closed = 1;
        }
        for (int a = 0; a < 3; a++) {
            if (p[a] < s[a][0]) s[a][0] = p[a];
            if (p[a] > s[a][1]) s[a][1] = p[a];
        }
    }
    const size_t limits[3] = {w, h, d};
    if (closed)
        for (int a = 0; a < 3; a++) {
            s[a][0] = s[a][0] > 0 ? s[a][0] - 1 : 0;
            s[a][1] = (s[a][1] + 1 < limits[a] - 1) ? s[a][1] + 1 : limits[a] - 1;
        }
    for (size_t i = 0; i < f->n; i++) {
        size_t p[3] = {i % w, (i / w) % h, i / slice};
        int inside = closed;
        for (int a = 0; a < 3 && inside; a++)
            if (!(s[a][0] <= p[a] && p[a] <= s[a][1])) inside = 0;
        if (!inside) f->held[i] = 1;
    }
    if (closed) {
        for (size_t z = s[2][0]; z <= s[2][1]; z++)
            for (size_t y = s[1][0]; y <= s[1][1]; y++)
                for (size_t x = s[0][0]; x <= s[0][1]; x++) {
                    int outer = x == s[0][0] || x == s[0][1] || y == s[1][0] || y == s[1][1] ||
                                (d > 1 && (z == s[2][0] || z == s[2][1]));
                    if (outer) reach(f, z * slice + y * w + x, is_open);
                }
    }
    while (f->frontier_len > 0) {
        size_t i = f->frontier[--f->frontier_len];
        size_t x = i % w, y = (i / w) % h, z = i / slice;
        if (x > 0) reach(f, i - 1, is_open);
        if (x + 1 < w) reach(f, i + 1, is_open);
        if (y > 0) reach(f, i - w, is_open);
        if (y + 1 < h) reach(f, i + w, is_open);
        if (z > 0) reach(f, i - slice, is_open);
        if (z + 1 < d) reach(f, i + slice, is_open);
    }
    for (size_t i = 0; i < f->n; i++)
        if (f->held[i]) f->values[i] = value;
}

/* pde.rs:651-667 */
static double laplacian(const orf_field *f, size_t i) {
    const size_t w = f->w, h = f->h, slice = w * h;
    size_t x = i % w, y = (i / w) % h, z = i / slice;
    double c = f->values[i];
#define AT(j, there) ((there) ? f->values[(j)] : c)
    double lap = AT(i - 1, x > 0) + AT(i + 1, x + 1 < w) + AT(i - w, y > 0) + AT(i + w, y + 1 < h) -
                 4.0 * c;
    if (f->d > 1) lap += AT(i - slice, z > 0) + AT(i + slice, z + 1 < f->d) - 2.0 * c;
#undef AT
    return lap;
}

/* pde.rs:673-701 */
static int check_step(const orf_field *f, double dt) {
    if (isnan(dt) || dt <= 0.0) return ORF_ERR_CONFIG;
    double limit = f->d > 1 ? 0.25 * 2.0 / 3.0 : 0.25;
    if (f->diff * dt > limit) return ORF_ERR_CONFIG;
    if (f->decay * dt > 1.0) return ORF_ERR_CONFIG;
    return ORF_OK;
}

/* pde.rs:705-728 */
int orf_step(orf_field *f, double dt) {
    int rc = check_step(f, dt);
    if (rc) return rc;
    for (size_t i = 0; i < f->n; i++) {
        double c = f->values[i];
        if (f->held[i]) { f->next[i] = c; continue; }
        double lap = laplacian(f, i);
        double change = f->diff * lap - f->decay * c + f->source + f->production[i] - f->uptake[i];
        f->next[i] = floor0(c + dt * change);
    }
    double *t = f->values; f->values = f->next; f->next = t;
    return ORF_OK;
}

/* coupling.rs:26-39 */
void orf_apply_uptake(orf_field *f, const uint32_t *grid, const uint8_t *consumes, size_t ncells,
                      double rate) {
    orf_clear_uptake(f);
    for (size_t i = 0; i < f->n; i++) {
        uint32_t id = grid[i];
        if (id == 0 || id == OBSTACLE_ID) continue;
        if (id < ncells && consumes[id]) orf_set_uptake_at(f, i, rate);
    }
}

void orf_apply_uptake_per_cell(orf_field *f, const uint32_t *grid, const uint8_t *consumes,
                               const double *per_cell, size_t ncells) {
    orf_clear_uptake(f);
    for (size_t i = 0; i < f->n; i++) {
        uint32_t id = grid[i];
        if (id == 0 || id == OBSTACLE_ID) continue;
        if (id < ncells && consumes[id] && per_cell[id] > 0.0) orf_set_uptake_at(f, i, per_cell[id]);
    }
}

/* coupling.rs:47-53 */
void orf_apply_background_uptake(orf_field *f, const uint32_t *grid, double rate) {
    for (size_t i = 0; i < f->n; i++)
        if (grid[i] == 0) orf_set_uptake_at(f, i, rate);
}

/* coupling.rs:61-73 */
void orf_apply_production(orf_field *f, const uint32_t *grid, const double *per_cell,
                          size_t ncells) {
    orf_clear_production(f);
    for (size_t i = 0; i < f->n; i++) {
        uint32_t id = grid[i];
        if (id == 0 || id == OBSTACLE_ID) continue;
        if (id < ncells && per_cell[id] > 0.0) orf_set_production_at(f, i, per_cell[id]);
    }
}

/* coupling.rs:164-182 */
void orf_mean_per_cell(const orf_field *f, const uint32_t *grid, size_t ncells, double *out) {
    uint32_t *count = (uint32_t *)calloc(ncells ? ncells : 1, sizeof(uint32_t));
    for (size_t k = 0; k < ncells; k++) out[k] = 0.0;
    for (size_t i = 0; i < f->n; i++) {
        uint32_t id = grid[i];
        if (id == 0 || id == OBSTACLE_ID) continue;
        if (id < ncells) { out[id] += f->values[i]; count[id] += 1; }
    }
    for (size_t k = 0; k < ncells; k++) out[k] = count[k] == 0 ? 0.0 : out[k] / (double)count[k];
    free(count);
}

/* Residual of one free voxel: r = b + D*sum - A_ii*c, with b = (source + p) - u,
 * and the held neighbours' part of D*sum in *held_part. */
static double voxel_residual(const orf_field *f, const double *c, size_t i, double *held_part,
                             double *aii_out) {
    const size_t w = f->w, h = f->h, slice = w * h;
    size_t x = i % w, y = (i / w) % h, z = i / slice;
    double sum = 0.0, hsum = 0.0, real = 0.0;
#define NB(j) do { sum += c[(j)]; real += 1.0; if (f->held[(j)]) hsum += c[(j)]; } while (0)
    if (x > 0) NB(i - 1);
    if (x + 1 < f->w) NB(i + 1);
    if (y > 0) NB(i - w);
    if (y + 1 < f->h) NB(i + w);
    if (z > 0) NB(i - slice);
    if (z + 1 < f->d) NB(i + slice);
#undef NB
    double aii = real * f->diff + f->decay;
    double b = (f->source + f->production[i]) - f->uptake[i];
    double r = (b + f->diff * sum) - aii * c[i];
    if (held_part) *held_part = b + f->diff * hsum;
    if (aii_out) *aii_out = aii;
    return r;
}

static double project(double r, double c, double aii) {
    if (aii > 0.0 && !(c + r / aii > 0.0)) return -aii * c;
    return r;
}

void orf_residual_vector(const orf_field *f, double *out) {
    for (size_t i = 0; i < f->n; i++) {
        if (f->held[i]) { out[i] = 0.0; continue; }
        double aii;
        double r = voxel_residual(f, f->values, i, NULL, &aii);
        out[i] = project(r, f->values[i], aii);
    }
}

double orf_rel_residual(const orf_field *f, double *r2o, double *b2o) {
    double r2 = 0.0, b2 = 0.0;
    for (size_t i = 0; i < f->n; i++) {
        if (f->held[i]) continue;
        double beff, aii;
        double r = voxel_residual(f, f->values, i, &beff, &aii);
        double rt = project(r, f->values[i], aii);
        r2 += rt * rt;
        b2 += beff * beff;
    }
    if (r2o) *r2o = r2;
    if (b2o) *b2o = b2;
    if (b2 == 0.0) return sqrt(r2);
    return sqrt(r2) / sqrt(b2);
}

int orf_relax_residual(orf_field *f, double rel_tol, size_t max_sweeps, double omega, int threads,
                       size_t *done) {
    uint32_t span[6];
    *done = 0;
    if (!orf_free_span(f, span)) return ORF_OK;
    if (isnan(omega)) omega = orf_reference_omega(f, span);
    memcpy(f->next, f->values, f->n * sizeof(double));
    for (size_t k = 0;; k++) {
        if (orf_rel_residual(f, NULL, NULL) <= rel_tol) { *done = k; return ORF_OK; }
        if (k == max_sweeps) { *done = k; return ORF_ERR_NOT_SETTLED; }
        orf_sor_sweep(f, omega, span, threads);
    }
}

/* Adjoint: red-black SOR on A_FF lambda = rhs_F with lambda = 0 on held voxels. */
int orf_adjoint(const orf_field *f, const double *target, double omega, double rel_tol,
                size_t max_sweeps, const uint32_t *grid, size_t ncells, double *lambda_out,
                double *dl_dd, double *dl_dg, double *dl_du, size_t *sweeps) {
    const size_t n = f->n, w = f->w, h = f->h, slice = w * h;
    const double D = f->diff;
    double *lam = (double *)calloc(n, sizeof(double));
    double *rhs = (double *)calloc(n, sizeof(double));
    double rhs2 = 0.0;
    for (size_t i = 0; i < n; i++)
        if (!f->held[i]) { rhs[i] = f->values[i] - target[i]; rhs2 += rhs[i] * rhs[i]; }
    uint32_t span[6];
    int rc = ORF_OK;
    size_t k = 0;
    if (orf_free_span(f, span)) {
        if (isnan(omega)) omega = orf_reference_omega(f, span);
        for (k = 0;; k++) {
            double r2 = 0.0;
            for (size_t i = 0; i < n; i++) {
                if (f->held[i]) continue;
                size_t x = i % w, y = (i / w) % h, z = i / slice;
                double sum = 0.0, real = 0.0;
                if (x > 0) { sum += lam[i - 1]; real += 1.0; }
                if (x + 1 < w) { sum += lam[i + 1]; real += 1.0; }
                if (y > 0) { sum += lam[i - w]; real += 1.0; }
                if (y + 1 < h) { sum += lam[i + w]; real += 1.0; }
                if (z > 0) { sum += lam[i - slice]; real += 1.0; }
                if (z + 1 < f->d) { sum += lam[i + slice]; real += 1.0; }
                double r = (rhs[i] + D * sum) - (real * D + f->decay) * lam[i];
                r2 += r * r;
            }
            if (rhs2 == 0.0 || sqrt(r2) <= rel_tol * sqrt(rhs2)) break;
            if (k == max_sweeps) { rc = ORF_ERR_NOT_SETTLED; break; }
            for (uint32_t colour = 0; colour < 2; colour++)
                for (size_t i = 0; i < n; i++) {
                    if (f->held[i]) continue;
                    size_t x = i % w, y = (i / w) % h, z = i / slice;
                    if ((x + y + z) % 2 != colour) continue;
                    double sum = 0.0, real = 0.0;
                    if (x > 0) { sum += lam[i - 1]; real += 1.0; }
                    if (x + 1 < w) { sum += lam[i + 1]; real += 1.0; }
                    if (y > 0) { sum += lam[i - w]; real += 1.0; }
                    if (y + 1 < h) { sum += lam[i + w]; real += 1.0; }
                    if (z > 0) { sum += lam[i - slice]; real += 1.0; }
                    if (z + 1 < f->d) { sum += lam[i + slice]; real += 1.0; }
                    double bal = (D * sum + rhs[i]) / (real * D + f->decay);
                    lam[i] = (1.0 - omega) * lam[i] + omega * bal;
                }
        }
    }
    if (sweeps) *sweeps = k;
    double gd = 0.0, gg = 0.0;
    if (dl_du) for (size_t c = 0; c < ncells; c++) dl_du[c] = 0.0;
    for (size_t i = 0; i < n; i++) {
        const double *c = f->values;
        if (f->held[i]) { gg += c[i] - target[i]; continue; }
        size_t x = i % w, y = (i / w) % h, z = i / slice;
        double dsum = 0.0, nheld = 0.0;
#define NB(j) do { dsum += c[(j)] - c[i]; if (f->held[(j)]) nheld += 1.0; } while (0)
        if (x > 0) NB(i - 1);
        if (x + 1 < w) NB(i + 1);
        if (y > 0) NB(i - w);
        if (y + 1 < h) NB(i + w);
        if (z > 0) NB(i - slice);
        if (z + 1 < f->d) NB(i + slice);
#undef NB
        gd += lam[i] * dsum;
        gg += D * lam[i] * nheld;
        if (dl_du && grid) {
            uint32_t id = grid[i];
            /* only voxels that take up at their cell's rate depend on it */
            if (id != 0 && id != OBSTACLE_ID && id < ncells && f->uptake[i] > 0.0)
                dl_du[id] -= lam[i];
        }
    }
    if (dl_dd) *dl_dd = gd;
    if (dl_dg) *dl_dg = gg;
    if (lambda_out) memcpy(lambda_out, lam, n * sizeof(double));
    free(lam);
    free(rhs);
    return rc;
}
