/*
 * oracle.h - C API of the CPU restatement (test infrastructure).
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, on the CPU and from the
 * reference sources, Abeille's multigroup transport path so that the HIP
 * implementation can be checked against it.  Only tests/, the smoke check
 * and bench.py's CPU-baseline leg may load it; the product never does.
 *
 * It consumes the same flat description as the product (include/abx.h).
 * `portable_math` selects the deterministic log/sin/cos shared with the GPU
 * (1) or glibc's (0, the reference's own libm).
 */
#ifndef ORACLE_H
#define ORACLE_H

#include "../include/abx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_problem orc_problem;

const char* orc_last_error(void);
int orc_create(const abx_problem_desc* desc, int portable_math,
               orc_problem** out);
void orc_destroy(orc_problem* p);

int orc_sample_sources(orc_problem* p, uint64_t first_history, size_t n,
                       abx_particle* out);

/* One Transporter::transport call.  trace may be NULL. */
int orc_transport(orc_problem* p, const abx_particle* bank, size_t n,
                  double kcol, int converged, int nthreads,
                  abx_banked* fission_out, size_t cap, size_t* n_out,
                  abx_scores* scores, abx_diag* diag, abx_trace_rec* trace,
                  int cap_per_history, int32_t* n_rec);

int orc_tally_size(orc_problem* p, int tally, size_t* n_bins);
int orc_tally_fetch(orc_problem* p, int tally, double* gen_out);
int orc_tally_clear(orc_problem* p);
/* MeshTally::record_generation(multiplier) of every mesh tally. */
int orc_tally_record(orc_problem* p, double multiplier);

/* Whole FixedSource / PowerIterator run. */
int orc_run(orc_problem* p, int ngenerations, int nignored, int nthreads,
            double* kcol_hist, abx_run_stats* stats);
int orc_tally_stats(orc_problem* p, int tally, double* avg, double* std_err);
/* Per-generation values of the last run (abx_gen_quantity order). */
int orc_gen_values(orc_problem* p, int quantity, double* out, size_t cap, size_t* n);
int orc_bank_size(orc_problem* p, size_t* n);
int orc_bank_fetch(orc_problem* p, double* out9);

/* Geometry probe: locate points (Tracker construction) -> cell / material
 * index (-1 if lost). */
int orc_locate(orc_problem* p, const double* r, const double* u, size_t n,
               int32_t* cell, int32_t* material);
/* Nearest boundary (Tracker::get_nearest_boundary) from located points. */
int orc_nearest_boundary(orc_problem* p, const double* r, const double* u,
                         size_t n, double* dist, int32_t* surface,
                         int32_t* btype);

/* Approximate mesh cancellation of a bank, in place. */
int orc_cancel(orc_problem* p, abx_banked* bank, size_t n);

/* Majorant per group (src/majorant.cpp:133-176). */
int orc_majorant(orc_problem* p, double* maj);

/* Math used by the restatement. */
/* parents of the last transport's fission sites (exact cancellation) */
int orc_parents_fetch(orc_problem* p, abx_parent_info* out, size_t cap, size_t* n);
/* one exact / basic-exact cancellation pass over a bank with its parents:
 * out = the bank with cancelled weights, then the uniform particles */
int orc_exact_cancel(orc_problem* p, const abx_banked* bank, const abx_parent_info* par, size_t n,
                     abx_banked* out, size_t cap, size_t* n_out);
int orc_rng_reset(orc_problem* p); /* settings::rng back to the seed */
uint64_t orc_sobol_matrix(unsigned dim, unsigned k);
double orc_sobol(uint64_t index, unsigned dim);
double orc_log(int portable, double x);
double orc_exp(int portable, double x);
double orc_sin(int portable, double x);
double orc_cos(int portable, double x);

#ifdef __cplusplus
}
#endif

#endif
