/*
 * oracle.h - C API of the FP64 CPU oracle (a restatement of the magnum.af reference arithmetic).
 *
 * TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg are
 * the only callers.  Parameters use the mafx.h POD types so one set of inputs drives both the
 * oracle and the GPU library.  All arrays are host arrays in the mafx.h layout.
 */
#ifndef ORACLE_H
#define ORACLE_H
#include "../include/mafx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_sim oracle_sim;

/* State(mesh, Ms, m) semantics (state.cpp:71-111): m is normalised (zero vectors stay zero);
 * when Ms_field == NULL and any |m| == 0, an Ms_field = Ms * (|m| != 0) is built. */
oracle_sim* oracle_create(const mafx_mesh* mesh, double Ms, const double* Ms_field, const mafx_term_desc* terms,
                          int nterms, const double* m0, int nthreads);
void oracle_destroy(oracle_sim* s);
int oracle_add_term(oracle_sim* s, const mafx_term_desc* term, int nthreads);
int oracle_set_m(oracle_sim* s, const double* m, int normalize);
int oracle_get_m(const oracle_sim* s, double* m);
int oracle_set_t(oracle_sim* s, double t);
double oracle_get_t(const oracle_sim* s);
int64_t oracle_get_steps(const oracle_sim* s);
int oracle_has_ms_field(const oracle_sim* s);
int oracle_heff(oracle_sim* s, uint32_t mask, double* h);
int oracle_energy(oracle_sim* s, uint32_t mask, double* E_terms, double* E_total);
int oracle_mean_m(const oracle_sim* s, double out[3]);

/* LLGIntegrator(alpha, terms, scheme, Controller(hmin, hmax, atol, rtol), dissipation_term_only).
 * normalize / analytic_rhs exist for the AdaptiveRungeKutta analytic-ODE tests (analytic_rhs 1:
 * y' = t sqrt(y) as in the reference's test; 2: y' = cos(t) y). */
int oracle_llg_init(oracle_sim* s, double alpha, const char* scheme, double hmin, double hmax, double atol,
                    double rtol, int dissipation_term_only, int normalize, int analytic_rhs);
int oracle_llg_set_alpha(oracle_sim* s, double alpha);
/* LLGIntegrator(af::array alpha_field, ...) (llg_integrator.cpp:18-38): alpha [N], NULL for the scalar. */
int oracle_llg_set_alpha_field(oracle_sim* s, const double* alpha);
/* One AdaptiveRungeKutta::step; outputs the accepted h, its error norm, counters. */
int oracle_llg_step(oracle_sim* s, double* h_used, double* err, double* h_next, int64_t* accepted,
                    int64_t* rejected, int64_t* heff_evals);
int oracle_llg_integrate_dense(oracle_sim* s, double time, double dt_write, double* out, int64_t max_rows,
                               int64_t* rows);
/* One attempt at step size dt without the controller: returns sumbk (3N) and the error norm. */
int oracle_llg_attempt(oracle_sim* s, double dt, double* dm, double* err);

int oracle_lbfgs(oracle_sim* s, double tol, int64_t maxiter, double torque_tol, double* E, int64_t* iters,
                 int64_t* evals, int* reason);

/* oracle_lbfgs without the torque stop, recording one row per line search: (f(0), g(0).s, step,
 * f(step), g(step).s, cvsrch info code, evaluations), 7 doubles per row. */
int oracle_lbfgs_trace(oracle_sim* s, double tol, int64_t maxiter, double* E, double* rows_out, int64_t max_rows,
                       int64_t* rows);
/* Controller counters: accepted, rejected, hmin, hmax, maxscale, minscale (controller.hpp:43-49). */
int oracle_llg_counters(const oracle_sim* s, int64_t out[6]);

/* Real-space Newell tensor on the expanded grid, [6][n2e][n1e][n0e] (demag_field.cpp:104-142). */
int oracle_demag_tensor(const mafx_mesh* mesh, double* N, int nthreads);
double oracle_newell_nxx(int ix, int iy, int iz, double dx, double dy, double dz);
double oracle_newell_nxy(int ix, int iy, int iz, double dx, double dy, double dz);
/* Unnormalised complex DFT of n interleaved (re, im) pairs in place; sign -1 forward. */
int oracle_fft(double* data, int64_t n, int sign);

#ifdef __cplusplus
}
#endif
#endif
