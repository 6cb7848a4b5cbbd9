// Integrator and minimizer state (llg.cpp, lbfgs.cpp).
#pragma once
#include <string>

#include "mafx_internal.hpp"

namespace mafx {

// controller.hpp:11-61 (constants of Numerical Recipes 3rd ed., as in the reference)
struct Controller {
    double hmin{1e-15}, hmax{3.5e-10}, atol{1e-6}, rtol{1e-6};
    double beta = 0.4 / 5.0;
    double alpha = 0.2 - 0.75 * beta;
    double headroom = 0.9, minscale = 0.2, maxscale = 10.;
    bool reject{false};
    double errold{1.0e-4};
    double hnext{0.0};
    int64_t counter_reject{0}, counter_accepted{0}, counter_hmax{0}, counter_hmin{0}, counter_maxscale{0},
        counter_minscale{0};
    bool success(double err, double& h);
};

} // namespace mafx

struct mafx_llg {
    mafx_ctx* ctx{nullptr};
    double alpha{0};
    std::string scheme{"RKF45"};
    bool dissipation_only{false};
    mafx::Controller ctrl;
    double h{1.01e-15}; // adaptive_runge_kutta.hpp:39
    double h_last{0}, err_last{0};
    int64_t heff_evals{0}, steps{0};
    mafx::DevBuf k[14]; // k[1..stages], stages <= 13 (DP78)
    mafx::DevBuf mtmp, mtmp2, sumbk;
    mafx::DevBuf alpha_field; // N f64 (LLGIntegrator(af::array alpha_field, ...)), used when has_alpha_field
    bool has_alpha_field{false};
};

namespace mafx {
// One attempt from (m, t) with step dt; stage i evaluates the field at time t + c_i dt.
double attempt(mafx_llg* L, const double* m, double t, double dt, const char* scheme);
// f evaluations per attempt of a scheme ("RKF45", "DP45", "BS23", "BS45", "DP78")
int scheme_stages(const char* scheme);
void llg_step(mafx_llg* L, double* m, double* t);
int64_t llg_integrate_dense(mafx_llg* L, double* m, double* t, double time_in_s, double dt_write, double* out,
                            int64_t max_rows);
int lbfgs_minimize(mafx_ctx* ctx, double* m, double tolerance, int64_t max_iter, int verbose, double torque_tol,
                   double* E_out, mafx_lbfgs_stats* stats);
} // namespace mafx
