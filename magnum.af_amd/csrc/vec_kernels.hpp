// Host launchers of the FP64 vector kernels (vec.hip) used by the integrator and the minimizer.
#pragma once
#include "mafx_internal.hpp"

namespace mafx {

// Largest stage count of the adaptive schemes (DP78); stage vectors are indexed 1..kMaxStages.
constexpr int kMaxStages = 13;

// sumbk = sum_j b_j k_j and the max-norm error of the embedded pair; k, b, e are indexed 1..ns.
// rkf_form selects rk_error = sumbk - sum_j e_j k_j (adaptive_runge_kutta.cpp:105-106) instead of
// rk_error = sum_j e_j k_j (:169).  The k_j are double*, or float* when k_f32 (sums in FP64).
double rk_error(mafx_ctx* ctx, const double* m, const void* const* k, const double* b, const double* e, int ns,
                bool rkf_form, double atol, double rtol, double* sumbk, bool k_f32 = false);
void accept(mafx_ctx* ctx, double* m, const double* dm, bool normalize_m, bool zero_safe);
void axpy(mafx_ctx* ctx, double* y, double a, const double* x, int64_t n);
void scale(mafx_ctx* ctx, double* y, double a, int64_t n);
void sub(mafx_ctx* ctx, double* out, const double* a, const double* b, int64_t n);
void copy(mafx_ctx* ctx, double* dst, const double* src, int64_t n);
int64_t reduction_partials_needed(int64_t n);
// Device-side scalars (no host synchronisation): *out_dev = a . b and max |a| with the kernels and
// reduction tree of dot() / max_abs(), y += (*a_dev) x, and one read of n device doubles.
void dot_dev(mafx_ctx* ctx, const double* a, const double* b, int64_t n, double* out_dev);
void max_abs_dev(mafx_ctx* ctx, const double* a, int64_t n, double* out_dev);
// out_dev[0] = a . b and out_dev[1] = max |a| in one pass over a, each bit for bit as dot_dev /
// max_abs_dev form it
void dot_max_dev(mafx_ctx* ctx, const double* a, const double* b, int64_t n, double* out_dev);
void axpy_dev(mafx_ctx* ctx, double* y, const double* a_dev, const double* x, int64_t n);
void read_device(mafx_ctx* ctx, const double* src_dev, int n, double* dst_host);
// LBFGS fused vector work, bit for bit the unfused sequence's results:
// lbfgs_fused: q_out = (q_in + (*coef_dev) v) [* scale] (negated when `negate`), *out_dev = dnext . q
//   (before negation), with dot()'s tree; coef_dev null skips the axpy.  q_out may alias q_in.
// lbfgs_pair: s = m - x_old, y = g - g_old, out_dev[0..4] = y.s, y.y, s.s, max|s|, max|m|.
// neg_copy: y = -x.
// mode 1 / 2 also forms the next step's scalars from the dot product d = *out_dev, as lbfgs_alpha
// (alpha_i = rho d, coef = -alpha_i) / lbfgs_beta (coef = alpha_i - rho d) would, in the same launch
// as the reduction's last stage.  lbfgs_dot_alpha: *dot_out = a . b with dot()'s tree, then mode 1.
// Launches the final reduction a two-loop step left pending (lbfgs_fused, lbfgs_dot_alpha), if any:
// the next two-loop pass forms it itself when it uses that step's coefficient; every other user of
// the partials buffer or of the step's scalars flushes it first.
void lbfgs_flush(mafx_ctx* ctx);
void lbfgs_fused(mafx_ctx* ctx, const double* q_in, const double* coef_dev, const double* v, double scale,
                 bool do_scale, const double* dnext, double* q_out, bool negate, int64_t n, double* out_dev,
                 int mode = 0, double rho = 0.0, double* alpha_i = nullptr, double* coef_out = nullptr);
void lbfgs_dot_alpha(mafx_ctx* ctx, const double* a, const double* b, int64_t n, double* dot_out, double rho,
                     double* alpha_i, double* coef);
void lbfgs_pair(mafx_ctx* ctx, const double* m, const double* x_old, const double* g, const double* g_old,
                double* s, double* y, int64_t n, double* out_dev);
void neg_copy(mafx_ctx* ctx, double* y, const double* x, int64_t n);
// <m> of mafx_mean_m with its partial sums in a static device array: no allocation, no
// synchronisation, result in out_dev[0..2] (device memory), launched on `stream`.
void mean_m_noalloc(const double* m, int64_t N, double* out_dev, hipStream_t stream);

} // namespace mafx
