// LLG adaptive Runge-Kutta integrator on the GPU.  The step-size controller runs on the host on
// one FP64 scalar per attempt (the error norm, one 8-byte device-to-host read); every 3N-sized
// operation runs on the device.
//
// Replaces magnumaf/src/integrators/adaptive_runge_kutta.cpp:39-173, controller.cpp:36-95 and
// llg_integrator.cpp:11-104.
#include <array>
#include <cmath>
#include <limits>
#include <string>

#include "mafx_internal.hpp"
#include "mafx_llg.hpp"
#include "vec_kernels.hpp"

namespace mafx {

// controller.cpp:36-95 (Numerical Recipes PI control; clamped steps are accepted)
bool Controller::success(const double err, double& h) {
    double scale = std::numeric_limits<double>::quiet_NaN();
    if (err <= 1.0) {
        if (err == 0.0) {
            scale = maxscale;
        } else {
            scale = headroom * std::pow(err, -alpha) * std::pow(errold, beta);
            if (scale < minscale) {
                scale = minscale;
                counter_maxscale++;
            }
            if (scale > maxscale) {
                scale = maxscale;
                counter_minscale++;
            }
        }
        hnext = reject ? h * std::min(scale, 1.0) : h * scale;
        if (hnext <= hmin) {
            hnext = hmin;
            counter_hmin++;
        } else if (hnext >= hmax) {
            hnext = hmax;
            counter_hmax++;
        }
        errold = std::max(err, 1.0e-4);
        reject = false;
        counter_accepted++;
        return true;
    }
    scale = std::max(headroom * std::pow(err, -alpha), minscale);
    h *= scale;
    if (h <= hmin) {
        h = hmin;
        counter_hmin++;
        return true;
    }
    if (h >= hmax) {
        h = hmax;
        counter_hmax++;
        return true;
    }
    reject = true;
    counter_reject++;
    return false;
}

namespace {

// Butcher tableaux exactly as the reference writes its coefficients.
struct Tableau {
    int stages;          // number of f evaluations
    double a[8][8];      // a[s][j], s = 2..stages
    double b[8];         // weights of the propagated solution (sumbk)
    double e[8];         // error weights
    bool rkf_form;       // rk_error = sumbk - sum e_j k_j
};

Tableau dp45() { // adaptive_runge_kutta.cpp:117-128
    Tableau t{};
    t.stages = 7;
    t.rkf_form = false;
    t.a[2][1] = 0.2;
    t.a[3][1] = 3.0 / 40.0, t.a[3][2] = 9.0 / 40.0;
    t.a[4][1] = 44.0 / 45.0, t.a[4][2] = -56.0 / 15.0, t.a[4][3] = 32.0 / 9.0;
    t.a[5][1] = 19372.0 / 6561.0, t.a[5][2] = -25360.0 / 2187.0, t.a[5][3] = 64448.0 / 6561.0,
    t.a[5][4] = -212.0 / 729.0;
    t.a[6][1] = 9017.0 / 3168.0, t.a[6][2] = -355.0 / 33.0, t.a[6][3] = 46732.0 / 5247.0, t.a[6][4] = 49.0 / 176.0,
    t.a[6][5] = -5103.0 / 18656.0;
    t.a[7][1] = 35.0 / 384.0, t.a[7][2] = 0.0, t.a[7][3] = 500.0 / 1113.0, t.a[7][4] = 125.0 / 192.0,
    t.a[7][5] = -2187.0 / 6784.0, t.a[7][6] = 11.0 / 84.0;
    for (int j = 1; j <= 6; ++j) t.b[j] = t.a[7][j];
    t.b[7] = 0.0;
    t.e[1] = 71.0 / 57600.0, t.e[2] = 0.0, t.e[3] = -71.0 / 16695.0, t.e[4] = 71.0 / 1920.0,
    t.e[5] = -17253.0 / 339200.0, t.e[6] = 22.0 / 525.0, t.e[7] = -1.0 / 40.0;
    return t;
}

Tableau rkf45() { // adaptive_runge_kutta.cpp:75-110 (subtractions written as negative weights)
    Tableau t{};
    t.stages = 6;
    t.rkf_form = true;
    t.a[2][1] = 1. / 4.;
    t.a[3][1] = 3. / 32., t.a[3][2] = 9 / 32.;
    t.a[4][1] = 1932. / 2197., t.a[4][2] = -7200. / 2197., t.a[4][3] = 7296. / 2197.;
    t.a[5][1] = 439. / 216., t.a[5][2] = -8., t.a[5][3] = 3680. / 513., t.a[5][4] = -845. / 4104.;
    t.a[6][1] = -8. / 27., t.a[6][2] = 2., t.a[6][3] = -3544. / 2565., t.a[6][4] = 1859. / 4104.,
    t.a[6][5] = -11. / 40.;
    t.b[1] = 16. / 135., t.b[2] = 0.0, t.b[3] = 6656. / 12825., t.b[4] = 28561. / 56430., t.b[5] = -9. / 50.,
    t.b[6] = 2. / 55.;
    t.e[1] = 25. / 216., t.e[2] = 0.0, t.e[3] = 1408. / 2565., t.e[4] = 2197. / 4104., t.e[5] = -1. / 5.,
    t.e[6] = 0.0;
    return t;
}

// The reference's first stage of RKF45 is written `state.m + 1./4.*k1` and its stage-4/5/6 sums
// with '-' for the negative entries; x - (c k) == x + ((-c) k) exactly, so the negative weights
// reproduce its rounding.

} // namespace

// One attempt of the scheme at step dt from m; returns the error norm, sumbk in llg->sumbk.
double attempt(mafx_llg* L, const double* m, double dt, const char* scheme) {
    mafx_ctx* ctx = L->ctx;
    const Tableau T = std::string(scheme) == "DP45" ? dp45() : rkf45();
    const int64_t len = 3 * ctx->N;
    EvalArgs ea;
    ea.mode = OutMode::LLG;
    ea.alpha = L->alpha;
    ea.dt = dt;
    ea.dissipation_only = L->dissipation_only;
    // stage 1: k1 = dt f(m) (recomputed every attempt: normalize_ is true, adaptive_runge_kutta.cpp:132)
    ea.m = m;
    ea.out = L->k[1].p;
    evaluate(ctx, ea);
    L->heff_evals++;
    for (int s = 2; s <= T.stages; ++s) {
        const double* v[8];
        double c[8];
        for (int j = 1; j < s; ++j) {
            v[j - 1] = L->k[j].as<double>();
            c[j - 1] = T.a[s][j];
        }
        lincomb(ctx, L->mtmp.as<double>(), m, s - 1, v, c, len);
        ea.m = L->mtmp.as<double>();
        ea.out = L->k[s].p;
        evaluate(ctx, ea);
        L->heff_evals++;
    }
    const double* kp[8] = {nullptr};
    for (int j = 1; j <= T.stages; ++j) kp[j] = L->k[j].as<double>();
    return rk_error(ctx, m, kp, T.b, T.e, T.stages, T.rkf_form, L->ctrl.atol, L->ctrl.rtol, L->sumbk.as<double>());
}

// adaptive_runge_kutta.cpp:39-71
void llg_step(mafx_llg* L, double* m, double* t) {
    double err = 0;
    do {
        err = attempt(L, m, L->h, L->scheme.c_str());
        if (!std::isfinite(err)) throw Error(MAFX_ERR_NUMERIC, "non-finite RK error norm");
    } while (!L->ctrl.success(err, L->h));
    *t += L->h;
    L->h_last = L->h;
    L->err_last = err;
    L->h = L->ctrl.hnext;
    accept(L->ctx, m, L->sumbk.as<double>(), true, L->ctx->has_ms_field);
    L->steps++;
}

// llg_integrator.cpp:68-104: the sample at each write time comes from a backwards RKF45 step
// (no renormalisation) from the state that passed it.
int64_t llg_integrate_dense(mafx_llg* L, double* m, double* t, double time_in_s, double dt_write, double* out,
                            int64_t max_rows) {
    mafx_ctx* ctx = L->ctx;
    int64_t rows = 0;
    auto emit = [&](const double* mm, double tt) {
        double mean[3];
        mean_m(ctx, mm, mean);
        if (rows < max_rows) {
            out[4 * rows + 0] = tt;
            out[4 * rows + 1] = mean[0];
            out[4 * rows + 2] = mean[1];
            out[4 * rows + 3] = mean[2];
        }
        rows++;
    };
    const double t0 = *t;
    const double tf = t0 + time_in_s;
    emit(m, *t);
    double t_next = t0 + dt_write;
    const int64_t len = 3 * ctx->N;
    while (*t < tf) {
        llg_step(L, m, t);
        if (*t >= t_next) {
            const double back = t_next - *t;
            attempt(L, m, back, "RKF45");
            if (t_next >= tf - 1e-20) {
                accept(ctx, m, L->sumbk.as<double>(), false, false);
                *t += back;
                emit(m, *t);
                break;
            }
            copy(ctx, L->mtmp2.as<double>(), m, len);
            accept(ctx, L->mtmp2.as<double>(), L->sumbk.as<double>(), false, false);
            emit(L->mtmp2.as<double>(), *t + back);
            t_next += dt_write;
            while (t_next < *t) t_next += dt_write;
        }
    }
    return rows;
}

} // namespace mafx
