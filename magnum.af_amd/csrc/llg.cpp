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
    int stages;                             // number of f evaluations
    double a[kMaxStages + 1][kMaxStages + 1]; // a[s][j], s = 2..stages
    double c[kMaxStages + 1];               // stage times t + c[s] h (the reference's tempstate.t)
    double b[kMaxStages + 1];               // weights of the propagated solution (sumbk)
    double e[kMaxStages + 1];               // error weights
    bool rkf_form;                          // rk_error = sumbk - sum e_j k_j
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
    t.c[2] = 0.2, t.c[3] = 0.3, t.c[4] = 0.8, t.c[5] = 8.0 / 9.0, t.c[6] = 1, t.c[7] = 1;
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
    t.c[2] = 1. / 4., t.c[3] = 3. / 8., t.c[4] = 12. / 13., t.c[5] = 1., t.c[6] = 1. / 2.;
    return t;
}

// The reference's first stage of RKF45 is written `state.m + 1./4.*k1` and its stage-4/5/6 sums
// with '-' for the negative entries; x - (c k) == x + ((-c) k) exactly, so the negative weights
// reproduce its rounding.

Tableau bs23() { // adaptive_runge_kutta.cpp:496-529
    // The reference keeps f (not dt f) in its stages and forms m + dt (1/2 k1), m + dt 3/4 k2,
    // sumbk = dt (2/9 k1 + 1/3 k2 + 4/9 k3) and rk_error = sumbk - dt (7/24 k1 + ... + 1/8 k4).
    // Here every stage is k = dt f like the other schemes: the same values up to rounding.
    Tableau t{};
    t.stages = 4;
    t.rkf_form = true;
    t.a[2][1] = 1. / 2.;
    t.a[3][2] = 3. / 4.;
    t.a[4][1] = 2. / 9., t.a[4][2] = 1. / 3., t.a[4][3] = 4. / 9.;
    t.b[1] = 2. / 9., t.b[2] = 1. / 3., t.b[3] = 4. / 9.;
    t.e[1] = 7. / 24., t.e[2] = 1. / 4., t.e[3] = 1. / 3., t.e[4] = 1. / 8.;
    t.c[2] = 1. / 2., t.c[3] = 3. / 4., t.c[4] = 1.;
    return t;
}

Tableau bs45() { // adaptive_runge_kutta.cpp:176-282 (eight stages, the last one at m + sumbk)
    Tableau t{};
    t.stages = 8;
    t.rkf_form = true;
    t.a[2][1] = 1.0 / 6.0;
    t.a[3][1] = 2.0 / 27.0, t.a[3][2] = 4.0 / 27.0;
    t.a[4][1] = 183.0 / 1372.0, t.a[4][2] = -162.0 / 343.0, t.a[4][3] = 1053.0 / 1372.0;
    t.a[5][1] = 68.0 / 297.0, t.a[5][2] = -4.0 / 11.0, t.a[5][3] = 42.0 / 143.0, t.a[5][4] = 1960.0 / 3861.0;
    t.a[6][1] = 597.0 / 22528.0, t.a[6][2] = 81.0 / 352.0, t.a[6][3] = 63099.0 / 585728.0,
    t.a[6][4] = 58653.0 / 366080.0, t.a[6][5] = 4617.0 / 20480.0;
    t.a[7][1] = 174197.0 / 959244.0, t.a[7][2] = -30942.0 / 79937.0, t.a[7][3] = 8152137.0 / 19744439.0,
    t.a[7][4] = 666106.0 / 1039181.0, t.a[7][5] = -29421.0 / 29068.0, t.a[7][6] = 482048.0 / 414219.0;
    t.a[8][1] = 587.0 / 8064.0, t.a[8][3] = 4440339.0 / 15491840.0, t.a[8][4] = 24353.0 / 124800.0,
    t.a[8][5] = 387.0 / 44800.0, t.a[8][6] = 2152.0 / 5985.0, t.a[8][7] = 7267.0 / 94080.0;
    for (int j = 1; j <= 7; ++j) t.b[j] = t.a[8][j];
    t.e[1] = 2479.0 / 34992.0, t.e[3] = 123.0 / 416.0, t.e[4] = 612941.0 / 3411720.0, t.e[5] = 43.0 / 1440.0,
    t.e[6] = 2272.0 / 6561.0, t.e[7] = 79937.0 / 1113912.0, t.e[8] = 3293.0 / 556956.0;
    t.c[2] = 1.0 / 6.0, t.c[3] = 2.0 / 9.0, t.c[4] = 3.0 / 7.0, t.c[5] = 2.0 / 3.0, t.c[6] = 3.0 / 4.0, t.c[7] = 1.0,
    t.c[8] = 1.0;
    return t;
}

Tableau dp78() { // adaptive_runge_kutta.cpp:285-493 (advances with the order-8 weights bhat)
    Tableau t{};
    t.stages = 13;
    t.rkf_form = true;
    auto& a = t.a;
    a[2][1] = 5.55555555555555555555555555556e-2;
    a[3][1] = 2.08333333333333333333333333333e-2, a[3][2] = 6.25e-2;
    a[4][1] = 3.125e-2, a[4][3] = 9.375e-2;
    a[5][1] = 3.125e-1, a[5][3] = -1.171875e0, a[5][4] = 1.171875e0;
    a[6][1] = 3.75e-2, a[6][4] = 1.875e-1, a[6][5] = 1.5e-1;
    a[7][1] = 4.79101371111111111111111111111e-2, a[7][4] = 1.12248712777777777777777777778e-1,
    a[7][5] = -2.55056737777777777777777777778e-2, a[7][6] = 1.28468238888888888888888888889e-2;
    a[8][1] = 1.6917989787292281181431107136e-2, a[8][4] = 3.87848278486043169526545744159e-1,
    a[8][5] = 3.59773698515003278967008896348e-2, a[8][6] = 1.96970214215666060156715256072e-1,
    a[8][7] = -1.72713852340501838761392997002e-1;
    a[9][1] = 6.90957533591923006485645489846e-2, a[9][4] = -6.34247976728854151882807874972e-1,
    a[9][5] = -1.61197575224604080366876923982e-1, a[9][6] = 1.38650309458825255419866950133e-1,
    a[9][7] = 9.4092861403575626972423968413e-1, a[9][8] = 2.11636326481943981855372117132e-1;
    a[10][1] = 1.83556996839045385489806023537e-1, a[10][4] = -2.46876808431559245274431575997e0,
    a[10][5] = -2.91286887816300456388002572804e-1, a[10][6] = -2.6473020233117375688439799466e-2,
    a[10][7] = 2.84783876419280044916451825422e0, a[10][8] = 2.81387331469849792539403641827e-1,
    a[10][9] = 1.23744899863314657627030212664e-1;
    a[11][1] = -1.21542481739588805916051052503e0, a[11][4] = 1.66726086659457724322804132886e1,
    a[11][5] = 9.15741828416817960595718650451e-1, a[11][6] = -6.05660580435747094755450554309e0,
    a[11][7] = -1.60035735941561781118417064101e1, a[11][8] = 1.4849303086297662557545391898e1,
    a[11][9] = -1.33715757352898493182930413962e1, a[11][10] = 5.13418264817963793317325361166e0;
    a[12][1] = 2.58860916438264283815730932232e-1, a[12][4] = -4.77448578548920511231011750971e0,
    a[12][5] = -4.3509301377703250944070041181e-1, a[12][6] = -3.04948333207224150956051286631e0,
    a[12][7] = 5.57792003993609911742367663447e0, a[12][8] = 6.15583158986104009733868912669e0,
    a[12][9] = -5.06210458673693837007740643391e0, a[12][10] = 2.19392617318067906127491429047e0,
    a[12][11] = 1.34627998659334941535726237887e-1;
    a[13][1] = 8.22427599626507477963168204773e-1, a[13][4] = -1.16586732572776642839765530355e1,
    a[13][5] = -7.57622116690936195881116154088e-1, a[13][6] = 7.13973588159581527978269282765e-1,
    a[13][7] = 1.20757749868900567395661704486e1, a[13][8] = -2.12765911392040265639082085897e0,
    a[13][9] = 1.99016620704895541832807169835e0, a[13][10] = -2.34286471544040292660294691857e-1,
    a[13][11] = 1.7589857770794226507310510589e-1;
    t.b[1] = 4.17474911415302462220859284685e-2, t.b[6] = -5.54523286112393089615218946547e-2,
    t.b[7] = 2.39312807201180097046747354249e-1, t.b[8] = 7.0351066940344302305804641089e-1,
    t.b[9] = -7.59759613814460929884487677085e-1, t.b[10] = 6.60563030922286341461378594838e-1,
    t.b[11] = 1.58187482510123335529614838601e-1, t.b[12] = -2.38109538752862804471863555306e-1, t.b[13] = 2.5e-1;
    t.e[1] = 2.9553213676353496981964883112e-2, t.e[6] = -8.28606276487797039766805612689e-1,
    t.e[7] = 3.11240900051118327929913751627e-1, t.e[8] = 2.46734519059988698196468570407e0,
    t.e[9] = -2.54694165184190873912738007542e0, t.e[10] = 1.44354858367677524030187495069e0,
    t.e[11] = 7.94155958811272872713019541622e-2, t.e[12] = 4.44444444444444444444444444445e-2;
    t.c[2] = 5.55555555555555555555555555556e-2, t.c[3] = 8.33333333333333333333333333334e-2, t.c[4] = 1.25e-1,
    t.c[5] = 3.125e-1, t.c[6] = 3.75e-1, t.c[7] = 1.475e-1, t.c[8] = 4.65e-1, t.c[9] = 5.64865451382259575398358501426e-1,
    t.c[10] = 6.5e-1, t.c[11] = 9.24656277640504446745013574318e-1, t.c[12] = 1.0, t.c[13] = 1.0;
    return t;
}

Tableau tableau_of(const std::string& s) {
    if (s == "DP45") return dp45();
    if (s == "BS23") return bs23();
    if (s == "BS45") return bs45();
    if (s == "DP78") return dp78();
    return rkf45();
}

} // namespace

int scheme_stages(const char* scheme) { return tableau_of(scheme).stages; }

// One attempt of the scheme at step dt from (m, t); returns the error norm, sumbk in llg->sumbk.
double attempt(mafx_llg* L, const double* m, double t, double dt, const char* scheme) {
    mafx_ctx* ctx = L->ctx;
    const Tableau T = tableau_of(scheme);
    const int64_t len = 3 * ctx->N;
    EvalArgs ea;
    ea.mode = OutMode::LLG;
    ea.alpha = L->alpha;
    ea.alpha_field = L->has_alpha_field ? L->alpha_field.as<double>() : nullptr;
    ea.dt = dt;
    ea.dissipation_only = L->dissipation_only;
    // The FP32-field path stores its stage derivatives k_j = dt f(m_j) in FP32: they carry the
    // FP32 field's ~1e-6 relative error already, and every sum over them (stage inputs, sumbk, the
    // error norm) stays FP64 on the FP64 state.  Halves the k_j traffic of every stage.
    const bool k32 = ctx->precision == MAFX_F32_FIELD && ctx->rk_k32;
    ea.k_f32 = k32;
    // Schemes of up to kMaxRkFused + 1 stages form each stage input m + sum_j a_sj k_j in the
    // per-cell pass that produces k_(s-1) (the stage combination of adaptive_runge_kutta.cpp:
    // 138-164 with k_(s-1) still in registers, so no pass re-reads it); the inputs alternate
    // between two buffers, since a pass still reads its own input's neighbours.  Results are the
    // separate pass's to the last bit.  DP78, and meshes below 2^20 cells (where the per-cell
    // pass has too few lanes to hide the extra loads: 256 x 128 x 4 runs 8% slower fused), form
    // them in a pass of their own.
    const bool fused = T.stages <= kMaxRkFused + 1 && ctx->N >= ctx->rk_fuse_min_cells;
    double* stage_in[2] = {L->mtmp.as<double>(), L->mtmp2.as<double>()};
    auto next_input = [&](int s) { // s: the stage whose input the pass being set up writes
        ea.rk_out = nullptr;
        if (!fused || s > T.stages) return;
        ea.rk_base = m;
        ea.rk_n = s - 2;
        for (int j = 1; j < s - 1; ++j) ea.rk_k[j - 1] = L->k[j].p;
        for (int j = 1; j < s; ++j) ea.rk_c[j - 1] = T.a[s][j];
        ea.rk_out = stage_in[s & 1];
    };
    // stage 1: k1 = dt f(m) (recomputed every attempt: normalize_ is true, adaptive_runge_kutta.cpp:132)
    ea.m = m;
    ea.out = L->k[1].p;
    ctx->time = t;
    next_input(2);
    evaluate(ctx, ea);
    L->heff_evals++;
    for (int s = 2; s <= T.stages; ++s) {
        if (!fused) {
            const void* v[kMaxStages];
            double c[kMaxStages];
            for (int j = 1; j < s; ++j) {
                v[j - 1] = L->k[j].p;
                c[j - 1] = T.a[s][j];
            }
            lincomb(ctx, stage_in[s & 1], m, s - 1, v, c, len, k32);
        }
        ea.m = stage_in[s & 1];
        ea.out = L->k[s].p;
        ctx->time = t + T.c[s] * dt;
        next_input(s + 1);
        evaluate(ctx, ea);
        L->heff_evals++;
    }
    const void* kp[kMaxStages + 1] = {nullptr};
    for (int j = 1; j <= T.stages; ++j) kp[j] = L->k[j].p;
    return rk_error(ctx, m, kp, T.b, T.e, T.stages, T.rkf_form, L->ctrl.atol, L->ctrl.rtol, L->sumbk.as<double>(), k32);
}

// adaptive_runge_kutta.cpp:39-71
void llg_step(mafx_llg* L, double* m, double* t) {
    double err = 0;
    do {
        err = attempt(L, m, *t, L->h, L->scheme.c_str());
        if (!std::isfinite(err)) throw Error(MAFX_ERR_NUMERIC, "non-finite RK error norm");
    } while (!L->ctrl.success(err, L->h));
    *t += L->h;
    L->ctx->time = *t;
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
            attempt(L, m, *t, back, "RKF45");
            ctx->time = *t;
            if (t_next >= tf - 1e-20) {
                accept(ctx, m, L->sumbk.as<double>(), false, false);
                *t += back;
                ctx->time = *t;
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
