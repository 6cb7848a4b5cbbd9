// Oracle solvers: FP64 CPU restatement of the reference's adaptive Runge-Kutta LLG integrator,
// step-size controller and LBFGS minimizer.  Test infrastructure only (see oracle_internal.hpp).
#include "oracle_internal.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <limits>
#include <stdexcept>

namespace oracle {

// ------------------------------------------------------------------ controller.cpp:36-95
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
    } else if (h >= hmax) {
        h = hmax;
        counter_hmax++;
        return true;
    }
    reject = true;
    counter_reject++;
    return false;
}

// ------------------------------------------------------------------ llg_integrator.cpp:18-38
void LLG::f(const State& s, Vec& out) const {
    ++f_calls;
    const int64_t N = static_cast<int64_t>(s.m.size() / 3);
    if (analytic_rhs == 1) { // adaptive_runge_kutta test callback: t * sqrt(m)
        out.resize(s.m.size());
        for (size_t i = 0; i < s.m.size(); ++i) out[i] = s.t * std::sqrt(s.m[i]);
        return;
    }
    if (analytic_rhs == 2) { // cos(t) m
        out.resize(s.m.size());
        for (size_t i = 0; i < s.m.size(); ++i) out[i] = std::cos(s.t) * s.m[i];
        return;
    }
    Vec h;
    heff(*terms, s, h);
    const_cast<LLG*>(this)->heff_evals++;
    Vec mxh;
    cross4(s.m, h, mxh, N);
    Vec mxmxh;
    cross4(s.m, mxh, mxmxh, N);
    out.resize(3 * N);
    if (!alpha_field.empty()) { // llg_integrator.cpp:22-24,30-32: alpha tiled over the components
        for (int64_t i = 0; i < 3 * N; ++i) {
            const double al = alpha_field[i % N];
            const double a2 = std::pow(al, 2.0);
            const double damp = -al * gamma_ / (1. + a2) * mxmxh[i];
            out[i] = dissipation_term_only ? damp : -gamma_ / (1. + a2) * mxh[i] + damp;
        }
        return;
    }
    const double a2 = std::pow(alpha, 2.0);
    const double cprec = -gamma_ / (1. + a2);          // equations.hpp:19-21
    const double cdamp = -alpha * gamma_ / (1. + a2);  // equations.hpp:24-27
    for (int64_t i = 0; i < 3 * N; ++i) {
        const double damp = cdamp * mxmxh[i];
        out[i] = dissipation_term_only ? damp : cprec * mxh[i] + damp;
    }
}

namespace {
Vec axpy_sum(const Vec& base, std::initializer_list<std::pair<double, const Vec*>> terms) {
    Vec r(base);
    for (auto& t : terms)
        for (size_t i = 0; i < r.size(); ++i) r[i] = r[i] + t.first * (*t.second)[i];
    return r;
}
Vec scaled(double a, const Vec& v) {
    Vec r(v.size());
    for (size_t i = 0; i < v.size(); ++i) r[i] = a * v[i];
    return r;
}
// err = max|rk_error / (atol + rtol |max(m, m + sumbk)|)|  (adaptive_runge_kutta.cpp:108,171)
double err_norm(const Controller& c, const Vec& m, const Vec& sumbk, const Vec& rk_error) {
    double e = 0;
    for (size_t i = 0; i < m.size(); ++i) {
        const double mm = std::max(m[i], m[i] + sumbk[i]);
        e = std::max(e, std::abs(rk_error[i] / c.givescale(mm)));
    }
    return e;
}
} // namespace

// ------------------------------------------------------------------ adaptive_runge_kutta.cpp:75-110
Vec LLG::RKF45(const State& state, const double dt, double& err) const {
    State ts = state;
    Vec f1;
    f(state, f1);
    const Vec k1 = scaled(dt, f1);
    ts.t = state.t + 1. / 4. * dt;
    ts.m = axpy_sum(state.m, {{1. / 4., &k1}});
    Vec tmp;
    f(ts, tmp);
    const Vec k2 = scaled(dt, tmp);
    ts.t = state.t + 3. / 8. * dt;
    ts.m = axpy_sum(state.m, {{3. / 32., &k1}, {9 / 32., &k2}});
    f(ts, tmp);
    const Vec k3 = scaled(dt, tmp);
    ts.t = state.t + 12. / 13. * dt;
    ts.m = axpy_sum(state.m, {{1932. / 2197., &k1}, {-7200. / 2197., &k2}, {7296. / 2197., &k3}});
    f(ts, tmp);
    const Vec k4 = scaled(dt, tmp);
    ts.t = state.t + dt;
    ts.m = axpy_sum(state.m, {{439. / 216., &k1}, {-8., &k2}, {3680. / 513., &k3}, {-845. / 4104., &k4}});
    f(ts, tmp);
    const Vec k5 = scaled(dt, tmp);
    ts.t = state.t + 1. / 2. * dt;
    ts.m = axpy_sum(state.m,
                    {{-8. / 27., &k1}, {2., &k2}, {-3544. / 2565., &k3}, {1859. / 4104., &k4}, {-11. / 40., &k5}});
    f(ts, tmp);
    const Vec k6 = scaled(dt, tmp);
    const size_t n = state.m.size();
    Vec sumbk(n), rk_error(n);
    for (size_t i = 0; i < n; ++i) {
        sumbk[i] = 16. / 135. * k1[i] + 6656. / 12825. * k3[i] + 28561. / 56430. * k4[i] - 9. / 50. * k5[i] +
                   2. / 55. * k6[i];
        rk_error[i] =
            sumbk[i] - (25. / 216. * k1[i] + 1408. / 2565. * k3[i] + 2197. / 4104. * k4[i] - 1. / 5. * k5[i]);
    }
    err = err_norm(controller, state.m, sumbk, rk_error);
    return sumbk;
}

// ------------------------------------------------------------------ adaptive_runge_kutta.cpp:113-173
Vec LLG::DP45(const State& state, const double dt, double& err) {
    std::array<std::array<double, 7>, 8> a{};
    std::array<double, 8> e{};
    std::array<double, 8> c{};
    c[2] = 0.2, c[3] = 0.3, c[4] = 0.8, c[5] = 8.0 / 9.0, c[6] = 1, c[7] = 1;
    e[1] = 71.0 / 57600.0, e[3] = -71.0 / 16695.0, e[4] = 71.0 / 1920.0, e[5] = -17253.0 / 339200.0,
    e[6] = 22.0 / 525.0, e[7] = -1.0 / 40.0;
    a[2][1] = 0.2;
    a[3][1] = 3.0 / 40.0, a[3][2] = 9.0 / 40.0;
    a[4][1] = 44.0 / 45.0, a[4][2] = -56.0 / 15.0, a[4][3] = 32.0 / 9.0;
    a[5][1] = 19372.0 / 6561.0, a[5][2] = -25360.0 / 2187.0, a[5][3] = 64448.0 / 6561.0, a[5][4] = -212.0 / 729.0;
    a[6][1] = 9017.0 / 3168.0, a[6][2] = -355.0 / 33.0, a[6][3] = 46732.0 / 5247.0, a[6][4] = 49.0 / 176.0,
    a[6][5] = -5103.0 / 18656.0;
    a[7][1] = 35.0 / 384.0, a[7][3] = 500.0 / 1113.0, a[7][4] = 125.0 / 192.0, a[7][5] = -2187.0 / 6784.0,
    a[7][6] = 11.0 / 84.0;

    State ts = state;
    std::array<Vec, 8> k;
    Vec tmp;
    if (controller.reject || normalize || state.steps == 0) {
        f(ts, tmp);
        k[1] = scaled(dt, tmp);
    } else {
        k[1] = k_fsal;
    }
    for (int s = 2; s <= 7; ++s) {
        ts.t = state.t + c[s] * dt;
        ts.m = state.m;
        for (int j = 1; j < s; ++j)
            for (size_t i = 0; i < ts.m.size(); ++i) ts.m[i] = ts.m[i] + a[s][j] * k[j][i];
        f(ts, tmp);
        k[s] = scaled(dt, tmp);
    }
    k_fsal = k[7];
    const size_t n = state.m.size();
    Vec sumbk(n), rk_error(n);
    for (size_t i = 0; i < n; ++i) {
        double sb = a[7][1] * k[1][i];
        for (int j = 2; j <= 6; ++j) sb = sb + a[7][j] * k[j][i];
        sumbk[i] = sb;
        double re = e[1] * k[1][i];
        for (int j = 2; j <= 7; ++j) re = re + e[j] * k[j][i];
        rk_error[i] = re;
    }
    err = err_norm(controller, state.m, sumbk, rk_error);
    return sumbk;
}

// ------------------------------------------------------------------ adaptive_runge_kutta.cpp:496-529
// Bogacki-Shampine 3(2).  The stage values are f (not dt*f); k1 is always recomputed because
// the reference's step_calls_ counter is never advanced.  Used by the analytic-ODE tests only.
Vec LLG::BS23(const State& state, const double dt, double& err) {
    State ts = state;
    Vec k1, k2, k3;
    f(ts, k1);
    const size_t n = state.m.size();
    ts.t = state.t + 1. / 2. * dt;
    ts.m.resize(n);
    for (size_t i = 0; i < n; ++i) ts.m[i] = state.m[i] + dt * (1. / 2. * k1[i]);
    f(ts, k2);
    ts.t = state.t + 3. / 4. * dt;
    for (size_t i = 0; i < n; ++i) ts.m[i] = state.m[i] + dt * 3. / 4. * k2[i];
    f(ts, k3);
    Vec sumbk(n);
    for (size_t i = 0; i < n; ++i) sumbk[i] = dt * (2. / 9. * k1[i] + 1. / 3. * k2[i] + 4. / 9. * k3[i]);
    ts.t = state.t + dt;
    for (size_t i = 0; i < n; ++i) ts.m[i] = state.m[i] + sumbk[i];
    f(ts, k_fsal);
    Vec rk_error(n);
    for (size_t i = 0; i < n; ++i)
        rk_error[i] = sumbk[i] - dt * (7. / 24. * k1[i] + 1. / 4. * k2[i] + 1. / 3. * k3[i] + 1. / 8. * k_fsal[i]);
    err = err_norm(controller, state.m, sumbk, rk_error);
    return sumbk;
}

Vec LLG::tableau(const State& state, const double dt, double& err, const int ns, const double (*a)[14],
                 const double* c, const double* bs, const double* be) const {
    const size_t n = state.m.size();
    State ts = state;
    std::vector<Vec> k(ns + 1);
    Vec tmp;
    for (int s = 1; s <= ns; ++s) {
        ts.t = state.t + c[s] * dt;
        ts.m = state.m;
        for (int j = 1; j < s; ++j)
            for (size_t i = 0; i < n; ++i) ts.m[i] = ts.m[i] + a[s][j] * k[j][i];
        f(ts, tmp);
        k[s] = scaled(dt, tmp);
    }
    Vec sumbk(n), rk_error(n);
    for (size_t i = 0; i < n; ++i) {
        double sb = bs[1] * k[1][i], se = be[1] * k[1][i];
        for (int j = 2; j <= ns; ++j) {
            sb = sb + bs[j] * k[j][i]; // zero weights add +0 and leave the sum unchanged
            se = se + be[j] * k[j][i];
        }
        sumbk[i] = sb;
        rk_error[i] = sb - se;
    }
    err = err_norm(controller, state.m, sumbk, rk_error);
    return sumbk;
}

// ------------------------------------------------------------------ adaptive_runge_kutta.cpp:176-282
// Bogacki-Shampine 5(4), eight stages, the last one at m + sumbk.  k1 is always recomputed (the
// reference reuses k_FSAL only when its step_calls_ counter is non-zero, and it is never advanced).
Vec LLG::BS45(const State& state, const double dt, double& err) {
    double a[14][14] = {};
    double b[14] = {}, c[14] = {}, bs[14] = {};
    a[2][1] = 1.0 / 6.0;
    a[3][1] = 2.0 / 27.0, a[3][2] = 4.0 / 27.0;
    a[4][1] = 183.0 / 1372.0, a[4][2] = -162.0 / 343.0, a[4][3] = 1053.0 / 1372.0;
    a[5][1] = 68.0 / 297.0, a[5][2] = -4.0 / 11.0, a[5][3] = 42.0 / 143.0, a[5][4] = 1960.0 / 3861.0;
    a[6][1] = 597.0 / 22528.0, a[6][2] = 81.0 / 352.0, a[6][3] = 63099.0 / 585728.0, a[6][4] = 58653.0 / 366080.0,
    a[6][5] = 4617.0 / 20480.0;
    a[7][1] = 174197.0 / 959244.0, a[7][2] = -30942.0 / 79937.0, a[7][3] = 8152137.0 / 19744439.0,
    a[7][4] = 666106.0 / 1039181.0, a[7][5] = -29421.0 / 29068.0, a[7][6] = 482048.0 / 414219.0;
    a[8][1] = 587.0 / 8064.0, a[8][2] = 0.0, a[8][3] = 4440339.0 / 15491840.0, a[8][4] = 24353.0 / 124800.0,
    a[8][5] = 387.0 / 44800.0, a[8][6] = 2152.0 / 5985.0, a[8][7] = 7267.0 / 94080.0;
    c[2] = 1.0 / 6.0, c[3] = 2.0 / 9.0, c[4] = 3.0 / 7.0, c[5] = 2.0 / 3.0, c[6] = 3.0 / 4.0, c[7] = 1.0, c[8] = 1.0;
    b[1] = 2479.0 / 34992.0, b[2] = 0.0, b[3] = 123.0 / 416.0, b[4] = 612941.0 / 3411720.0, b[5] = 43.0 / 1440.0,
    b[6] = 2272.0 / 6561.0, b[7] = 79937.0 / 1113912.0, b[8] = 3293.0 / 556956.0;
    for (int j = 1; j <= 7; ++j) bs[j] = a[8][j];
    // the reference's sumbk runs over k1..k7; bs[8] = 0 adds 0 * k8, which leaves it unchanged
    return tableau(state, dt, err, 8, a, c, bs, b);
}

// ------------------------------------------------------------------ adaptive_runge_kutta.cpp:285-493
// Dormand-Prince 8(7), thirteen stages; advances with the order-8 weights bhat.
Vec LLG::DP78(const State& state, const double dt, double& err) const {
    double a[14][14] = {};
    double b[14] = {}, bh[14] = {}, c[14] = {};
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
    bh[1] = 4.17474911415302462220859284685e-2, bh[6] = -5.54523286112393089615218946547e-2,
    bh[7] = 2.39312807201180097046747354249e-1, bh[8] = 7.0351066940344302305804641089e-1,
    bh[9] = -7.59759613814460929884487677085e-1, bh[10] = 6.60563030922286341461378594838e-1,
    bh[11] = 1.58187482510123335529614838601e-1, bh[12] = -2.38109538752862804471863555306e-1, bh[13] = 2.5e-1;
    b[1] = 2.9553213676353496981964883112e-2, b[6] = -8.28606276487797039766805612689e-1,
    b[7] = 3.11240900051118327929913751627e-1, b[8] = 2.46734519059988698196468570407e0,
    b[9] = -2.54694165184190873912738007542e0, b[10] = 1.44354858367677524030187495069e0,
    b[11] = 7.94155958811272872713019541622e-2, b[12] = 4.44444444444444444444444444445e-2;
    c[2] = 5.55555555555555555555555555556e-2, c[3] = 8.33333333333333333333333333334e-2, c[4] = 1.25e-1,
    c[5] = 3.125e-1, c[6] = 3.75e-1, c[7] = 1.475e-1, c[8] = 4.65e-1, c[9] = 5.64865451382259575398358501426e-1,
    c[10] = 6.5e-1, c[11] = 9.24656277640504446745013574318e-1, c[12] = 1.0, c[13] = 1.0;
    return tableau(state, dt, err, 13, a, c, bh, b);
}

// ------------------------------------------------------------------ adaptive_runge_kutta.cpp:39-71
void LLG::step(State& s) {
    Vec mtemp;
    double err = 0;
    do {
        if (scheme == "RKF45")
            mtemp = RKF45(s, h, err);
        else if (scheme == "DP45")
            mtemp = DP45(s, h, err);
        else if (scheme == "BS23")
            mtemp = BS23(s, h, err);
        else if (scheme == "BS45")
            mtemp = BS45(s, h, err);
        else if (scheme == "DP78")
            mtemp = DP78(s, h, err);
        else
            throw std::runtime_error("oracle: unsupported scheme " + scheme);
    } while (!controller.success(err, h));
    s.t += h;
    h_last = h;
    err_last = err;
    h = controller.hnext;
    for (size_t i = 0; i < s.m.size(); ++i) s.m[i] += mtemp[i];
    if (normalize) {
        const int64_t N = static_cast<int64_t>(s.m.size() / 3);
        if (!s.has_Ms_field())
            normalize_even_zero_vectors(s.m, N);
        else
            normalize_handle_zero_vectors(s.m, N);
    }
    s.steps++;
}

// ------------------------------------------------------------------ llg_integrator.cpp:68-104
std::vector<std::array<double, 4>> LLG::integrate_dense(State& s, double time, double dt_write) {
    std::vector<std::array<double, 4>> rows;
    auto emit = [&](const State& st) {
        double mm[3];
        mean_m(st, mm);
        rows.push_back({st.t, mm[0], mm[1], mm[2]});
    };
    const double t0 = s.t;
    const double tf = t0 + time;
    emit(s);
    double t_next = t0 + dt_write;
    while (s.t < tf) {
        step(s);
        if (s.t >= t_next) {
            double dummy = 0;
            const double back = t_next - s.t;
            if (t_next >= tf - 1e-20) {
                Vec dm = RKF45(s, back, dummy);
                for (size_t i = 0; i < s.m.size(); ++i) s.m[i] += dm[i];
                s.t += back;
                emit(s);
                break;
            } else {
                State cur = s;
                Vec dm = RKF45(cur, back, dummy);
                for (size_t i = 0; i < cur.m.size(); ++i) cur.m[i] += dm[i];
                cur.t += back;
                emit(cur);
            }
            t_next += dt_write;
            while (t_next < s.t) t_next += dt_write;
        }
    }
    return rows;
}

// ------------------------------------------------------------------ lbfgs_minimizer.cpp
namespace {

// m x (m x H) (lbfgs_minimizer.cpp:33-36)
Vec gradient(const State& s, const Vec& h) {
    const int64_t N = s.mesh.n();
    Vec a, b;
    cross4(s.m, h, a, N);
    cross4(s.m, a, b, N);
    return b;
}

// max|m x H| / max|H| (build-side torque criterion; the reference has none)
double torque_rel(const State& s, const Vec& h) {
    Vec mxh;
    cross4(s.m, h, mxh, s.mesh.n());
    const double hm = max_4d_abs(h);
    return hm > 0 ? max_4d_abs(mxh) / hm : 0.0;
}

struct EvalCtx {
    const Terms* terms;
    int64_t evals{0};
    Vec last_h;
    std::pair<double, Vec> energy_and_gradient(const State& s) { // :38-42
        ++evals;
        const double e = heff_and_energy(*terms, s, last_h);
        return {e, gradient(s, last_h)};
    }
};

double mydot(const Vec& a, const Vec& b) {
    double s = 0;
    const size_t n = a.size() / 3;
    // af::sum over x, y, z then component: per-component partial sums then total
    for (int c = 0; c < 3; ++c) {
        double sc = 0;
        for (size_t i = 0; i < n; ++i) sc += a[c * n + i] * b[c * n + i];
        s += sc;
    }
    return s;
}
double mynorm(const Vec& a) { return std::sqrt(mydot(a, a)); }

// :362-505
int cstep(double& stx, double& fx, double& dx, double& sty, double& fy, double& dy, double& stp, double& fp,
          double& dp, bool& brackt, double& stpmin, double& stpmax, int& info) {
    info = 0;
    bool bound = false;
    if ((brackt & ((stp <= std::min(stx, sty)) | (stp >= std::max(stx, sty)))) | (dx * (stp - stx) >= 0.0) |
        (stpmax < stpmin))
        return -1;
    const double sgnd = dp * (dx / std::fabs(dx));
    double stpf = 0, stpc = 0, stpq = 0;
    if (fp > fx) {
        info = 1;
        bound = true;
        const double theta = 3. * (fx - fp) / (stp - stx) + dx + dp;
        const double s = std::max(theta, std::max(dx, dp));
        double gamma = s * std::sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
        if (stp < stx) gamma = -gamma;
        const double p = (gamma - dx) + theta;
        const double q = ((gamma - dx) + gamma) + dp;
        const double r = p / q;
        stpc = stx + r * (stp - stx);
        stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2.) * (stp - stx);
        stpf = std::fabs(stpc - stx) < std::fabs(stpq - stx) ? stpc : stpc + (stpq - stpc) / 2;
        brackt = true;
    } else if (sgnd < 0.0) {
        info = 2;
        bound = false;
        const double theta = 3 * (fx - fp) / (stp - stx) + dx + dp;
        const double s = std::max(theta, std::max(dx, dp));
        double gamma = s * std::sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
        if (stp > stx) gamma = -gamma;
        const double p = (gamma - dp) + theta;
        const double q = ((gamma - dp) + gamma) + dx;
        const double r = p / q;
        stpc = stp + r * (stx - stp);
        stpq = stp + (dp / (dp - dx)) * (stx - stp);
        stpf = std::fabs(stpc - stp) > std::fabs(stpq - stp) ? stpc : stpq;
        brackt = true;
    } else if (std::fabs(dp) < std::fabs(dx)) {
        info = 3;
        bound = true;
        const double theta = 3 * (fx - fp) / (stp - stx) + dx + dp;
        const double s = std::max(theta, std::max(dx, dp));
        double gamma = s * std::sqrt(std::max(0., (theta / s) * (theta / s) - (dx / s) * (dp / s)));
        if (stp > stx) gamma = -gamma;
        const double p = (gamma - dp) + theta;
        const double q = (gamma + (dx - dp)) + gamma;
        const double r = p / q;
        if ((r < 0.0) & (gamma != 0.0))
            stpc = stp + r * (stx - stp);
        else if (stp > stx)
            stpc = stpmax;
        else
            stpc = stpmin;
        stpq = stp + (dp / (dp - dx)) * (stx - stp);
        if (brackt)
            stpf = std::fabs(stp - stpc) < std::fabs(stp - stpq) ? stpc : stpq;
        else
            stpf = std::fabs(stp - stpc) > std::fabs(stp - stpq) ? stpc : stpq;
    } else {
        info = 4;
        bound = false;
        if (brackt) {
            const double theta = 3 * (fp - fy) / (sty - stp) + dy + dp;
            const double s = std::max(theta, std::max(dy, dp));
            double gamma = s * std::sqrt((theta / s) * (theta / s) - (dy / s) * (dp / s));
            if (stp > sty) gamma = -gamma;
            const double p = (gamma - dp) + theta;
            const double q = ((gamma - dp) + gamma) + dy;
            const double r = p / q;
            stpc = stp + r * (sty - stp);
            stpf = stpc;
        } else if (stp > stx) {
            stpf = stpmax;
        } else {
            stpf = stpmin;
        }
    }
    if (fp > fx) {
        sty = stp;
        fy = fp;
        dy = dp;
    } else {
        if (sgnd < 0.0) {
            sty = stx;
            fy = fx;
            dy = dx;
        }
        stx = stp;
        fx = fp;
        dx = dp;
    }
    stpf = std::min(stpmax, stpf);
    stpf = std::max(stpmin, stpf);
    stp = stpf;
    if (brackt & bound) {
        if (sty > stx)
            stp = std::min(stx + 0.66 * (sty - stx), stp);
        else
            stp = std::max(stx + 0.66 * (sty - stx), stp);
    }
    return 0;
}

// :213-360
int cvsrch(EvalCtx& ev, State& state, const Vec& wa, double& f, Vec& g, double& stp, const Vec& s,
           const double tolf, LbfgsTraceRow* trace) {
    int info = 0;
    int infoc = 1;
    const double xtol = 1e-15;
    const double ftol = 1.0e-4;
    const double gtol = 0.9;
    const double eps = tolf;
    const double stpmin = 1e-15;
    const double stpmax = 1e15;
    const double xtrapf = 4;
    const int maxfev = 20;
    int nfev = 0;
    const double dginit = mydot(g, s);
    if (dginit >= 0.0) return -1;
    bool brackt = false;
    bool stage1 = true;
    const double finit = f;
    const double dgtest = ftol * dginit;
    double width = stpmax - stpmin;
    double width1 = 2 * width;
    double stx = 0.0, fx = finit, dgx = dginit;
    double sty = 0.0, fy = finit, dgy = dginit;
    double stmin = std::numeric_limits<double>::quiet_NaN();
    double stmax = std::numeric_limits<double>::quiet_NaN();
    const int64_t N = state.mesh.n();
    while (true) {
        if (brackt) {
            stmin = std::min(stx, sty);
            stmax = std::max(stx, sty);
        } else {
            stmin = stx;
            stmax = stp + xtrapf * (stp - stx);
        }
        stp = std::max(stp, stpmin);
        stp = std::min(stp, stpmax);
        if ((brackt && ((stp <= stmin) | (stp >= stmax))) | (nfev >= maxfev - 1) | (infoc == 0) |
            (brackt & (stmax - stmin <= xtol * stmax)))
            stp = stx;
        for (size_t i = 0; i < state.m.size(); ++i) state.m[i] = wa[i] + stp * s[i];
        normalize_handle_zero_vectors(state.m, N);
        std::tie(f, g) = ev.energy_and_gradient(state);
        nfev++;
        const double dg = mydot(g, s);
        const double ftest1 = finit + stp * dgtest;
        const double ftest2 = finit + eps * std::fabs(finit);
        const double ft = 2 * ftol - 1;
        if ((brackt & ((stp <= stmin) | (stp >= stmax))) | (infoc == 0)) info = 6;
        if ((stp == stpmax) & (f <= ftest2) & (dg <= dgtest)) info = 5;
        if ((stp == stpmin) & ((f > ftest2) | (dg >= dgtest))) info = 4;
        if (nfev >= maxfev) info = 3;
        if (brackt & (stmax - stmin <= xtol * stmax)) info = 2;
        if ((f <= ftest1) & (std::fabs(dg) <= gtol * (-dginit))) info = 1;
        if (((f <= ftest2) && (ft * dginit >= dg)) && (std::fabs(dg) <= gtol * (-dginit))) info = 1;
        if (info != 0) {
            if (trace) *trace = LbfgsTraceRow{finit, dginit, stp, f, dg, static_cast<double>(info),
                                              static_cast<double>(nfev)};
            return -1;
        }
        if (stage1 & ((f <= ftest2) && (ft * dginit >= dg)) & (dg >= std::min(ftol, gtol) * dginit)) stage1 = false;
        if (stage1 & (f <= fx) & (!((f <= ftest2) && (ft * dginit >= dg)))) {
            double fm = f - stp * dgtest;
            double fxm = fx - stx * dgtest;
            double fym = fy - sty * dgtest;
            double dgm = dg - dgtest;
            double dgxm = dgx - dgtest;
            double dgym = dgy - dgtest;
            cstep(stx, fxm, dgxm, sty, fym, dgym, stp, fm, dgm, brackt, stmin, stmax, infoc);
            fx = fxm + stx * dgtest;
            fy = fym + sty * dgtest;
            dgx = dgxm + dgtest;
            dgy = dgym + dgtest;
        } else {
            double dgv = dg;
            cstep(stx, fx, dgx, sty, fy, dgy, stp, f, dgv, brackt, stmin, stmax, infoc);
        }
        if (brackt) {
            if (std::fabs(sty - stx) >= 0.66 * width1) stp = stx + 0.5 * (sty - stx);
            width1 = width;
            width = std::fabs(sty - stx);
        }
    }
}

} // namespace

// :55-198
LbfgsResult lbfgs_minimize(const Terms& terms, State& state, double tolerance, int64_t maxIter, int verbose,
                           double torque_tol, std::vector<LbfgsTraceRow>* trace) {
    (void)verbose;
    LbfgsResult res;
    EvalCtx ev{&terms, 0, {}};
    const double eps = 2.22e-16;
    const double eps2 = std::sqrt(eps);
    const double epsr = std::pow(eps, 0.9);
    const double tolf2 = std::sqrt(tolerance);
    const double tolf3 = std::pow(tolerance, 0.3333333333333333333333333);
    auto [f, grad] = ev.energy_and_gradient(state);
    double gradNorm = max_4d_abs(grad);
    if (gradNorm < (epsr * (1 + std::fabs(f)))) {
        res.E = f;
        res.evals = ev.evals;
        res.stop_reason = 1;
        return res;
    }
    const size_t m = 5;
    const size_t n = state.m.size();
    std::array<Vec, m> sV, yV;
    for (auto& v : sV) v.assign(n, 0.0);
    for (auto& v : yV) v.assign(n, 0.0);
    std::array<double, m> alpha{};
    Vec q(n), s(n), y(n), grad_old(n), x_old = state.m;
    double f_old = 0.0;
    size_t iter = 0, globIter = 0;
    double H0k = 1;
    do {
        f_old = f;
        x_old = state.m;
        grad_old = grad;
        q = grad;
        const int k = static_cast<int>(std::min(m, iter));
        for (int i = k - 1; i >= 0; i--) {
            const double rho = 1.0 / mydot(sV[i], yV[i]);
            alpha[i] = rho * mydot(sV[i], q);
            for (size_t j = 0; j < n; ++j) q[j] -= alpha[i] * yV[i][j];
        }
        for (size_t j = 0; j < n; ++j) q[j] = H0k * q[j];
        for (int i = 0; i < k; i++) {
            const double rho = 1.0 / mydot(sV[i], yV[i]);
            const double beta = rho * mydot(yV[i], q);
            for (size_t j = 0; j < n; ++j) q[j] += sV[i][j] * (alpha[i] - beta);
        }
        const double phiPrime0 = -mydot(grad, q);
        if (phiPrime0 > 0) {
            q = grad;
            iter = 0;
        }
        Vec dir(n);
        for (size_t j = 0; j < n; ++j) dir[j] = -q[j];
        double rate = 1.0;
        LbfgsTraceRow row{};
        cvsrch(ev, state, x_old, f, grad, rate, dir, tolerance, trace ? &row : nullptr);
        if (trace) trace->push_back(row);
        const double f1 = 1 + std::fabs(f);
        gradNorm = max_4d_abs(grad);
        if (gradNorm < (epsr * f1)) {
            res.stop_reason = 1;
            globIter++;
            break;
        }
        for (size_t j = 0; j < n; ++j) s[j] = state.m[j] - x_old[j];
        if (((f_old - f) < (tolerance * f1)) && (max_4d_abs(s) < (tolf2 * (1 + max_4d_abs(state.m)))) &&
            (gradNorm <= (tolf3 * f1))) {
            res.stop_reason = 2;
            globIter++;
            break;
        }
        if (torque_tol > 0 && torque_rel(state, ev.last_h) <= torque_tol) {
            res.stop_reason = 3;
            globIter++;
            break;
        }
        for (size_t j = 0; j < n; ++j) y[j] = grad[j] - grad_old[j];
        const double ys = mydot(y, s);
        if (ys <= eps2 * mynorm(y) * mynorm(s)) {
            // Dennis-Schnabel 9.4.1: skip the update
        } else {
            if (iter < m) {
                sV[iter] = s;
                yV[iter] = y;
            } else {
                std::rotate(sV.begin(), sV.begin() + 1, sV.end());
                sV[m - 1] = s;
                std::rotate(yV.begin(), yV.begin() + 1, yV.end());
                yV[m - 1] = y;
            }
            H0k = ys / mydot(y, y);
            iter++;
        }
        globIter++;
    } while (globIter < static_cast<size_t>(maxIter));
    res.E = f;
    res.iterations = static_cast<int64_t>(globIter);
    res.evals = ev.evals;
    return res;
}

} // namespace oracle
