// LBFGS minimizer on the GPU: L-BFGS (history 5) with the More-Thuente line search on the sphere.
// Control flow and scalar arithmetic run on the host exactly as the reference's; every 3N-sized
// vector lives on the device (gradient m x (m x H) and the energy come from one fused field pass).
//
// Replaces magnumaf/src/solvers/lbfgs_minimizer.cpp:33-505.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <limits>

#include "mafx_internal.hpp"
#include "mafx_llg.hpp"
#include "vec_kernels.hpp"

namespace mafx {

namespace {

struct Lbfgs {
    mafx_ctx* ctx;
    int64_t n3;
    int64_t evals{0};
    double last_torque{0}, last_hmax{0};

    // EnergyAndGradient (lbfgs_minimizer.cpp:38-42): one field pass writes g and the energy partials
    double energy_and_gradient(const double* m, double* g) {
        EvalArgs ea;
        ea.m = m;
        ea.mode = OutMode::GRAD;
        ea.out = g;
        ea.want_energy = true;
        evaluate(ctx, ea);
        double E = 0;
        finish_energy(ctx, ~0u, nullptr, &E);
        last_torque = ctx->last_max_torque;
        last_hmax = ctx->last_max_h;
        ++evals;
        return E;
    }
    // The line search's trial: energy and gradient, plus g . s and max |g| reduced on the device and
    // read back with the energy partials in one transfer (one host synchronisation per trial).
    // launch_trial() queues the device work and the transfer; after the stream is synchronised,
    // read_trial() reads the values back and count_trial() makes them the search's latest trial.
    struct Trial {
        double stp, f, dg, gmax, torque, hmax;
    };
    int trial_ns{0};
    void launch_trial(const double* m, double* g, const double* s) {
        EvalArgs ea;
        ea.m = m;
        ea.mode = OutMode::GRAD;
        ea.out = g;
        ea.want_energy = true;
        evaluate(ctx, ea);
        trial_ns = finish_energy_launch(ctx);
        double* sc = ctx->scal.as<double>();
        dot_max_dev(ctx, g, s, n3, sc + trial_ns);
        MAFX_HIP(hipMemcpyAsync(ctx->host_scal, sc, sizeof(double) * (trial_ns + 2), hipMemcpyDeviceToHost,
                                ctx->stream));
    }
    Trial read_trial(double stp) {
        Trial t{};
        t.stp = stp;
        finish_energy_combine(ctx, ~0u, nullptr, &t.f);
        t.dg = ctx->host_scal[trial_ns];
        t.gmax = ctx->host_scal[trial_ns + 1];
        t.torque = ctx->last_max_torque;
        t.hmax = ctx->last_max_h;
        return t;
    }
    void count_trial(const Trial& t) {
        last_torque = t.torque;
        last_hmax = t.hmax;
        ++evals;
    }
    double energy_gradient_dg(const double* m, double* g, const double* s, double& dg, double& gmax) {
        launch_trial(m, g, s);
        MAFX_HIP(hipStreamSynchronize(ctx->stream));
        const Trial t = read_trial(0.0);
        count_trial(t);
        dg = t.dg;
        gmax = t.gmax;
        return t.f;
    }
    double mydot(const double* a, const double* b) { return dot(ctx, a, b, n3); }
    double mynorm(const double* a) { return std::sqrt(mydot(a, a)); }
    double maxabs(const double* a) { return max_abs(ctx, a, n3); }
};

// lbfgs_minimizer.cpp:362-505
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

// lbfgs_minimizer.cpp:213-360; m is overwritten with normalize(wa + stp s) at every trial.
// dginit = g . s is handed in by the caller (the two-loop's phiPrime0, the same dot product with
// the sign flipped, exactly); gmax returns max |g| of the last trial's gradient.
int cvsrch(Lbfgs& L, double* m, const double* wa, double& f, double* g, double& stp, const double* s,
           const double tolf, const double dginit, double& gmax, const Lbfgs::Trial* pre = nullptr) {
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
        double dg = 0;
        if (pre && nfev == 0 && stp == pre->stp) {
            // the first trial, evaluated before the search began (m and g hold it)
            L.count_trial(*pre);
            f = pre->f;
            dg = pre->dg;
            gmax = pre->gmax;
        } else {
            line_point(L.ctx, m, wa, stp, s);
            f = L.energy_gradient_dg(m, g, s, dg, gmax);
        }
        nfev++;
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
        if (info != 0) return -1;
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
            double fv = f, dgv = dg;
            cstep(stx, fx, dgx, sty, fy, dgy, stp, fv, dgv, brackt, stmin, stmax, infoc);
        }
        if (brackt) {
            if (std::fabs(sty - stx) >= 0.66 * width1) stp = stx + 0.5 * (sty - stx);
            width1 = width;
            width = std::fabs(sty - stx);
        }
    }
}

} // namespace

// lbfgs_minimizer.cpp:55-198
int lbfgs_minimize(mafx_ctx* ctx, double* m, double tolerance, int64_t max_iter, int verbose, double torque_tol,
                   double* E_out, mafx_lbfgs_stats* stats) {
    Lbfgs L{ctx, 3 * ctx->N};
    const int64_t n3 = L.n3;
    const size_t bytes = sizeof(double) * n3;
    constexpr size_t hist = 5;
    // Buffers.  The reference copies x_old = m and grad_old = grad at the top of every iteration
    // and copies each new pair into its history; here the iterate, the gradient and the history
    // slots are pointers that change roles, so no vector is copied.  mcur / gcur hold the current
    // iterate and gradient, mold / gold the previous ones; the history is a ring of hist + 1 slots
    // whose spare slot receives the next pair.
    DevBuf xbuf, gbuf0, gbuf1, q, dir, xspare, gspare;
    for (auto* b : {&xbuf, &gbuf0, &gbuf1, &q, &dir, &xspare, &gspare}) b->alloc(bytes);
    std::array<DevBuf, hist + 1> sB, yB;
    for (size_t i = 0; i <= hist; ++i) {
        sB[i].alloc(bytes); // read only after being written (k <= iter)
        yB[i].alloc(bytes);
    }
    std::array<double*, hist + 1> sV, yV; // [0, hist) stored pairs, oldest first; [hist] the spare
    for (size_t i = 0; i <= hist; ++i) {
        sV[i] = sB[i].as<double>();
        yV[i] = yB[i].as<double>();
    }
    double* mcur = m;
    double* mold = xbuf.as<double>();
    double* gcur = gbuf0.as<double>();
    double* gold = gbuf1.as<double>();
    // the line search's first trial is evaluated into these before the search starts
    double* mspec = xspare.as<double>();
    double* gspec = gspare.as<double>();
    constexpr int kPhiSlot = 32; // host_scal slot of the search direction's slope
    // rho_i = 1 / (s_i . y_i) of the stored pairs: the reference recomputes the dot product at every
    // use (lbfgs_minimizer.cpp:88,98); it is the y . s computed when the pair was stored (the dot
    // is commutative and deterministic), kept here.  Device scalars of the two-loop recursion:
    // [0, hist) alpha_i, [hist] the current dot, [hist + 1] the axpy coefficient, [hist + 2, +5)
    // y . s, y . y, s . s, max|s|, max|m| of a new pair.
    std::array<double, hist> rho{};
    DevBuf dsc;
    dsc.alloc(sizeof(double) * 16);
    double* ds = dsc.as<double>();
    double* d_dot = ds + hist;
    double* d_coef = ds + hist + 1;
    // A two-loop step's final reduction may be left pending on a context (lbfgs_flush); it points
    // into dsc, so none may outlive this call: discarded on every exit.
    struct PendingGuard {
        mafx_ctx* a;
        mafx_ctx* b;
        ~PendingGuard() {
            a->lbfgs_pend.mode = -1;
            if (b) b->lbfgs_pend.mode = -1;
        }
    } pending_guard{ctx, ctx->lbfgs_finish};
    ctx->lbfgs_pend.mode = -1;
    const double eps = 2.22e-16;
    const double eps2 = std::sqrt(eps);
    const double epsr = std::pow(eps, 0.9);
    const double tolf2 = std::sqrt(tolerance);
    const double tolf3 = std::pow(tolerance, 0.3333333333333333333333333);
    int reason = 0;
    int stalled = 0; // consecutive iterations without an energy decrease (torque_tol > 0 only)
    // a stalled run moves to the finish context once (mafx_lbfgs_set_finish_context)
    mafx_ctx* finish_ctx = torque_tol > 0 ? ctx->lbfgs_finish : nullptr;
    std::FILE* const conv = ctx->lbfgs_convergence;
    size_t finish_from = 0; // globIter at the move; 0 = no move
    bool move_now = false;
    double f = L.energy_and_gradient(mcur, gcur);
    double gradNorm = L.maxabs(gcur);
    size_t globIter = 0;
    auto finish = [&](int r) {
        if (mcur != m) copy(ctx, m, mcur, n3);
        MAFX_HIP(hipStreamSynchronize(ctx->stream));
        if (E_out) *E_out = f;
        if (stats) {
            stats->iterations = static_cast<int64_t>(globIter);
            stats->energy_evals = L.evals;
            stats->final_energy = f;
            stats->final_torque_rel = L.last_hmax > 0 ? L.last_torque / L.last_hmax : 0.0;
            stats->stop_reason = r;
            stats->finish_iterations = finish_from > 0 ? static_cast<int64_t>(globIter - finish_from) : 0;
        }
        return MAFX_OK;
    };
    // lbfgs_minimizer.cpp:67-69,142-146: verbose > 0 prints the start energy, verbose > 1 one
    // line per iteration (iteration, energy, max |gradient|, evaluations so far, step length)
    if (verbose > 0) {
        std::printf("f= %.17g\n", f);
        std::fflush(stdout);
    }
    if (gradNorm < (epsr * (1 + std::fabs(f)))) return finish(1);
    double f_old = 0.0;
    size_t iter = 0;
    double H0k = 1;
    do {
        f_old = f;
        // x_old = m, grad_old = grad: the current iterate and gradient become the old ones, and the
        // line search writes the new ones into the other buffers
        std::swap(mcur, mold);
        std::swap(gcur, gold);
        const int k = static_cast<int>(std::min(hist, iter));
        // The two-loop recursion q = H grad (lbfgs_minimizer.cpp:86-108) with its scalars on the
        // device: alpha_i = rho_i s_i . q, q -= alpha_i y_i; q *= H0k; beta = rho_i y_i . q,
        // q += (alpha_i - beta) s_i; then searchDir = -q and phiPrime0 = -grad . q.  Each axpy runs
        // fused with the dot product that follows it, the first one reads grad in place of the copy
        // q = grad, and the last one writes -q straight into the search direction.
        if (k == 0) {
            lbfgs_fused(ctx, gold, nullptr, nullptr, H0k, true, gold, dir.as<double>(), true, n3, d_dot);
        } else {
            // each reduction's last stage also forms the next step's alpha_i / coefficient
            lbfgs_dot_alpha(ctx, sV[k - 1], gold, n3, d_dot, rho[k - 1], ds + (k - 1), d_coef);
            for (int i = k - 1; i >= 1; i--)
                lbfgs_fused(ctx, i == k - 1 ? gold : q.as<double>(), d_coef, yV[i], 1.0, false, sV[i - 1],
                            q.as<double>(), false, n3, d_dot, 1, rho[i - 1], ds + (i - 1), d_coef);
            lbfgs_fused(ctx, k == 1 ? gold : q.as<double>(), d_coef, yV[0], H0k, true, yV[0], q.as<double>(), false,
                        n3, d_dot, 2, rho[0], ds + 0, d_coef);
            for (int i = 0; i < k; i++) {
                const bool last = i == k - 1;
                lbfgs_fused(ctx, q.as<double>(), d_coef, sV[i], 1.0, false, last ? gold : yV[i + 1],
                            last ? dir.as<double>() : q.as<double>(), last, n3, d_dot, last ? 0 : 2,
                            last ? 0.0 : rho[i + 1], last ? nullptr : ds + (i + 1), last ? nullptr : d_coef);
            }
        }
        // cvsrch's first trial is at step length 1 whenever dir is a descent direction (rate = 1
        // and the step bounds, lbfgs_minimizer.cpp:205-206,253-282): it is
        // evaluated now, into the spare buffers, and read back with the slope phi in one transfer.
        // A direction that is not a descent direction discards it.
        double rate = 1.0;
        line_point(ctx, mspec, mold, rate, dir.as<double>());
        L.launch_trial(mspec, gspec, dir.as<double>());
        MAFX_HIP(hipMemcpyAsync(ctx->host_scal + kPhiSlot, d_dot, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        MAFX_HIP(hipStreamSynchronize(ctx->stream));
        const double phiPrime0 = -ctx->host_scal[kPhiSlot];
        // the line search's g . dir for dir = -q is phiPrime0 exactly (every product and partial
        // sum only changes sign); after the reset to q = grad it is computed
        double dginit = phiPrime0;
        double gmax = -1.0;
        if (phiPrime0 < 0) {
            const Lbfgs::Trial first = L.read_trial(rate);
            std::swap(mcur, mspec);
            std::swap(gcur, gspec);
            cvsrch(L, mcur, mold, f, gcur, rate, dir.as<double>(), tolerance, dginit, gmax, &first);
        } else {
            if (phiPrime0 > 0) {
                neg_copy(ctx, dir.as<double>(), gold, n3);
                iter = 0;
                dginit = L.mydot(gold, dir.as<double>());
            }
            cvsrch(L, mcur, mold, f, gcur, rate, dir.as<double>(), tolerance, dginit, gmax);
        }
        const double f1 = 1 + std::fabs(f);
        gradNorm = gmax >= 0.0 ? gmax : L.maxabs(gcur);
        if (verbose > 1) {
            std::printf("bb> %zu %.17g  %.17g %lld %.17g\n", globIter, f, gradNorm, static_cast<long long>(L.evals),
                        rate);
            std::fflush(stdout);
        }
        if (gradNorm < (epsr * f1)) {
            globIter++;
            reason = 1;
            break;
        }
        // s = m - x_old and y = grad - grad_old into the spare history slot, with max|s|, max|m| for
        // the stop test and y . s, |y|^2, |s|^2 for the update (lbfgs_minimizer.cpp:137-170), in one
        // pass and one transfer
        double pr[5];
        lbfgs_pair(ctx, mcur, mold, gcur, gold, sV[hist], yV[hist], n3, ds + hist + 2);
        read_device(ctx, ds + hist + 2, 5, pr);
        const double ys = pr[0], yy = pr[1], ss = pr[2], smax = pr[3], mmax = pr[4];
        // lbfgs_minimizer.cpp:148-153: the three stop ratios, one line per iteration
        if (conv) {
            std::fprintf(conv, "%.27g\t%.27g\t%.27g\n", (f_old - f) / (tolerance * f1),
                         smax / (tolf2 * (1 + mmax)), gradNorm / (tolf3 * f1));
            std::fflush(conv);
        }
        if (((f_old - f) < (tolerance * f1)) && (smax < (tolf2 * (1 + mmax))) && (gradNorm <= (tolf3 * f1))) {
            globIter++;
            reason = 2;
            break;
        }
        if (torque_tol > 0 && L.last_hmax > 0 && L.last_torque / L.last_hmax <= torque_tol) {
            globIter++;
            reason = 3;
            break;
        }
        // with the torque stop, an iteration that cannot lower the energy any more (its change is
        // below the rounding of the energy sum, as on the FP32-field path near the minimum) five
        // times in a row ends the run instead of spending the remaining iterations on it
        stalled = (torque_tol > 0 && f >= f_old) ? stalled + 1 : 0;
        if (stalled >= 5) {
            if (finish_ctx && finish_from == 0) {
                move_now = true; // after this iteration's curvature pair is stored
            } else {
                globIter++;
                reason = 5;
                break;
            }
        }
        if (ys <= eps2 * std::sqrt(yy) * std::sqrt(ss)) {
            // Dennis-Schnabel 9.4.1: skip the update (the spare slot is simply reused)
        } else {
            // store the pair: into slot iter while the history fills, else drop the oldest
            // (lbfgs_minimizer.cpp:160-176; the reference rotates the vectors, here the pointers)
            size_t slot = iter;
            if (iter >= hist) {
                std::rotate(sV.begin(), sV.begin() + 1, sV.begin() + hist);
                std::rotate(yV.begin(), yV.begin() + 1, yV.begin() + hist);
                std::rotate(rho.begin(), rho.begin() + 1, rho.end());
                slot = hist - 1;
            }
            std::swap(sV[slot], sV[hist]);
            std::swap(yV[slot], yV[hist]);
            rho[slot] = 1.0 / ys;
            H0k = ys / yy;
            iter++;
        }
        globIter++;
        if (move_now) {
            // the same iterate, history and H0 scaling on the finish context: its energy and
            // gradient replace the stalled context's, and the next iteration starts from them
            move_now = false;
            MAFX_HIP(hipStreamSynchronize(ctx->stream));
            ctx = finish_ctx;
            L.ctx = finish_ctx;
            f = L.energy_and_gradient(mcur, gcur);
            stalled = 0;
            finish_from = globIter;
            if (verbose > 1) {
                std::printf("bb> finish context from iteration %zu: f= %.17g\n", globIter, f);
                std::fflush(stdout);
            }
        }
    } while (globIter < static_cast<size_t>(max_iter));
    // lbfgs_minimizer.cpp:181-192: the closing report after the loop (the gradient stop returns
    // before it); its max |m x (m x H)| is evaluated, as there, on a state whose m is the last gradient
    if (globIter >= static_cast<size_t>(max_iter) && max_iter > 99 && verbose > 0)
        std::printf("WARNING : maximum number of iterations exceeded in LBFGS\n");
    if (verbose > 0 && reason != 1) {
        EvalArgs ea;
        ea.m = gcur;
        ea.mode = OutMode::GRAD;
        ea.out = q.as<double>();
        evaluate(ctx, ea);
        const double mxh = L.maxabs(q.as<double>());
        std::printf("aa> %zu %.27g  %.27g\n", globIter, f - f_old, mxh);
        std::fflush(stdout);
    }
    return finish(reason);
}

} // namespace mafx
