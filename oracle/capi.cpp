// C API of the oracle (see oracle.h).  Test infrastructure only.
#include "oracle.h"

#include <cmath>
#include <cstring>
#include <exception>
#include <string>

#include "oracle_internal.hpp"

using namespace oracle;

struct oracle_sim {
    State state;
    Terms terms;
    LLG llg;
    bool llg_ready{false};
};

namespace {

std::unique_ptr<Term> build_term(const Mesh& g, const mafx_term_desc& d, int nthreads) {
    const int64_t N = g.n();
    auto hext = [&]() {
        Vec h(3 * N);
        if (d.array0) {
            std::memcpy(h.data(), d.array0, sizeof(double) * 3 * N);
        } else {
            for (int c = 0; c < 3; ++c)
                for (int64_t i = 0; i < N; ++i) h[c * N + i] = d.field[c];
        }
        return h;
    };
    switch (d.kind) {
    case MAFX_DEMAG: return make_demag(g, nthreads);
    case MAFX_EXCHANGE: return make_exchange(d.value);
    case MAFX_DMI: return make_dmi(d.value, d.axis, d.array0 ? Vec(d.array0, d.array0 + N) : Vec{});
    case MAFX_UNIAXIAL:
        return make_uniaxial(d.value, d.axis, d.array0 ? Vec(d.array0, d.array0 + N) : Vec{},
                             d.array1 ? Vec(d.array1, d.array1 + 3 * N) : Vec{});
    case MAFX_ZEEMAN:
        if (d.callback)
            return make_zeeman_callback(d.callback, d.callback_user, (d.flags & MAFX_TERM_HOMOGENEOUS_CALLBACK) != 0,
                                        false);
        return make_zeeman(hext());
    case MAFX_RKKY: {
        Vec rk(d.array0, d.array0 + N), A(d.array1, d.array1 + N);
        std::vector<uint32_t> idx;
        if (d.index0) idx.assign(d.index0, d.index0 + N);
        return make_rkky(g, rk, A, idx);
    }
    case MAFX_ATOM_DIPOLE: return make_atom_dipole(g);
    case MAFX_ATOM_EXCHANGE: return make_atom_exchange(d.value);
    case MAFX_ATOM_DMI: return make_atom_dmi(d.value, d.axis);
    case MAFX_ATOM_UNIAXIAL: return make_atom_uniaxial(d.value, d.axis);
    case MAFX_ATOM_ZEEMAN:
        if (d.callback)
            return make_zeeman_callback(d.callback, d.callback_user, (d.flags & MAFX_TERM_HOMOGENEOUS_CALLBACK) != 0,
                                        true);
        return make_atom_zeeman(hext());
    default: return nullptr;
    }
}

} // namespace

extern "C" {

oracle_sim* oracle_create(const mafx_mesh* mesh, double Ms, const double* Ms_field, const mafx_term_desc* terms,
                          int nterms, const double* m0, int nthreads) {
    try {
        auto* s = new oracle_sim;
        Mesh g{mesh->nx, mesh->ny, mesh->nz, mesh->dx, mesh->dy, mesh->dz};
        s->state.mesh = g;
        const int64_t N = g.n();
        s->state.m.assign(m0, m0 + 3 * N);
        normalize_handle_zero_vectors(s->state.m, N); // state.cpp:92
        s->state.Ms = Ms;
        if (Ms_field) {
            s->state.Ms_field.assign(Ms_field, Ms_field + N);
            s->state.Ms = std::nan("");
        } else {
            bool any_zero = false; // state.cpp:71-89
            for (int64_t i = 0; i < N; ++i) {
                const double n = std::sqrt(s->state.m[i] * s->state.m[i] + s->state.m[N + i] * s->state.m[N + i] +
                                           s->state.m[2 * N + i] * s->state.m[2 * N + i]);
                if (n == 0) any_zero = true;
            }
            if (any_zero) {
                s->state.Ms_field.assign(N, 0.0);
                for (int64_t i = 0; i < N; ++i) {
                    const double n = std::sqrt(s->state.m[i] * s->state.m[i] + s->state.m[N + i] * s->state.m[N + i] +
                                               s->state.m[2 * N + i] * s->state.m[2 * N + i]);
                    s->state.Ms_field[i] = Ms * (n != 0 ? 1.0 : 0.0);
                }
            }
        }
        for (int i = 0; i < nterms; ++i) {
            auto t = build_term(g, terms[i], nthreads);
            if (!t) {
                delete s;
                return nullptr;
            }
            s->terms.push_back(std::move(t));
        }
        return s;
    } catch (...) {
        return nullptr;
    }
}

void oracle_destroy(oracle_sim* s) { delete s; }

int oracle_add_term(oracle_sim* s, const mafx_term_desc* term, int nthreads) {
    auto t = build_term(s->state.mesh, *term, nthreads);
    if (!t) return 1;
    s->terms.push_back(std::move(t));
    return 0;
}

int oracle_set_m(oracle_sim* s, const double* m, int normalize) {
    s->state.m.assign(m, m + 3 * s->state.mesh.n());
    if (normalize) normalize_handle_zero_vectors(s->state.m, s->state.mesh.n());
    return 0;
}
int oracle_get_m(const oracle_sim* s, double* m) {
    std::memcpy(m, s->state.m.data(), sizeof(double) * s->state.m.size());
    return 0;
}
int oracle_set_t(oracle_sim* s, double t) {
    s->state.t = t;
    return 0;
}
double oracle_get_t(const oracle_sim* s) { return s->state.t; }
int64_t oracle_get_steps(const oracle_sim* s) { return s->state.steps; }
int oracle_has_ms_field(const oracle_sim* s) { return s->state.has_Ms_field() ? 1 : 0; }

int oracle_heff(oracle_sim* s, uint32_t mask, double* h) {
    try {
        const int64_t n3 = 3 * s->state.mesh.n();
        Vec acc, t;
        bool first = true;
        for (size_t i = 0; i < s->terms.size(); ++i) {
            if (!(mask & (1u << i))) continue;
            s->terms[i]->H(s->state, t);
            if (first) {
                acc = t;
                first = false;
            } else {
                for (int64_t j = 0; j < n3; ++j) acc[j] = acc[j] + t[j];
            }
        }
        if (first) acc.assign(n3, 0.0);
        std::memcpy(h, acc.data(), sizeof(double) * n3);
        return 0;
    } catch (...) {
        return 1;
    }
}

int oracle_energy(oracle_sim* s, uint32_t mask, double* E_terms, double* E_total) {
    try {
        double tot = 0;
        for (size_t i = 0; i < s->terms.size(); ++i) {
            double e = 0;
            if (mask & (1u << i)) e = s->terms[i]->E(s->state);
            if (E_terms) E_terms[i] = e;
            tot = tot + e;
        }
        if (E_total) *E_total = tot;
        return 0;
    } catch (...) {
        return 1;
    }
}

int oracle_mean_m(const oracle_sim* s, double out[3]) {
    mean_m(s->state, out);
    return 0;
}

int oracle_llg_init(oracle_sim* s, double alpha, const char* scheme, double hmin, double hmax, double atol,
                    double rtol, int dissipation_term_only, int normalize, int analytic_rhs) {
    s->llg = LLG{};
    s->llg.terms = &s->terms;
    s->llg.alpha = alpha;
    s->llg.scheme = scheme;
    s->llg.controller.hmin = hmin;
    s->llg.controller.hmax = hmax;
    s->llg.controller.atol = atol;
    s->llg.controller.rtol = rtol;
    s->llg.dissipation_term_only = dissipation_term_only != 0;
    s->llg.normalize = normalize != 0;
    s->llg.analytic_rhs = analytic_rhs;
    s->llg_ready = true;
    return 0;
}

int oracle_llg_set_alpha(oracle_sim* s, double alpha) {
    s->llg.alpha = alpha;
    return 0;
}

int oracle_llg_set_alpha_field(oracle_sim* s, const double* alpha) {
    if (alpha)
        s->llg.alpha_field.assign(alpha, alpha + s->state.mesh.n());
    else
        s->llg.alpha_field.clear();
    return 0;
}

int oracle_llg_step(oracle_sim* s, double* h_used, double* err, double* h_next, int64_t* accepted,
                    int64_t* rejected, int64_t* heff_evals) {
    if (!s->llg_ready) return 4;
    try {
        s->llg.step(s->state);
        if (h_used) *h_used = s->llg.h_last;
        if (err) *err = s->llg.err_last;
        if (h_next) *h_next = s->llg.h;
        if (accepted) *accepted = s->llg.controller.counter_accepted;
        if (rejected) *rejected = s->llg.controller.counter_reject;
        if (heff_evals) *heff_evals = s->llg.heff_evals;
        return 0;
    } catch (...) {
        return 1;
    }
}

int oracle_llg_integrate_dense(oracle_sim* s, double time, double dt_write, double* out, int64_t max_rows,
                               int64_t* rows) {
    if (!s->llg_ready) return 4;
    auto r = s->llg.integrate_dense(s->state, time, dt_write);
    const int64_t n = std::min<int64_t>(static_cast<int64_t>(r.size()), max_rows);
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 4; ++c) out[4 * i + c] = r[i][c];
    *rows = n;
    return 0;
}

int oracle_llg_attempt(oracle_sim* s, double dt, double* dm, double* err) {
    if (!s->llg_ready) return 4;
    double e = 0;
    Vec r;
    if (s->llg.scheme == "DP45")
        r = s->llg.DP45(s->state, dt, e);
    else if (s->llg.scheme == "BS23")
        r = s->llg.BS23(s->state, dt, e);
    else if (s->llg.scheme == "BS45")
        r = s->llg.BS45(s->state, dt, e);
    else if (s->llg.scheme == "DP78")
        r = s->llg.DP78(s->state, dt, e);
    else
        r = s->llg.RKF45(s->state, dt, e);
    std::memcpy(dm, r.data(), sizeof(double) * r.size());
    *err = e;
    return 0;
}

int oracle_lbfgs(oracle_sim* s, double tol, int64_t maxiter, double torque_tol, double* E, int64_t* iters,
                 int64_t* evals, int* reason) {
    try {
        auto r = lbfgs_minimize(s->terms, s->state, tol, maxiter, 0, torque_tol);
        if (E) *E = r.E;
        if (iters) *iters = r.iterations;
        if (evals) *evals = r.evals;
        if (reason) *reason = r.stop_reason;
        return 0;
    } catch (...) {
        return 1;
    }
}

int oracle_lbfgs_trace(oracle_sim* s, double tol, int64_t maxiter, double* E, double* rows_out, int64_t max_rows,
                       int64_t* rows) {
    try {
        std::vector<LbfgsTraceRow> tr;
        auto r = lbfgs_minimize(s->terms, s->state, tol, maxiter, 0, 0.0, &tr);
        if (E) *E = r.E;
        const int64_t n = std::min<int64_t>(static_cast<int64_t>(tr.size()), max_rows);
        for (int64_t i = 0; i < n; ++i) {
            const LbfgsTraceRow& t = tr[i];
            const double v[7] = {t.finit, t.dginit, t.stp, t.f, t.dg, t.info, t.nfev};
            for (int c = 0; c < 7; ++c) rows_out[7 * i + c] = v[c];
        }
        *rows = n;
        return 0;
    } catch (...) {
        return 1;
    }
}

int oracle_llg_counters(const oracle_sim* s, int64_t out[6]) {
    const Controller& c = s->llg.controller;
    out[0] = c.counter_accepted;
    out[1] = c.counter_reject;
    out[2] = c.counter_hmin;
    out[3] = c.counter_hmax;
    out[4] = c.counter_maxscale;
    out[5] = c.counter_minscale;
    return 0;
}

int oracle_demag_tensor(const mafx_mesh* mesh, double* N, int nthreads) {
    Mesh g{mesh->nx, mesh->ny, mesh->nz, mesh->dx, mesh->dy, mesh->dz};
    Vec v;
    demag_tensor_real(g, v, nthreads);
    std::memcpy(N, v.data(), sizeof(double) * v.size());
    return 0;
}

double oracle_newell_nxx(int ix, int iy, int iz, double dx, double dy, double dz) {
    return newell_Nxx(ix, iy, iz, dx, dy, dz);
}
double oracle_newell_nxy(int ix, int iy, int iz, double dx, double dy, double dz) {
    return newell_Nxy(ix, iy, iz, dx, dy, dz);
}

int oracle_fft(double* data, int64_t n, int sign) {
    fft_line(reinterpret_cast<cplx*>(data), n, sign);
    return 0;
}

} // extern "C"
