// Oracle internals: an FP64 CPU restatement of the magnum.af reference arithmetic on the
// effective-field / integrator / minimizer path.
//
// THIS IS TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker or the CPU baseline, never as the
// product path.  Reference citations are relative to the magnum.af source tree.
//
// Layout of vector fields: v[c*N + x + nx*(y + ny*z)], the column-major af::array
// [nx, ny, nz, 3] (magnumaf/include/magnumaf/mesh.hpp:34).
#pragma once
#include <complex>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../include/mafx.h"

namespace oracle {

using cplx = std::complex<double>;
using Vec = std::vector<double>;

// magnumaf/include/magnumaf/constants.hpp:9-14
constexpr double kPi = 3.14159265358979323846;
constexpr double mu0 = 4e-7 * kPi;
constexpr double gamma_ = 1.760859644e11 * mu0;
constexpr double mu_b = 9.274009994e-24;

struct Mesh {
    int64_t nx{0}, ny{0}, nz{0};
    double dx{0}, dy{0}, dz{0};
    int64_t n() const { return nx * ny * nz; }
    int64_t at(int64_t x, int64_t y, int64_t z) const { return x + nx * (y + ny * z); }
};

// State (state.hpp / state.cpp:71-111): m, Ms or Ms_field, t, steps.
struct State {
    Mesh mesh;
    Vec m;             // 3N
    double Ms{0};      // used when Ms_field is empty
    Vec Ms_field;      // N, empty when not set
    double t{0};
    int64_t steps{0};
    bool has_Ms_field() const { return !Ms_field.empty(); }
};

// ---------------------------------------------------------------- FFT (fft.cpp)
// Unnormalised complex DFT along one axis of a [n2][n1][n0] complex array.
// sign = -1 forward, +1 backward (no 1/n scaling).
void fft_axis(std::vector<cplx>& a, int64_t n0, int64_t n1, int64_t n2, int axis, int sign);
// In-place 1-D transform of a contiguous line (used by the unit tests).
void fft_line(cplx* a, int64_t n, int sign);
// 3-D transforms over every axis whose length is > 1.
void fft3(std::vector<cplx>& a, int64_t n0, int64_t n1, int64_t n2, int sign);

// ---------------------------------------------------------------- util (terms.cpp)
void normalize_even_zero_vectors(Vec& m, int64_t N);   // util.cpp:34
void normalize_handle_zero_vectors(Vec& m, int64_t N); // util.cpp:37-42
double full_inner_product(const Vec& a, const Vec& b); // util.cpp:30-32
double max_4d_abs(const Vec& a);                       // math.cpp:6-8
void cross4(const Vec& a, const Vec& b, Vec& c, int64_t N); // math.cpp:10-19
void mean_m(const State& s, double out[3]);            // state.cpp:32-45

// ---------------------------------------------------------------- field terms
struct Term {
    virtual ~Term() = default;
    virtual void H(const State& s, Vec& h) const = 0;
    // Default: micro energy -mu0/2 Ms sum(h.m) dV  (micro_term.hpp:18-30)
    virtual double E(const State& s, const Vec& h) const;
    double E(const State& s) const {
        Vec h;
        H(s, h);
        return E(s, h);
    }
};

// Newell tensor, reference operation order (demag_field.cpp:20-121).
double newell_f(double x, double y, double z);
double newell_g(double x, double y, double z);
double newell_Nxx(int ix, int iy, int iz, double dx, double dy, double dz);
double newell_Nxy(int ix, int iy, int iz, double dx, double dy, double dz);
// Real-space tensor [6][n2e][n1e][n0e] on the expanded grid (demag_field.cpp:104-142).
void demag_tensor_real(const Mesh& mesh, Vec& N, int nthreads);

std::unique_ptr<Term> make_demag(const Mesh& mesh, int nthreads);
std::unique_ptr<Term> make_exchange(double A);
std::unique_ptr<Term> make_dmi(double D, const double axis[3], const Vec& Dfield); // Dfield N or empty
// Kfield N or empty, axis_field 3N or empty
std::unique_ptr<Term> make_uniaxial(double Ku1, const double axis[3], const Vec& Kfield, const Vec& axis_field);
std::unique_ptr<Term> make_zeeman(const Vec& h_ext); // 3N
// time-dependent ExternalField / AtomisticExternalField (external_field.cpp:13-17,28-36)
std::unique_ptr<Term> make_zeeman_callback(mafx_field_fn cb, void* user, bool homogeneous, bool atomistic);
std::unique_ptr<Term> make_rkky(const Mesh& mesh, const Vec& rkky, const Vec& A,
                                const std::vector<uint32_t>& idx);
std::unique_ptr<Term> make_atom_dipole(const Mesh& mesh);
std::unique_ptr<Term> make_atom_exchange(double J);
std::unique_ptr<Term> make_atom_dmi(double D, const double axis[3]);
std::unique_ptr<Term> make_atom_uniaxial(double K, const double axis[3]);
std::unique_ptr<Term> make_atom_zeeman(const Vec& h_ext);

using Terms = std::vector<std::unique_ptr<Term>>;
void heff(const Terms& terms, const State& s, Vec& h);          // field_term.hpp:116-123
double eeff(const Terms& terms, const State& s);                // field_term.hpp:180-183
// Heff and energy in one pass, term-wise (field_term.hpp:196-200)
double heff_and_energy(const Terms& terms, const State& s, Vec& h);

// ---------------------------------------------------------------- solvers (solvers.cpp)
struct Controller { // controller.hpp:11-61, controller.cpp:36-95
    double hmin{1e-15}, hmax{3.5e-10}, atol{1e-6}, rtol{1e-6};
    double beta = 0.4 / 5.0;
    double alpha = 0.2 - 0.75 * beta;
    double headroom = 0.9, minscale = 0.2, maxscale = 10.;
    bool reject{false};
    double errold{1.0e-4};
    double hnext{0.0};
    int64_t counter_reject{0}, counter_accepted{0}, counter_hmax{0}, counter_hmin{0},
        counter_maxscale{0}, counter_minscale{0};
    bool success(double err, double& h);
    double givescale(double v) const { return atol + rtol * std::abs(v); }
};

struct LLG { // llg_integrator.hpp / adaptive_runge_kutta.hpp
    const Terms* terms{nullptr};
    double alpha{0};
    Vec alpha_field; // N, empty for the scalar alpha (llg_integrator.hpp:32-40)
    std::string scheme{"RKF45"};
    Controller controller;
    bool normalize{true};
    bool dissipation_term_only{false};
    double h{1.01e-15};
    Vec k_fsal;
    int64_t heff_evals{0};
    double h_last{0}, err_last{0};
    // optional analytic right-hand side for the integrator unit tests: 1 is the reference's
    // y' = t sqrt(y) (test/adaptive_runge_kutta.cpp:9-17), 2 is y' = cos(t) y, y = y0 exp(sin t)
    int analytic_rhs{0};

    void f(const State& s, Vec& out) const; // llg_integrator.cpp:18-38
    void step(State& s);                    // adaptive_runge_kutta.cpp:39-71
    Vec RKF45(const State& s, double dt, double& err) const;
    Vec DP45(const State& s, double dt, double& err);
    Vec BS23(const State& s, double dt, double& err);
    Vec BS45(const State& s, double dt, double& err);
    Vec DP78(const State& s, double dt, double& err) const;
    // explicit embedded scheme in the k = dt f form: stage inputs m + a[s][1] k1 + ..., sumbk =
    // sum bs_j k_j, rk_error = sumbk - sum be_j k_j (all sums left to right)
    Vec tableau(const State& s, double dt, double& err, int ns, const double (*a)[14], const double* c,
                const double* bs, const double* be) const;
    // llg_integrator.cpp:68-104: returns rows (t, <mx>, <my>, <mz>)
    std::vector<std::array<double, 4>> integrate_dense(State& s, double time, double dt_write);
    mutable int64_t f_calls{0};
};

struct LbfgsResult {
    double E{0};
    int64_t iterations{0}, evals{0};
    int stop_reason{0};
};
// One line search as cvsrch ends it: f(0), g(0).s, the returned step, f and g.s there, its info
// code (1 = the sufficient-decrease and curvature tests hold) and the evaluation count.
struct LbfgsTraceRow {
    double finit, dginit, stp, f, dg, info, nfev;
};
// lbfgs_minimizer.cpp:55-505; torque_tol > 0 adds the build-side torque stop; trace, when given,
// receives one row per line search.
LbfgsResult lbfgs_minimize(const Terms& terms, State& s, double tolerance, int64_t maxIter,
                           int verbose, double torque_tol, std::vector<LbfgsTraceRow>* trace = nullptr);

} // namespace oracle
