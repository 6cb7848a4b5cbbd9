// Oracle field terms: FP64 CPU restatement of the reference's effective-field terms and energies.
// Test infrastructure only (see oracle_internal.hpp).
#include "oracle_internal.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <thread>

namespace oracle {

// ------------------------------------------------------------------ util
// util.cpp:34  a / |a| (zero vectors give NaN, as in the reference)
void normalize_even_zero_vectors(Vec& m, int64_t N) {
    for (int64_t i = 0; i < N; ++i) {
        const double n = std::sqrt(m[i] * m[i] + m[N + i] * m[N + i] + m[2 * N + i] * m[2 * N + i]);
        m[i] /= n;
        m[N + i] /= n;
        m[2 * N + i] /= n;
    }
}
// util.cpp:37-42  a / |a|, zero vectors stay zero
void normalize_handle_zero_vectors(Vec& m, int64_t N) {
    for (int64_t i = 0; i < N; ++i) {
        const double n = std::sqrt(m[i] * m[i] + m[N + i] * m[N + i] + m[2 * N + i] * m[2 * N + i]);
        for (int c = 0; c < 3; ++c) m[c * N + i] = n != 0 ? m[c * N + i] / n : 0.0;
    }
}
double full_inner_product(const Vec& a, const Vec& b) { // util.cpp:30-32
    long double s = 0;
    for (size_t i = 0; i < a.size(); ++i) s += static_cast<long double>(a[i]) * b[i];
    return static_cast<double>(s);
}
double max_4d_abs(const Vec& a) { // math.cpp:6-8
    double r = 0;
    for (double v : a) r = std::max(r, std::abs(v));
    return r;
}
void cross4(const Vec& a, const Vec& b, Vec& c, int64_t N) { // math.cpp:10-19
    c.resize(3 * N);
    for (int64_t i = 0; i < N; ++i) {
        const double a0 = a[i], a1 = a[N + i], a2 = a[2 * N + i];
        const double b0 = b[i], b1 = b[N + i], b2 = b[2 * N + i];
        c[i] = a1 * b2 - a2 * b1;
        c[N + i] = a2 * b0 - a0 * b2;
        c[2 * N + i] = a0 * b1 - a1 * b0;
    }
}
void mean_m(const State& s, double out[3]) { // state.cpp:32-45
    const int64_t N = s.mesh.n();
    int64_t nonzero = 0;
    for (int64_t i = 0; i < N; ++i) {
        const double n = std::sqrt(s.m[i] * s.m[i] + s.m[N + i] * s.m[N + i] + s.m[2 * N + i] * s.m[2 * N + i]);
        if (n != 0) ++nonzero;
    }
    for (int c = 0; c < 3; ++c) {
        double sum = 0;
        for (int64_t i = 0; i < N; ++i) sum += s.m[c * N + i];
        out[c] = sum / static_cast<double>(nonzero);
    }
}

namespace {

int64_t nexp_xy(int64_t n) { return 2 * n; }              // demag_field.cpp:14-15
int64_t nexp_z(int64_t n) { return n == 1 ? 1 : 2 * n; } // demag_field.cpp:16

// 3x3x3 spatial convolution with zero padding, output the size of the input
// (af::convolve(..., AF_CONV_DEFAULT, AF_CONV_SPATIAL)): out[x] = sum_i f[i] in[x + 1 - i].
// filt is f[i + 3 (j + 3 k)] for one component.
void conv3(const Mesh& g, const double* in, const double* filt, double* out) {
    for (int64_t z = 0; z < g.nz; ++z)
        for (int64_t y = 0; y < g.ny; ++y)
            for (int64_t x = 0; x < g.nx; ++x) {
                double acc = 0;
                for (int k = 0; k < 3; ++k) {
                    const int64_t zz = z + 1 - k;
                    if (zz < 0 || zz >= g.nz) continue;
                    for (int j = 0; j < 3; ++j) {
                        const int64_t yy = y + 1 - j;
                        if (yy < 0 || yy >= g.ny) continue;
                        for (int i = 0; i < 3; ++i) {
                            const int64_t xx = x + 1 - i;
                            if (xx < 0 || xx >= g.nx) continue;
                            const double w = filt[i + 3 * (j + 3 * k)];
                            if (w != 0.0) acc += w * in[g.at(xx, yy, zz)];
                        }
                    }
                }
                out[g.at(x, y, z)] = acc;
            }
}

// 1/(mu0 Ms) style coefficient application with the Ms_field branch: cells where Ms_field==0
// are set to 0 (af::replace(heff, Ms_field != 0, 0)).
void scale_by_ms(const State& s, double numerator, Vec& h) {
    const int64_t N = s.mesh.n();
    if (!s.has_Ms_field()) {
        const double c = numerator / (mu0 * s.Ms);
        for (auto& v : h) v = c * v;
    } else {
        for (int c = 0; c < 3; ++c)
            for (int64_t i = 0; i < N; ++i) {
                const double ms = s.Ms_field[i];
                h[c * N + i] = ms != 0 ? numerator / (mu0 * ms) * h[c * N + i] : 0.0;
            }
    }
}

std::array<double, 3> normalized(const double a[3]) { // util.hpp:62-69
    const double n = std::sqrt(std::pow(a[0], 2) + std::pow(a[1], 2) + std::pow(a[2], 2));
    return {a[0] / n, a[1] / n, a[2] / n};
}

// sum over all entries of (Ms or Ms_field) * h * m, as af::sum(af::sum(af::sum(af::sum(.,0),1),2),3)
double sum_hm(const State& s, const Vec& h, bool with_ms_field) {
    const int64_t N = s.mesh.n();
    double total = 0;
    for (int c = 0; c < 3; ++c) {
        double sc = 0;
        for (int64_t i = 0; i < N; ++i) {
            const double w = with_ms_field ? s.Ms_field[i] * h[c * N + i] : h[c * N + i];
            sc += w * s.m[c * N + i];
        }
        total += sc;
    }
    return total;
}

// micro energy -mu0/2 Ms sum(h.m) dV, Ms_field variant (micro_term.hpp:18-30); `factor` is
// 1/2 for the linear terms and 1 for Zeeman (external_field.cpp:38-48).
double micro_energy(const State& s, const Vec& h, double factor) {
    const Mesh& g = s.mesh;
    if (!s.has_Ms_field())
        return -mu0 * factor * s.Ms * sum_hm(s, h, false) * g.dx * g.dy * g.dz;
    return -mu0 * factor * sum_hm(s, h, true) * g.dx * g.dy * g.dz;
}

} // namespace

double Term::E(const State& s, const Vec& h) const { return micro_energy(s, h, 0.5); }

// ------------------------------------------------------------------ Newell tensor
// demag_field.cpp:20-40; operation order kept so that the far-field rounding (which
// dominates N at large offsets) matches the reference bit for bit.
double newell_f(double x, double y, double z) {
    x = std::fabs(x);
    y = std::fabs(y);
    z = std::fabs(z);
    const double R = std::sqrt(std::pow(x, 2) + std::pow(y, 2) + std::pow(z, 2));
    const double xx = std::pow(x, 2);
    const double yy = std::pow(y, 2);
    const double zz = std::pow(z, 2);
    double result = 1.0 / 6.0 * (2.0 * xx - yy - zz) * R;
    if (xx + zz > 0) result += y / 2.0 * (zz - xx) * std::asinh(y / (std::sqrt(xx + zz)));
    if (xx + yy > 0) result += z / 2.0 * (yy - xx) * std::asinh(z / (std::sqrt(xx + yy)));
    if (x * R > 0) result += -x * y * z * std::atan(y * z / (x * R));
    return result;
}

// demag_field.cpp:42-69
double newell_g(const double x, const double y, double z) {
    z = std::fabs(z);
    const double R = std::sqrt(std::pow(x, 2) + std::pow(y, 2) + std::pow(z, 2));
    const double xx = std::pow(x, 2);
    const double yy = std::pow(y, 2);
    const double zz = std::pow(z, 2);
    double result = -x * y * R / 3.0;
    if (xx + yy > 0) result += x * y * z * std::asinh(z / (std::sqrt(xx + yy)));
    if (yy + zz > 0) result += y / 6.0 * (3.0 * zz - yy) * std::asinh(x / (std::sqrt(yy + zz)));
    if (xx + zz > 0) result += x / 6.0 * (3.0 * zz - xx) * std::asinh(y / (std::sqrt(xx + zz)));
    if (z * R > 0) result += -std::pow(z, 3) / 6.0 * std::atan(x * y / (z * R));
    if (y * R != 0) result += -z * yy / 2.0 * std::atan(x * z / (y * R));
    if (x * R != 0) result += -z * xx / 2.0 * std::atan(y * z / (x * R));
    return result;
}

namespace {
// The 27 cell-pair offsets in the order the reference writes its second-difference sum
// (demag_field.cpp:75-83): weight 8, -4 (six faces), 2 (twelve edges), -1 (eight corners).
struct Tap {
    double w;
    int sx, sy, sz;
};
const Tap kTaps[27] = {
    {8, 0, 0, 0},   {-4, 1, 0, 0},  {-4, -1, 0, 0},  {-4, 0, 1, 0},  {-4, 0, -1, 0},  {-4, 0, 0, 1},
    {-4, 0, 0, -1}, {2, 1, 1, 0},   {2, 1, -1, 0},   {2, -1, 1, 0},  {2, -1, -1, 0},  {2, 1, 0, 1},
    {2, 1, 0, -1},  {2, -1, 0, 1},  {2, -1, 0, -1},  {2, 0, 1, 1},   {2, 0, 1, -1},   {2, 0, -1, 1},
    {2, 0, -1, -1}, {-1, 1, 1, 1},  {-1, 1, 1, -1},  {-1, 1, -1, 1}, {-1, 1, -1, -1}, {-1, -1, 1, 1},
    {-1, -1, 1, -1}, {-1, -1, -1, 1}, {-1, -1, -1, -1}};

template <class F>
double second_difference(F fn, int ix, int iy, int iz, double dx, double dy, double dz) {
    const double x = dx * ix;
    const double y = dy * iy;
    const double z = dz * iz;
    double result = 0;
    for (int t = 0; t < 27; ++t) {
        const Tap& p = kTaps[t];
        const double xs = p.sx == 0 ? x : p.sx > 0 ? x + dx : x - dx;
        const double ys = p.sy == 0 ? y : p.sy > 0 ? y + dy : y - dy;
        const double zs = p.sz == 0 ? z : p.sz > 0 ? z + dz : z - dz;
        const double v = fn(xs, ys, zs);
        if (t == 0)
            result = 8.0 * v;
        else if (p.w < 0)
            result = result - (-p.w) * v;
        else
            result = result + p.w * v;
    }
    return -result / (4.0 * kPi * dx * dy * dz);
}
} // namespace

double newell_Nxx(int ix, int iy, int iz, double dx, double dy, double dz) { // :71-85
    return second_difference(newell_f, ix, iy, iz, dx, dy, dz);
}
double newell_Nxy(int ix, int iy, int iz, double dx, double dy, double dz) { // :87-102
    return second_difference(newell_g, ix, iy, iz, dx, dy, dz);
}

// demag_field.cpp:104-142.  Output layout N[c][i2][i1][i0] (x fastest).
void demag_tensor_real(const Mesh& mesh, Vec& N, int nthreads) {
    const int64_t n0 = nexp_xy(mesh.nx), n1 = nexp_xy(mesh.ny), n2 = nexp_z(mesh.nz);
    const int64_t P = n0 * n1 * n2;
    N.assign(6 * P, 0.0);
    auto work = [&](int64_t i0_begin, int64_t i0_end) {
        for (int64_t i0 = i0_begin; i0 < i0_end; ++i0) {
            const int j0 = static_cast<int>((i0 + n0 / 2) % n0 - n0 / 2);
            for (int64_t i1 = 0; i1 < n1; ++i1) {
                const int j1 = static_cast<int>((i1 + n1 / 2) % n1 - n1 / 2);
                for (int64_t i2 = 0; i2 < n2; ++i2) {
                    const int j2 = static_cast<int>((i2 + n2 / 2) % n2 - n2 / 2);
                    const int64_t at = i0 + n0 * (i1 + n1 * i2);
                    N[0 * P + at] = newell_Nxx(j0, j1, j2, mesh.dx, mesh.dy, mesh.dz);
                    N[1 * P + at] = newell_Nxy(j0, j1, j2, mesh.dx, mesh.dy, mesh.dz);
                    N[2 * P + at] = newell_Nxy(j0, j2, j1, mesh.dx, mesh.dz, mesh.dy);
                    N[3 * P + at] = newell_Nxx(j1, j2, j0, mesh.dy, mesh.dz, mesh.dx);
                    N[4 * P + at] = newell_Nxy(j1, j2, j0, mesh.dy, mesh.dz, mesh.dx);
                    N[5 * P + at] = newell_Nxx(j2, j0, j1, mesh.dz, mesh.dx, mesh.dy);
                }
            }
        }
    };
    if (nthreads < 1) nthreads = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) {
        const int64_t b = n0 * t / nthreads, e = n0 * (t + 1) / nthreads;
        th.emplace_back(work, b, e);
    }
    for (auto& t : th) t.join();
}

namespace {

// Shared FFT-convolution of a 6-component symmetric tensor with a 3-vector source
// (demag_field.cpp:215-264, atomistic_dipole_dipole_field.cpp:17-53).
struct TensorConv {
    int64_t nx, ny, nz, n0, n1, n2, P;
    std::vector<cplx> Nfft[6];
    void init(const Mesh& g, const Vec& Nreal) {
        nx = g.nx;
        ny = g.ny;
        nz = g.nz;
        n0 = nexp_xy(nx);
        n1 = nexp_xy(ny);
        n2 = nexp_z(nz);
        P = n0 * n1 * n2;
        for (int c = 0; c < 6; ++c) {
            Nfft[c].assign(P, cplx(0, 0));
            for (int64_t i = 0; i < P; ++i) Nfft[c][i] = cplx(Nreal[c * P + i], 0.0);
            fft3(Nfft[c], n0, n1, n2, -1);
        }
    }
    // h = crop(IFFT(N^ . FFT(pad(src)))) / P
    void apply(const Vec& src, Vec& h) const {
        const int64_t N = nx * ny * nz;
        std::vector<cplx> S[3];
        for (int c = 0; c < 3; ++c) {
            S[c].assign(P, cplx(0, 0));
            for (int64_t z = 0; z < nz; ++z)
                for (int64_t y = 0; y < ny; ++y)
                    for (int64_t x = 0; x < nx; ++x)
                        S[c][x + n0 * (y + n1 * z)] = cplx(src[c * N + x + nx * (y + ny * z)], 0.0);
            fft3(S[c], n0, n1, n2, -1);
        }
        // component map of the symmetric product (demag_field.cpp:236-249)
        static const int idx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
        std::vector<cplx> Hc[3];
        for (int i = 0; i < 3; ++i) {
            Hc[i].resize(P);
            for (int64_t k = 0; k < P; ++k)
                Hc[i][k] = Nfft[idx[i][0]][k] * S[0][k] + Nfft[idx[i][1]][k] * S[1][k] +
                           Nfft[idx[i][2]][k] * S[2][k];
            fft3(Hc[i], n0, n1, n2, +1);
        }
        h.assign(3 * N, 0.0);
        const double inv = 1.0 / static_cast<double>(P);
        for (int c = 0; c < 3; ++c)
            for (int64_t z = 0; z < nz; ++z)
                for (int64_t y = 0; y < ny; ++y)
                    for (int64_t x = 0; x < nx; ++x)
                        h[c * N + x + nx * (y + ny * z)] = Hc[c][x + n0 * (y + n1 * z)].real() * inv;
    }
};

// ------------------------------------------------------------------ micro terms
struct Demag : Term { // demag_field.cpp:211-264
    TensorConv tc;
    Demag(const Mesh& g, int nthreads) {
        Vec Nr;
        demag_tensor_real(g, Nr, nthreads);
        tc.init(g, Nr);
    }
    void H(const State& s, Vec& h) const override {
        const int64_t N = s.mesh.n();
        Vec M(3 * N);
        for (int c = 0; c < 3; ++c)
            for (int64_t i = 0; i < N; ++i)
                M[c * N + i] = (s.has_Ms_field() ? s.Ms_field[i] : s.Ms) * s.m[c * N + i];
        tc.apply(M, h);
    }
};

void exchange_filter(const Mesh& g, double f[27]) { // exchange_field.cpp:116-125
    for (int i = 0; i < 27; ++i) f[i] = 0;
    f[1 + 3 * (1 + 3 * 1)] = -2. / std::pow(g.dx, 2.) - 2. / std::pow(g.dy, 2.) - 2. / std::pow(g.dz, 2.);
    f[0 + 3 * (1 + 3 * 1)] = 1 / std::pow(g.dx, 2);
    f[2 + 3 * (1 + 3 * 1)] = 1 / std::pow(g.dx, 2);
    f[1 + 3 * (0 + 3 * 1)] = 1 / std::pow(g.dy, 2);
    f[1 + 3 * (2 + 3 * 1)] = 1 / std::pow(g.dy, 2);
    f[1 + 3 * (1 + 3 * 0)] = 1 / std::pow(g.dz, 2);
    f[1 + 3 * (1 + 3 * 2)] = 1 / std::pow(g.dz, 2);
}

struct Exchange : Term {
    double A;
    explicit Exchange(double a) : A(a) {}
    // _convolve_old_impl_exchange_H_in_Apm, exchange_field.cpp:115-134
    void H(const State& s, Vec& h) const override {
        const Mesh& g = s.mesh;
        const int64_t N = g.n();
        double f[27];
        exchange_filter(g, f);
        h.assign(3 * N, 0.0);
        for (int c = 0; c < 3; ++c) conv3(g, s.m.data() + c * N, f, h.data() + c * N);
        scale_by_ms(s, 2. * A, h);
    }
    // exchange_field.cpp:14-60: energy from the Neumann-corrected field h_withedges
    double E(const State& s, const Vec&) const override {
        const Mesh& g = s.mesh;
        const int64_t N = g.n();
        double f[27];
        exchange_filter(g, f);
        Vec h(3 * N, 0.0);
        for (int c = 0; c < 3; ++c) {
            double* hc = h.data() + c * N;
            const double* mc = s.m.data() + c * N;
            conv3(g, mc, f, hc);
            for (int64_t z = 0; z < g.nz; ++z)
                for (int64_t y = 0; y < g.ny; ++y) {
                    hc[g.at(0, y, z)] += mc[g.at(0, y, z)] / std::pow(g.dx, 2);
                    hc[g.at(g.nx - 1, y, z)] += mc[g.at(g.nx - 1, y, z)] / std::pow(g.dx, 2);
                }
            for (int64_t z = 0; z < g.nz; ++z)
                for (int64_t x = 0; x < g.nx; ++x) {
                    hc[g.at(x, 0, z)] += mc[g.at(x, 0, z)] / std::pow(g.dy, 2);
                    hc[g.at(x, g.ny - 1, z)] += mc[g.at(x, g.ny - 1, z)] / std::pow(g.dy, 2);
                }
            for (int64_t y = 0; y < g.ny; ++y)
                for (int64_t x = 0; x < g.nx; ++x) {
                    hc[g.at(x, y, 0)] += mc[g.at(x, y, 0)] / std::pow(g.dz, 2);
                    hc[g.at(x, y, g.nz - 1)] += mc[g.at(x, y, g.nz - 1)] / std::pow(g.dz, 2);
                }
        }
        scale_by_ms(s, 2. * A, h);
        return micro_energy(s, h, 0.5);
    }
};

// Central-difference filter of the DMI terms (dmi_field.cpp:53-64, atomistic_dmi_field.cpp:25-36):
// component c differentiates along axis c; `unit` selects the atomistic unit differences.
void dmi_filters(const Mesh& g, bool unit, double f[3][27]) {
    for (int c = 0; c < 3; ++c)
        for (int i = 0; i < 27; ++i) f[c][i] = 0;
    const double ax = unit ? 1.0 : 1 / (2. * g.dx);
    const double ay = unit ? 1.0 : 1 / (2. * g.dy);
    const double az = unit ? 1.0 : 1 / (2. * g.dz);
    f[0][0 + 3 * (1 + 3 * 1)] = ax;
    f[0][2 + 3 * (1 + 3 * 1)] = unit ? -1.0 : -1 / (2. * g.dx);
    f[1][1 + 3 * (0 + 3 * 1)] = ay;
    f[1][1 + 3 * (2 + 3 * 1)] = unit ? -1.0 : -1 / (2. * g.dy);
    f[2][1 + 3 * (1 + 3 * 0)] = az;
    f[2][1 + 3 * (1 + 3 * 2)] = unit ? -1.0 : -1 / (2. * g.dz);
}

// (div m) n - grad(n.m) with zero padding (dmi_field.cpp:66-86)
void dmi_core(const State& s, const std::array<double, 3>& n, bool unit, Vec& out) {
    const Mesh& g = s.mesh;
    const int64_t N = g.n();
    double f[3][27];
    dmi_filters(g, unit, f);
    Vec first(3 * N), div(N, 0.0);
    for (int c = 0; c < 3; ++c) conv3(g, s.m.data() + c * N, f[c], first.data() + c * N);
    for (int64_t i = 0; i < N; ++i) div[i] = first[i] + first[N + i] + first[2 * N + i];
    Vec nm(N);
    for (int64_t i = 0; i < N; ++i) nm[i] = n[0] * s.m[i] + n[1] * s.m[N + i] + n[2] * s.m[2 * N + i];
    Vec second(3 * N);
    for (int c = 0; c < 3; ++c) conv3(g, nm.data(), f[c], second.data() + c * N);
    out.resize(3 * N);
    for (int c = 0; c < 3; ++c)
        for (int64_t i = 0; i < N; ++i) out[c * N + i] = n[c] * div[i] - second[c * N + i];
}

// Micro coefficient numerator(i) / (mu0 Ms) per cell, with the Ms_field branch that replaces cells
// of Ms == 0 by 0 (dmi_field.cpp:97-109, uniaxial_anisotropy_field.cpp:78-88): the spatially
// varying D_constants / Ku1_field variants of scale_by_ms.
void scale_by_ms_field(const State& s, const Vec& numer, Vec& h) {
    const int64_t N = s.mesh.n();
    for (int c = 0; c < 3; ++c)
        for (int64_t i = 0; i < N; ++i) {
            if (!s.has_Ms_field()) {
                h[c * N + i] = numer[i] / (mu0 * s.Ms) * h[c * N + i];
            } else {
                const double ms = s.Ms_field[i];
                h[c * N + i] = ms != 0 ? numer[i] / (mu0 * ms) * h[c * N + i] : 0.0;
            }
        }
}

struct Dmi : Term { // dmi_field.cpp:41-110 (scalar D, or the D_constants field of :17-19)
    double D;
    std::array<double, 3> axis;
    Vec Dfield; // N, empty for the scalar D
    Dmi(double d, const double a[3], const Vec& df) : D(d), axis{a[0], a[1], a[2]}, Dfield(df) {}
    void H(const State& s, Vec& h) const override {
        const double norm = std::sqrt(std::pow(axis[0], 2) + std::pow(axis[1], 2) + std::pow(axis[2], 2));
        const std::array<double, 3> n{axis[0] / norm, axis[1] / norm, axis[2] / norm};
        dmi_core(s, n, false, h);
        if (Dfield.empty()) {
            scale_by_ms(s, 2. * D, h);
        } else {
            Vec numer(Dfield.size());
            for (size_t i = 0; i < numer.size(); ++i) numer[i] = 2. * Dfield[i];
            scale_by_ms_field(s, numer, h);
        }
    }
};

// uniaxial_anisotropy_field.cpp:11-25,58-89: scalar Ku1 or Ku1_field, a uniform axis (normalised
// in the constructor) or an axis field (normalised per cell by util::normalize, a / |a|).
struct Uniaxial : Term {
    double Ku1;
    std::array<double, 3> eu;
    Vec Kfield; // N, empty for the scalar Ku1
    Vec eufield; // 3N, empty for the uniform axis
    Uniaxial(double k, const double a[3], const Vec& kf, const Vec& af) : Ku1(k), eu(normalized(a)), Kfield(kf) {
        if (!af.empty()) {
            const int64_t N = static_cast<int64_t>(af.size() / 3);
            eufield.resize(af.size());
            for (int64_t i = 0; i < N; ++i) {
                const double n = std::sqrt(af[i] * af[i] + af[N + i] * af[N + i] + af[2 * N + i] * af[2 * N + i]);
                for (int c = 0; c < 3; ++c) eufield[c * N + i] = af[c * N + i] / n;
            }
        }
    }
    void H(const State& s, Vec& h) const override {
        const int64_t N = s.mesh.n();
        h.resize(3 * N);
        for (int64_t i = 0; i < N; ++i) {
            double e[3];
            for (int c = 0; c < 3; ++c) e[c] = eufield.empty() ? eu[c] : eufield[c * N + i];
            const double an = e[0] * s.m[i] + e[1] * s.m[N + i] + e[2] * s.m[2 * N + i];
            for (int c = 0; c < 3; ++c) h[c * N + i] = e[c] * an;
        }
        if (Kfield.empty()) {
            scale_by_ms(s, 2. * Ku1, h);
        } else {
            Vec numer(Kfield.size());
            for (size_t i = 0; i < numer.size(); ++i) numer[i] = 2. * Kfield[i];
            scale_by_ms_field(s, numer, h);
        }
    }
};

// external_field.cpp:28-48: a constant field, or the callback of :13-17 evaluated at the State's
// time (here the callback writes the field into a host buffer, 3N values or one homogeneous 3-vector)
struct Zeeman : Term {
    Vec hext;
    mafx_field_fn cb{nullptr};
    void* cb_user{nullptr};
    bool cb_homogeneous{false};
    explicit Zeeman(const Vec& h) : hext(h) {}
    void H(const State& s, Vec& h) const override {
        if (!cb) {
            h = hext;
            return;
        }
        const int64_t N = s.mesh.n();
        h.assign(3 * N, 0.0);
        if (cb_homogeneous) {
            double v[3] = {0, 0, 0};
            if (cb(cb_user, s.t, s.m.data(), v, nullptr) != 0) throw std::runtime_error("field callback failed");
            for (int c = 0; c < 3; ++c)
                for (int64_t i = 0; i < N; ++i) h[c * N + i] = v[c];
        } else if (cb(cb_user, s.t, s.m.data(), h.data(), nullptr) != 0) {
            throw std::runtime_error("field callback failed");
        }
    }
    double E(const State& s, const Vec& h) const override { return micro_energy(s, h, 1.0); }
};

// RKKYExchangeField::calc_COO_matrix (rkky_exchange_field.cpp:195-337) applied as a sparse
// product (:19-29).  Entries are assembled in the reference's loop order.
struct Rkky : Term {
    std::vector<std::vector<std::pair<int64_t, double>>> rows; // per row: (col, value)
    Rkky(const Mesh& g, const Vec& rk, const Vec& A, const std::vector<uint32_t>& idx) {
        const int64_t N = g.n();
        rows.assign(3 * N, {});
        auto st = [&](int64_t x, int64_t y, int64_t z) { return g.at(x, y, z); };
        auto fi = [&](int64_t x, int64_t y, int64_t z, int c) { return st(x, y, z) + N * c; };
        auto index_of = [&](int64_t at) -> uint32_t { return idx.empty() ? 0u : idx[at]; };
        for (int im = 0; im < 3; ++im)
            for (int64_t i2 = 0; i2 < g.nz; ++i2)
                for (int64_t i1 = 0; i1 < g.ny; ++i1)
                    for (int64_t i0 = 0; i0 < g.nx; ++i0) {
                        const int64_t ind = fi(i0, i1, i2, im);
                        auto& row = rows[ind];
                        const double A_i = A[st(i0, i1, i2)];
                        auto inplane = [&](int64_t j0, int64_t j1) {
                            const double A_j = A[st(j0, j1, i2)];
                            if (A_i != 0)
                                row.push_back({fi(j0, j1, i2, im), (2. * A_i) / (mu0 * std::pow(g.dx, 2)) * 2. *
                                                                      A_j / (A_j + A_i)});
                        };
                        if (i0 < g.nx - 1) inplane(i0 + 1, i1);
                        if (i0 > 0) inplane(i0 - 1, i1);
                        if (i1 < g.ny - 1) inplane(i0, i1 + 1); // the reference uses dx^2 here (:243)
                        if (i1 > 0) inplane(i0, i1 - 1);        // and here (:254)
                        auto zcoupling = [&](int64_t j2) {
                            const double R_i = rk[st(i0, i1, i2)];
                            const double R_j = rk[st(i0, i1, j2)];
                            const uint32_t I_i = index_of(st(i0, i1, i2));
                            const uint32_t I_j = index_of(st(i0, i1, j2));
                            if ((I_i == I_j) && (R_i != 0) && (R_j != 0)) {
                                row.push_back({fi(i0, i1, j2, im), (2. * R_i) / (mu0 * std::pow(g.dz, 2))});
                            } else {
                                const double A_j = A[st(i0, i1, j2)];
                                if (A_i != 0)
                                    row.push_back({fi(i0, i1, j2, im), (2. * A_i) / (mu0 * std::pow(g.dz, 2)) *
                                                                           2. * A_j / (A_j + A_i)});
                            }
                        };
                        if (i2 < g.nz - 1) zcoupling(i2 + 1);
                        if (i2 > 0) zcoupling(i2 - 1);
                    }
    }
    void H(const State& s, Vec& h) const override {
        const int64_t N = s.mesh.n();
        h.assign(3 * N, 0.0);
        for (int64_t r = 0; r < 3 * N; ++r) {
            double acc = 0;
            for (auto& e : rows[r]) acc += e.second * s.m[e.first];
            h[r] = acc;
        }
        if (!s.has_Ms_field()) {
            for (auto& v : h) v = v / s.Ms;
        } else {
            for (int c = 0; c < 3; ++c)
                for (int64_t i = 0; i < N; ++i)
                    h[c * N + i] = s.Ms_field[i] != 0 ? h[c * N + i] / s.Ms_field[i] : 0.0;
        }
    }
};

// ------------------------------------------------------------------ atomistic terms
double atom_energy(const State& s, const Vec& h, double factor) { // atom_term.hpp:12-15
    return -mu0 * factor * s.Ms * sum_hm(s, h, false);
}

struct AtomDipole : Term { // atomistic_dipole_dipole_field.cpp:14-113
    TensorConv tc;
    explicit AtomDipole(const Mesh& g) {
        const int64_t n0 = nexp_xy(g.nx), n1 = nexp_xy(g.ny), n2 = nexp_z(g.nz), P = n0 * n1 * n2;
        Vec N(6 * P, 0.0);
        for (int64_t i0 = 0; i0 < n0; ++i0) {
            const int64_t j0 = (i0 + n0 / 2) % n0 - n0 / 2;
            for (int64_t i1 = 0; i1 < n1; ++i1) {
                const int64_t j1 = (i1 + n1 / 2) % n1 - n1 / 2;
                for (int64_t i2 = 0; i2 < n2; ++i2) {
                    const int64_t j2 = (i2 + n2 / 2) % n2 - n2 / 2;
                    const int64_t at = i0 + n0 * (i1 + n1 * i2);
                    const double rx = j0 * g.dx, ry = j1 * g.dy, rz = j2 * g.dz;
                    const double r = std::sqrt(std::pow(rx, 2) + std::pow(ry, 2) + std::pow(rz, 2));
                    if (r == 0.) continue;
                    N[0 * P + at] = 3. * rx * rx / std::pow(r, 5) - 1. / std::pow(r, 3);
                    N[1 * P + at] = 3. * rx * ry / std::pow(r, 5);
                    N[2 * P + at] = 3. * rx * rz / std::pow(r, 5);
                    N[3 * P + at] = 3. * ry * ry / std::pow(r, 5) - 1. / std::pow(r, 3);
                    N[4 * P + at] = 3. * ry * rz / std::pow(r, 5);
                    N[5 * P + at] = 3. * rz * rz / std::pow(r, 5) - 1. / std::pow(r, 3);
                }
            }
        }
        const double s4pi = 1. / (4. * kPi);
        for (auto& v : N) v *= s4pi;
        tc.init(g, N);
    }
    void H(const State& s, Vec& h) const override {
        tc.apply(s.m, h);
        for (auto& v : h) v = s.Ms * v;
    }
    double E(const State& s, const Vec& h) const override { return atom_energy(s, h, 0.5); }
};

struct AtomExchange : Term { // atomistic_exchange_field.cpp:8-40
    double J;
    explicit AtomExchange(double j) : J(j) {}
    void H(const State& s, Vec& h) const override {
        const Mesh& g = s.mesh;
        const int64_t N = g.n();
        double f[27] = {0};
        f[0 + 3 * (1 + 3 * 1)] = 1.;
        f[2 + 3 * (1 + 3 * 1)] = 1.;
        f[1 + 3 * (0 + 3 * 1)] = 1.;
        f[1 + 3 * (2 + 3 * 1)] = 1.;
        f[1 + 3 * (1 + 3 * 0)] = 1.;
        f[1 + 3 * (1 + 3 * 2)] = 1.;
        h.assign(3 * N, 0.0);
        for (int c = 0; c < 3; ++c) conv3(g, s.m.data() + c * N, f, h.data() + c * N);
        const double coef = J / (mu0 * s.Ms);
        for (auto& v : h) v = coef * v;
    }
    double E(const State& s, const Vec& h) const override { return atom_energy(s, h, 0.5); }
};

struct AtomDmi : Term { // atomistic_dmi_field.cpp:6-49
    double D;
    std::array<double, 3> n;
    AtomDmi(double d, const double a[3]) : D(d), n(normalized(a)) {}
    void H(const State& s, Vec& h) const override {
        dmi_core(s, n, true, h);
        const double coef = D / (mu0 * s.Ms);
        for (auto& v : h) v = coef * v;
    }
    double E(const State& s, const Vec& h) const override { return atom_energy(s, h, 0.5); }
};

struct AtomUniaxial : Term { // atomistic_uniaxial_anisotropy_field.cpp:6-34
    double K;
    std::array<double, 3> eu;
    AtomUniaxial(double k, const double a[3]) : K(k), eu(normalized(a)) {}
    void H(const State& s, Vec& h) const override {
        const int64_t N = s.mesh.n();
        h.resize(3 * N);
        const double coef = 2 * K / (mu0 * s.Ms);
        for (int64_t i = 0; i < N; ++i) {
            const double an = eu[0] * s.m[i] + eu[1] * s.m[N + i] + eu[2] * s.m[2 * N + i];
            for (int c = 0; c < 3; ++c) h[c * N + i] = coef * an * eu[c];
        }
    }
    double E(const State& s, const Vec& h) const override { return atom_energy(s, h, 0.5); }
};

struct AtomZeeman : Zeeman { // atomistic_external_field.cpp:5-11 (the field of ExternalField)
    explicit AtomZeeman(const Vec& h) : Zeeman(h) {}
    double E(const State& s, const Vec& h) const override { return atom_energy(s, h, 1.0); }
};

} // namespace

std::unique_ptr<Term> make_demag(const Mesh& g, int nthreads) { return std::make_unique<Demag>(g, nthreads); }
std::unique_ptr<Term> make_exchange(double A) { return std::make_unique<Exchange>(A); }
std::unique_ptr<Term> make_dmi(double D, const double axis[3], const Vec& Dfield) {
    return std::make_unique<Dmi>(D, axis, Dfield);
}
std::unique_ptr<Term> make_uniaxial(double K, const double axis[3], const Vec& Kfield, const Vec& axis_field) {
    return std::make_unique<Uniaxial>(K, axis, Kfield, axis_field);
}
std::unique_ptr<Term> make_zeeman(const Vec& h) { return std::make_unique<Zeeman>(h); }
std::unique_ptr<Term> make_zeeman_callback(mafx_field_fn cb, void* user, bool homogeneous, bool atomistic) {
    auto z = std::make_unique<Zeeman>(Vec{});
    z->cb = cb;
    z->cb_user = user;
    z->cb_homogeneous = homogeneous;
    if (!atomistic) return z;
    auto a = std::make_unique<AtomZeeman>(Vec{});
    a->cb = cb;
    a->cb_user = user;
    a->cb_homogeneous = homogeneous;
    return a;
}
std::unique_ptr<Term> make_rkky(const Mesh& g, const Vec& rk, const Vec& A, const std::vector<uint32_t>& idx) {
    return std::make_unique<Rkky>(g, rk, A, idx);
}
std::unique_ptr<Term> make_atom_dipole(const Mesh& g) { return std::make_unique<AtomDipole>(g); }
std::unique_ptr<Term> make_atom_exchange(double J) { return std::make_unique<AtomExchange>(J); }
std::unique_ptr<Term> make_atom_dmi(double D, const double a[3]) { return std::make_unique<AtomDmi>(D, a); }
std::unique_ptr<Term> make_atom_uniaxial(double K, const double a[3]) { return std::make_unique<AtomUniaxial>(K, a); }
std::unique_ptr<Term> make_atom_zeeman(const Vec& h) { return std::make_unique<AtomZeeman>(h); }

// field_term.hpp:116-123 (sequential accumulation in term order)
void heff(const Terms& terms, const State& s, Vec& h) {
    if (terms.empty()) throw std::runtime_error("heff: no terms");
    terms[0]->H(s, h);
    Vec t;
    for (size_t i = 1; i < terms.size(); ++i) {
        terms[i]->H(s, t);
        for (size_t j = 0; j < h.size(); ++j) h[j] = h[j] + t[j];
    }
}
double eeff(const Terms& terms, const State& s) { // field_term.hpp:180-183
    double e = 0;
    for (auto& t : terms) e = e + t->E(s);
    return e;
}
double heff_and_energy(const Terms& terms, const State& s, Vec& h) { // field_term.hpp:196-200
    if (terms.empty()) throw std::runtime_error("heff_and_energy: no terms");
    terms[0]->H(s, h);
    double e = terms[0]->E(s, h);
    Vec t;
    for (size_t i = 1; i < terms.size(); ++i) {
        terms[i]->H(s, t);
        e = e + terms[i]->E(s, t);
        for (size_t j = 0; j < h.size(); ++j) h[j] = h[j] + t[j];
    }
    return e;
}

} // namespace oracle
