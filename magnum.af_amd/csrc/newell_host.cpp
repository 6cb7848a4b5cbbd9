// Host setup of the real-space demag and dipole tensors.
//
// The Newell tensor is evaluated on the CPU in exactly the reference's operation order
// (magnumaf/src/field_terms/micro/demag_field.cpp:20-142): at offsets of a few hundred cells the
// 27-term second difference of f/g cancels to below the rounding of its terms, so the far-field
// values are rounding noise whose exact bits depend on operation order and on the libm.  The
// reference's own FP64 result carries that noise; evaluating in the same order with the same libm
// (glibc) and without FMA contraction (this file is built with g++ -ffp-contract=off, as the
// reference's -O3 x86-64 build has no FMA) reproduces it bit for bit.  Threads split the
// expanded x range, as demag_field.cpp:125-137 does.
#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "../../include/mafx.h"

namespace mafx {

namespace {

// demag_field.cpp:20-40
double nf(double x, double y, double z) {
    x = std::fabs(x);
    y = std::fabs(y);
    z = std::fabs(z);
    const double R = std::sqrt(std::pow(x, 2) + std::pow(y, 2) + std::pow(z, 2));
    const double xx = std::pow(x, 2);
    const double yy = std::pow(y, 2);
    const double zz = std::pow(z, 2);
    double r = 1.0 / 6.0 * (2.0 * xx - yy - zz) * R;
    if (xx + zz > 0) r += y / 2.0 * (zz - xx) * std::asinh(y / (std::sqrt(xx + zz)));
    if (xx + yy > 0) r += z / 2.0 * (yy - xx) * std::asinh(z / (std::sqrt(xx + yy)));
    if (x * R > 0) r += -x * y * z * std::atan(y * z / (x * R));
    return r;
}

// demag_field.cpp:42-69
double ng(const double x, const double y, double z) {
    z = std::fabs(z);
    const double R = std::sqrt(std::pow(x, 2) + std::pow(y, 2) + std::pow(z, 2));
    const double xx = std::pow(x, 2);
    const double yy = std::pow(y, 2);
    const double zz = std::pow(z, 2);
    double r = -x * y * R / 3.0;
    if (xx + yy > 0) r += x * y * z * std::asinh(z / (std::sqrt(xx + yy)));
    if (yy + zz > 0) r += y / 6.0 * (3.0 * zz - yy) * std::asinh(x / (std::sqrt(yy + zz)));
    if (xx + zz > 0) r += x / 6.0 * (3.0 * zz - xx) * std::asinh(y / (std::sqrt(xx + zz)));
    if (z * R > 0) r += -std::pow(z, 3) / 6.0 * std::atan(x * y / (z * R));
    if (y * R != 0) r += -z * yy / 2.0 * std::atan(x * z / (y * R));
    if (x * R != 0) r += -z * xx / 2.0 * std::atan(y * z / (x * R));
    return r;
}

// Second difference over the 27 neighbouring cell pairs, accumulated left to right in the order
// of demag_field.cpp:75-83 (centre, 6 faces, 12 edges, 8 corners).  The shifted coordinates are
// formed as x + dx / x - dx from x = dx * ix, like the reference.
template <double (*F)(double, double, double)>
double second_diff(int ix, int iy, int iz, double dx, double dy, double dz) {
    const double x = dx * ix, y = dy * iy, z = dz * iz;
    const double xp = x + dx, xm = x - dx, yp = y + dy, ym = y - dy, zp = z + dz, zm = z - dz;
    double s = 8.0 * F(x, y, z);
    s = s - 4.0 * F(xp, y, z);
    s = s - 4.0 * F(xm, y, z);
    s = s - 4.0 * F(x, yp, z);
    s = s - 4.0 * F(x, ym, z);
    s = s - 4.0 * F(x, y, zp);
    s = s - 4.0 * F(x, y, zm);
    s = s + 2.0 * F(xp, yp, z);
    s = s + 2.0 * F(xp, ym, z);
    s = s + 2.0 * F(xm, yp, z);
    s = s + 2.0 * F(xm, ym, z);
    s = s + 2.0 * F(xp, y, zp);
    s = s + 2.0 * F(xp, y, zm);
    s = s + 2.0 * F(xm, y, zp);
    s = s + 2.0 * F(xm, y, zm);
    s = s + 2.0 * F(x, yp, zp);
    s = s + 2.0 * F(x, yp, zm);
    s = s + 2.0 * F(x, ym, zp);
    s = s + 2.0 * F(x, ym, zm);
    s = s - 1.0 * F(xp, yp, zp);
    s = s - 1.0 * F(xp, yp, zm);
    s = s - 1.0 * F(xp, ym, zp);
    s = s - 1.0 * F(xp, ym, zm);
    s = s - 1.0 * F(xm, yp, zp);
    s = s - 1.0 * F(xm, yp, zm);
    s = s - 1.0 * F(xm, ym, zp);
    s = s - 1.0 * F(xm, ym, zm);
    return -s / (4.0 * 3.14159265358979323846 * dx * dy * dz);
}

int64_t exp_xy(int64_t n) { return 2 * n; }
int64_t exp_z(int64_t n) { return n == 1 ? 1 : 2 * n; }

} // namespace

// [6][n2][n1][n0] with the wrapped index j = (i + n/2) % n - n/2 (demag_field.cpp:104-121)
void newell_tensor_host(const mafx_mesh& mesh, double* N, int nthreads) {
    const int64_t n0 = exp_xy(mesh.nx), n1 = exp_xy(mesh.ny), n2 = exp_z(mesh.nz);
    const int64_t P = n0 * n1 * n2;
    const double dx = mesh.dx, dy = mesh.dy, dz = mesh.dz;
    auto work = [&](int64_t b, int64_t e) {
        for (int64_t i0 = b; i0 < e; ++i0) {
            const int j0 = static_cast<int>((i0 + n0 / 2) % n0 - n0 / 2);
            for (int64_t i1 = 0; i1 < n1; ++i1) {
                const int j1 = static_cast<int>((i1 + n1 / 2) % n1 - n1 / 2);
                for (int64_t i2 = 0; i2 < n2; ++i2) {
                    const int j2 = static_cast<int>((i2 + n2 / 2) % n2 - n2 / 2);
                    const int64_t at = i0 + n0 * (i1 + n1 * i2);
                    N[0 * P + at] = second_diff<nf>(j0, j1, j2, dx, dy, dz);
                    N[1 * P + at] = second_diff<ng>(j0, j1, j2, dx, dy, dz);
                    N[2 * P + at] = second_diff<ng>(j0, j2, j1, dx, dz, dy);
                    N[3 * P + at] = second_diff<nf>(j1, j2, j0, dy, dz, dx);
                    N[4 * P + at] = second_diff<ng>(j1, j2, j0, dy, dz, dx);
                    N[5 * P + at] = second_diff<nf>(j2, j0, j1, dz, dx, dy);
                }
            }
        }
    };
    if (nthreads < 1) nthreads = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
    nthreads = static_cast<int>(std::min<int64_t>(nthreads, n0));
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) th.emplace_back(work, n0 * t / nthreads, n0 * (t + 1) / nthreads);
    for (auto& t : th) t.join();
}

// The non-negative octant of the same tensor, [6][m2][m1][m0] with m0 = nx, m1 = ny, m2 = nz (or
// 1 for a film), entries at offsets j >= 0 in the same operation order; and its extension to the
// expanded grid by the parities of the components (xx, yy, zz even; xy odd in x and y; xz odd in x
// and z; yz odd in y and z), with the planes j = -n/2, which have no mirror, set to zero.  The
// extension differs from newell_tensor_host only at negative offsets, by the summation order of
// the 27 terms (|delta| ~ 1e-10 of an entry); it costs an eighth of the evaluations.
void newell_octant_host(const mafx_mesh& mesh, double* O, int nthreads) {
    const int64_t m0 = mesh.nx, m1 = mesh.ny, m2 = mesh.nz;
    const int64_t Q = m0 * m1 * m2;
    const double dx = mesh.dx, dy = mesh.dy, dz = mesh.dz;
    auto work = [&](int64_t b, int64_t e) {
        for (int64_t j0 = b; j0 < e; ++j0)
            for (int64_t j1 = 0; j1 < m1; ++j1)
                for (int64_t j2 = 0; j2 < m2; ++j2) {
                    const int a = static_cast<int>(j0), c = static_cast<int>(j1), d = static_cast<int>(j2);
                    const int64_t at = j0 + m0 * (j1 + m1 * j2);
                    O[0 * Q + at] = second_diff<nf>(a, c, d, dx, dy, dz);
                    O[1 * Q + at] = second_diff<ng>(a, c, d, dx, dy, dz);
                    O[2 * Q + at] = second_diff<ng>(a, d, c, dx, dz, dy);
                    O[3 * Q + at] = second_diff<nf>(c, d, a, dy, dz, dx);
                    O[4 * Q + at] = second_diff<ng>(c, d, a, dy, dz, dx);
                    O[5 * Q + at] = second_diff<nf>(d, a, c, dz, dx, dy);
                }
    };
    if (nthreads < 1) nthreads = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
    nthreads = static_cast<int>(std::min<int64_t>(nthreads, m0));
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) th.emplace_back(work, m0 * t / nthreads, m0 * (t + 1) / nthreads);
    for (auto& t : th) t.join();
}

void newell_octant_extend(const mafx_mesh& mesh, const double* O, double* N) {
    const int64_t n0 = exp_xy(mesh.nx), n1 = exp_xy(mesh.ny), n2 = exp_z(mesh.nz);
    const int64_t m0 = mesh.nx, m1 = mesh.ny, m2 = mesh.nz;
    const int64_t P = n0 * n1 * n2, Q = m0 * m1 * m2;
    // odd[c][axis]: component c changes sign with the offset along axis
    static const bool odd[6][3] = {{false, false, false}, {true, true, false}, {true, false, true},
                                   {false, false, false}, {false, true, true}, {false, false, false}};
    for (int64_t i2 = 0; i2 < n2; ++i2) {
        const int64_t j2 = (i2 + n2 / 2) % n2 - n2 / 2;
        for (int64_t i1 = 0; i1 < n1; ++i1) {
            const int64_t j1 = (i1 + n1 / 2) % n1 - n1 / 2;
            for (int64_t i0 = 0; i0 < n0; ++i0) {
                const int64_t j0 = (i0 + n0 / 2) % n0 - n0 / 2;
                const int64_t at = i0 + n0 * (i1 + n1 * i2);
                const bool unpaired = (n0 > 1 && j0 == -n0 / 2) || (n1 > 1 && j1 == -n1 / 2) || (n2 > 1 && j2 == -n2 / 2);
                const int64_t src = (j0 < 0 ? -j0 : j0) + m0 * ((j1 < 0 ? -j1 : j1) + m1 * (j2 < 0 ? -j2 : j2));
                for (int c = 0; c < 6; ++c) {
                    if (unpaired) {
                        N[c * P + at] = 0.0;
                        continue;
                    }
                    const bool neg = (odd[c][0] && j0 < 0) != (odd[c][1] && j1 < 0) != (odd[c][2] && j2 < 0);
                    const double v = O[c * Q + src];
                    N[c * P + at] = neg ? -v : v;
                }
            }
        }
    }
}

// atomistic_dipole_dipole_field.cpp:55-113: N = (3 r r^T / r^5 - I / r^3) / (4 pi), N(0) = 0
void dipole_tensor_host(const mafx_mesh& mesh, double* N) {
    const int64_t n0 = exp_xy(mesh.nx), n1 = exp_xy(mesh.ny), n2 = exp_z(mesh.nz);
    const int64_t P = n0 * n1 * n2;
    for (int64_t i0 = 0; i0 < n0; ++i0) {
        const int64_t j0 = (i0 + n0 / 2) % n0 - n0 / 2;
        for (int64_t i1 = 0; i1 < n1; ++i1) {
            const int64_t j1 = (i1 + n1 / 2) % n1 - n1 / 2;
            for (int64_t i2 = 0; i2 < n2; ++i2) {
                const int64_t j2 = (i2 + n2 / 2) % n2 - n2 / 2;
                const int64_t at = i0 + n0 * (i1 + n1 * i2);
                const double rx = j0 * mesh.dx, ry = j1 * mesh.dy, rz = j2 * mesh.dz;
                const double r = std::sqrt(std::pow(rx, 2) + std::pow(ry, 2) + std::pow(rz, 2));
                double v[6] = {0, 0, 0, 0, 0, 0};
                if (r != 0.) {
                    v[0] = 3. * rx * rx / std::pow(r, 5) - 1. / std::pow(r, 3);
                    v[1] = 3. * rx * ry / std::pow(r, 5);
                    v[2] = 3. * rx * rz / std::pow(r, 5);
                    v[3] = 3. * ry * ry / std::pow(r, 5) - 1. / std::pow(r, 3);
                    v[4] = 3. * ry * rz / std::pow(r, 5);
                    v[5] = 3. * rz * rz / std::pow(r, 5) - 1. / std::pow(r, 3);
                }
                const double s = 1. / (4. * 3.14159265358979323846);
                for (int c = 0; c < 6; ++c) N[c * P + at] = v[c] * s;
            }
        }
    }
}

} // namespace mafx
