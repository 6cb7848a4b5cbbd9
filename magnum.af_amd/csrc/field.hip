// Effective-field evaluation: demag tensor setup, the zero-padded FFT convolution (its last pass,
// the inverse x transform with C2R + crop + 1/P, writes the convolution field), and one per-cell
// kernel that adds every local term in term order, then writes H, the LLG right-hand side, or the
// LBFGS gradient, and accumulates per-term energy partial sums.  Running the inverse x transform
// and the stencils as two kernels lets each hold its best occupancy (an LDS transform and a
// register-heavy stencil in one kernel left two waves per SIMD); the extra field round trip costs
// less than that loss at 1024 x 1024 x 4.
//
// Replaces (reference paths relative to magnumaf/src):
//   field_terms/micro/demag_field.cpp:104-264, exchange_field.cpp:14-60,115-157,
//   dmi_field.cpp:41-110, uniaxial_anisotropy_field.cpp:58-89, external_field.cpp:28-48,
//   rkky_exchange_field.cpp:19-29,195-337, field_terms/atom/*.cpp,
//   include/magnumaf/field_terms/field_term.hpp:116-150, micro/micro_term.hpp:18-30,
//   include/magnumaf/equations.hpp:19-33, solvers/lbfgs_minimizer.cpp:33-42.
#include <cmath>
#include <cstring>
#include <initializer_list>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "fft_device.hpp"
#include "fft_kernels.hpp"
#include "mafx_internal.hpp"
#include "vec_kernels.hpp"
#include "tensor_cache.hpp"

namespace mafx {

constexpr int kMaxTerms = 8;
constexpr int kBlock = 256;

struct EpiTerm {
    int kind;
    double c0;  // coefficient with the scalar Ms (host-computed in the reference's order)
    double num; // numerator for the per-cell Ms_field variant: coef = num / (mu0 Ms_i)
    double n[3];
    double f[3];
    const double* zee;
    const double* rk;
    const double* aex;
    const uint32_t* ridx;
    const void* rkw; // 6N RKKY weights per neighbour in the stencils' type T, or null
    const double* pf; // N: D_constants (DMI) or Ku1_field (UNIAXIAL), or null
    const double* ax; // 3N: normalised Ku1_axis_field (UNIAXIAL), or null
};

struct EpiParams {
    int64_t nx, ny, nz, N;
    const double* m;
    const double* ms; // Ms_field or null
    double Ms;
    double ex_w[3], ex_self; // exchange filter (exchange_field.cpp:116-125)
    double dmi_c[3];         // 1 / (2 d) central difference weights (dmi_field.cpp:53-64)
    double rk_pref[3];       // mu0 * d^2 of the RKKY couplings: dx^2 for x AND y, dz^2 for z
    int nterms;
    int fft_slot; // term index fed by the FFT convolution, -1 if none
    int needs_nb; // some selected term reads the six neighbours (loaded with the centre cell)
    int has_fields; // some selected term has a parameter field (run-time term loop only)
    EpiTerm t[kMaxTerms];
    int mode;
    void* out;
    int out_f32;
    double alpha, dt;
    const double* alpha_f; // N cell-wise damping or null
    int diss_only;
    int want_energy;
    double* partials; // [nslots][nblocks]
    int64_t nblocks;
    const void* hd; // FFT-convolution field [3][N] of type T for the per-cell kernel, or null
    // LLG mode: next RK stage input (EvalArgs::rk_*), rk_out null when not fused
    const double* rk_base;
    const void* rk_k[kMaxRkFused]; // double, or float when k_f32
    int k_f32;                     // k written (and rk_k read) as float32
    double rk_c[kMaxRkFused]; // weights of rk_k
    double rk_clast;          // weight of this pass's own k
    int rk_n;
    double* rk_out;
};

namespace {

// Cell arrays are read and written as (uniform plane base, 32-bit byte offset): the address
// then lives in one VGPR next to an SGPR base (global_load/store with saddr) instead of a
// 64-bit VGPR pair per access.  Valid while one component plane is below 4 GiB (N < 2^29).
template <class U> __device__ __forceinline__ U ldb(const U* base, uint32_t boff) {
    return *reinterpret_cast<const U*>(reinterpret_cast<const char*>(base) + boff);
}
template <class U> __device__ __forceinline__ void stb(U* base, uint32_t boff, U v) {
    *reinterpret_cast<U*>(reinterpret_cast<char*>(base) + boff) = v;
}

template <class T> struct V3 {
    T x, y, z;
};

// Sum over the block of `nv` per-thread values with a fixed tree, written to partials[s][blk].
__device__ void block_reduce_store(double* vals, int nv, const int* is_max, double* partials, int64_t nblocks) {
    __shared__ double red[16][kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int s = 0; s < nv; ++s) {
        double v = vals[s];
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_xor(v, off, 64);
            v = is_max[s] ? fmax(v, o) : v + o;
        }
        if (lane == 0) red[s][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int s = 0; s < nv; ++s) {
            double v = red[s][0];
            for (int w = 1; w < kBlock / 64; ++w) v = is_max[s] ? fmax(v, red[s][w]) : v + red[s][w];
            partials[s * nblocks + blockIdx.x] = v;
        }
    }
}

// All local terms and the output stage for one cell.  hd holds the FFT-convolution field of the
// cell (ignored when P.fft_slot < 0).  With RED (energies or the gradient mode's maxima wanted)
// acc[] collects the energy partials (slots 0..nterms-1), acc[nterms] = max |m x H|,
// acc[nterms+1] = max |H| (gradient mode); without RED acc is not touched, which keeps the
// accumulators out of the register budget of the plain field and LLG evaluations.
// PROG = 0 reads the term kinds from P at run time; otherwise it lists them at compile time
// (prog_of: count in bits 0-3, kind of term t in bits 4 + 4t ..), so the term loop unrolls and
// each term compiles to its own branch-free code.
constexpr uint64_t prog_of(std::initializer_list<int> kinds) {
    uint64_t p = kinds.size();
    int t = 0;
    for (int k : kinds) p |= static_cast<uint64_t>(k) << (4 + 4 * t++);
    return p;
}
constexpr int prog_count(uint64_t prog) { return static_cast<int>(prog & 15u); }
constexpr int prog_kind(uint64_t prog, int t) { return static_cast<int>((prog >> (4 + 4 * t)) & 15u); }

template <int I, int N, class F> __device__ __forceinline__ void static_for(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

template <class T, bool RED, uint64_t PROG>
__device__ __forceinline__ void cell_terms(const EpiParams& P, int64_t x, int64_t y, int64_t z, double mc0,
                                           double mc1, double mc2, const V3<T> (&nb)[6], const T hd[3], double* acc);

template <class T, bool RED, uint64_t PROG = 0>
__device__ __forceinline__ void cell_epilogue(const EpiParams& P, int64_t x, int64_t y, int64_t z, const T hd[3],
                                              double* acc) {
    const int64_t N = P.N;
    const int64_t i = x + P.nx * (y + P.ny * z);
    const double* mp[3] = {P.m, P.m + N, P.m + 2 * N};
    const uint32_t o8 = static_cast<uint32_t>(i) * 8u;
    const double mc0 = ldb(mp[0], o8), mc1 = ldb(mp[1], o8), mc2 = ldb(mp[2], o8);

    // Neighbours with zero padding (af::convolve AF_CONV_DEFAULT semantics), x-, x+, y-, y+,
    // z-, z+.  A missing neighbour reads the centre cell and is then replaced by 0, so all 18
    // loads are issued together, without branches, with the centre cell's: one memory round
    // trip per cell.  P.needs_nb is set whenever a selected term reads nb.
    V3<T> nb[6];
    if (P.needs_nb) {
        const bool ok[6] = {x > 0, x + 1 < P.nx, y > 0, y + 1 < P.ny, z > 0, z + 1 < P.nz};
        const uint32_t sy = static_cast<uint32_t>(P.nx) * 8u, sz = static_cast<uint32_t>(P.nx * P.ny) * 8u;
        const uint32_t jo[6] = {ok[0] ? o8 - 8u : o8, ok[1] ? o8 + 8u : o8, ok[2] ? o8 - sy : o8,
                                ok[3] ? o8 + sy : o8, ok[4] ? o8 - sz : o8, ok[5] ? o8 + sz : o8};
        double v[6][3];
#pragma unroll
        for (int k = 0; k < 6; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) v[k][c] = ldb(mp[c], jo[k]);
#pragma unroll
        for (int k = 0; k < 6; ++k)
            nb[k] = ok[k] ? V3<T>{static_cast<T>(v[k][0]), static_cast<T>(v[k][1]), static_cast<T>(v[k][2])}
                          : V3<T>{0, 0, 0};
    } else {
#pragma unroll
        for (int k = 0; k < 6; ++k) nb[k] = V3<T>{0, 0, 0};
    }
    cell_terms<T, RED, PROG>(P, x, y, z, mc0, mc1, mc2, nb, hd, acc);
}

// The local terms and the output stage of cell (x, y, z) from its magnetisation mc (FP64), its six
// neighbours nb (x-, x+, y-, y+, z-, z+, zero where the mesh ends) and the convolution field hd.
template <class T, bool RED, uint64_t PROG>
__device__ __forceinline__ void cell_terms(const EpiParams& P, int64_t x, int64_t y, int64_t z, double mc0,
                                           double mc1, double mc2, const V3<T> (&nb)[6], const T hd[3], double* acc) {
    const int64_t N = P.N;
    const int64_t i = x + P.nx * (y + P.ny * z);
    const uint32_t o8 = static_cast<uint32_t>(i) * 8u;
    const T m0 = static_cast<T>(mc0), m1 = static_cast<T>(mc1), m2 = static_cast<T>(mc2);
    const double msi = P.ms ? P.ms[i] : P.Ms;
    T H0 = 0, H1 = 0, H2 = 0;
    auto term = [&](int t, int kind) __attribute__((always_inline)) {
        const EpiTerm& tm = P.t[t];
        T h0 = 0, h1 = 0, h2 = 0;
        T e_extra = 0; // exchange energy boundary correction (h_withedges - h) . m
        // micro coefficient with the Ms_field branch (cells with Ms == 0 give 0); the compile-time
        // term programs are only used for lists without parameter fields
        auto micro_coef = [&](T& c) -> bool {
            if (prog_count(PROG) == 0 && tm.pf) { // D_constants / Ku1_field: 2 X_i / (mu0 Ms)
                const double num = 2. * tm.pf[i];
                if (!P.ms) {
                    c = static_cast<T>(num / (kMu0 * P.Ms));
                    return true;
                }
                if (msi == 0) return false;
                c = static_cast<T>(num / (kMu0 * msi));
                return true;
            }
            if (!P.ms) {
                c = static_cast<T>(tm.c0);
                return true;
            }
            if (msi == 0) return false;
            c = static_cast<T>(tm.num / (kMu0 * msi));
            return true;
        };
        switch (kind) {
        case MAFX_DEMAG:
        case MAFX_ATOM_DIPOLE:
            h0 = hd[0];
            h1 = hd[1];
            h2 = hd[2];
            break;
        case MAFX_EXCHANGE: {
            T c;
            if (micro_coef(c)) {
                const T wx = static_cast<T>(P.ex_w[0]), wy = static_cast<T>(P.ex_w[1]), wz = static_cast<T>(P.ex_w[2]);
                const T sf = static_cast<T>(P.ex_self);
                const T l0 = sf * m0 + wx * (nb[0].x + nb[1].x) + wy * (nb[2].x + nb[3].x) + wz * (nb[4].x + nb[5].x);
                const T l1 = sf * m1 + wx * (nb[0].y + nb[1].y) + wy * (nb[2].y + nb[3].y) + wz * (nb[4].y + nb[5].y);
                const T l2 = sf * m2 + wx * (nb[0].z + nb[1].z) + wy * (nb[2].z + nb[3].z) + wz * (nb[4].z + nb[5].z);
                h0 = c * l0;
                h1 = c * l1;
                h2 = c * l2;
                if (RED && P.want_energy) {
                    // Neumann correction of exchange_field.cpp:44-51: + m / d^2 on every boundary face
                    T corr = 0;
                    const T mm = m0 * m0 + m1 * m1 + m2 * m2;
                    if (x == 0) corr += wx * mm;
                    if (x == P.nx - 1) corr += wx * mm;
                    if (y == 0) corr += wy * mm;
                    if (y == P.ny - 1) corr += wy * mm;
                    if (z == 0) corr += wz * mm;
                    if (z == P.nz - 1) corr += wz * mm;
                    e_extra = c * corr;
                }
            }
            break;
        }
        case MAFX_DMI:
        case MAFX_ATOM_DMI: {
            T c;
            bool ok;
            T cx, cy, cz;
            if (kind == MAFX_DMI) {
                ok = micro_coef(c);
                cx = static_cast<T>(P.dmi_c[0]);
                cy = static_cast<T>(P.dmi_c[1]);
                cz = static_cast<T>(P.dmi_c[2]);
            } else {
                ok = true;
                c = static_cast<T>(tm.c0);
                cx = cy = cz = 1;
            }
            if (ok) {
                const T n0 = static_cast<T>(tm.n[0]), n1 = static_cast<T>(tm.n[1]), n2 = static_cast<T>(tm.n[2]);
                // div m with central differences (f[i+1] - f[i-1]) / (2 d)
                const T div = (nb[1].x - nb[0].x) * cx + (nb[3].y - nb[2].y) * cy + (nb[5].z - nb[4].z) * cz;
                auto ndot = [&](const V3<T>& v) { return n0 * v.x + n1 * v.y + n2 * v.z; };
                const T g0 = (ndot(nb[1]) - ndot(nb[0])) * cx;
                const T g1 = (ndot(nb[3]) - ndot(nb[2])) * cy;
                const T g2 = (ndot(nb[5]) - ndot(nb[4])) * cz;
                h0 = c * (n0 * div - g0);
                h1 = c * (n1 * div - g1);
                h2 = c * (n2 * div - g2);
            }
            break;
        }
        case MAFX_UNIAXIAL: {
            T c;
            if (micro_coef(c)) {
                T u0 = static_cast<T>(tm.n[0]), u1 = static_cast<T>(tm.n[1]), u2 = static_cast<T>(tm.n[2]);
                if (prog_count(PROG) == 0 && tm.ax) {
                    u0 = static_cast<T>(tm.ax[i]);
                    u1 = static_cast<T>(tm.ax[N + i]);
                    u2 = static_cast<T>(tm.ax[2 * N + i]);
                }
                const T an = u0 * m0 + u1 * m1 + u2 * m2;
                h0 = c * (u0 * an);
                h1 = c * (u1 * an);
                h2 = c * (u2 * an);
            }
            break;
        }
        case MAFX_ATOM_UNIAXIAL: {
            const T u0 = static_cast<T>(tm.n[0]), u1 = static_cast<T>(tm.n[1]), u2 = static_cast<T>(tm.n[2]);
            const T an = u0 * m0 + u1 * m1 + u2 * m2;
            const T c = static_cast<T>(tm.c0);
            h0 = c * an * u0;
            h1 = c * an * u1;
            h2 = c * an * u2;
            break;
        }
        case MAFX_ZEEMAN:
        case MAFX_ATOM_ZEEMAN:
            if (tm.zee) {
                h0 = static_cast<T>(tm.zee[i]);
                h1 = static_cast<T>(tm.zee[N + i]);
                h2 = static_cast<T>(tm.zee[2 * N + i]);
            } else {
                h0 = static_cast<T>(tm.f[0]);
                h1 = static_cast<T>(tm.f[1]);
                h2 = static_cast<T>(tm.f[2]);
            }
            break;
        case MAFX_ATOM_EXCHANGE: {
            const T c = static_cast<T>(tm.c0);
            h0 = c * (nb[0].x + nb[1].x + nb[2].x + nb[3].x + nb[4].x + nb[5].x);
            h1 = c * (nb[0].y + nb[1].y + nb[2].y + nb[3].y + nb[4].y + nb[5].y);
            h2 = c * (nb[0].z + nb[1].z + nb[2].z + nb[3].z + nb[4].z + nb[5].z);
            break;
        }
        case MAFX_RKKY: {
            // rkky_exchange_field.cpp:195-337 entries, applied as a stencil, then / Ms (:19-29)
            double a0 = 0, a1 = 0, a2 = 0;
            bool pre = false;
            // the neighbour weights precomputed once per term (k_rkky_weights), summed in the
            // order of the in-kernel form below (an absent neighbour's weight is 0)
            if (tm.rkw) {
                pre = true;
                const T* wp = static_cast<const T*>(tm.rkw);
                const uint32_t ot = static_cast<uint32_t>(i) * static_cast<uint32_t>(sizeof(T));
                constexpr int ord[6] = {1, 0, 3, 2, 5, 4};
                T w[6];
#pragma unroll
                for (int k = 0; k < 6; ++k) w[k] = ldb(wp + ord[k] * N, ot);
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    const double wk = static_cast<double>(w[k]);
                    a0 += wk * static_cast<double>(nb[ord[k]].x);
                    a1 += wk * static_cast<double>(nb[ord[k]].y);
                    a2 += wk * static_cast<double>(nb[ord[k]].z);
                }
            }
            const int64_t sxy = P.nx * P.ny;
            const double Ai = pre ? 0.0 : tm.aex[i];
            if (!pre && Ai != 0) {
                auto inplane = [&](bool ok, int64_t j, int k, double pref) {
                    if (!ok) return;
                    const double Aj = tm.aex[j];
                    const double w = (2. * Ai) / pref * 2. * Aj / (Aj + Ai);
                    a0 += w * static_cast<double>(nb[k].x);
                    a1 += w * static_cast<double>(nb[k].y);
                    a2 += w * static_cast<double>(nb[k].z);
                };
                inplane(x + 1 < P.nx, i + 1, 1, P.rk_pref[0]);
                inplane(x > 0, i - 1, 0, P.rk_pref[0]);
                inplane(y + 1 < P.ny, i + P.nx, 3, P.rk_pref[1]);
                inplane(y > 0, i - P.nx, 2, P.rk_pref[1]);
            }
            auto zc = [&](bool ok, int64_t j, int k) {
                if (!ok) return;
                const double Ri = tm.rk[i], Rj = tm.rk[j];
                const uint32_t Ii = tm.ridx ? tm.ridx[i] : 0u, Ij = tm.ridx ? tm.ridx[j] : 0u;
                double w;
                if ((Ii == Ij) && (Ri != 0) && (Rj != 0)) {
                    w = (2. * Ri) / P.rk_pref[2];
                } else {
                    if (Ai == 0) return;
                    const double Aj = tm.aex[j];
                    w = (2. * Ai) / P.rk_pref[2] * 2. * Aj / (Aj + Ai);
                }
                a0 += w * static_cast<double>(nb[k].x);
                a1 += w * static_cast<double>(nb[k].y);
                a2 += w * static_cast<double>(nb[k].z);
            };
            if (!pre) {
                zc(z + 1 < P.nz, i + sxy, 5);
                zc(z > 0, i - sxy, 4);
            }
            if (!P.ms) {
                h0 = static_cast<T>(a0 / P.Ms);
                h1 = static_cast<T>(a1 / P.Ms);
                h2 = static_cast<T>(a2 / P.Ms);
            } else if (msi != 0) {
                h0 = static_cast<T>(a0 / msi);
                h1 = static_cast<T>(a1 / msi);
                h2 = static_cast<T>(a2 / msi);
            }
            break;
        }
        default: break;
        }
        if (t == 0) {
            H0 = h0;
            H1 = h1;
            H2 = h2;
        } else {
            H0 = H0 + h0;
            H1 = H1 + h1;
            H2 = H2 + h2;
        }
        if (RED && P.want_energy) {
            const double w = (P.ms && kind < MAFX_ATOM_DIPOLE) ? msi : 1.0;
            acc[t] += w * (static_cast<double>(h0) * mc0 + static_cast<double>(h1) * mc1 +
                           static_cast<double>(h2) * mc2 + static_cast<double>(e_extra));
        }
    };
    if constexpr (prog_count(PROG) > 0) {
        static_for<0, prog_count(PROG)>([&](auto tc) { term(tc.value, prog_kind(PROG, tc.value)); });
    } else {
        for (int t = 0; t < P.nterms; ++t) term(t, P.t[t].kind);
    }
    const double Hx = static_cast<double>(H0), Hy = static_cast<double>(H1), Hz = static_cast<double>(H2);
    switch (P.mode) {
    case static_cast<int>(OutMode::H):
        if (P.out_f32) {
            float* o = static_cast<float*>(P.out);
            const uint32_t o4 = static_cast<uint32_t>(i) * 4u;
            stb(o, o4, static_cast<float>(H0));
            stb(o + N, o4, static_cast<float>(H1));
            stb(o + 2 * N, o4, static_cast<float>(H2));
        } else {
            double* o = static_cast<double*>(P.out);
            stb(o, o8, Hx);
            stb(o + N, o8, Hy);
            stb(o + 2 * N, o8, Hz);
        }
        break;
    case static_cast<int>(OutMode::LLG): {
        // equations.hpp:19-33 with math::cross4 (math.cpp:10-19); k = dt * f (adaptive_runge_kutta.cpp)
        const double a0 = mc1 * Hz - mc2 * Hy, a1 = mc2 * Hx - mc0 * Hz, a2 = mc0 * Hy - mc1 * Hx;
        const double b0 = mc1 * a2 - mc2 * a1, b1 = mc2 * a0 - mc0 * a2, b2 = mc0 * a1 - mc1 * a0;
        // llg_integrator.cpp:18-38: scalar alpha, or the cell's entry of alpha_field
        const double al = P.alpha_f ? P.alpha_f[i] : P.alpha;
        const double a2p = al * al;
        const double cp = -kGamma / (1. + a2p);
        const double cd = -al * kGamma / (1. + a2p);
        double k[3];
        if (P.diss_only) {
            k[0] = P.dt * (cd * b0);
            k[1] = P.dt * (cd * b1);
            k[2] = P.dt * (cd * b2);
        } else {
            k[0] = P.dt * (cp * a0 + cd * b0);
            k[1] = P.dt * (cp * a1 + cd * b1);
            k[2] = P.dt * (cp * a2 + cd * b2);
        }
        if (P.k_f32) {
            // stored as float32; the next stage input below uses the stored value
            float* o = static_cast<float*>(P.out);
            const uint32_t o4 = static_cast<uint32_t>(i) * 4u;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float kf = static_cast<float>(k[c]);
                k[c] = static_cast<double>(kf);
                stb(o + c * N, o4, kf);
            }
        } else {
            double* o = static_cast<double*>(P.out);
            stb(o, o8, k[0]);
            stb(o + N, o8, k[1]);
            stb(o + 2 * N, o8, k[2]);
        }
        if (P.rk_out) {
            // the next stage's input m + sum_j a_j k_j (adaptive_runge_kutta.cpp:138-164), j in
            // order with this stage's k last, each term added as the stage combination adds it
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int64_t cN = c * N; // component plane (8 N bytes < 2^32 each)
                double s = ldb(P.rk_base + cN, o8);
#pragma unroll
                for (int j = 0; j < kMaxRkFused; ++j) {
                    if (j < P.rk_n && P.rk_c[j] != 0.0) {
                        const double kj = P.k_f32 ? static_cast<double>(ldb(static_cast<const float*>(P.rk_k[j]) + cN, o8 / 2u))
                                                  : ldb(static_cast<const double*>(P.rk_k[j]) + cN, o8);
                        s = s + P.rk_c[j] * kj;
                    }
                }
                if (P.rk_clast != 0.0) s = s + P.rk_clast * k[c];
                stb(P.rk_out + cN, o8, s);
            }
        }
        break;
    }
    case static_cast<int>(OutMode::GRAD): {
        const double a0 = mc1 * Hz - mc2 * Hy, a1 = mc2 * Hx - mc0 * Hz, a2 = mc0 * Hy - mc1 * Hx;
        double* o = static_cast<double*>(P.out);
        stb(o, o8, mc1 * a2 - mc2 * a1);
        stb(o + N, o8, mc2 * a0 - mc0 * a2);
        stb(o + 2 * N, o8, mc0 * a1 - mc1 * a0);
        if constexpr (RED) {
            acc[P.nterms] = fmax(acc[P.nterms], fmax(fabs(a0), fmax(fabs(a1), fabs(a2))));
            acc[P.nterms + 1] = fmax(acc[P.nterms + 1], fmax(fabs(Hx), fmax(fabs(Hy), fabs(Hz))));
        }
        break;
    }
    default: break;
    }
}

template <bool RED> __device__ void finish_block(const EpiParams& P, double* acc) {
    if constexpr (!RED) return;
    int is_max[kMaxTerms + 2];
    for (int s = 0; s < P.nterms; ++s) is_max[s] = 0;
    is_max[P.nterms] = 1;
    is_max[P.nterms + 1] = 1;
    block_reduce_store(acc, P.nterms + 2, is_max, P.partials, P.nblocks);
}

// One thread per cell; the FFT-convolution field, if any, comes from P.hd ([3][N], type T).
// ROWS maps work-group b (after the XCD renumbering) to a tile of kBlock cells of row
// b / bpr in the z-fastest row order (row = y nz + z), so the y and z stencil neighbours of a
// tile are read behind the same L2; otherwise cells are taken in memory order.
// MINW is the waves per SIMD the register budget must allow: 5 (96 VGPRs, no spills) for the
// FP32 field and LLG evaluations, which run 7 times per RK attempt; 1 (uncapped) otherwise.
template <class T, class TH, bool RED, bool ROWS, int MINW, uint64_t PROG = 0>
__global__ void __launch_bounds__(kBlock, MINW) k_epi_cells(EpiParams P, FastDiv nxd, FastDiv nyd, FastDiv nzd,
                                                            FastDiv bprd) {
    double acc[kMaxTerms + 2];
    for (int s = 0; s < kMaxTerms + 2; ++s) acc[s] = 0;
    const int nx = static_cast<int>(P.nx), ny = static_cast<int>(P.ny), nz = static_cast<int>(P.nz);
    int x, y, z;
    bool valid;
    if constexpr (ROWS) {
        const int lb = static_cast<int>(xcd_swizzle(blockIdx.x, gridDim.x));
        const int row = static_cast<int>(fdiv(lb, bprd)), xc = lb - row * static_cast<int>(bprd.d);
        y = static_cast<int>(fdiv(row, nzd));
        z = row - y * nz;
        x = xc * kBlock + static_cast<int>(threadIdx.x);
        valid = x < nx;
    } else {
        const int i = static_cast<int>(blockIdx.x) * kBlock + static_cast<int>(threadIdx.x);
        valid = i < P.N;
        const int yz = static_cast<int>(fdiv(i, nxd));
        x = i - yz * nx;
        z = static_cast<int>(fdiv(yz, nyd));
        y = yz - z * ny;
    }
    if (valid) {
        T hd[3] = {0, 0, 0};
        if (P.hd) {
            const TH* h = static_cast<const TH*>(P.hd);
            const uint32_t ob = static_cast<uint32_t>(x + nx * (y + ny * z)) * static_cast<uint32_t>(sizeof(TH));
            hd[0] = static_cast<T>(ldb(h, ob));
            hd[1] = static_cast<T>(ldb(h + P.N, ob));
            hd[2] = static_cast<T>(ldb(h + 2 * P.N, ob));
        }
        cell_epilogue<T, RED, PROG>(P, x, y, z, hd, acc);
    }
    finish_block<RED>(P, acc);
}

// Per-cell stage with z-column lanes, for rows of at least one work-group (LLG outputs, field
// outputs of films thicker than k_epi_col4 takes, and with RED the energy sums and the LBFGS
// gradient's maxima, reduced per work-group as in k_epi_cells).  Work-group g (after the XCD renumbering) takes cells
// x0 .. x0 + kBlock - 1 of row y at every z, and lane t walks its column z = 0 .. nz - 1: the z
// neighbours are the lane's previous and next cells, the x neighbours come from an LDS row of the
// group's cells at the same z (one halo cell each side, loaded by the edge lanes), so only the
// centre, the next z and the two y neighbours are read from memory, 9 loads per cell instead
// of 21.  Zero padding as in cell_epilogue.
template <class T, class TH, bool RED, uint64_t PROG, int MINW>
__global__ void __launch_bounds__(kBlock, MINW) k_epi_zcol(EpiParams P, FastDiv bprd) {
    __shared__ V3<T> row[kBlock + 2];
    const int nx = static_cast<int>(P.nx), ny = static_cast<int>(P.ny), nz = static_cast<int>(P.nz);
    const int lb = static_cast<int>(xcd_swizzle(blockIdx.x, gridDim.x));
    const int y = static_cast<int>(fdiv(lb, bprd)), xc = lb - y * static_cast<int>(bprd.d);
    const int t = static_cast<int>(threadIdx.x), x0 = xc * kBlock, x = x0 + t;
    const bool valid = x < nx;
    const double* mp[3] = {P.m, P.m + P.N, P.m + 2 * P.N};
    const uint32_t sy = static_cast<uint32_t>(nx) * 8u, sz = static_cast<uint32_t>(nx) * static_cast<uint32_t>(ny) * 8u;
    const bool oky0 = y > 0, oky1 = y + 1 < ny;                  // uniform over the group
    const bool hl = t == 0 && x0 > 0, hr = t == kBlock - 1 && x0 + kBlock < nx;
    // byte offset of (x, y, z = 0); lanes past the row end read the row's last cell (unused)
    uint32_t o = (static_cast<uint32_t>(y) * static_cast<uint32_t>(nx) + static_cast<uint32_t>(valid ? x : nx - 1)) * 8u;
    auto ld3 = [&](uint32_t off, double (&v)[3]) {
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] = ldb(mp[c], off);
    };
    auto to_t = [](const double (&v)[3]) { return V3<T>{static_cast<T>(v[0]), static_cast<T>(v[1]), static_cast<T>(v[2])}; };
    double cur[3];
    ld3(o, cur);
    V3<T> prev{0, 0, 0};
    double acc[RED ? kMaxTerms + 2 : 1] = {};
    for (int z = 0; z < nz; ++z) {
        double nxt[3] = {0, 0, 0}, ym[3] = {0, 0, 0}, yp[3] = {0, 0, 0}, hv[3] = {0, 0, 0};
        if (z + 1 < nz) ld3(o + sz, nxt);
        if (oky0) ld3(o - sy, ym);
        if (oky1) ld3(o + sy, yp);
        if (hl) ld3(o - 8u, hv);
        if (hr) ld3(o + 8u, hv);
        T hd[3] = {0, 0, 0};
        if (P.hd) {
            const TH* h = static_cast<const TH*>(P.hd);
            const uint32_t ob = (o / 8u) * static_cast<uint32_t>(sizeof(TH));
            hd[0] = static_cast<T>(ldb(h, ob));
            hd[1] = static_cast<T>(ldb(h + P.N, ob));
            hd[2] = static_cast<T>(ldb(h + 2 * P.N, ob));
        }
        __syncthreads(); // the previous z's row readers are done
        row[t + 1] = valid ? to_t(cur) : V3<T>{0, 0, 0};
        if (hl) row[0] = to_t(hv);
        if (t == 0 && x0 == 0) row[0] = V3<T>{0, 0, 0};
        if (hr) row[kBlock + 1] = to_t(hv);
        __syncthreads();
        if (valid) {
            const V3<T> nb[6] = {row[t], (x + 1 < nx) ? row[t + 2] : V3<T>{0, 0, 0}, to_t(ym), to_t(yp), prev, to_t(nxt)};
            cell_terms<T, RED, PROG>(P, x, y, z, cur[0], cur[1], cur[2], nb, hd, acc);
        }
        prev = to_t(cur);
#pragma unroll
        for (int c = 0; c < 3; ++c) cur[c] = nxt[c];
        o += sz;
    }
    finish_block<RED>(P, acc); // every lane of the group gets here
}

// Per-cell stage for films of at most four layers on 2-D tiles: a work-group of kTileX x kTileY
// lanes takes cells x0 .. x0 + 63 of rows y0 .. y0 + 3, and each lane loads its whole z column
// (every layer's cell and convolution field) at once, so a wave has all of its reads in flight
// together instead of one z step's.  The layers go to an LDS tile with a one-cell halo, each halo
// cell loaded by one lane: tile row 0 loads the row below the tile, row 1 the row above, and
// eight lanes of row 2 the columns left and right of it.  One barrier; zero padding as in
// cell_epilogue.
constexpr int kTileX = 64, kTileY = kBlock / kTileX, kColZ = 4;
static_assert(kTileY >= 3 && 2 * kTileY <= kTileX, "halo roles need three tile rows and 2 kTileY lanes");
template <class T, class TH, uint64_t PROG, int MINW>
__global__ void __launch_bounds__(kBlock, MINW) k_epi_col4(EpiParams P, FastDiv bxd) {
    __shared__ V3<T> tile[kColZ][kTileY + 2][kTileX + 2];
    const int nx = static_cast<int>(P.nx), ny = static_cast<int>(P.ny), nz = static_cast<int>(P.nz);
    const int lb = static_cast<int>(xcd_swizzle(blockIdx.x, gridDim.x));
    const int by = static_cast<int>(fdiv(lb, bxd)), bx = lb - by * static_cast<int>(bxd.d);
    const int tx = static_cast<int>(threadIdx.x) & (kTileX - 1), ty = static_cast<int>(threadIdx.x) / kTileX;
    const int x0 = bx * kTileX, y0 = by * kTileY, x = x0 + tx, y = y0 + ty;
    const bool valid = x < nx && y < ny;
    // halo role: this lane loads mesh cell (hx, hy) into tile slot [hr][hc] (0 outside the mesh)
    bool role = true;
    int hx = x, hy = 0, hr = 0, hc = tx + 1;
    if (ty == 0) {
        hy = y0 - 1;
    } else if (ty == 1) {
        hy = y0 + kTileY;
        hr = kTileY + 1;
    } else if (ty == 2 && tx < 2 * kTileY) {
        const int r = tx % kTileY;
        hy = y0 + r;
        hr = r + 1;
        hx = tx < kTileY ? x0 - 1 : x0 + kTileX;
        hc = tx < kTileY ? 0 : kTileX + 1;
    } else {
        role = false;
    }
    const bool hload = role && hx >= 0 && hx < nx && hy >= 0 && hy < ny;
    const double* mp[3] = {P.m, P.m + P.N, P.m + 2 * P.N};
    const uint32_t sz = static_cast<uint32_t>(nx) * static_cast<uint32_t>(ny) * 8u;
    const uint32_t o = valid ? (static_cast<uint32_t>(y) * static_cast<uint32_t>(nx) + static_cast<uint32_t>(x)) * 8u : 0u;
    const uint32_t oh = hload ? (static_cast<uint32_t>(hy) * static_cast<uint32_t>(nx) + static_cast<uint32_t>(hx)) * 8u : 0u;
    auto to_t = [](const double (&v)[3]) { return V3<T>{static_cast<T>(v[0]), static_cast<T>(v[1]), static_cast<T>(v[2])}; };
    double c[kColZ][3] = {}, h[kColZ][3] = {};
    T hd[kColZ][3] = {};
#pragma unroll
    for (int z = 0; z < kColZ; ++z) {
        if (z < nz) {
            if (valid) {
#pragma unroll
                for (int k = 0; k < 3; ++k) c[z][k] = ldb(mp[k], o + z * sz);
            }
            if (hload) {
#pragma unroll
                for (int k = 0; k < 3; ++k) h[z][k] = ldb(mp[k], oh + z * sz);
            }
            if (P.hd && valid) {
                const TH* hp = static_cast<const TH*>(P.hd);
                const uint32_t ob = (o / 8u + z * (sz / 8u)) * static_cast<uint32_t>(sizeof(TH));
#pragma unroll
                for (int k = 0; k < 3; ++k) hd[z][k] = static_cast<T>(ldb(hp + k * P.N, ob));
            }
        }
    }
    const V3<T> zero{0, 0, 0};
#pragma unroll
    for (int z = 0; z < kColZ; ++z) {
        if (z < nz) {
            tile[z][ty + 1][tx + 1] = valid ? to_t(c[z]) : zero;
            if (role) tile[z][hr][hc] = hload ? to_t(h[z]) : zero;
        }
    }
    __syncthreads();
    if (!valid) return;
    double acc[1] = {0};
#pragma unroll
    for (int z = 0; z < kColZ; ++z) {
        if (z < nz) {
            const V3<T> nb[6] = {tile[z][ty + 1][tx], tile[z][ty + 1][tx + 2], tile[z][ty][tx + 1], tile[z][ty + 2][tx + 1],
                                 z > 0 ? tile[z > 0 ? z - 1 : 0][ty + 1][tx + 1] : zero,
                                 z + 1 < nz ? tile[z + 1 < kColZ ? z + 1 : z][ty + 1][tx + 1] : zero};
            cell_terms<T, false, PROG>(P, x, y, z, c[z][0], c[z][1], c[z][2], nb, hd[z], acc);
        }
    }
}

// Fixed-order reduction of partials[s][0..nb) into out[s].
__global__ void k_reduce_slots(const double* __restrict__ partials, int64_t nb, int nslots, int nsum,
                               double* __restrict__ out) {
    __shared__ double red[kBlock];
    const int s = blockIdx.x;
    if (s >= nslots) return;
    const bool is_max = s >= nsum;
    double v = is_max ? 0.0 : 0.0;
    for (int64_t b = threadIdx.x; b < nb; b += kBlock) {
        const double p = partials[s * nb + b];
        v = is_max ? fmax(v, p) : v + p;
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = is_max ? fmax(red[threadIdx.x], red[threadIdx.x + w])
                                                       : red[threadIdx.x] + red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[s] = red[0];
}

template <class C> C* spec_of(mafx_ctx* ctx) { return ctx->spec.as<C>(); }

// RKKY stencil weights of cell i towards its six neighbours (x-, x+, y-, y+, z-, z+; 0 where the
// neighbour is outside the mesh or the entry is absent), formed in FP64 exactly as the per-cell
// in-kernel form does (rkky_exchange_field.cpp:195-337) and stored in the stencils' type W.
template <class W>
__global__ void __launch_bounds__(256) k_rkky_weights(const double* __restrict__ rk, const double* __restrict__ aex,
                                                      const uint32_t* __restrict__ ridx, int64_t nx, int64_t ny,
                                                      int64_t nz, double pxy, double py, double pz,
                                                      W* __restrict__ w) {
    const int64_t N = nx * ny * nz, sxy = nx * ny;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= N) return;
    const int64_t z = i / sxy, r = i - z * sxy, y = r / nx, x = r - y * nx;
    const double Ai = aex[i];
    auto inplane = [&](bool ok, int64_t j, double pref) -> double {
        if (!ok || Ai == 0) return 0.0;
        const double Aj = aex[j];
        return (2. * Ai) / pref * 2. * Aj / (Aj + Ai);
    };
    auto zc = [&](bool ok, int64_t j) -> double {
        if (!ok) return 0.0;
        const double Ri = rk[i], Rj = rk[j];
        const uint32_t Ii = ridx ? ridx[i] : 0u, Ij = ridx ? ridx[j] : 0u;
        if ((Ii == Ij) && (Ri != 0) && (Rj != 0)) return (2. * Ri) / pz;
        if (Ai == 0) return 0.0;
        const double Aj = aex[j];
        return (2. * Ai) / pz * 2. * Aj / (Aj + Ai);
    };
    const double wk[6] = {inplane(x > 0, i - 1, pxy),      inplane(x + 1 < nx, i + 1, pxy),
                          inplane(y > 0, i - nx, py),      inplane(y + 1 < ny, i + nx, py),
                          zc(z > 0, i - sxy),              zc(z + 1 < nz, i + sxy)};
#pragma unroll
    for (int k = 0; k < 6; ++k) w[k * N + i] = static_cast<W>(wk[k]);
}

} // namespace

void rkky_weights(mafx_ctx* ctx, TermDev& t) {
    const mafx_mesh& g = ctx->mesh;
    const int64_t N = ctx->N;
    const bool f32 = stencils_f32(ctx->precision);
    t.rkw.alloc((f32 ? sizeof(float) : sizeof(double)) * 6 * N);
    // the same prefactors as make_params: mu0 dx^2 couples x AND y neighbours, mu0 dz^2 the layers
    const double pxy = kMu0 * std::pow(g.dx, 2), pz = kMu0 * std::pow(g.dz, 2);
    const uint32_t* ridx = t.has_ridx ? t.ridx.as<uint32_t>() : nullptr;
    if (f32)
        hipLaunchKernelGGL(k_rkky_weights<float>, dim3(grid_for(N, 256)), dim3(256), 0, ctx->stream,
                           t.rkky.as<double>(), t.aex.as<double>(), ridx, g.nx, g.ny, g.nz, pxy, pxy, pz,
                           t.rkw.as<float>());
    else
        hipLaunchKernelGGL(k_rkky_weights<double>, dim3(grid_for(N, 256)), dim3(256), 0, ctx->stream,
                           t.rkky.as<double>(), t.aex.as<double>(), ridx, g.nx, g.ny, g.nz, pxy, pxy, pz,
                           t.rkw.as<double>());
    MAFX_HIP(hipGetLastError());
    MAFX_HIP(hipStreamSynchronize(ctx->stream));
}

namespace {

EpiParams make_params(mafx_ctx* ctx, const EvalArgs& a, int64_t nblocks) {
    EpiParams P{};
    const mafx_mesh& g = ctx->mesh;
    P.nx = g.nx;
    P.ny = g.ny;
    P.nz = g.nz;
    P.N = ctx->N;
    P.m = a.m;
    P.ms = ctx->has_ms_field ? ctx->ms_field.as<double>() : nullptr;
    P.Ms = ctx->Ms;
    P.ex_w[0] = 1 / std::pow(g.dx, 2);
    P.ex_w[1] = 1 / std::pow(g.dy, 2);
    P.ex_w[2] = 1 / std::pow(g.dz, 2);
    P.ex_self = -2. / std::pow(g.dx, 2.) - 2. / std::pow(g.dy, 2.) - 2. / std::pow(g.dz, 2.);
    P.dmi_c[0] = 1 / (2. * g.dx);
    P.dmi_c[1] = 1 / (2. * g.dy);
    P.dmi_c[2] = 1 / (2. * g.dz);
    P.rk_pref[0] = kMu0 * std::pow(g.dx, 2);
    P.rk_pref[1] = kMu0 * std::pow(g.dx, 2); // the reference couples y-neighbours with dx^2
    P.rk_pref[2] = kMu0 * std::pow(g.dz, 2);
    P.nterms = 0;
    P.fft_slot = -1;
    for (size_t k = 0; k < ctx->terms.size(); ++k) {
        if (!(a.mask & (1u << k))) continue;
        const TermDev& td = ctx->terms[k];
        EpiTerm& e = P.t[P.nterms];
        e.kind = td.kind;
        for (int c = 0; c < 3; ++c) {
            e.n[c] = td.axis[c];
            e.f[c] = td.field[c];
        }
        e.zee = td.has_zee_array ? td.zee.as<double>() : nullptr;
        e.rk = td.rkky.as<double>();
        e.aex = td.aex.as<double>();
        e.ridx = td.has_ridx ? td.ridx.as<uint32_t>() : nullptr;
        e.rkw = td.rkw.p;
        e.pf = td.pfield.p ? td.pfield.as<double>() : nullptr;
        e.ax = td.axfield.p ? td.axfield.as<double>() : nullptr;
        if (e.pf || e.ax) P.has_fields = 1;
        const double Ms = ctx->Ms;
        switch (td.kind) {
        case MAFX_EXCHANGE: e.num = 2. * td.value; e.c0 = (2. * td.value) / (kMu0 * Ms); break;
        case MAFX_DMI: e.num = 2. * td.value; e.c0 = (2. * td.value) / (kMu0 * Ms); break;
        case MAFX_UNIAXIAL: e.num = 2. * td.value; e.c0 = 2. * td.value / (kMu0 * Ms); break;
        case MAFX_ATOM_EXCHANGE: e.c0 = td.value / (kMu0 * Ms); break;
        case MAFX_ATOM_DMI: e.c0 = td.value / (kMu0 * Ms); break;
        case MAFX_ATOM_UNIAXIAL: e.c0 = 2 * td.value / (kMu0 * Ms); break;
        default: break;
        }
        if (static_cast<int>(k) == ctx->demag_term) P.fft_slot = P.nterms;
        if (td.kind == MAFX_EXCHANGE || td.kind == MAFX_DMI || td.kind == MAFX_ATOM_DMI ||
            td.kind == MAFX_ATOM_EXCHANGE || td.kind == MAFX_RKKY)
            P.needs_nb = 1;
        P.nterms++;
    }
    P.mode = static_cast<int>(a.mode);
    P.out = a.out;
    P.out_f32 = a.out_f32 ? 1 : 0;
    P.alpha = a.alpha;
    P.alpha_f = a.alpha_field;
    P.dt = a.dt;
    P.diss_only = a.dissipation_only ? 1 : 0;
    P.want_energy = (a.want_energy || a.mode == OutMode::ENERGY) ? 1 : 0;
    P.partials = ctx->partials.as<double>();
    P.nblocks = nblocks;
    P.rk_out = a.mode == OutMode::LLG ? a.rk_out : nullptr;
    P.rk_base = a.rk_base;
    P.k_f32 = a.k_f32 ? 1 : 0;
    P.rk_n = a.rk_n;
    for (int j = 0; j < kMaxRkFused; ++j) {
        P.rk_k[j] = a.rk_k[j];
        P.rk_c[j] = j < a.rk_n ? a.rk_c[j] : 0.0;
    }
    P.rk_clast = a.rk_n >= 0 && a.rk_n <= kMaxRkFused ? a.rk_c[a.rk_n] : 0.0;
    return P;
}

// C: complex type of the FFT convolution; TH: its field between the inverse x transform and the
// per-cell kernel; T: arithmetic of the local terms (stencils) in the per-cell kernel.
template <class C, class TH, class T> void evaluate_impl(mafx_ctx* ctx, const EvalArgs& a) {
    const mafx_mesh& g = ctx->mesh;
    const bool fft = ctx->demag_term >= 0 && (a.mask & (1u << ctx->demag_term));
    // cell and row indices are 32-bit inside the kernels, and the per-cell kernel addresses a
    // component plane with 32-bit byte offsets (8 N < 2^32)
    require(ctx->N < (int64_t{1} << 29), "meshes of 2^29 cells or more are not supported");
    const int64_t nrows = g.ny * g.nz;
    // per-cell kernel: row tiles when rows are at least one work-group long
    const bool rowmap = g.nx >= kBlock;
    const int64_t bpr = (g.nx + kBlock - 1) / kBlock;
    const int64_t nblocks = rowmap ? nrows * bpr : (ctx->N + kBlock - 1) / kBlock;
    require((ctx->partials.bytes / sizeof(double)) >= static_cast<size_t>((kMaxTerms + 2) * nblocks),
            "partials buffer too small");
    EpiParams P = make_params(ctx, a, nblocks);
    const double field_bytes = (a.mode == OutMode::H) ? (a.out_f32 ? 12.0 : 24.0)
                               : (a.mode == OutMode::ENERGY) ? 0.0 : 24.0;
    if (fft) {
        C* spec = spec_of<C>(ctx);
        const int64_t hxs = ctx->hxs, n1 = ctx->n1, n2 = ctx->n2, nzs = ctx->nzs;
        const int64_t sa = nzs * n1 * hxs;
        // Ms * m (micro; Ms_field per cell) or m (atomistic dipole) -> x half spectrum
        const double* ms = (!ctx->atom_dipole && ctx->has_ms_field) ? ctx->ms_field.as<double>() : nullptr;
        const double scale = ctx->atom_dipole ? 1.0 : (ctx->has_ms_field ? 1.0 : ctx->Ms);
        fft_x_forward<C>(ctx, a.m, ms, scale, g.nx, g.ny, g.nz, 3, g.ny, g.nz, spec, nzs);
        // column tiles: rows ky are kTile elements apart, planes n1 hxs
        fft_y_split<C>(ctx, spec, sa, 3, n1 * hxs, static_cast<int>(g.nz), kTile, false);
        // tensor product: complex reference spectrum (FP64) or its real symmetric quarter (FP32)
        // (plus, MAFX_TERM_FULL_TENSOR, the FP16 residual that makes it the reference's spectrum)
        TensorRes res{};
        const TensorRes* resp = nullptr;
        if (demag_full_c32(ctx->precision, ctx->demag_flags)) {
            res.r = reinterpret_cast<const __half2*>(static_cast<const char*>(ctx->tensor.p) + ctx->res_off);
            for (int c = 0; c < 6; ++c) res.inv[c] = ctx->res_inv[c];
            resp = &res;
        }
        auto kmul = [&](int64_t planes) {
            if constexpr (std::is_same<C, float2>::value)
                kspace_mul_sym(ctx, spec, ctx->tensor.as<float>(), planes, resp);
            else
                kspace_mul<C>(ctx, spec, ctx->tensor.as<C>(), sa);
        };
        if (n2 == 1) {
            kmul(1);
        } else if (nzs == g.nz) {
            if constexpr (std::is_same<C, float2>::value)
                zfused_sym(ctx, spec, ctx->tensor.as<float>(), static_cast<int>(g.nz), resp);
            else
                zfused<C>(ctx, spec, ctx->tensor.as<C>(), static_cast<int>(g.nz));
        } else {
            fft_strided<C>(ctx, ctx->pz, spec, sa, 3, kTile, static_cast<int>(n1), n1 * hxs, static_cast<int>(g.nz),
                           static_cast<int>(n2), false, sizeof(C) == 8 ? "fft_z_fwd_f32" : "fft_z_fwd_f64");
            kmul(n2);
            fft_strided<C>(ctx, ctx->pz, spec, sa, 3, kTile, static_cast<int>(n1), n1 * hxs, static_cast<int>(n2),
                           static_cast<int>(g.nz), true, sizeof(C) == 8 ? "fft_z_inv_f32" : "fft_z_inv_f64");
        }
        fft_y_split<C>(ctx, spec, sa, 3, n1 * hxs, static_cast<int>(g.nz), kTile, true);
        const double P_total = static_cast<double>(ctx->n0) * n1 * n2;
        const double hscale = (ctx->atom_dipole ? ctx->Ms : 1.0) / P_total;
        TH* hd = ctx->hdem.as<TH>();
        fft_x_inverse<C, TH>(ctx, spec, nzs, hd, hscale);
        P.hd = hd;
    }
    {
        // per-cell kernel: local terms, plus the FFT field read back from P.hd
        const double hd_bytes = P.hd ? 3.0 * sizeof(TH) : 0.0;
        ProfScope ps(ctx, P.hd ? (sizeof(T) == 4 ? "epilogue_cells_f32" : "epilogue_cells_f64")
                               : (sizeof(T) == 4 ? "local_cells_f32" : "local_cells_f64"),
                     static_cast<double>(ctx->N) * (24.0 + field_bytes + hd_bytes));
        const bool red = P.want_energy || P.mode == static_cast<int>(OutMode::GRAD);
        constexpr int MW = sizeof(T) == 4 ? 5 : 1;
        auto kern = red ? (rowmap ? k_epi_cells<T, TH, true, true, 1> : k_epi_cells<T, TH, true, false, 1>)
                        : (rowmap ? k_epi_cells<T, TH, false, true, MW> : k_epi_cells<T, TH, false, false, MW>);
        // compile-time term programs of the common micromagnetic term lists; any other list runs
        // the run-time term loop
        uint64_t prog = P.has_fields ? 0 : static_cast<uint64_t>(P.nterms);
        for (int t = 0; t < P.nterms && !P.has_fields; ++t) prog |= static_cast<uint64_t>(P.t[t].kind) << (4 + 4 * t);
        constexpr uint64_t kP1 = prog_of({MAFX_DEMAG, MAFX_EXCHANGE, MAFX_DMI, MAFX_UNIAXIAL});
        constexpr uint64_t kP2 = prog_of({MAFX_DEMAG, MAFX_EXCHANGE, MAFX_ZEEMAN});
        constexpr uint64_t kP3 = prog_of({MAFX_DEMAG, MAFX_RKKY, MAFX_DMI, MAFX_UNIAXIAL, MAFX_ZEEMAN});
        constexpr uint64_t kP4 = prog_of({MAFX_DEMAG, MAFX_EXCHANGE, MAFX_DMI, MAFX_UNIAXIAL, MAFX_ZEEMAN});
        if (red && rowmap && P.needs_nb && g.nx % kBlock == 0) {
            // energy sums / LBFGS gradient on z-column lanes: one partial per work-group and slot
            using KZ = void (*)(EpiParams, FastDiv);
            KZ kz = k_epi_zcol<T, TH, true, 0, 1>;
            if (prog == kP1) kz = k_epi_zcol<T, TH, true, kP1, 1>;
            if (prog == kP2) kz = k_epi_zcol<T, TH, true, kP2, 1>;
            if (prog == kP3) kz = k_epi_zcol<T, TH, true, kP3, 1>;
            if (prog == kP4) kz = k_epi_zcol<T, TH, true, kP4, 1>;
            const int64_t ng = g.ny * bpr;
            require(ng <= nblocks, "z-column partials exceed the per-cell partials");
            EpiParams Pz = P;
            Pz.nblocks = ng;
            hipLaunchKernelGGL(kz, dim3(static_cast<unsigned>(ng)), dim3(kBlock), 0, ctx->stream, Pz,
                               make_fastdiv(static_cast<uint32_t>(bpr)));
            MAFX_HIP(hipGetLastError());
            ctx->scal_nblocks = ng;
            ctx->scal_nterms = P.nterms;
            return;
        }
        if (!red) {
            if (prog == kP1) kern = rowmap ? k_epi_cells<T, TH, false, true, MW, kP1> : k_epi_cells<T, TH, false, false, MW, kP1>;
            if (prog == kP2) kern = rowmap ? k_epi_cells<T, TH, false, true, MW, kP2> : k_epi_cells<T, TH, false, false, MW, kP2>;
            if (prog == kP3) kern = rowmap ? k_epi_cells<T, TH, false, true, MW, kP3> : k_epi_cells<T, TH, false, false, MW, kP3>;
            if (prog == kP4) kern = rowmap ? k_epi_cells<T, TH, false, true, MW, kP4> : k_epi_cells<T, TH, false, false, MW, kP4>;
            // z-column lanes for rows of whole work-groups when a term reads the neighbours
            if (rowmap && P.needs_nb && g.nx % kBlock == 0) {
                using KZ = void (*)(EpiParams, FastDiv);
                // the run-time term loop and the RKKY list need up to 128 VGPRs without spills
                constexpr int MWW = sizeof(T) == 4 ? 4 : 1;
                // the other programs fit 80 VGPRs without spills: 6 waves per SIMD, measured
                // 64.4 -> 61.6 us at 1024 x 1024 x 4 against 5 (8 waves spill)
                constexpr int MWZ = sizeof(T) == 4 ? 6 : 1;
                // field output of films up to four layers: all of a column's loads at once, measured
                // 54.2 -> 45.9 us (FP32) and 69.3 -> 60.6 us (FP64) at 1024 x 1024 x 4; LLG stages,
                // whose reads of the earlier stages dominate, keep the z-column lanes (133 against
                // 135.5 us); 5 waves per SIMD hold the column without spills
                constexpr int MWC = sizeof(T) == 4 ? 5 : 1;
                if (g.nz <= kColZ && P.mode == static_cast<int>(OutMode::H)) {
                    KZ kt = k_epi_col4<T, TH, 0, MWW>;
                    if (prog == kP1) kt = k_epi_col4<T, TH, kP1, MWC>;
                    if (prog == kP2) kt = k_epi_col4<T, TH, kP2, MWC>;
                    if (prog == kP3) kt = k_epi_col4<T, TH, kP3, MWW>;
                    if (prog == kP4) kt = k_epi_col4<T, TH, kP4, MWC>;
                    const int64_t bxn = (g.nx + kTileX - 1) / kTileX, byn = (g.ny + kTileY - 1) / kTileY;
                    hipLaunchKernelGGL(kt, dim3(static_cast<unsigned>(bxn * byn)), dim3(kBlock), 0, ctx->stream, P,
                                       make_fastdiv(static_cast<uint32_t>(bxn)));
                    MAFX_HIP(hipGetLastError());
                    ctx->scal_nblocks = nblocks;
                    ctx->scal_nterms = P.nterms;
                    return;
                }
                KZ kz = k_epi_zcol<T, TH, false, 0, MWW>;
                if (prog == kP1) kz = k_epi_zcol<T, TH, false, kP1, MWZ>;
                if (prog == kP2) kz = k_epi_zcol<T, TH, false, kP2, MWZ>;
                if (prog == kP3) kz = k_epi_zcol<T, TH, false, kP3, MWW>;
                if (prog == kP4) kz = k_epi_zcol<T, TH, false, kP4, MWZ>;
                const int64_t ng = g.ny * bpr;
                hipLaunchKernelGGL(kz, dim3(static_cast<unsigned>(ng)), dim3(kBlock), 0, ctx->stream, P,
                                   make_fastdiv(static_cast<uint32_t>(bpr)));
                MAFX_HIP(hipGetLastError());
                ctx->scal_nblocks = nblocks;
                ctx->scal_nterms = P.nterms;
                return;
            }
        }
        hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(nblocks)), dim3(kBlock), 0, ctx->stream, P,
                           make_fastdiv(static_cast<uint32_t>(g.nx)), make_fastdiv(static_cast<uint32_t>(g.ny)),
                           make_fastdiv(static_cast<uint32_t>(g.nz)), make_fastdiv(static_cast<uint32_t>(bpr)));
        MAFX_HIP(hipGetLastError());
    }
    ctx->scal_nblocks = nblocks;
    ctx->scal_nterms = P.nterms;
}

} // namespace

void evaluate(mafx_ctx* ctx, const EvalArgs& a) {
    lbfgs_flush(ctx); // the per-cell pass overwrites the reduction partials
    require(ctx->terms.size() <= static_cast<size_t>(kMaxTerms), "at most 8 terms per context");
    // time-dependent external fields (external_field.cpp:28-36): the callback fills the term's
    // field for this evaluation's time and magnetisation
    for (size_t k = 0; k < ctx->terms.size(); ++k) {
        TermDev& td = ctx->terms[k];
        if (!td.callback || !(a.mask & (1u << k))) continue;
        int rc;
        if (td.callback_homogeneous) {
            double v[3] = {0, 0, 0};
            rc = td.callback(td.callback_user, ctx->time, a.m, v, ctx->stream);
            for (int c = 0; c < 3; ++c) td.field[c] = v[c];
        } else if (td.callback_host) {
            const size_t bytes = sizeof(double) * 3 * static_cast<size_t>(ctx->N);
            MAFX_HIP(hipMemcpyAsync(td.cb_m.data(), a.m, bytes, hipMemcpyDeviceToHost, ctx->stream));
            MAFX_HIP(hipStreamSynchronize(ctx->stream));
            rc = td.callback(td.callback_user, ctx->time, td.cb_m.data(), td.cb_h.data(), nullptr);
            if (rc == 0) {
                MAFX_HIP(hipMemcpyAsync(td.zee.p, td.cb_h.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
                MAFX_HIP(hipStreamSynchronize(ctx->stream));
            }
        } else {
            rc = td.callback(td.callback_user, ctx->time, a.m, td.zee.as<double>(), ctx->stream);
        }
        require(rc == 0, "external field callback failed (returned " + std::to_string(rc) + ")");
    }
    switch (ctx->precision) {
    case MAFX_F64: evaluate_impl<double2, double, double>(ctx, a); break;
    case MAFX_F32_FIELD: evaluate_impl<float2, float, float>(ctx, a); break;
    case MAFX_F32_DEMAG: evaluate_impl<float2, float, double>(ctx, a); break;
    case MAFX_F32_STENCILS: evaluate_impl<double2, double, float>(ctx, a); break;
    default: throw Error(MAFX_ERR_ARG, "unknown precision");
    }
}

// Reduces the partials of the last evaluate() (energy slots and, in gradient mode, the two
// maxima) to host values.  E_terms is indexed by the context's term order.
int finish_energy_launch(mafx_ctx* ctx) {
    const int nt = ctx->scal_nterms;
    const int nslots = nt + 2;
    hipLaunchKernelGGL(k_reduce_slots, dim3(nslots), dim3(kBlock), 0, ctx->stream, ctx->partials.as<double>(),
                       ctx->scal_nblocks, nslots, nt, ctx->scal.as<double>());
    MAFX_HIP(hipGetLastError());
    return nslots;
}

void finish_energy(mafx_ctx* ctx, uint32_t mask, double* E_terms, double* E_total) {
    const int nslots = finish_energy_launch(ctx);
    MAFX_HIP(hipMemcpyAsync(ctx->host_scal, ctx->scal.p, sizeof(double) * nslots, hipMemcpyDeviceToHost, ctx->stream));
    MAFX_HIP(hipStreamSynchronize(ctx->stream));
    finish_energy_combine(ctx, mask, E_terms, E_total);
}

void finish_energy_combine(mafx_ctx* ctx, uint32_t mask, double* E_terms, double* E_total) {
    const int nt = ctx->scal_nterms;
    const mafx_mesh& g = ctx->mesh;
    double total = 0;
    int slot = 0;
    for (size_t k = 0; k < ctx->terms.size(); ++k) {
        double e = 0;
        if (mask & (1u << k)) {
            const double S = ctx->host_scal[slot++];
            const int kind = ctx->terms[k].kind;
            const bool zee = kind == MAFX_ZEEMAN || kind == MAFX_ATOM_ZEEMAN;
            if (kind >= MAFX_ATOM_DIPOLE) { // atom_term.hpp:12-15, atomistic_external_field.cpp:5-11
                e = zee ? -kMu0 * ctx->Ms * S : -kMu0 / 2. * ctx->Ms * S;
            } else if (!ctx->has_ms_field) { // micro_term.hpp:18-30, external_field.cpp:38-48
                e = zee ? -kMu0 * ctx->Ms * S * g.dx * g.dy * g.dz : -kMu0 / 2. * ctx->Ms * S * g.dx * g.dy * g.dz;
            } else {
                e = zee ? -kMu0 * S * g.dx * g.dy * g.dz : -kMu0 / 2. * S * g.dx * g.dy * g.dz;
            }
        }
        if (E_terms) E_terms[k] = e;
        total = total + e;
    }
    if (E_total) *E_total = total;
    ctx->last_max_torque = ctx->host_scal[nt];
    ctx->last_max_h = ctx->host_scal[nt + 1];
}

// Process-wide cache of the host real-space tensor: the bit-exact Newell setup costs minutes of
// CPU at 1024 x 1024 x 4, so contexts of the same mesh (e.g. the FP64 and FP32 paths) share it.
namespace {
std::mutex g_tensor_mu;
std::map<std::tuple<int64_t, int64_t, int64_t, double, double, double, bool>, std::shared_ptr<std::vector<double>>>
    g_tensor_cache;

std::shared_ptr<std::vector<double>> host_tensor(const mafx_mesh& g, bool dipole, int64_t P) {
    const auto key = std::make_tuple(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, dipole);
    {
        std::lock_guard<std::mutex> lk(g_tensor_mu);
        auto it = g_tensor_cache.find(key);
        if (it != g_tensor_cache.end()) return it->second;
    }
    auto v = std::make_shared<std::vector<double>>(static_cast<size_t>(6 * P));
    if (dipole)
        dipole_tensor_host(g, v->data());
    else
        newell_tensor_host(g, v->data(), 0);
    std::lock_guard<std::mutex> lk(g_tensor_mu);
    if (g_tensor_cache.size() >= 2) g_tensor_cache.erase(g_tensor_cache.begin());
    g_tensor_cache[key] = v;
    return v;
}
} // namespace

// ---- real-space tensors on the GPU (FP32 field path) ---------------------------------------
// The same formulas as newell_host.cpp / demag_field.cpp:20-121 evaluated in FP64 on the device.
// The device libm rounds differently from glibc, so the far-field cancellation noise differs from
// the reference's in its last bits; the FP32 path (whose own rounding is ~1e-7) uses this, the FP64
// path keeps the bit-exact host tensor.
namespace {

__device__ double nf_dev(double x, double y, double z) { // demag_field.cpp:20-40
    x = fabs(x);
    y = fabs(y);
    z = fabs(z);
    const double xx = x * x, yy = y * y, zz = z * z;
    const double R = sqrt(xx + yy + zz);
    double r = 1.0 / 6.0 * (2.0 * xx - yy - zz) * R;
    if (xx + zz > 0) r += y / 2.0 * (zz - xx) * asinh(y / sqrt(xx + zz));
    if (xx + yy > 0) r += z / 2.0 * (yy - xx) * asinh(z / sqrt(xx + yy));
    if (x * R > 0) r += -x * y * z * atan(y * z / (x * R));
    return r;
}

__device__ double ng_dev(double x, double y, double z) { // demag_field.cpp:42-69
    z = fabs(z);
    const double xx = x * x, yy = y * y, zz = z * z;
    const double R = sqrt(xx + yy + zz);
    double r = -x * y * R / 3.0;
    if (xx + yy > 0) r += x * y * z * asinh(z / sqrt(xx + yy));
    if (yy + zz > 0) r += y / 6.0 * (3.0 * zz - yy) * asinh(x / sqrt(yy + zz));
    if (xx + zz > 0) r += x / 6.0 * (3.0 * zz - xx) * asinh(y / sqrt(xx + zz));
    if (z * R > 0) r += -zz * z / 6.0 * atan(x * y / (z * R));
    if (y * R != 0) r += -z * yy / 2.0 * atan(x * z / (y * R));
    if (x * R != 0) r += -z * xx / 2.0 * atan(y * z / (x * R));
    return r;
}

// 27-point second difference in the order of demag_field.cpp:75-83
template <bool G>
__device__ double second_diff_dev(int ix, int iy, int iz, double dx, double dy, double dz) {
    const double x = dx * ix, y = dy * iy, z = dz * iz;
    const double xs[3] = {x, x + dx, x - dx}, ys[3] = {y, y + dy, y - dy}, zs[3] = {z, z + dz, z - dz};
    auto F = [&](int a, int b, int c) { return G ? ng_dev(xs[a], ys[b], zs[c]) : nf_dev(xs[a], ys[b], zs[c]); };
    double s = 8.0 * F(0, 0, 0);
    s = s - 4.0 * F(1, 0, 0) - 4.0 * F(2, 0, 0) - 4.0 * F(0, 1, 0) - 4.0 * F(0, 2, 0) - 4.0 * F(0, 0, 1) -
        4.0 * F(0, 0, 2);
    s = s + 2.0 * F(1, 1, 0) + 2.0 * F(1, 2, 0) + 2.0 * F(2, 1, 0) + 2.0 * F(2, 2, 0) + 2.0 * F(1, 0, 1) +
        2.0 * F(1, 0, 2) + 2.0 * F(2, 0, 1) + 2.0 * F(2, 0, 2) + 2.0 * F(0, 1, 1) + 2.0 * F(0, 1, 2) +
        2.0 * F(0, 2, 1) + 2.0 * F(0, 2, 2);
    s = s - F(1, 1, 1) - F(1, 1, 2) - F(1, 2, 1) - F(1, 2, 2) - F(2, 1, 1) - F(2, 1, 2) - F(2, 2, 1) - F(2, 2, 2);
    return -s / (4.0 * 3.14159265358979323846 * dx * dy * dz);
}

// [6][n2][n1][n0] with the wrapped index j = (i + n/2) % n - n/2 (demag_field.cpp:104-121); one
// thread per position.  dipole: atomistic_dipole_dipole_field.cpp:55-113 instead.
__global__ void __launch_bounds__(256)
    k_real_tensor(double* __restrict__ N, int64_t n0, int64_t n1, int64_t n2, double dx, double dy, double dz,
                  int dipole) {
    const int64_t P = n0 * n1 * n2;
    const int64_t at = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (at >= P) return;
    const int64_t i0 = at % n0, r = at / n0, i1 = r % n1, i2 = r / n1;
    const int j0 = static_cast<int>((i0 + n0 / 2) % n0 - n0 / 2);
    const int j1 = static_cast<int>((i1 + n1 / 2) % n1 - n1 / 2);
    const int j2 = static_cast<int>((i2 + n2 / 2) % n2 - n2 / 2);
    if (!dipole) {
        N[0 * P + at] = second_diff_dev<false>(j0, j1, j2, dx, dy, dz);
        N[1 * P + at] = second_diff_dev<true>(j0, j1, j2, dx, dy, dz);
        N[2 * P + at] = second_diff_dev<true>(j0, j2, j1, dx, dz, dy);
        N[3 * P + at] = second_diff_dev<false>(j1, j2, j0, dy, dz, dx);
        N[4 * P + at] = second_diff_dev<true>(j1, j2, j0, dy, dz, dx);
        N[5 * P + at] = second_diff_dev<false>(j2, j0, j1, dz, dx, dy);
        return;
    }
    const double rx = j0 * dx, ry = j1 * dy, rz = j2 * dz;
    const double rr = sqrt(rx * rx + ry * ry + rz * rz);
    double v[6] = {0, 0, 0, 0, 0, 0};
    if (rr != 0.0) {
        const double r3 = rr * rr * rr, r5 = r3 * rr * rr;
        v[0] = 3.0 * rx * rx / r5 - 1.0 / r3;
        v[1] = 3.0 * rx * ry / r5;
        v[2] = 3.0 * rx * rz / r5;
        v[3] = 3.0 * ry * ry / r5 - 1.0 / r3;
        v[4] = 3.0 * ry * rz / r5;
        v[5] = 3.0 * rz * rz / r5 - 1.0 / r3;
    }
    const double s = 1.0 / (4.0 * 3.14159265358979323846);
    for (int c = 0; c < 6; ++c) N[c * P + at] = v[c] * s;
}

// byte offset of the FP16 residual in the MAFX_TERM_FULL_TENSOR tensor buffer
// [symmetric quarter f32][pad to 256 B][residual half2 [6][K]][1 / scale f32 x 6, pad to 8]
int64_t full32_res_offset(const mafx_ctx* ctx) {
    const int64_t nS = 6 * ctx->n2h * ctx->n1h * ctx->hxs;
    return (static_cast<int64_t>(sizeof(float)) * nS + 255) / 256 * 256;
}

} // namespace

// Demag / dipole tensor: the real-space tensor (FP64 path: on the host in the reference's
// arithmetic; FP32 path: on the GPU), then a full FP64 R2C transform on the GPU with the same
// kernels, stored as the complex half spectrum (FP64) or its real symmetric quarter (FP32).
void build_tensor_uncached(mafx_ctx* ctx);

// With MAFX_TERM_CACHING the stored tensor is first looked up in the on-disk cache
// (tensor_cache.hpp) and written there after a computation (demag_field.cpp:185-204).
void build_tensor(mafx_ctx* ctx) {
    const int64_t n0 = ctx->n0, n1 = ctx->n1, n2 = ctx->n2, hxs = ctx->hxs;
    const int64_t K = n2 * n1 * hxs;
    ctx->n1h = n1 / 2 + 1;
    ctx->n2h = n2 / 2 + 1;
    const bool caching = (ctx->demag_flags & MAFX_TERM_CACHING) != 0;
    const bool verbose = (ctx->demag_flags & MAFX_TERM_VERBOSE) != 0;
    const bool full32 = demag_full_c32(ctx->precision, ctx->demag_flags);
    const uint64_t stored =
        demag_f64(ctx->precision)
            ? sizeof(double2) * 6 * static_cast<uint64_t>(K)
        : full32 ? static_cast<uint64_t>(full32_res_offset(ctx)) + sizeof(__half2) * 6 * static_cast<uint64_t>(K) +
                       8 * sizeof(float)
                 : sizeof(float) * 6 * static_cast<uint64_t>(ctx->n2h * ctx->n1h * hxs);
    // the FP32 layout from the host-built and the FP64 layout from the GPU-built real-space tensor
    // are stored under keys of their own
    const TensorCacheKey key{n0, n1, n2, hxs, tensor_cache_layout(demag_f64(ctx->precision) ? MAFX_F64 : MAFX_F32_FIELD,
                                                                   ctx->demag_flags),
                             ctx->atom_dipole ? 1 : 0, ctx->mesh.dx, ctx->mesh.dy, ctx->mesh.dz};
    if (caching) {
        std::vector<char> host;
        if (tensor_cache_read(key, stored, host, verbose)) {
            ctx->tensor.alloc(stored);
            MAFX_HIP(hipMemcpy(ctx->tensor.p, host.data(), stored, hipMemcpyHostToDevice));
            if (full32) {
                ctx->res_off = full32_res_offset(ctx);
                std::memcpy(ctx->res_inv, host.data() + ctx->res_off + sizeof(__half2) * 6 * K, sizeof(ctx->res_inv));
            }
            ctx->tensor_source = 2;
            tensor_cache_shrink(verbose);
            return;
        }
    }
    build_tensor_uncached(ctx);
    ctx->tensor_source = 1;
    if (caching) {
        std::vector<char> host(stored);
        MAFX_HIP(hipMemcpy(host.data(), ctx->tensor.p, stored, hipMemcpyDeviceToHost));
        tensor_cache_write(key, host.data(), stored, verbose);
        tensor_cache_shrink(verbose);
    }
}

void build_tensor_uncached(mafx_ctx* ctx) {
    const int64_t n0 = ctx->n0, n1 = ctx->n1, n2 = ctx->n2, hxs = ctx->hxs;
    const int64_t P = n0 * n1 * n2, K = n2 * n1 * hxs;
    DevBuf Nd, T64;
    Nd.alloc(sizeof(double) * 6 * P);
    const bool f64 = demag_f64(ctx->precision);
    const bool full32 = demag_full_c32(ctx->precision, ctx->demag_flags);
    const mafx_mesh& g = ctx->mesh;
    // real-space tensor Nd [6][n2][n1][n0] -> its complex half spectrum T [6][K] (FP64, GPU)
    auto transform = [&](DevBuf& T) {
        T.alloc(sizeof(double2) * 6 * K);
        MAFX_HIP(hipMemsetAsync(T.p, 0, T.bytes, ctx->stream)); // padding columns stay defined
        double2* t = T.as<double2>();
        fft_x_forward<double2>(ctx, Nd.as<double>(), nullptr, 1.0, n0, n1, n2, 6, n1, n2, t, n2);
        // column tiles (fft_device.hpp): lines along y step kTile, lines along z one plane
        fft_strided<double2>(ctx, ctx->py, t, K, 6, n1 * hxs, static_cast<int>(n2), kTile, static_cast<int>(n1),
                             static_cast<int>(n1), false, "tensor_fft_y");
        if (n2 > 1)
            fft_strided<double2>(ctx, ctx->pz, t, K, 6, kTile, static_cast<int>(n1), n1 * hxs, static_cast<int>(n2),
                                 static_cast<int>(n2), false, "tensor_fft_z");
        MAFX_HIP(hipStreamSynchronize(ctx->stream));
    };
    auto octant = [&] {
        // the reference's tensor on the non-negative octant, extended by parity (an eighth of the
        // host evaluations; the FP32 path's symmetric spectrum needs no more)
        std::vector<double> oct(static_cast<size_t>(6 * g.nx * g.ny * g.nz));
        newell_octant_host(g, oct.data(), 0);
        std::vector<double> full(static_cast<size_t>(6 * P));
        newell_octant_extend(g, oct.data(), full.data());
        MAFX_HIP(hipMemcpy(Nd.p, full.data(), sizeof(double) * 6 * P, hipMemcpyHostToDevice));
    };
    if (ctx->demag_flags & MAFX_TERM_DEVICE_TENSOR) {
        // the formulas evaluated on the GPU (fast; its own far-field rounding noise)
        hipLaunchKernelGGL(k_real_tensor, dim3(grid_for(P, 256)), dim3(256), 0, ctx->stream, Nd.as<double>(), n0, n1,
                           n2, g.dx, g.dy, g.dz, ctx->atom_dipole ? 1 : 0);
        MAFX_HIP(hipGetLastError());
    } else if (ctx->atom_dipole || (ctx->demag_flags & MAFX_TERM_HOST_TENSOR) || full32 ||
               (f64 && !(ctx->demag_flags & MAFX_TERM_OCTANT_TENSOR))) {
        // the reference's tensor, bit for bit, over the whole expanded grid
        auto Nh = host_tensor(ctx->mesh, ctx->atom_dipole, P);
        MAFX_HIP(hipMemcpy(Nd.p, Nh->data(), sizeof(double) * 6 * P, hipMemcpyHostToDevice));
    } else {
        octant();
    }
    if (full32) {
        // the reference's spectrum T_ref = S + R: S the symmetric quarter of the parity-symmetric
        // tensor (the octant extension, or the GPU tensor's symmetric form), R the residual.  The
        // unpaired planes (distance -n/2 along an axis) never reach the cropped field, so they are
        // zeroed in T_ref as in S: otherwise R would carry their spectrum (at nz = 4 the plane z =
        // -4 cells is near field), whose FP16 rounding does reach it.
        zero_unpaired_planes(ctx, Nd.as<double>(), 6);
        transform(T64); // T_ref from the whole tensor
        DevBuf Ts;
        if (ctx->demag_flags & MAFX_TERM_DEVICE_TENSOR) {
            MAFX_HIP(hipStreamSynchronize(ctx->stream));
        } else {
            octant();
        }
        zero_unpaired_planes(ctx, Nd.as<double>(), 6);
        transform(Ts);
        Nd.release();
        ctx->res_off = full32_res_offset(ctx);
        ctx->tensor.alloc(static_cast<size_t>(ctx->res_off) + sizeof(__half2) * 6 * K + 8 * sizeof(float));
        float* S = ctx->tensor.as<float>();
        tensor_sym_f32(ctx, Ts.as<double2>(), S);
        float inv[6];
        tensor_residual_f16(ctx, T64.as<double2>(), S,
                            reinterpret_cast<__half2*>(static_cast<char*>(ctx->tensor.p) + ctx->res_off), inv);
        float* tail = reinterpret_cast<float*>(static_cast<char*>(ctx->tensor.p) + ctx->res_off +
                                               sizeof(__half2) * 6 * K);
        MAFX_HIP(hipMemcpyAsync(tail, inv, sizeof(inv), hipMemcpyHostToDevice, ctx->stream));
        MAFX_HIP(hipStreamSynchronize(ctx->stream));
        for (int c = 0; c < 6; ++c) ctx->res_inv[c] = inv[c];
        return;
    }
    if (!f64) zero_unpaired_planes(ctx, Nd.as<double>(), 6);
    transform(T64);
    Nd.release();
    if (f64) {
        std::swap(ctx->tensor.p, T64.p);
        std::swap(ctx->tensor.bytes, T64.bytes);
    } else {
        ctx->tensor.alloc(sizeof(float) * 6 * ctx->n2h * ctx->n1h * hxs);
        tensor_sym_f32(ctx, T64.as<double2>(), ctx->tensor.as<float>());
        MAFX_HIP(hipStreamSynchronize(ctx->stream));
    }
}

} // namespace mafx
