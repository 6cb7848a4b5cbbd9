// FP64 vector kernels of the integrator and minimizer: deterministic two-pass reductions
// (fixed block count and fixed tree, so results are reproducible run to run), normalisation,
// <m>, the LLG right-hand side, RK stage combinations, the embedded error norm and the accept
// update.  Reference citations are relative to magnumaf/src.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <type_traits>

#include "mafx_internal.hpp"
#include "vec_kernels.hpp"

namespace mafx {

namespace {

constexpr int kB = 256;
constexpr int kMaxRedBlocks = 1024;

unsigned red_blocks(int64_t n) {
    int64_t b = (n + kB - 1) / kB;
    if (b > kMaxRedBlocks) b = kMaxRedBlocks;
    return static_cast<unsigned>(b < 1 ? 1 : b);
}

__device__ double block_sum(double v) {
    __shared__ double red[kB / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = red[0];
    for (int w = 1; w < kB / 64; ++w) s += red[w];
    return s;
}
__device__ double block_max(double v) {
    __shared__ double red[kB / 64];
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = red[0];
    for (int w = 1; w < kB / 64; ++w) s = fmax(s, red[w]);
    return s;
}

__global__ void __launch_bounds__(kB) k_dot(const double* __restrict__ a, const double* __restrict__ b, int64_t n,
                                            double* __restrict__ part) {
    double s = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * kB)
        s += a[i] * b[i];
    s = block_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// k_dot's a . b and k_maxabs' max |a| in one pass over a: partials [0][block] and [1][block], each
// formed as the two kernels form theirs
__global__ void __launch_bounds__(kB) k_dot_maxabs(const double* __restrict__ a, const double* __restrict__ b,
                                                  int64_t n, double* __restrict__ part) {
    double s = 0, mx = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * kB) {
        const double ai = a[i];
        s += ai * b[i];
        mx = fmax(mx, fabs(ai));
    }
    s = block_sum(s);
    mx = block_max(mx);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s;
        part[gridDim.x + blockIdx.x] = mx;
    }
}

__global__ void __launch_bounds__(kB) k_maxabs(const double* __restrict__ a, int64_t n, double* __restrict__ part) {
    double s = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * kB)
        s = fmax(s, fabs(a[i]));
    s = block_max(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// final pass over `nb` partials of `nslots` slots laid out [slot][nb]; slots >= nsum take the max
__global__ void __launch_bounds__(kB) k_final(const double* __restrict__ part, int nb, int nslots, int nsum,
                                             double* __restrict__ out) {
    const int s = blockIdx.x;
    if (s >= nslots) return;
    const bool mx = s >= nsum;
    double v = 0;
    for (int i = threadIdx.x; i < nb; i += kB) v = mx ? fmax(v, part[s * nb + i]) : v + part[s * nb + i];
    v = mx ? block_max(v) : block_sum(v);
    if (threadIdx.x == 0) out[s] = v;
}

// util.cpp:34 (a / |a|) and util.cpp:37-42 (zero vectors stay zero)
__global__ void __launch_bounds__(kB) k_normalize(double* __restrict__ m, int64_t N, int zero_safe) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (i >= N) return;
    const double a = m[i], b = m[N + i], c = m[2 * N + i];
    const double n = sqrt(a * a + b * b + c * c);
    if (zero_safe && n == 0) {
        m[i] = 0;
        m[N + i] = 0;
        m[2 * N + i] = 0;
    } else {
        m[i] = a / n;
        m[N + i] = b / n;
        m[2 * N + i] = c / n;
    }
}

// partials: [0..2] component sums, [3] count of |m| != 0  (state.cpp:32-45)
__global__ void __launch_bounds__(kB) k_mean(const double* __restrict__ m, int64_t N, double* __restrict__ part) {
    double s0 = 0, s1 = 0, s2 = 0, cnt = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < N; i += static_cast<int64_t>(gridDim.x) * kB) {
        const double a = m[i], b = m[N + i], c = m[2 * N + i];
        s0 += a;
        s1 += b;
        s2 += c;
        if (sqrt(a * a + b * b + c * c) != 0) cnt += 1;
    }
    s0 = block_sum(s0);
    s1 = block_sum(s1);
    s2 = block_sum(s2);
    cnt = block_sum(cnt);
    if (threadIdx.x == 0) {
        part[0 * gridDim.x + blockIdx.x] = s0;
        part[1 * gridDim.x + blockIdx.x] = s1;
        part[2 * gridDim.x + blockIdx.x] = s2;
        part[3 * gridDim.x + blockIdx.x] = cnt;
    }
}

// k = dt * (-gamma/(1+a^2) m x h - a gamma/(1+a^2) m x (m x h))  (equations.hpp:19-33)
__global__ void __launch_bounds__(kB) k_llg_rhs(const double* __restrict__ m, const double* __restrict__ h,
                                               double* __restrict__ k, int64_t N, double alpha, double dt) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (i >= N) return;
    const double m0 = m[i], m1 = m[N + i], m2 = m[2 * N + i];
    const double H0 = h[i], H1 = h[N + i], H2 = h[2 * N + i];
    const double a0 = m1 * H2 - m2 * H1, a1 = m2 * H0 - m0 * H2, a2 = m0 * H1 - m1 * H0;
    const double b0 = m1 * a2 - m2 * a1, b1 = m2 * a0 - m0 * a2, b2 = m0 * a1 - m1 * a0;
    const double cp = -kGamma / (1. + alpha * alpha), cd = -alpha * kGamma / (1. + alpha * alpha);
    k[i] = dt * (cp * a0 + cd * b0);
    k[N + i] = dt * (cp * a1 + cd * b1);
    k[2 * N + i] = dt * (cp * a2 + cd * b2);
}

struct LinArgs {
    const void* v[kMaxStages]; // K* (double, or float for FP32 stage derivatives)
    double c[kMaxStages];
    int n;
};

// out = ((base + c0 v0) + c1 v1) + ...  (the element-wise evaluation order of the reference's
// stage inputs, adaptive_runge_kutta.cpp:139-165), in FP64 whatever the type K of the v_j.  N is
// the vector count, so the loads of one element are issued together; terms with a zero weight
// are skipped (x + 0 v == x for finite v).
template <int N, class K>
__global__ void __launch_bounds__(kB) k_lincomb(double* __restrict__ out, const double* __restrict__ base, LinArgs a,
                                               int64_t len) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (i >= len) return;
    double v[N > 0 ? N : 1];
#pragma unroll
    for (int j = 0; j < N; ++j) v[j] = a.c[j] != 0.0 ? static_cast<double>(static_cast<const K*>(a.v[j])[i]) : 0.0;
    double s = base[i];
#pragma unroll
    for (int j = 0; j < N; ++j)
        if (a.c[j] != 0.0) s = s + a.c[j] * v[j];
    out[i] = s;
}

// x = normalize_handle_zero_vectors(a + s b)  (lbfgs_minimizer.cpp:277-278)
__global__ void __launch_bounds__(kB) k_line_point(double* __restrict__ x, const double* __restrict__ a, double s,
                                                  const double* __restrict__ b, int64_t N) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (i >= N) return;
    const double p = a[i] + s * b[i], q = a[N + i] + s * b[N + i], r = a[2 * N + i] + s * b[2 * N + i];
    const double n = sqrt(p * p + q * q + r * r);
    x[i] = n != 0 ? p / n : 0.0;
    x[N + i] = n != 0 ? q / n : 0.0;
    x[2 * N + i] = n != 0 ? r / n : 0.0;
}

struct ErrArgs {
    const void* k[kMaxStages + 1]; // k[1..ns] as K* (double, or float for FP32 stage derivatives)
    double b[kMaxStages + 1], e[kMaxStages + 1];
    int ns;
    double atol, rtol;
};

// sumbk = sum b_j k_j, rk_error = sum e_j k_j (left to right, adaptive_runge_kutta.cpp:105-106,168-169),
// err = max |rk_error / (atol + rtol |max(m, m + sumbk)|)| (:108,171) as per-block maxima.
// For RKF45 the reference forms rk_error = sumbk - (sum c_j k_j); `e` then holds c_j and
// RKF selects that shape.  NS is the stage count, so one element's stage loads are issued
// together; stages whose b and e weights are both zero are not read.  The sums are FP64 whatever
// the type K of the stored k_j.
template <int NS, bool RKF, class K, int P>
__global__ void __launch_bounds__(kB) k_rk_error(const double* __restrict__ m, ErrArgs a, double* __restrict__ sumbk,
                                                int64_t len, double* __restrict__ part) {
    // Units of P consecutive elements (P = 2: 8-byte loads of float k_j, 16-byte loads of m and
    // stores of sumbk), U units per lane and trip with all their loads issued before any is used
    // (the grid is one resident round of groups, so each lane walks many units); the P - 1
    // elements at most past the last whole unit go to the first lane.  A max does not depend on
    // the order its elements are taken in.
    constexpr int U = NS <= 8 ? 2 : 1;
    using KP = typename std::conditional<P == 2, typename std::conditional<sizeof(K) == 4, float2, double2>::type, K>::type;
    using MP = typename std::conditional<P == 2, double2, double>::type;
    const int64_t nunit = len / P;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kB;
    double emax = 0;
    // one element: sumbk and the error ratio from its stage derivatives
    auto elem = [&](const K (&kv)[NS + 1], double mi) -> double {
        double k[NS + 1];
#pragma unroll
        for (int j = 1; j <= NS; ++j) k[j] = static_cast<double>(kv[j]);
        double sb = a.b[1] * k[1];
#pragma unroll
        for (int j = 2; j <= NS; ++j)
            if (a.b[j] != 0.0) sb = sb + a.b[j] * k[j];
        double re;
        if constexpr (RKF) {
            double t = a.e[1] * k[1];
#pragma unroll
            for (int j = 2; j <= NS; ++j)
                if (a.e[j] != 0.0) t = t + a.e[j] * k[j];
            re = sb - t;
        } else {
            re = a.e[1] * k[1];
#pragma unroll
            for (int j = 2; j <= NS; ++j)
                if (a.e[j] != 0.0) re = re + a.e[j] * k[j];
        }
        const double mm = fmax(mi, mi + sb);
        emax = fmax(emax, fabs(re / (a.atol + a.rtol * fabs(mm))));
        return sb;
    };
    for (int64_t q0 = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; q0 < nunit; q0 += U * stride) {
        K kv[U][P][NS + 1]; // held in the stored type until used
        double mi[U][P];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t q = q0 + u * stride;
            const bool in = q < nunit;
#pragma unroll
            for (int j = 1; j <= NS; ++j) {
                const bool rd = in && (a.b[j] != 0.0 || a.e[j] != 0.0);
                if constexpr (P == 2) {
                    const KP v = rd ? static_cast<const KP*>(a.k[j])[q] : KP{0, 0};
                    kv[u][0][j] = v.x;
                    kv[u][1][j] = v.y;
                } else {
                    kv[u][0][j] = rd ? static_cast<const K*>(a.k[j])[q] : K(0);
                }
            }
            if constexpr (P == 2) {
                const MP v = in ? reinterpret_cast<const MP*>(m)[q] : MP{0, 0};
                mi[u][0] = v.x;
                mi[u][1] = v.y;
            } else {
                mi[u][0] = in ? m[q] : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t q = q0 + u * stride;
            if (q >= nunit) break;
            double sb[P];
#pragma unroll
            for (int p = 0; p < P; ++p) sb[p] = elem(kv[u][p], mi[u][p]);
            if constexpr (P == 2)
                reinterpret_cast<MP*>(sumbk)[q] = MP{sb[0], sb[1]};
            else
                sumbk[q] = sb[0];
        }
    }
    if (P > 1 && blockIdx.x == 0 && threadIdx.x == 0) {
        for (int64_t i = nunit * P; i < len; ++i) {
            K kv[NS + 1];
#pragma unroll
            for (int j = 1; j <= NS; ++j)
                kv[j] = (a.b[j] != 0.0 || a.e[j] != 0.0) ? static_cast<const K*>(a.k[j])[i] : K(0);
            sumbk[i] = elem(kv, m[i]);
        }
    }
    emax = block_max(emax);
    if (threadIdx.x == 0) part[blockIdx.x] = emax;
}

// m += dm, then normalise (adaptive_runge_kutta.cpp:59-66)
__global__ void __launch_bounds__(kB) k_accept(double* __restrict__ m, const double* __restrict__ dm, int64_t N,
                                              int normalize, int zero_safe) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (i >= N) return;
    double a = m[i] + dm[i], b = m[N + i] + dm[N + i], c = m[2 * N + i] + dm[2 * N + i];
    if (normalize) {
        const double n = sqrt(a * a + b * b + c * c);
        if (zero_safe && n == 0) {
            a = b = c = 0;
        } else {
            a = a / n;
            b = b / n;
            c = c / n;
        }
    }
    m[i] = a;
    m[N + i] = b;
    m[2 * N + i] = c;
}

__global__ void __launch_bounds__(kB) k_axpy(double* __restrict__ y, double a, const double* __restrict__ x, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (i < n) y[i] = y[i] + a * x[i];
}

__global__ void __launch_bounds__(kB) k_scale(double* __restrict__ y, double a, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (i < n) y[i] = a * y[i];
}

__global__ void __launch_bounds__(kB) k_sub(double* __restrict__ out, const double* __restrict__ a,
                                           const double* __restrict__ b, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (i < n) out[i] = a[i] - b[i];
}

double read_scalar(mafx_ctx* ctx, int nslots) {
    MAFX_HIP(hipMemcpyAsync(ctx->host_scal, ctx->scal.p, sizeof(double) * nslots, hipMemcpyDeviceToHost, ctx->stream));
    MAFX_HIP(hipStreamSynchronize(ctx->stream));
    return ctx->host_scal[0];
}

} // namespace

double dot(mafx_ctx* ctx, const double* a, const double* b, int64_t n) {
    lbfgs_flush(ctx);
    const unsigned nb = red_blocks(n);
    {
        ProfScope ps(ctx, "dot_f64", 16.0 * n);
        hipLaunchKernelGGL(k_dot, dim3(nb), dim3(kB), 0, ctx->stream, a, b, n, ctx->partials.as<double>());
        hipLaunchKernelGGL(k_final, dim3(1), dim3(kB), 0, ctx->stream, ctx->partials.as<double>(),
                           static_cast<int>(nb), 1, 1, ctx->scal.as<double>());
    }
    MAFX_HIP(hipGetLastError());
    return read_scalar(ctx, 1);
}

double max_abs(mafx_ctx* ctx, const double* a, int64_t n) {
    lbfgs_flush(ctx);
    const unsigned nb = red_blocks(n);
    {
        ProfScope ps(ctx, "maxabs_f64", 8.0 * n);
        hipLaunchKernelGGL(k_maxabs, dim3(nb), dim3(kB), 0, ctx->stream, a, n, ctx->partials.as<double>());
        hipLaunchKernelGGL(k_final, dim3(1), dim3(kB), 0, ctx->stream, ctx->partials.as<double>(),
                           static_cast<int>(nb), 1, 0, ctx->scal.as<double>());
    }
    MAFX_HIP(hipGetLastError());
    return read_scalar(ctx, 1);
}

void normalize(mafx_ctx* ctx, double* m, bool zero_safe) {
    ProfScope ps(ctx, "normalize_f64", 48.0 * ctx->N);
    hipLaunchKernelGGL(k_normalize, dim3(grid_for(ctx->N, kB)), dim3(kB), 0, ctx->stream, m, ctx->N,
                       zero_safe ? 1 : 0);
    MAFX_HIP(hipGetLastError());
}

void mean_m(mafx_ctx* ctx, const double* m, double out[3]) {
    const unsigned nb = red_blocks(ctx->N);
    hipLaunchKernelGGL(k_mean, dim3(nb), dim3(kB), 0, ctx->stream, m, ctx->N, ctx->partials.as<double>());
    hipLaunchKernelGGL(k_final, dim3(4), dim3(kB), 0, ctx->stream, ctx->partials.as<double>(), static_cast<int>(nb), 4,
                       4, ctx->scal.as<double>());
    MAFX_HIP(hipGetLastError());
    read_scalar(ctx, 4);
    for (int c = 0; c < 3; ++c) out[c] = ctx->host_scal[c] / ctx->host_scal[3];
}

// ------------------------------------------------------------------ <m> without allocation
// mafx_mean_m_noalloc's block partials [4][kMaxRedBlocks] (component sums, non-zero count) live in
// this file-scope array, one copy per GPU for the life of the process.  Host code never passes the
// array's name to a kernel: the host-side symbol is not a device address.  The device address is
// looked up once per device with hipGetSymbolAddress, with that device current, and kept in
// g_noalloc_ptr.  Every slot a final pass reads is written by the same call's first pass, so
// nothing needs clearing; concurrent calls on one GPU share the array and race.
static __device__ __align__(256) double scratch[4 * kMaxRedBlocks];
namespace {
constexpr int kNoallocMaxDevices = 64;
std::mutex g_noalloc_mu;
double* g_noalloc_ptr[kNoallocMaxDevices] = {};

double* noalloc_scratch() {
    int dev = -1;
    MAFX_HIP(hipGetDevice(&dev));
    require(dev >= 0 && dev < kNoallocMaxDevices, "mafx_mean_m_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (g_noalloc_ptr[dev] == nullptr) {
        void* ptr = nullptr;
        MAFX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        require(ptr != nullptr, "mafx_mean_m_noalloc: no device address for the scratch array");
        g_noalloc_ptr[dev] = static_cast<double*>(ptr);
    }
    return g_noalloc_ptr[dev];
}

// the four sums of k_final's tree (same partial order and block_sum), then <m> = sum / count
__global__ void __launch_bounds__(kB) k_mean_final(const double* __restrict__ part, int nb, double* __restrict__ out) {
    double s[4];
    for (int c = 0; c < 4; ++c) {
        double v = 0;
        for (int i = threadIdx.x; i < nb; i += kB) v = v + part[c * nb + i];
        s[c] = block_sum(v);
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; ++c) out[c] = s[c] / s[3];
}
} // namespace

void mean_m_noalloc(const double* m, int64_t N, double* out_dev, hipStream_t stream) {
    double* part = noalloc_scratch();
    const unsigned nb = red_blocks(N);
    hipLaunchKernelGGL(k_mean, dim3(nb), dim3(kB), 0, stream, m, N, part);
    hipLaunchKernelGGL(k_mean_final, dim3(1), dim3(kB), 0, stream, part, static_cast<int>(nb), out_dev);
    MAFX_HIP(hipGetLastError());
}

namespace {
__global__ void k_axpy_dev(double* __restrict__ y, const double* __restrict__ a, const double* __restrict__ x, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (i < n) y[i] = y[i] + (*a) * x[i];
}
} // namespace

void dot_dev(mafx_ctx* ctx, const double* a, const double* b, int64_t n, double* out_dev) {
    lbfgs_flush(ctx);
    const unsigned nb = red_blocks(n);
    ProfScope ps(ctx, "dot_f64", 16.0 * n);
    hipLaunchKernelGGL(k_dot, dim3(nb), dim3(kB), 0, ctx->stream, a, b, n, ctx->partials.as<double>());
    hipLaunchKernelGGL(k_final, dim3(1), dim3(kB), 0, ctx->stream, ctx->partials.as<double>(), static_cast<int>(nb),
                       1, 1, out_dev);
    MAFX_HIP(hipGetLastError());
}

void dot_max_dev(mafx_ctx* ctx, const double* a, const double* b, int64_t n, double* out_dev) {
    lbfgs_flush(ctx);
    const unsigned nb = red_blocks(n);
    ProfScope ps(ctx, "dot_max_f64", 16.0 * n);
    hipLaunchKernelGGL(k_dot_maxabs, dim3(nb), dim3(kB), 0, ctx->stream, a, b, n, ctx->partials.as<double>());
    hipLaunchKernelGGL(k_final, dim3(2), dim3(kB), 0, ctx->stream, ctx->partials.as<double>(), static_cast<int>(nb),
                       2, 1, out_dev);
    MAFX_HIP(hipGetLastError());
}
void max_abs_dev(mafx_ctx* ctx, const double* a, int64_t n, double* out_dev) {
    lbfgs_flush(ctx);
    const unsigned nb = red_blocks(n);
    ProfScope ps(ctx, "maxabs_f64", 8.0 * n);
    hipLaunchKernelGGL(k_maxabs, dim3(nb), dim3(kB), 0, ctx->stream, a, n, ctx->partials.as<double>());
    hipLaunchKernelGGL(k_final, dim3(1), dim3(kB), 0, ctx->stream, ctx->partials.as<double>(), static_cast<int>(nb),
                       1, 0, out_dev);
    MAFX_HIP(hipGetLastError());
}

namespace {
// One step of the LBFGS two-loop recursion with the next step's dot product: qi = q_in + (*coef) v
// (no axpy when coef is null), then qi = scale qi when do_scale, stored as q_out = qi or, negate,
// q_out = -qi; and partials of dnext . qi.  Every expression and the element-to-thread mapping are
// those of k_axpy_dev, k_scale and k_dot, so the results equal the unfused sequence bit for bit.
// PMODE >= 1: the coefficient is the previous step's, formed here from its partials ppart[0..pnb)
// with k_final_coef<PMODE>'s tree and expressions (every work-group forms the same value; group 0
// also stores the dot product, alpha_i and the coefficient as k_final_coef would).  PMODE = -1:
// the coefficient is read from coef (none when coef is null).
template <int PMODE>
__global__ void __launch_bounds__(kB) k_lbfgs_fused(const double* q_in, const double* __restrict__ coef,
                                                   const double* __restrict__ v, double scale, int do_scale,
                                                   const double* __restrict__ dnext, double* q_out, int negate,
                                                   int64_t n, double* __restrict__ part, const double* ppart, int pnb,
                                                   double prho, double* palpha, double* pcoef, double* pdot) {
    double c = 0;
    bool axpy = coef != nullptr;
    if constexpr (PMODE >= 1) {
        axpy = true;
        double d = 0;
        for (int i = threadIdx.x; i < pnb; i += kB) d = d + ppart[i];
        d = block_sum(d);
        const bool g0 = blockIdx.x == 0 && threadIdx.x == 0;
        if constexpr (PMODE == 1) {
            const double a = prho * d;
            c = -a;
            if (g0) *palpha = a;
        }
        if constexpr (PMODE == 2) {
#pragma clang fp contract(off)
            const double beta = prho * d;
            c = *palpha - beta;
        }
        if (g0) {
            *pdot = d;
            *pcoef = c;
        }
    } else {
        if (axpy) c = *coef;
    }
    double s = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * kB) {
        double qi = q_in[i];
        if (axpy) qi = qi + c * v[i];
        if (do_scale) qi = scale * qi;
        q_out[i] = negate ? -qi : qi;
        s += dnext[i] * qi;
    }
    s = block_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// The new curvature pair: s = m - x_old, y = g - g_old, with partials [slot][block] of y.s, y.y,
// s.s (sums) and max|s|, max|m| (maxima), each as k_sub, k_dot and k_maxabs form them.
__global__ void __launch_bounds__(kB) k_lbfgs_pair(const double* __restrict__ m, const double* __restrict__ xo,
                                                  const double* __restrict__ g, const double* __restrict__ go,
                                                  double* __restrict__ s_out, double* __restrict__ y_out, int64_t n,
                                                  double* __restrict__ part) {
    double ys = 0, yy = 0, ss = 0, smax = 0, mmax = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * kB) {
        const double mi = m[i];
        const double si = mi - xo[i];
        const double yi = g[i] - go[i];
        s_out[i] = si;
        y_out[i] = yi;
        ys += yi * si;
        yy += yi * yi;
        ss += si * si;
        smax = fmax(smax, fabs(si));
        mmax = fmax(mmax, fabs(mi));
    }
    ys = block_sum(ys);
    yy = block_sum(yy);
    ss = block_sum(ss);
    smax = block_max(smax);
    mmax = block_max(mmax);
    if (threadIdx.x == 0) {
        const int64_t nb = gridDim.x;
        part[0 * nb + blockIdx.x] = ys;
        part[1 * nb + blockIdx.x] = yy;
        part[2 * nb + blockIdx.x] = ss;
        part[3 * nb + blockIdx.x] = smax;
        part[4 * nb + blockIdx.x] = mmax;
    }
}

// k_final's one-slot sum, then (MODE 1) alpha_i = rho d, coef = -alpha_i or (MODE 2) coef =
// alpha_i - rho d with beta = rho d rounded first (the reference's host arithmetic rounds beta before
// the subtraction; a fused multiply-subtract changes the iterates by an ulp): the two-loop
// step's scalars without a launch of their own
template <int MODE>
__global__ void __launch_bounds__(kB) k_final_coef(const double* __restrict__ part, int nb, double* __restrict__ dot_out,
                                                  double rho, double* alpha_i, double* coef) {
    double v = 0;
    for (int i = threadIdx.x; i < nb; i += kB) v = v + part[i];
    v = block_sum(v);
    if (threadIdx.x == 0) {
        *dot_out = v;
        if constexpr (MODE == 1) {
            const double a = rho * v;
            *alpha_i = a;
            *coef = -a;
        }
        if constexpr (MODE == 2) {
#pragma clang fp contract(off)
            const double beta = rho * v;
            *coef = *alpha_i - beta;
        }
    }
}

__global__ void __launch_bounds__(kB) k_neg(double* __restrict__ y, const double* __restrict__ x, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (i < n) y[i] = -x[i];
}
} // namespace

void lbfgs_flush(mafx_ctx* ctx) {
    auto& p = ctx->lbfgs_pend;
    if (p.mode < 0) return;
    const int mode = p.mode;
    p.mode = -1;
    double* part = const_cast<double*>(p.part);
    switch (mode) {
    case 1:
        hipLaunchKernelGGL(k_final_coef<1>, dim3(1), dim3(kB), 0, ctx->stream, part, p.nb, p.dot, p.rho, p.alpha_i, p.coef);
        break;
    case 2:
        hipLaunchKernelGGL(k_final_coef<2>, dim3(1), dim3(kB), 0, ctx->stream, part, p.nb, p.dot, p.rho, p.alpha_i, p.coef);
        break;
    default:
        hipLaunchKernelGGL(k_final_coef<0>, dim3(1), dim3(kB), 0, ctx->stream, part, p.nb, p.dot, p.rho, p.alpha_i, p.coef);
        break;
    }
    MAFX_HIP(hipGetLastError());
}

namespace {
// The half of the partials buffer the pending step does not hold ([0, nb) or [nb, 2 nb); the
// buffer holds at least 10 red_blocks(3 N) doubles).
double* free_half(mafx_ctx* ctx, int nb) {
    double* base = ctx->partials.as<double>();
    return ctx->lbfgs_pend.mode >= 0 && ctx->lbfgs_pend.part == base ? base + nb : base;
}
void set_pending(mafx_ctx* ctx, int mode, const double* part, int nb, double rho, double* alpha_i, double* coef,
                 double* dot) {
    ctx->lbfgs_pend = mafx_ctx::LbfgsPending{mode, part, nb, rho, alpha_i, coef, dot};
}
} // namespace

void lbfgs_fused(mafx_ctx* ctx, const double* q_in, const double* coef_dev, const double* v, double scale,
                 bool do_scale, const double* dnext, double* q_out, bool negate, int64_t n, double* out_dev, int mode,
                 double rho, double* alpha_i, double* coef_out) {
    const unsigned nb = red_blocks(n);
    ProfScope ps(ctx, "lbfgs_twoloop_f64", 8.0 * n * (coef_dev ? 4 : 3));
    const auto pend = ctx->lbfgs_pend;
    // the previous step's coefficient is formed in this pass when it is the one this step uses
    const bool chain = pend.mode >= 1 && coef_dev != nullptr && coef_dev == pend.coef;
    if (!chain) lbfgs_flush(ctx);
    require(2 * static_cast<int64_t>(nb) <= static_cast<int64_t>(ctx->partials.bytes / sizeof(double)),
            "lbfgs: partials buffer too small");
    double* part = free_half(ctx, static_cast<int>(nb));
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(nb), dim3(kB), 0, ctx->stream, q_in, coef_dev, v, scale, do_scale ? 1 : 0, dnext,
                           q_out, negate ? 1 : 0, n, part, pend.part, pend.nb, pend.rho, pend.alpha_i, pend.coef, pend.dot);
    };
    if (chain && pend.mode == 1)
        go(k_lbfgs_fused<1>);
    else if (chain && pend.mode == 2)
        go(k_lbfgs_fused<2>);
    else
        go(k_lbfgs_fused<-1>);
    MAFX_HIP(hipGetLastError());
    set_pending(ctx, mode, part, static_cast<int>(nb), rho, alpha_i, coef_out, out_dev);
}

void lbfgs_dot_alpha(mafx_ctx* ctx, const double* a, const double* b, int64_t n, double* dot_out, double rho,
                     double* alpha_i, double* coef) {
    lbfgs_flush(ctx);
    const unsigned nb = red_blocks(n);
    ProfScope ps(ctx, "dot_f64", 16.0 * n);
    double* part = ctx->partials.as<double>();
    hipLaunchKernelGGL(k_dot, dim3(nb), dim3(kB), 0, ctx->stream, a, b, n, part);
    MAFX_HIP(hipGetLastError());
    set_pending(ctx, 1, part, static_cast<int>(nb), rho, alpha_i, coef, dot_out);
}

void lbfgs_pair(mafx_ctx* ctx, const double* m, const double* x_old, const double* g, const double* g_old,
                double* s, double* y, int64_t n, double* out_dev) {
    lbfgs_flush(ctx);
    const unsigned nb = red_blocks(n);
    ProfScope ps(ctx, "lbfgs_pair_f64", 48.0 * n);
    hipLaunchKernelGGL(k_lbfgs_pair, dim3(nb), dim3(kB), 0, ctx->stream, m, x_old, g, g_old, s, y, n,
                       ctx->partials.as<double>());
    hipLaunchKernelGGL(k_final, dim3(5), dim3(kB), 0, ctx->stream, ctx->partials.as<double>(), static_cast<int>(nb),
                       5, 3, out_dev);
    MAFX_HIP(hipGetLastError());
}

void neg_copy(mafx_ctx* ctx, double* y, const double* x, int64_t n) {
    hipLaunchKernelGGL(k_neg, dim3(grid_for(n, kB)), dim3(kB), 0, ctx->stream, y, x, n);
    MAFX_HIP(hipGetLastError());
}

void axpy_dev(mafx_ctx* ctx, double* y, const double* a_dev, const double* x, int64_t n) {
    hipLaunchKernelGGL(k_axpy_dev, dim3(grid_for(n, kB)), dim3(kB), 0, ctx->stream, y, a_dev, x, n);
    MAFX_HIP(hipGetLastError());
}

void read_device(mafx_ctx* ctx, const double* src_dev, int n, double* dst_host) {
    lbfgs_flush(ctx);
    MAFX_HIP(hipMemcpyAsync(ctx->host_scal, src_dev, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    MAFX_HIP(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; ++i) dst_host[i] = ctx->host_scal[i];
}

void llg_rhs(mafx_ctx* ctx, double alpha, double dt, const double* m, const double* h, double* k) {
    hipLaunchKernelGGL(k_llg_rhs, dim3(grid_for(ctx->N, kB)), dim3(kB), 0, ctx->stream, m, h, k, ctx->N, alpha, dt);
    MAFX_HIP(hipGetLastError());
}

template <class K>
void lincomb_k(mafx_ctx* ctx, double* out, const double* base, const LinArgs& a, int n, int64_t len) {
    const dim3 g(grid_for(len, kB)), b(kB);
    switch (n) {
    case 0: hipLaunchKernelGGL((k_lincomb<0, K>), g, b, 0, ctx->stream, out, base, a, len); break;
    case 1: hipLaunchKernelGGL((k_lincomb<1, K>), g, b, 0, ctx->stream, out, base, a, len); break;
    case 2: hipLaunchKernelGGL((k_lincomb<2, K>), g, b, 0, ctx->stream, out, base, a, len); break;
    case 3: hipLaunchKernelGGL((k_lincomb<3, K>), g, b, 0, ctx->stream, out, base, a, len); break;
    case 4: hipLaunchKernelGGL((k_lincomb<4, K>), g, b, 0, ctx->stream, out, base, a, len); break;
    case 5: hipLaunchKernelGGL((k_lincomb<5, K>), g, b, 0, ctx->stream, out, base, a, len); break;
    case 6: hipLaunchKernelGGL((k_lincomb<6, K>), g, b, 0, ctx->stream, out, base, a, len); break;
    case 7: hipLaunchKernelGGL((k_lincomb<7, K>), g, b, 0, ctx->stream, out, base, a, len); break;
    case 8: hipLaunchKernelGGL((k_lincomb<8, K>), g, b, 0, ctx->stream, out, base, a, len); break;
    case 9: hipLaunchKernelGGL((k_lincomb<9, K>), g, b, 0, ctx->stream, out, base, a, len); break;
    case 10: hipLaunchKernelGGL((k_lincomb<10, K>), g, b, 0, ctx->stream, out, base, a, len); break;
    case 11: hipLaunchKernelGGL((k_lincomb<11, K>), g, b, 0, ctx->stream, out, base, a, len); break;
    default: hipLaunchKernelGGL((k_lincomb<12, K>), g, b, 0, ctx->stream, out, base, a, len); break;
    }
}

void lincomb(mafx_ctx* ctx, double* out, const double* base, int n, const void* const* v, const double* c,
             int64_t len, bool v_f32) {
    require(n >= 0 && n <= kMaxStages - 1, "lincomb: at most 12 vectors");
    LinArgs a{};
    a.n = n;
    for (int j = 0; j < n; ++j) {
        a.v[j] = v[j];
        a.c[j] = c[j];
    }
    int nz = 0;
    for (int j = 0; j < n; ++j) nz += c[j] != 0.0;
    ProfScope ps(ctx, v_f32 ? "rk_stage_k32" : "rk_stage_f64", (v_f32 ? 4.0 * nz + 16.0 : 8.0 * (nz + 2)) * len);
    if (v_f32)
        lincomb_k<float>(ctx, out, base, a, n, len);
    else
        lincomb_k<double>(ctx, out, base, a, n, len);
    MAFX_HIP(hipGetLastError());
}

void line_point(mafx_ctx* ctx, double* x, const double* a, double s, const double* b) {
    ProfScope ps(ctx, "line_point_f64", 72.0 * ctx->N);
    hipLaunchKernelGGL(k_line_point, dim3(grid_for(ctx->N, kB)), dim3(kB), 0, ctx->stream, x, a, s, b, ctx->N);
    MAFX_HIP(hipGetLastError());
}

template <class K, int P>
void rk_error_k(mafx_ctx* ctx, const double* m, const ErrArgs& a, int ns, bool rkf_form, double* sumbk, int64_t len,
                unsigned nb) {
    double* part = ctx->partials.as<double>();
    const dim3 g(nb), bl(kB);
    auto go = [&](auto kern) { hipLaunchKernelGGL(kern, g, bl, 0, ctx->stream, m, a, sumbk, len, part); };
    if (rkf_form) {
        switch (ns) {
        case 1: go(k_rk_error<1, true, K, P>); break;
        case 2: go(k_rk_error<2, true, K, P>); break;
        case 3: go(k_rk_error<3, true, K, P>); break;
        case 4: go(k_rk_error<4, true, K, P>); break;
        case 5: go(k_rk_error<5, true, K, P>); break;
        case 6: go(k_rk_error<6, true, K, P>); break;
        case 7: go(k_rk_error<7, true, K, P>); break;
        case 8: go(k_rk_error<8, true, K, P>); break;
        default: go(k_rk_error<13, true, K, P>); break;
        }
    } else {
        switch (ns) {
        case 1: go(k_rk_error<1, false, K, P>); break;
        case 2: go(k_rk_error<2, false, K, P>); break;
        case 3: go(k_rk_error<3, false, K, P>); break;
        case 4: go(k_rk_error<4, false, K, P>); break;
        case 5: go(k_rk_error<5, false, K, P>); break;
        case 6: go(k_rk_error<6, false, K, P>); break;
        default: go(k_rk_error<7, false, K, P>); break;
        }
    }
}

double rk_error(mafx_ctx* ctx, const double* m, const void* const* k, const double* b, const double* e, int ns,
                bool rkf_form, double atol, double rtol, double* sumbk, bool k_f32) {
    ErrArgs a{};
    a.ns = ns;
    for (int j = 1; j <= ns; ++j) {
        a.k[j] = k[j];
        a.b[j] = b[j];
        a.e[j] = e[j];
    }
    a.atol = atol;
    a.rtol = rtol;
    const int64_t len = 3 * ctx->N;
    require((ns >= 1 && ns <= 8) || ns == 13, "rk_error: 1..8 or 13 stages");
    // float stage derivatives in pairs when every array is aligned for it
    bool pairs = k_f32 && reinterpret_cast<uintptr_t>(m) % 16 == 0 && reinterpret_cast<uintptr_t>(sumbk) % 16 == 0;
    for (int j = 1; j <= ns && pairs; ++j) pairs = reinterpret_cast<uintptr_t>(k[j]) % 8 == 0;
    // one resident round of the kernel (6 groups per CU at <= 80 VGPRs, 4 for the pair form at
    // <= 128), each lane taking two units per trip; never more partials than the context's buffer
    // holds (50 red_blocks)
    const int64_t per_lane = pairs ? 4 : 2;
    const unsigned nb = static_cast<unsigned>(std::max<int64_t>(
        1, std::min<int64_t>({(len + per_lane * kB - 1) / (per_lane * kB), int64_t{pairs ? 4 : 6} * ctx->num_cus,
                              int64_t{50} * red_blocks(len)})));
    require(ns <= 7 || rkf_form, "rk_error: more than 7 stages only in the sumbk - sum e k form");
    int nread = 0;
    for (int j = 1; j <= ns; ++j) nread += (b[j] != 0.0 || e[j] != 0.0);
    {
        ProfScope ps(ctx, k_f32 ? "rk_error_k32" : "rk_error_f64", (k_f32 ? 4.0 * nread + 16.0 : 8.0 * (nread + 2)) * len);
        if (pairs)
            rk_error_k<float, 2>(ctx, m, a, ns, rkf_form, sumbk, len, nb);
        else if (k_f32)
            rk_error_k<float, 1>(ctx, m, a, ns, rkf_form, sumbk, len, nb);
        else
            rk_error_k<double, 1>(ctx, m, a, ns, rkf_form, sumbk, len, nb);
        hipLaunchKernelGGL(k_final, dim3(1), dim3(kB), 0, ctx->stream, ctx->partials.as<double>(),
                           static_cast<int>(nb), 1, 0, ctx->scal.as<double>());
    }
    MAFX_HIP(hipGetLastError());
    return read_scalar(ctx, 1);
}

void accept(mafx_ctx* ctx, double* m, const double* dm, bool normalize_m, bool zero_safe) {
    ProfScope ps(ctx, "rk_accept_f64", 72.0 * ctx->N);
    hipLaunchKernelGGL(k_accept, dim3(grid_for(ctx->N, kB)), dim3(kB), 0, ctx->stream, m, dm, ctx->N,
                       normalize_m ? 1 : 0, zero_safe ? 1 : 0);
    MAFX_HIP(hipGetLastError());
}

void axpy(mafx_ctx* ctx, double* y, double a, const double* x, int64_t n) {
    hipLaunchKernelGGL(k_axpy, dim3(grid_for(n, kB)), dim3(kB), 0, ctx->stream, y, a, x, n);
    MAFX_HIP(hipGetLastError());
}
void scale(mafx_ctx* ctx, double* y, double a, int64_t n) {
    hipLaunchKernelGGL(k_scale, dim3(grid_for(n, kB)), dim3(kB), 0, ctx->stream, y, a, n);
    MAFX_HIP(hipGetLastError());
}
void sub(mafx_ctx* ctx, double* out, const double* a, const double* b, int64_t n) {
    hipLaunchKernelGGL(k_sub, dim3(grid_for(n, kB)), dim3(kB), 0, ctx->stream, out, a, b, n);
    MAFX_HIP(hipGetLastError());
}
void copy(mafx_ctx* ctx, double* dst, const double* src, int64_t n) {
    MAFX_HIP(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
}

// the widest reductions over n entries use 5 slots (lbfgs_pair)
int64_t reduction_partials_needed(int64_t n) { return static_cast<int64_t>(red_blocks(n)) * 5; }

} // namespace mafx
