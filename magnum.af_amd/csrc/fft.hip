// FFT plans and axis kernels of the zero-padded demag convolution
// (replaces af::fftR2C / af::fftC2R in demag_field.cpp:229-257 and
// atomistic_dipole_dipole_field.cpp:20-51).
//
// Spectrum layout: [comp][kz][ky][kx] complex with kx in [0, hx), hx = n0/2 + 1 (R2C half
// spectrum along x), ky in [0, n1), kz in [0, n2).  Only the rows that can hold non-zero data are
// touched: the forward x pass writes rows y < ny, z < nz; the forward y pass reads y < ny and
// writes every ky for z < nz; the z pass reads z < nz and writes every kz; the inverse passes
// mirror this and stop at the cropped extent.
#include <cmath>

#include "fft_device.hpp"
#include "fft_kernels.hpp"
#include "mafx_internal.hpp"

namespace mafx {

void FftPlan::build(int len) {
    L = len;
    nstages = 0;
    int n = len;
    auto push = [&](int r) {
        require(nstages < 16, "FFT plan: too many stages");
        radix[nstages++] = r;
    };
    while (n % 8 == 0) {
        push(8);
        n /= 8;
    }
    while (n % 4 == 0) {
        push(4);
        n /= 4;
    }
    while (n % 2 == 0) {
        push(2);
        n /= 2;
    }
    for (int p : {3, 5, 7, 11, 13}) {
        while (n % p == 0) {
            push(p);
            n /= p;
        }
    }
    require(n == 1, "FFT length " + std::to_string(len) +
                        " has a prime factor > 13 (supported radices: 2, 3, 4, 5, 7, 8, 11, 13)");
    std::vector<double2> t64(len);
    std::vector<float2> t32(len);
    const long double two_pi = 6.283185307179586476925286766559L;
    for (int t = 0; t < len; ++t) {
        const long double a = two_pi * static_cast<long double>(t) / static_cast<long double>(len);
        const double c = static_cast<double>(std::cos(a)), s = -static_cast<double>(std::sin(a));
        t64[t] = double2{c, s};
        t32[t] = float2{static_cast<float>(std::cos(a)), -static_cast<float>(std::sin(a))};
    }
    tw64.alloc(sizeof(double2) * len);
    tw32.alloc(sizeof(float2) * len);
    MAFX_HIP(hipMemcpy(tw64.p, t64.data(), sizeof(double2) * len, hipMemcpyHostToDevice));
    MAFX_HIP(hipMemcpy(tw32.p, t32.data(), sizeof(float2) * len, hipMemcpyHostToDevice));
}

namespace {

FftDesc desc_of(const FftPlan& p) {
    FftDesc d{};
    d.L = p.L;
    d.nstages = p.nstages;
    for (int i = 0; i < p.nstages; ++i) d.radix[i] = p.radix[i];
    return d;
}

template <class C> const C* tw_of(const FftPlan& p);
template <> const double2* tw_of<double2>(const FftPlan& p) { return p.tw64.as<double2>(); }
template <> const float2* tw_of<float2>(const FftPlan& p) { return p.tw32.as<float2>(); }

template <class C> constexpr int lds_elems() { return 65536 / static_cast<int>(sizeof(C)); }

// Forward real-to-complex transform along x of rows (c, z, y), zero-padded from Xi to L.
// in[((c Zi + z) Yi + y) Xi + x] * scale * (ms ? ms[(z Yi + y) Xi + x] : 1)
template <class C>
__global__ void __launch_bounds__(kFftBlock)
    k_fft_x_fwd(const double* __restrict__ in, const double* __restrict__ ms, double scale, int64_t Xi, int64_t Yi,
                int64_t Zi, int64_t Yr, int64_t Zr, int64_t rows, int W, FftDesc d, const C* __restrict__ tw,
                C* __restrict__ out, int64_t hx, int64_t n1, int64_t n2) {
    using T = typename real_of<C>::type;
    extern __shared__ unsigned char smem_raw[];
    C* buf = reinterpret_cast<C*>(smem_raw);
    const int L = d.L;
    const int64_t row0 = static_cast<int64_t>(blockIdx.x) * W;
    for (int e = threadIdx.x; e < L * W; e += kFftBlock) {
        const int w = e % W, pos = e / W;
        const int64_t r = row0 + w;
        T v = 0;
        if (r < rows && pos < Xi) {
            const int64_t y = r % Yr, zc = r / Yr, z = zc % Zr, c = zc / Zr;
            const int64_t cell = (z * Yi + y) * Xi + pos;
            double val = in[(c * Zi * Yi) * Xi + cell] * scale;
            if (ms) val *= ms[cell];
            v = static_cast<T>(val);
        }
        buf[pos * W + w] = C{v, 0};
    }
    __syncthreads();
    fft_lds<C, lds_elems<C>()>(buf, d, W, tw, false);
    for (int e = threadIdx.x; e < hx * W; e += kFftBlock) {
        const int w = e % W, k = e / W;
        const int64_t r = row0 + w;
        if (r < rows) {
            const int64_t y = r % Yr, zc = r / Yr, z = zc % Zr, c = zc / Zr;
            out[((c * n2 + z) * n1 + y) * hx + k] = buf[k * W + w];
        }
    }
}

// Complex transform of strided lines: line (tile, b, a) covers columns kx0 .. kx0+W-1 and holds
// data[a sa + b sb + kx + pos * stride].  Inputs pos >= n_in are taken as zero; outputs pos >= n_out
// are not written.
template <class C>
__global__ void __launch_bounds__(kFftBlock)
    k_fft_strided(C* __restrict__ data, int64_t sa, int64_t sb, int64_t stride, int64_t hx, int n_in, int n_out,
                  int W, FftDesc d, const C* __restrict__ tw, int inv) {
    extern __shared__ unsigned char smem_raw[];
    C* buf = reinterpret_cast<C*>(smem_raw);
    const int L = d.L;
    const int64_t kx0 = static_cast<int64_t>(blockIdx.x) * W;
    C* base = data + static_cast<int64_t>(blockIdx.z) * sa + static_cast<int64_t>(blockIdx.y) * sb + kx0;
    const int wv = static_cast<int>(hx - kx0 < W ? hx - kx0 : W);
    for (int e = threadIdx.x; e < L * W; e += kFftBlock) {
        const int w = e % W, pos = e / W;
        C v{0, 0};
        if (w < wv && pos < n_in) v = base[pos * stride + w];
        buf[pos * W + w] = v;
    }
    __syncthreads();
    fft_lds<C, lds_elems<C>()>(buf, d, W, tw, inv != 0);
    for (int e = threadIdx.x; e < n_out * W; e += kFftBlock) {
        const int w = e % W, pos = e / W;
        if (w < wv) base[pos * stride + w] = buf[pos * W + w];
    }
}

// H^ = N^ M^ per k-point with the symmetric tensor (xx, xy, xz, yy, yz, zz) stored as full complex
// (demag_field.cpp:236-249); in place on the spectrum.
template <class C>
__global__ void __launch_bounds__(256) k_kspace_mul(C* __restrict__ s, const C* __restrict__ t, int64_t K) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const C m0 = s[k], m1 = s[K + k], m2 = s[2 * K + k];
    const C nxx = t[k], nxy = t[K + k], nxz = t[2 * K + k], nyy = t[3 * K + k], nyz = t[4 * K + k],
            nzz = t[5 * K + k];
    s[k] = cadd(cadd(cmul(nxx, m0), cmul(nxy, m1)), cmul(nxz, m2));
    s[K + k] = cadd(cadd(cmul(nxy, m0), cmul(nyy, m1)), cmul(nyz, m2));
    s[2 * K + k] = cadd(cadd(cmul(nxz, m0), cmul(nyz, m1)), cmul(nzz, m2));
}

__global__ void k_c64_to_c32(const double2* __restrict__ a, float2* __restrict__ b, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) b[i] = float2{static_cast<float>(a[i].x), static_cast<float>(a[i].y)};
}

int lines_per_block(int L, int elems) {
    int W = elems / L;
    if (W < 1) W = 1;
    if (W > 64) W = 64;
    return W;
}

} // namespace

template <class C>
void fft_x_forward(mafx_ctx* ctx, const FftPlan& p, const double* in, const double* ms, double scale, int64_t Xi,
                   int64_t Yi, int64_t Zi, int ncomp, int64_t Yr, int64_t Zr, C* out) {
    const int L = p.L;
    require(L <= lds_elems<C>(), "x transform length exceeds the LDS budget");
    const int W = lines_per_block(L, lds_elems<C>());
    const int64_t rows = static_cast<int64_t>(ncomp) * Zr * Yr;
    const unsigned grid = static_cast<unsigned>((rows + W - 1) / W);
    const double bytes = static_cast<double>(rows) * (Xi * 8.0 + ctx->hx * sizeof(C));
    ProfScope ps(ctx, sizeof(C) == 8 ? "fft_x_fwd_f32" : "fft_x_fwd_f64", bytes);
    hipLaunchKernelGGL(k_fft_x_fwd<C>, dim3(grid), dim3(kFftBlock), W * L * sizeof(C), ctx->stream, in, ms, scale,
                       Xi, Yi, Zi, Yr, Zr, rows, W, desc_of(p), tw_of<C>(p), out, ctx->hx, ctx->n1, ctx->n2);
    MAFX_HIP(hipGetLastError());
}

template <class C>
void fft_strided(mafx_ctx* ctx, const FftPlan& p, C* data, int64_t sa, int na, int64_t sb, int nb, int64_t stride,
                 int n_in, int n_out, bool inv, const char* name) {
    const int L = p.L;
    require(L <= lds_elems<C>(), "strided transform length exceeds the LDS budget");
    const int W = lines_per_block(L, lds_elems<C>());
    const unsigned gx = static_cast<unsigned>((ctx->hx + W - 1) / W);
    const double bytes = static_cast<double>(na) * nb * ctx->hx * (n_in + n_out) * sizeof(C);
    ProfScope ps(ctx, name, bytes);
    hipLaunchKernelGGL(k_fft_strided<C>, dim3(gx, nb, na), dim3(kFftBlock), W * L * sizeof(C), ctx->stream, data, sa,
                       sb, stride, ctx->hx, n_in, n_out, W, desc_of(p), tw_of<C>(p), inv ? 1 : 0);
    MAFX_HIP(hipGetLastError());
}

template <class C> void kspace_mul(mafx_ctx* ctx, C* spec, const C* tensor, int64_t K) {
    ProfScope ps(ctx, sizeof(C) == 8 ? "kspace_mul_f32" : "kspace_mul_f64", static_cast<double>(K) * 12 * sizeof(C));
    hipLaunchKernelGGL(k_kspace_mul<C>, dim3(grid_for(K, 256)), dim3(256), 0, ctx->stream, spec, tensor, K);
    MAFX_HIP(hipGetLastError());
}

void convert_tensor_f32(mafx_ctx* ctx, const double2* a, float2* b, int64_t n) {
    hipLaunchKernelGGL(k_c64_to_c32, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, a, b, n);
    MAFX_HIP(hipGetLastError());
}

template void fft_x_forward<double2>(mafx_ctx*, const FftPlan&, const double*, const double*, double, int64_t,
                                     int64_t, int64_t, int, int64_t, int64_t, double2*);
template void fft_x_forward<float2>(mafx_ctx*, const FftPlan&, const double*, const double*, double, int64_t,
                                    int64_t, int64_t, int, int64_t, int64_t, float2*);
template void fft_strided<double2>(mafx_ctx*, const FftPlan&, double2*, int64_t, int, int64_t, int, int64_t, int,
                                   int, bool, const char*);
template void fft_strided<float2>(mafx_ctx*, const FftPlan&, float2*, int64_t, int, int64_t, int, int64_t, int, int,
                                  bool, const char*);
template void kspace_mul<double2>(mafx_ctx*, double2*, const double2*, int64_t);
template void kspace_mul<float2>(mafx_ctx*, float2*, const float2*, int64_t);

} // namespace mafx
