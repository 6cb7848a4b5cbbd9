// FFT plans and axis kernels of the zero-padded demag convolution
// (replaces af::fftR2C / af::fftC2R in demag_field.cpp:229-257 and
// atomistic_dipole_dipole_field.cpp:20-51).
//
// Spectrum layout: [comp][kz][ky][kx] complex with kx in [0, hx), hx = n0/2 + 1 (R2C half
// spectrum along x), ky in [0, n1), kz in [0, nzs) where nzs = nz when the z transform is fused
// into the tensor product (2 nz <= 16) and nzs = n2 otherwise.  Only rows that can hold non-zero
// data are touched: the forward x pass writes rows y < ny, z < nz; the forward y pass reads
// y < ny and writes every ky for z < nz; the fused z kernel reads and writes z < nz; the inverse
// passes stop at the cropped extent.
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

FftDesc desc_of(const FftPlan& p) {
    FftDesc d{};
    d.L = p.L;
    d.nstages = p.nstages;
    for (int i = 0; i < p.nstages; ++i) d.radix[i] = p.radix[i];
    return d;
}

namespace {

template <class C> const C* tw_of(const FftPlan& p);
template <> const double2* tw_of<double2>(const FftPlan& p) { return p.tw64.as<double2>(); }
template <> const float2* tw_of<float2>(const FftPlan& p) { return p.tw32.as<float2>(); }

template <class C> constexpr int lds_elems() { return 65536 / static_cast<int>(sizeof(C)); }

// 512-thread work-groups for the axis transforms: W L <= 64 KiB of complex data spread over twice
// the lanes halves the butterflies each lane keeps in registers across a stage.
constexpr int kAxisBlock = 512;

// Forward real-to-complex transform along x of rows (c, z, y), zero-padded from Xi to L.
// in[((c Zi + z) Yi + y) Xi + x] * scale * (ms ? ms[(z Yi + y) Xi + x] : 1); rows land at
// out[((c n2o + z) n1 + y) hx + k].
template <class C, bool POW2>
__global__ void __launch_bounds__(kAxisBlock)
    k_fft_x_fwd(const double* __restrict__ in, const double* __restrict__ ms, double scale, int64_t Xi, int64_t Yi,
                int64_t Zi, int64_t Yr, int64_t Zr, int64_t rows, int W, FftDesc d, const C* __restrict__ tw,
                C* __restrict__ out, int64_t hx, int64_t n1, int64_t n2o) {
    using T = typename real_of<C>::type;
    extern __shared__ unsigned char smem_raw[];
    C* buf = reinterpret_cast<C*>(smem_raw);
    const int L = d.L;
    const int64_t row0 = static_cast<int64_t>(blockIdx.x) * W;
    for (int e = threadIdx.x; e < L * W; e += kAxisBlock) {
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
    fft_lds<C, lds_elems<C>(), POW2, kAxisBlock>(buf, d, W, tw, false);
    for (int e = threadIdx.x; e < hx * W; e += kAxisBlock) {
        const int w = e % W, k = e / W;
        const int64_t r = row0 + w;
        if (r < rows) {
            const int64_t y = r % Yr, zc = r / Yr, z = zc % Zr, c = zc / Zr;
            out[((c * n2o + z) * n1 + y) * hx + k] = buf[k * W + w];
        }
    }
}

// Complex transform of strided lines: line (tile, b, a) covers columns kx0 .. kx0+W-1 and holds
// data[a sa + b sb + kx + pos * stride].  Inputs pos >= n_in are taken as zero; outputs pos >= n_out
// are not written.
template <class C, bool POW2>
__global__ void __launch_bounds__(kAxisBlock)
    k_fft_strided(C* __restrict__ data, int64_t sa, int64_t sb, int64_t stride, int64_t hx, int n_in, int n_out,
                  int W, FftDesc d, const C* __restrict__ tw, int inv) {
    extern __shared__ unsigned char smem_raw[];
    C* buf = reinterpret_cast<C*>(smem_raw);
    const int L = d.L;
    const int64_t kx0 = static_cast<int64_t>(blockIdx.x) * W;
    C* base = data + static_cast<int64_t>(blockIdx.z) * sa + static_cast<int64_t>(blockIdx.y) * sb + kx0;
    const int wv = static_cast<int>(hx - kx0 < W ? hx - kx0 : W);
    for (int e = threadIdx.x; e < L * W; e += kAxisBlock) {
        const int w = e % W, pos = e / W;
        C v{0, 0};
        if (w < wv && pos < n_in) v = base[pos * stride + w];
        buf[pos * W + w] = v;
    }
    __syncthreads();
    fft_lds<C, lds_elems<C>(), POW2, kAxisBlock>(buf, d, W, tw, inv != 0);
    for (int e = threadIdx.x; e < n_out * W; e += kAxisBlock) {
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

// Twiddles w_{2 NZ}^t, t < 2 NZ, passed by value.
template <class C> struct ZTw {
    C w[16];
};

// Fused z stage, one thread per (ky, kx) column: forward DFT of length 2 NZ of the NZ non-zero z
// planes, tensor product per kz, inverse DFT kept for z < NZ.  Streams over kz so only the NZ
// inputs and NZ outputs per component stay in registers.  Spectrum planes [3][NZ][plane],
// tensor [6][2 NZ][plane].
template <class C, int NZ>
__global__ void __launch_bounds__(256)
    k_zfused(C* __restrict__ s, const C* __restrict__ t, int64_t plane, ZTw<C> tw) {
    constexpr int NZ2 = 2 * NZ;
    const int64_t col = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (col >= plane) return;
    C m[3][NZ], h[3][NZ];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int z = 0; z < NZ; ++z) {
            m[c][z] = s[(c * NZ + z) * plane + col];
            h[c][z] = C{0, 0};
        }
#pragma unroll
    for (int k = 0; k < NZ2; ++k) {
        C X[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            C acc = m[c][0];
#pragma unroll
            for (int z = 1; z < NZ; ++z) acc = cadd(acc, cmul(m[c][z], tw.w[(k * z) % NZ2]));
            X[c] = acc;
        }
        const C nxx = t[(0 * NZ2 + k) * plane + col], nxy = t[(1 * NZ2 + k) * plane + col],
                nxz = t[(2 * NZ2 + k) * plane + col], nyy = t[(3 * NZ2 + k) * plane + col],
                nyz = t[(4 * NZ2 + k) * plane + col], nzz = t[(5 * NZ2 + k) * plane + col];
        const C H0 = cadd(cadd(cmul(nxx, X[0]), cmul(nxy, X[1])), cmul(nxz, X[2]));
        const C H1 = cadd(cadd(cmul(nxy, X[0]), cmul(nyy, X[1])), cmul(nyz, X[2]));
        const C H2 = cadd(cadd(cmul(nxz, X[0]), cmul(nyz, X[1])), cmul(nzz, X[2]));
#pragma unroll
        for (int z = 0; z < NZ; ++z) {
            const C w = conj_if(tw.w[(k * z) % NZ2], true);
            h[0][z] = cadd(h[0][z], z == 0 ? H0 : cmul(H0, w));
            h[1][z] = cadd(h[1][z], z == 0 ? H1 : cmul(H1, w));
            h[2][z] = cadd(h[2][z], z == 0 ? H2 : cmul(H2, w));
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int z = 0; z < NZ; ++z) s[(c * NZ + z) * plane + col] = h[c][z];
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
                   int64_t Yi, int64_t Zi, int ncomp, int64_t Yr, int64_t Zr, C* out, int64_t n2o) {
    const int L = p.L;
    require(L <= lds_elems<C>(), "x transform length exceeds the LDS budget");
    const int W = lines_per_block(L, lds_elems<C>());
    const int64_t rows = static_cast<int64_t>(ncomp) * Zr * Yr;
    const unsigned grid = static_cast<unsigned>((rows + W - 1) / W);
    const double bytes = static_cast<double>(rows) * (Xi * 8.0 + ctx->hx * sizeof(C));
    const FftDesc d = desc_of(p);
    ProfScope ps(ctx, sizeof(C) == 8 ? "fft_x_fwd_f32" : "fft_x_fwd_f64", bytes);
    if (is_pow2_plan(d))
        hipLaunchKernelGGL((k_fft_x_fwd<C, true>), dim3(grid), dim3(kAxisBlock), W * L * sizeof(C), ctx->stream, in,
                           ms, scale, Xi, Yi, Zi, Yr, Zr, rows, W, d, tw_of<C>(p), out, ctx->hx, ctx->n1, n2o);
    else
        hipLaunchKernelGGL((k_fft_x_fwd<C, false>), dim3(grid), dim3(kAxisBlock), W * L * sizeof(C), ctx->stream, in,
                           ms, scale, Xi, Yi, Zi, Yr, Zr, rows, W, d, tw_of<C>(p), out, ctx->hx, ctx->n1, n2o);
    MAFX_HIP(hipGetLastError());
}

template <class C>
void fft_strided(mafx_ctx* ctx, const FftPlan& p, C* data, int64_t sa, int na, int64_t sb, int nb, int64_t stride,
                 int n_in, int n_out, bool inv, const char* name) {
    const int L = p.L;
    require(L <= lds_elems<C>(), "strided transform length exceeds the LDS budget");
    require(nb <= 65535 && na <= 65535, "strided transform grid too large");
    const int W = lines_per_block(L, lds_elems<C>());
    const unsigned gx = static_cast<unsigned>((ctx->hx + W - 1) / W);
    const double bytes = static_cast<double>(na) * nb * ctx->hx * (n_in + n_out) * sizeof(C);
    const FftDesc d = desc_of(p);
    ProfScope ps(ctx, name, bytes);
    if (is_pow2_plan(d))
        hipLaunchKernelGGL((k_fft_strided<C, true>), dim3(gx, nb, na), dim3(kAxisBlock), W * L * sizeof(C),
                           ctx->stream, data, sa, sb, stride, ctx->hx, n_in, n_out, W, d, tw_of<C>(p), inv ? 1 : 0);
    else
        hipLaunchKernelGGL((k_fft_strided<C, false>), dim3(gx, nb, na), dim3(kAxisBlock), W * L * sizeof(C),
                           ctx->stream, data, sa, sb, stride, ctx->hx, n_in, n_out, W, d, tw_of<C>(p), inv ? 1 : 0);
    MAFX_HIP(hipGetLastError());
}

template <class C> void kspace_mul(mafx_ctx* ctx, C* spec, const C* tensor, int64_t K) {
    ProfScope ps(ctx, sizeof(C) == 8 ? "kspace_mul_f32" : "kspace_mul_f64", static_cast<double>(K) * 12 * sizeof(C));
    hipLaunchKernelGGL(k_kspace_mul<C>, dim3(grid_for(K, 256)), dim3(256), 0, ctx->stream, spec, tensor, K);
    MAFX_HIP(hipGetLastError());
}

template <class C> void zfused(mafx_ctx* ctx, C* spec, const C* tensor, int nz) {
    const int64_t plane = ctx->n1 * ctx->hx;
    ZTw<C> tw{};
    const int n2 = 2 * nz;
    const long double two_pi = 6.283185307179586476925286766559L;
    for (int t = 0; t < n2; ++t) {
        const long double a = two_pi * t / n2;
        using T = typename real_of<C>::type;
        tw.w[t] = C{static_cast<T>(std::cos(a)), static_cast<T>(-std::sin(a))};
    }
    const double bytes = static_cast<double>(plane) * (2.0 * 3 * nz + 6.0 * n2) * sizeof(C);
    ProfScope ps(ctx, sizeof(C) == 8 ? "zfused_mul_f32" : "zfused_mul_f64", bytes);
    const dim3 g(grid_for(plane, 256)), b(256);
    switch (nz) {
    case 2: hipLaunchKernelGGL((k_zfused<C, 2>), g, b, 0, ctx->stream, spec, tensor, plane, tw); break;
    case 3: hipLaunchKernelGGL((k_zfused<C, 3>), g, b, 0, ctx->stream, spec, tensor, plane, tw); break;
    case 4: hipLaunchKernelGGL((k_zfused<C, 4>), g, b, 0, ctx->stream, spec, tensor, plane, tw); break;
    case 5: hipLaunchKernelGGL((k_zfused<C, 5>), g, b, 0, ctx->stream, spec, tensor, plane, tw); break;
    case 6: hipLaunchKernelGGL((k_zfused<C, 6>), g, b, 0, ctx->stream, spec, tensor, plane, tw); break;
    case 7: hipLaunchKernelGGL((k_zfused<C, 7>), g, b, 0, ctx->stream, spec, tensor, plane, tw); break;
    case 8: hipLaunchKernelGGL((k_zfused<C, 8>), g, b, 0, ctx->stream, spec, tensor, plane, tw); break;
    default: throw Error(MAFX_ERR_ARG, "fused z stage supports 2 <= nz <= 8");
    }
    MAFX_HIP(hipGetLastError());
}

void convert_tensor_f32(mafx_ctx* ctx, const double2* a, float2* b, int64_t n) {
    hipLaunchKernelGGL(k_c64_to_c32, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, a, b, n);
    MAFX_HIP(hipGetLastError());
}

template void fft_x_forward<double2>(mafx_ctx*, const FftPlan&, const double*, const double*, double, int64_t,
                                     int64_t, int64_t, int, int64_t, int64_t, double2*, int64_t);
template void fft_x_forward<float2>(mafx_ctx*, const FftPlan&, const double*, const double*, double, int64_t,
                                    int64_t, int64_t, int, int64_t, int64_t, float2*, int64_t);
template void fft_strided<double2>(mafx_ctx*, const FftPlan&, double2*, int64_t, int, int64_t, int, int64_t, int,
                                   int, bool, const char*);
template void fft_strided<float2>(mafx_ctx*, const FftPlan&, float2*, int64_t, int, int64_t, int, int64_t, int, int,
                                  bool, const char*);
template void kspace_mul<double2>(mafx_ctx*, double2*, const double2*, int64_t);
template void kspace_mul<float2>(mafx_ctx*, float2*, const float2*, int64_t);
template void zfused<double2>(mafx_ctx*, double2*, const double2*, int);
template void zfused<float2>(mafx_ctx*, float2*, const float2*, int);

} // namespace mafx
