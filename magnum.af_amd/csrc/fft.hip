// FFT plans and axis kernels of the zero-padded demag convolution
// (replaces af::fftR2C / af::fftC2R in demag_field.cpp:229-257 and
// atomistic_dipole_dipole_field.cpp:20-51).
//
// Spectrum layout: planes [comp][kz] of n1 hxs complex values, kx in [0, hx), hx = n0/2 + 1 (R2C
// half spectrum along x), hxs >= hx padded to 16 columns; ky in [0, n1); kz in [0, nzs) where
// nzs = nz when the z transform is fused into the tensor product (2 <= nz <= 8) and nzs = n2
// otherwise.  Inside a plane the columns are grouped in tiles of 4 (tcol, fft_device.hpp):
// (ky, kx) sits at tcol(kx, 4 n1) + 4 ky.  The tensor spectra use the same planes.  Only rows that can hold non-zero data are touched: the forward x pass
// writes rows y < ny, z < nz; the forward y pass reads y < ny and writes every ky for z < nz; the
// fused z kernel reads and writes z < nz; the inverse passes stop at the cropped extent.
//
// Both x passes run a complex FFT of half the padded length on the packed sequence
// z[n] = x[2n] + i x[2n+1] (the standard real-FFT packing); the zero-padded y passes split the
// length-2M transform into its even and odd outputs, two length-M transforms.
#include <cmath>
#include <map>
#include <mutex>
#include <type_traits>

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
    int ns = 1;
    for (int i = 0; i < p.nstages; ++i) {
        d.radix[i] = p.radix[i];
        d.ns[i] = make_fastdiv(static_cast<uint32_t>(ns));
        d.twstep[i] = p.L / (ns * p.radix[i]);
        ns *= p.radix[i];
    }
    return d;
}

namespace {

// The reference spectrum's entry c at a k-point: the symmetric quarter's (signed) value plus the
// stored FP16 residual, unscaled.
__device__ __forceinline__ float2 res_entry(const TensorRes& res, int c, int64_t idx, float sym) {
    const float2 r = __half22float2(res.r[idx]);
    return float2{sym + r.x * res.inv[c], r.y * res.inv[c]};
}

// 512-thread work-groups for the axis transforms: W L <= 64 KiB of complex data spread over twice
// the lanes halves the butterflies each lane keeps in registers across a stage.
constexpr int kAxisBlock = 512;
// Minimum waves per SIMD requested for the power-of-two axis kernels: 4 caps them at 128 VGPRs, so
// two 512-lane work-groups (the 64 KiB LDS budget allows two) are resident per CU.
constexpr int kMinWaves = 4;

// Forward real-to-complex transform along x of rows (c, z, y): padded length L = 2 M, input
// x[t] = in[((c Zi + z) Yi + y) Xi + t] * scale * (ms ? ms[cell] : 1) for t < Xi, 0 beyond.
// Z = FFT_M(x[2n] + i x[2n+1]);  X[k] = E + w_L^k O,  E = (Z[k] + conj Z[M-k]) / 2,
// O = -i (Z[k] - conj Z[M-k]) / 2, k = 0..M, written to out[(c n2o + z) n1 hxs + 4 y + tcol(k)]
// (column tiles).  The input lanes walk along rows; the output lanes take the 4 columns of a
// tile for each of the block's rows in turn, so consecutive rows y give contiguous blocks.  The
// (c, z, y) decomposition of each of the block's W <= 64 rows is done once, into LDS.
template <class C, bool POW2, int ELEMS = lds_elems<C>(), int MINW = kMinWaves>
__global__ void __launch_bounds__(kAxisBlock, POW2 ? MINW : 1)
    k_fft_x_fwd(const double* __restrict__ in, const double* __restrict__ ms, double scale, int Xi, int64_t Yi,
                int64_t Zi, FastDiv Yr, FastDiv Zr, int rows, FastDiv Wd, FftDesc dh, FastDiv Md, FastDiv M1d,
                const C* __restrict__ twh, const C* __restrict__ twL, C* __restrict__ out, int64_t hxs, int64_t n1,
                int64_t n2o) {
    using T = typename real_of<C>::type;
    extern __shared__ unsigned char smem_raw[];
    __shared__ int64_t s_src[64], s_cell[64], s_dst[64];
    C* buf = reinterpret_cast<C*>(smem_raw);
    const int M = dh.L;
    const int W = static_cast<int>(Wd.d);
    const int row0 = static_cast<int>(blockIdx.x) * W;
    if (static_cast<int>(threadIdx.x) < W) {
        const int r = row0 + static_cast<int>(threadIdx.x);
        int64_t src = -1, cell0 = 0, dst = -1;
        if (r < rows) {
            const uint32_t zc = fdiv(r, Yr), y = r - zc * Yr.d;
            const uint32_t c = fdiv(zc, Zr), z = zc - c * Zr.d;
            cell0 = (z * Yi + y) * Xi;
            src = c * Zi * Yi * Xi + cell0;
            dst = (c * n2o + z) * n1 * hxs + kTile * static_cast<int64_t>(y);
        }
        s_src[threadIdx.x] = src;
        s_cell[threadIdx.x] = cell0;
        s_dst[threadIdx.x] = dst;
    }
    C* tws = buf + lds_padded(W * M); // stage twiddles, staged next to the data
    for (int t = threadIdx.x; t < M; t += kAxisBlock) tws[t] = twh[t];
    __syncthreads();
    // fixed trip count (ELEMS bounds M W), so a lane's loads are all issued before the first use
    constexpr int IT = (ELEMS + kAxisBlock - 1) / kAxisBlock;
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int e = static_cast<int>(threadIdx.x) + i * kAxisBlock;
        if (e >= M * W) break;
        const int w = static_cast<int>(fdiv(e, Md)), n = e - w * M;
        const int64_t so = s_src[w];
        T re = 0, im = 0;
        if (so >= 0) {
            const double* src = in + so;
            const int t0 = 2 * n, t1 = t0 + 1;
            if (ms) {
                const double* mrow = ms + s_cell[w];
                if (t0 < Xi) re = static_cast<T>(src[t0] * scale * mrow[t0]);
                if (t1 < Xi) im = static_cast<T>(src[t1] * scale * mrow[t1]);
            } else {
                if (t0 < Xi) re = static_cast<T>(src[t0] * scale);
                if (t1 < Xi) im = static_cast<T>(src[t1] * scale);
            }
        }
        buf[lds_pad(n * W + w)] = C{re, im};
    }
    __syncthreads();
    fft_lds<C, ELEMS, POW2, kAxisBlock>(buf, dh, Wd, tws, false);
    const T half = static_cast<T>(0.5);
    const int64_t n1T = kTile * n1;
    // e = (kt W + w) 4 + kk: column k = 4 kt + kk of row w
    const int ntile = (M + kTile) / kTile;
    for (int e = threadIdx.x; e < ntile * W * kTile; e += kAxisBlock) {
        const int kk = e & (kTile - 1), r = e >> 2;
        const int kt = static_cast<int>(fdiv(r, Wd)), w = r - kt * W;
        const int k = kt * kTile + kk;
        const int64_t dst = s_dst[w];
        if (dst < 0 || k > M) continue;
        const C Zk = buf[lds_pad((k == M ? 0 : k) * W + w)];
        const C Zm = buf[lds_pad((k == 0 ? 0 : M - k) * W + w)];
        const C E{half * (Zk.x + Zm.x), half * (Zk.y - Zm.y)};
        const C O{half * (Zk.y + Zm.y), -half * (Zk.x - Zm.x)};
        out[dst + tcol(k, n1T)] = cadd(E, cmul(twL[k], O));
    }
}

// Inverse real transform along x of spectrum rows (c, z, y), c < 3, z < Zr, y < Yr, read at
// spec[(c n2s + z) n1 hxs + 4 y + tcol(k)] (column tiles, the input lanes taking the 4 columns of
// a tile for each of the block's rows in turn): C2R through the half-length packing, an inverse
// FFT of length M = n0 / 2, the crop to x < Xo and the 1 / P scale,
// out[c No + (z Yr + y) Xo + x] = T(hscale * x_real[x]).
template <class C, class T, bool POW2, int ELEMS = lds_elems<C>(), int MINW = kMinWaves>
__global__ void __launch_bounds__(kAxisBlock, POW2 ? MINW : 1)
    k_fft_x_inv(const C* __restrict__ spec, int64_t hxs, int64_t n1, int64_t n2s, FastDiv Yr, FastDiv Zr, int rows,
                FastDiv Wd, FftDesc dh, FastDiv Md, FastDiv Mhd, const C* __restrict__ twh, const C* __restrict__ twL,
                T* __restrict__ out, int Xo, int64_t No, double hscale) {
    extern __shared__ unsigned char smem_raw[];
    __shared__ int64_t s_src[64], s_dst[64];
    C* buf = reinterpret_cast<C*>(smem_raw);
    const int M = dh.L;
    const int W = static_cast<int>(Wd.d);
    const int row0 = static_cast<int>(blockIdx.x) * W;
    if (static_cast<int>(threadIdx.x) < W) {
        const int r = row0 + static_cast<int>(threadIdx.x);
        int64_t src = -1, dst = -1;
        if (r < rows) {
            const uint32_t zc = fdiv(r, Yr), y = r - zc * Yr.d;
            const uint32_t c = fdiv(zc, Zr), z = zc - c * Zr.d;
            src = (c * n2s + z) * n1 * hxs + kTile * static_cast<int64_t>(y);
            dst = c * No + (static_cast<int64_t>(z) * Yr.d + y) * Xo;
        }
        s_src[threadIdx.x] = src;
        s_dst[threadIdx.x] = dst;
    }
    C* tws = buf + lds_padded(W * M); // stage twiddles, staged next to the data
    for (int t = threadIdx.x; t < M; t += kAxisBlock) tws[t] = twh[t];
    __syncthreads();
    // e = (kt W + w) 4 + kk: column k = 4 kt + kk of row w; rounding the M columns up to whole
    // tiles adds at most 3 W <= 192 elements to the M W <= ELEMS of the block
    constexpr int IT = (ELEMS + 3 * 64 + kAxisBlock - 1) / kAxisBlock;
    const int64_t n1T = kTile * n1;
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int e = static_cast<int>(threadIdx.x) + i * kAxisBlock;
        if (e >= ((M + kTile - 1) / kTile) * W * kTile) break;
        const int kk = e & (kTile - 1), r = e >> 2;
        const int kt = static_cast<int>(fdiv(r, Wd)), w = r - kt * W;
        const int k = kt * kTile + kk;
        const int64_t so = s_src[w];
        if (k < M)
            buf[lds_pad(k * W + w)] =
                so >= 0 ? c2r_combine(spec[so + tcol(k, n1T)], spec[so + tcol(M - k, n1T)], twL[k]) : C{0, 0};
    }
    __syncthreads();
    fft_lds<C, ELEMS, POW2, kAxisBlock>(buf, dh, Wd, tws, true);
    const int Mh = static_cast<int>(Mhd.d);
    for (int e = threadIdx.x; e < Mh * W; e += kAxisBlock) {
        const int w = static_cast<int>(fdiv(e, Mhd)), n = e - w * Mh;
        const int64_t dst = s_dst[w];
        if (dst < 0) continue;
        const C v = buf[lds_pad(n * W + w)];
        const int x = 2 * n;
        out[dst + x] = static_cast<T>(static_cast<double>(v.x) * hscale);
        if (x + 1 < Xo) out[dst + x + 1] = static_cast<T>(static_cast<double>(v.y) * hscale);
    }
}

// Resident 512-lane groups per CU of the persistent x forward transform (LDS: three 41 KiB tiles)
constexpr int kXGroupsPerCU = 3;
// Resident 1024-lane FP32 groups per CU of the persistent y forward transform (LDS: two 72 KiB tiles)
constexpr int kYGroupsPerCU = 2;

// Compile-time shaped forms of k_fft_x_fwd / k_fft_x_inv for the demag field: power-of-two
// M = nx = n0 / 2, rows of exactly M reals, W = 2^(LOGE - LOGM) rows per work-group taken from
// one (c, z) plane (ny % W == 0), so a group covers rows y0 .. y0 + W - 1 of one plane and every
// lane coordinate is a bit field of its index.  Group g = (c nz + z) (ny / W) + y0 / W.
// Packed input pairs (x[2n], x[2n+1]), n < M / 2, are single 16-byte loads; n >= M / 2 is the
// zero padding.  Spectrum rows are written and read as in the generic kernels: column k of row y
// at tcol(k, n1T) + 4 y, so the W rows of a column tile form 4 W contiguous elements.
template <class C, int LOGM, int LOGE, int BS, int MINW, bool MS>
__global__ void __launch_bounds__(BS, MINW)
    k_fft_x_fwd_ct(const double* __restrict__ in, const double* __restrict__ ms, double scale, int ny, FastDiv gyd,
                   FastDiv nzd, const C* __restrict__ twh, const C* __restrict__ twL, C* __restrict__ out,
                   int64_t plane, int64_t n2o, int n1T, int64_t N, int ngroups) {
    using T = typename real_of<C>::type;
    constexpr int M = 1 << LOGM, LOGW = LOGE - LOGM, W = 1 << LOGW, E = 1 << LOGE, MH = M / 2;
    static_assert(LOGW >= 0 && E % BS == 0 && (W * MH) % BS == 0, "whole lane loops");
    extern __shared__ unsigned char smem_raw[];
    C* buf = reinterpret_cast<C*>(smem_raw);
    C* tws = buf + lds_padded(E);
    for (int t = threadIdx.x; t < M / 2; t += BS) tws[t] = twh[t]; // half table: w^(t + M/2) = -w^t
    const int tid = static_cast<int>(threadIdx.x);
    constexpr int IT = W * MH / BS;
    // Persistent groups: group g takes row groups g, g + gridDim.x, ... and loads the next one's
    // input into registers while the current one is transformed, so each CU keeps loads in flight
    // through the butterfly stages.
    double2 p[IT], q[MS ? IT : 1];
    const auto cell_of = [&](int grp, uint32_t& c, uint32_t& z, int& y0) -> int64_t {
        const uint32_t zc = fdiv(static_cast<uint32_t>(grp), gyd), yg = static_cast<uint32_t>(grp) - zc * gyd.d;
        c = fdiv(zc, nzd);
        z = zc - c * nzd.d;
        y0 = static_cast<int>(yg) << LOGW;
        return (static_cast<int64_t>(z) * ny + y0) << LOGM; // cell of (x = 0, y0, z)
    };
    // RS: the first stage (radix 8, one butterfly per lane per NB0 round) runs from registers.
    // Positions >= M / 2 are the zero padding, so butterfly (j, w) has the non-zero inputs
    // j + r M / 8, r < 4, which are this lane's 4 pair loads; lanes take j fastest so a wave
    // reads 64 consecutive pairs of one row per input.
    using S0 = CtStage<LOGM, LOGW, 0>;
    constexpr bool RS = S0::R == 8 && S0::NB % BS == 0 && S0::NB / BS * 4 == IT;
    constexpr int NB0 = RS ? S0::NB / BS : 1, M8 = M / 8;
    if constexpr (RS) __syncthreads(); // tws is complete before any butterfly runs
    const auto slot = [&](int t, int i, int& w, int& o) { // load i of lane t: row w, offset o
        if constexpr (RS) {
            const int bt = t + (i >> 2) * BS, r = i & 3;
            const int j = bt & (M8 - 1);
            w = bt >> (LOGM - 3);
            o = (w << LOGM) + 2 * (j + r * M8);
        } else {
            const int g = t + i * BS, n = g & (MH - 1);
            w = g >> (LOGM - 1);
            o = (w << LOGM) + 2 * n;
        }
    };
    const auto fetch = [&](int grp) {
        uint32_t c, z;
        int y0;
        const int64_t cell0 = cell_of(grp, c, z, y0);
        const double* src = in + static_cast<int64_t>(c) * N + cell0;
        const int ft = opaque_int(tid);
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            int w, o;
            slot(ft, i, w, o);
            p[i] = *reinterpret_cast<const double2*>(src + o);
            if constexpr (MS) q[i] = *reinterpret_cast<const double2*>(ms + cell0 + o);
        }
    };
    // FP64 with an Ms field would spill its doubled prefetch: those groups take one row group
    // each and load it at the top
    constexpr bool PF = !(MS && sizeof(T) == 8);
    int grp = static_cast<int>(blockIdx.x);
    if constexpr (PF) fetch(grp);
    for (; grp < ngroups; grp += static_cast<int>(gridDim.x)) {
        if constexpr (!PF) fetch(grp);
        const int lt = opaque_int(tid);
        const auto val = [&](int i) {
            if constexpr (MS)
                return C{static_cast<T>(p[i].x * scale * q[i].x), static_cast<T>(p[i].y * scale * q[i].y)};
            else
                return C{static_cast<T>(p[i].x * scale), static_cast<T>(p[i].y * scale)};
        };
        if constexpr (RS) {
#pragma unroll
            for (int bi = 0; bi < NB0; ++bi) {
                const int bt = lt + bi * BS, j = bt & (M8 - 1), w = bt >> (LOGM - 3);
                const int b = (j << LOGW) + w; // the stage's butterfly index of (j, w)
                C v[8];
#pragma unroll
                for (int r = 0; r < 4; ++r) v[r] = val(bi * 4 + r);
#pragma unroll
                for (int r = 4; r < 8; ++r) v[r] = C{0, 0};
                S0::template bfly<C, false, true>(v, b, tws);
#pragma unroll
                for (int r = 0; r < 8; ++r) buf[lds_pad((S0::opos(b, r) << LOGW) + w)] = v[r];
            }
        } else {
#pragma unroll
            for (int i = 0; i < IT; ++i) {
                int w, o;
                slot(lt, i, w, o);
                const int n = (o - (w << LOGM)) >> 1;
                buf[lds_pad((n << LOGW) + w)] = val(i);
                buf[lds_pad(((n + MH) << LOGW) + w)] = C{0, 0};
            }
        }
        __syncthreads();
        const int next = grp + static_cast<int>(gridDim.x);
        if (PF && next < ngroups) fetch(next);
        ct_fft<C, LOGM, LOGW, BS, false, RS ? 1 : 0, ct_nstages(LOGM), true>(buf, tws, lt);
        // X[k] = E + w_L^k O, E = (Z[k] + conj Z[M-k]) / 2, O = -i (Z[k] - conj Z[M-k]) / 2, k <= M
        uint32_t c, z;
        int y0;
        cell_of(grp, c, z, y0);
        const T half = static_cast<T>(0.5);
        C* dst = out + (static_cast<int64_t>(c) * n2o + z) * plane + kTile * y0;
        constexpr int NOUT = (M / kTile + 1) * W * kTile, OT = (NOUT + BS - 1) / BS;
#pragma unroll
        for (int i = 0; i < OT; ++i) {
            const int g = lt + i * BS;
            if (g < NOUT) {
                const int kk = g & 3, w = (g >> 2) & (W - 1), kt = g >> (LOGW + 2), k = kt * kTile + kk;
                if (k <= M) {
                    const C Zk = buf[lds_pad(((k & (M - 1)) << LOGW) + w)];
                    const C Zm = buf[lds_pad((((M - k) & (M - 1)) << LOGW) + w)];
                    const C Ev{half * (Zk.x + Zm.x), half * (Zk.y - Zm.y)};
                    const C Ov{half * (Zk.y + Zm.y), -half * (Zk.x - Zm.x)};
                    dst[kt * n1T + kTile * w + kk] = cadd(Ev, cmul(twL[k], Ov));
                }
            }
        }
        __syncthreads(); // the next group's input overwrites buf
    }
}

template <class C, class T, int LOGM, int LOGE, int BS, int MINW>
__global__ void __launch_bounds__(BS, MINW)
    k_fft_x_inv_ct(const C* __restrict__ spec, int64_t plane, int64_t n2s, int n1T, int ny, FastDiv gyd, FastDiv nzd,
                   const C* __restrict__ twh, const C* __restrict__ twL, T* __restrict__ out, int64_t N, double hscale) {
    constexpr int M = 1 << LOGM, LOGW = LOGE - LOGM, W = 1 << LOGW, E = 1 << LOGE, MH = M / 2;
    static_assert(LOGW >= 0 && E % BS == 0 && (W * MH) % BS == 0, "whole lane loops");
    extern __shared__ unsigned char smem_raw[];
    C* buf = reinterpret_cast<C*>(smem_raw);
    C* tws = buf + lds_padded(E);
    for (int t = threadIdx.x; t < M / 2; t += BS) tws[t] = twh[t]; // half table: w^(t + M/2) = -w^t
    const uint32_t grp = blockIdx.x, zc = fdiv(grp, gyd), yg = grp - zc * gyd.d;
    const uint32_t c = fdiv(zc, nzd), z = zc - c * nzd.d;
    const int y0 = static_cast<int>(yg) << LOGW;
    const C* srow = spec + (static_cast<int64_t>(c) * n2s + z) * plane + kTile * y0;
    const int tid = static_cast<int>(threadIdx.x);
    // Zp[k] = c2r_combine(X[k], X[M - k]), k < M: g = (kt W + w) 4 + kk
    constexpr int IT = E / BS;
    C a[IT], b[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int g = tid + i * BS;
        const int kk = g & 3, w = (g >> 2) & (W - 1), kt = g >> (LOGW + 2), m = M - (kt * kTile + kk);
        a[i] = srow[kt * n1T + kTile * w + kk];
        b[i] = srow[(m >> 2) * n1T + kTile * w + (m & 3)];
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int g = tid + i * BS;
        const int kk = g & 3, w = (g >> 2) & (W - 1), kt = g >> (LOGW + 2), k = kt * kTile + kk;
        buf[lds_pad((k << LOGW) + w)] = c2r_combine(a[i], b[i], twL[k]);
    }
    __syncthreads();
    constexpr int NST = ct_nstages(LOGM);
    using SL = CtStage<LOGM, LOGW, NST - 1>;
    T* orow = out + static_cast<int64_t>(c) * N + ((static_cast<int64_t>(z) * ny + y0) << LOGM);
    using T2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
    if constexpr (SL::NB % BS != 0) { // fewer last-stage butterflies than lanes: all stages in LDS
        ct_fft<C, LOGM, LOGW, BS, true, 0, NST, true>(buf, tws);
        constexpr int OT = W * MH / BS;
#pragma unroll
        for (int i = 0; i < OT; ++i) {
            const int g = tid + i * BS, w = g >> (LOGM - 1), n = g & (MH - 1);
            const C v = buf[lds_pad((n << LOGW) + w)];
            *reinterpret_cast<T2*>(orow + (w << LOGM) + 2 * n) = T2{
                static_cast<T>(static_cast<double>(v.x) * hscale), static_cast<T>(static_cast<double>(v.y) * hscale)};
        }
        return;
    }
    ct_fft<C, LOGM, LOGW, BS, true, 0, NST - 1, true>(buf, tws);
    // x[2n] + i x[2n+1] = Z[n], n < M / 2 (x < nx = M), scaled by 1 / P.  The last stage runs from
    // registers: butterfly (j, w) has outputs j + r M / R, and only r < R / 2 (positions below
    // M / 2) survive the crop, so those go straight to memory; lanes take j fastest so a wave
    // writes consecutive pairs of one row.
    static_assert(SL::NS == (M >> SL::LOGR), "last stage shape");
    constexpr int NBL = SL::NB / BS, LOGJ = LOGM - SL::LOGR;
#pragma unroll
    for (int i = 0; i < NBL; ++i) {
        const int bt = tid + i * BS, j = bt & ((1 << LOGJ) - 1), w = bt >> LOGJ;
        const int b = (j << LOGW) + w;
        C v[SL::R];
#pragma unroll
        for (int r = 0; r < SL::R; ++r) v[r] = buf[lds_pad(b + r * SL::NB)];
        SL::template bfly<C, true, true>(v, b, tws);
#pragma unroll
        for (int r = 0; r < SL::R / 2; ++r) {
            const int n = j + r * SL::NS;
            *reinterpret_cast<T2*>(orow + (w << LOGM) + 2 * n) = T2{
                static_cast<T>(static_cast<double>(v[r].x) * hscale), static_cast<T>(static_cast<double>(v[r].y) * hscale)};
        }
    }
}

// Complex transform of strided lines: line (tile, b, a) covers columns kx0 .. kx0+W-1 and holds
// data[a sa + b sb + tcol(kx, n1T) + pos * stride] (column tiles: for lines along y, b is the
// plane and stride = 4; for lines along z, b sb = 4 ky and stride is the plane).  Inputs
// pos >= n_in are taken as zero; outputs pos >= n_out are not written.
template <class C, bool POW2>
__global__ void __launch_bounds__(kAxisBlock, POW2 ? kMinWaves : 1)
    k_fft_strided(C* __restrict__ data, int64_t sa, int64_t sb, int64_t stride, int64_t hx, int n_in, int n_out,
                  FastDiv Wd, FftDesc d, const C* __restrict__ tw, int inv, int64_t n1T) {
    extern __shared__ unsigned char smem_raw[];
    C* buf = reinterpret_cast<C*>(smem_raw);
    const int L = d.L;
    const int W = static_cast<int>(Wd.d);
    const uint3 blk = swizzled_block();
    const int64_t kx0 = static_cast<int64_t>(blk.x) * W;
    C* base = data + static_cast<int64_t>(blk.z) * sa + static_cast<int64_t>(blk.y) * sb;
    const int wv = static_cast<int>(hx - kx0 < W ? hx - kx0 : W);
    for (int e = threadIdx.x; e < L * W; e += kAxisBlock) {
        const int pos = static_cast<int>(fdiv(e, Wd)), w = e - pos * W;
        C v{0, 0};
        if (w < wv && pos < n_in) v = base[pos * stride + tcol(kx0 + w, n1T)];
        buf[lds_pad(pos * W + w)] = v;
    }
    __syncthreads();
    fft_lds<C, lds_elems<C>(), POW2, kAxisBlock>(buf, d, Wd, tw, inv != 0);
    for (int e = threadIdx.x; e < n_out * W; e += kAxisBlock) {
        const int pos = static_cast<int>(fdiv(e, Wd)), w = e - pos * W;
        if (w < wv) base[pos * stride + tcol(kx0 + w, n1T)] = buf[lds_pad(pos * W + w)];
    }
}

// Both halves of the zero-padded y transform in one LDS pass.  The work-group's W columns become
// 2W lines of one length-M transform: line w holds x[n] (its outputs are X[2k]) and line W + w
// holds x[n] w_L^n (outputs X[2k+1]), so the stage sequence, and its barriers, run once per
// column pair instead of twice.  With TWL the stage twiddles (M entries) are staged in LDS behind
// the data; the inputs are kept in registers only between their loads and the LDS writes.
template <class C, bool POW2, int BS, int ELEMS, int MINW, bool TWL>
__global__ void __launch_bounds__(BS, POW2 ? MINW : 1)
    k_fft_y_fwd2(C* __restrict__ data, int64_t sa, int64_t sb, int64_t stride, int64_t hx, FastDiv Wd, FastDiv W2d,
                 FftDesc dh, const C* __restrict__ twh, const C* __restrict__ twL, int64_t n1T) {
    extern __shared__ unsigned char smem_raw[];
    C* buf = reinterpret_cast<C*>(smem_raw);
    const int M = dh.L;
    const int W = static_cast<int>(Wd.d), W2 = 2 * W;
    C* tws = buf + lds_padded(W2 * M);
    if constexpr (TWL)
        for (int t = threadIdx.x; t < M; t += BS) tws[t] = twh[t];
    const C* tw = TWL ? tws : twh;
    const uint3 blk = swizzled_block();
    const int64_t kx0 = static_cast<int64_t>(blk.x) * W;
    C* base = data + static_cast<int64_t>(blk.z) * sa + static_cast<int64_t>(blk.y) * sb;
    const int wv = static_cast<int>(hx - kx0 < W ? hx - kx0 : W);
    // all of a lane's loads are issued before the first use (IT is a compile-time bound)
    constexpr int IT = (ELEMS / 2 + BS - 1) / BS;
    C v[IT], t[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int e = threadIdx.x + i * BS;
        if (e < M * W) {
            const int n = static_cast<int>(fdiv(e, Wd)), w = e - n * W;
            v[i] = w < wv ? base[n * stride + tcol(kx0 + w, n1T)] : C{0, 0};
            t[i] = twL[n];
        }
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int e = threadIdx.x + i * BS;
        if (e < M * W) {
            const int n = static_cast<int>(fdiv(e, Wd)), w = e - n * W;
            buf[lds_pad(n * W2 + w)] = v[i];
            buf[lds_pad(n * W2 + W + w)] = cmul(v[i], t[i]);
        }
    }
    __syncthreads();
    fft_lds<C, ELEMS, POW2, BS>(buf, dh, W2d, tw, false);
    for (int e = threadIdx.x; e < M * W2; e += BS) {
        const int k = static_cast<int>(fdiv(e, W2d)), l = e - k * W2;
        const int odd = l >= W ? 1 : 0, w = l - odd * W;
        if (w < wv) base[(2 * k + odd) * stride + tcol(kx0 + w, n1T)] = buf[lds_pad(e)];
    }
}

// Inverse counterpart: lines w and W + w take X[2k] and X[2k+1] of column w; after one inverse
// LDS pass x[n] = IFFT_M(X_even)[n] + w_L^-n IFFT_M(X_odd)[n] is written for n < M.
template <class C, bool POW2, int BS, int ELEMS, int MINW, bool TWL>
__global__ void __launch_bounds__(BS, POW2 ? MINW : 1)
    k_fft_y_inv2(C* __restrict__ data, int64_t sa, int64_t sb, int64_t stride, int64_t hx, FastDiv Wd, FastDiv W2d,
                 FftDesc dh, const C* __restrict__ twh, const C* __restrict__ twL, int64_t n1T) {
    extern __shared__ unsigned char smem_raw[];
    C* buf = reinterpret_cast<C*>(smem_raw);
    const int M = dh.L;
    const int W = static_cast<int>(Wd.d), W2 = 2 * W;
    C* tws = buf + lds_padded(W2 * M);
    if constexpr (TWL)
        for (int t = threadIdx.x; t < M; t += BS) tws[t] = twh[t];
    const C* tw = TWL ? tws : twh;
    const uint3 blk = swizzled_block();
    const int64_t kx0 = static_cast<int64_t>(blk.x) * W;
    C* base = data + static_cast<int64_t>(blk.z) * sa + static_cast<int64_t>(blk.y) * sb;
    const int wv = static_cast<int>(hx - kx0 < W ? hx - kx0 : W);
    constexpr int IT = (ELEMS + BS - 1) / BS;
    C v[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int e = threadIdx.x + i * BS;
        if (e < M * W2) {
            const int k = static_cast<int>(fdiv(e, W2d)), l = e - k * W2;
            const int odd = l >= W ? 1 : 0, w = l - odd * W;
            v[i] = w < wv ? base[(2 * k + odd) * stride + tcol(kx0 + w, n1T)] : C{0, 0};
        }
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int e = threadIdx.x + i * BS;
        if (e < M * W2) buf[lds_pad(e)] = v[i];
    }
    __syncthreads();
    fft_lds<C, ELEMS, POW2, BS>(buf, dh, W2d, tw, true);
    for (int e = threadIdx.x; e < M * W; e += BS) {
        const int n = static_cast<int>(fdiv(e, Wd)), w = e - n * W;
        if (w < wv)
            base[n * stride + tcol(kx0 + w, n1T)] =
                cadd(buf[lds_pad(n * W2 + w)], cmul(conj_if(twL[n], true), buf[lds_pad(n * W2 + W + w)]));
    }
}

// Compile-time shaped forms of k_fft_y_fwd2 / k_fft_y_inv2 for power-of-two M = 2^LOGM with
// 2^LOGE complex values per LDS tile: W = 2^(LOGE - LOGM - 1) >= 4 columns, i.e. W / 4 whole
// column tiles, per work-group.  Row y of column kx0 + w sits at tile w / 4, offset 4 y + w % 4,
// and a tile spans n1T = 4 n1 = 8 M elements.  Line l = odd W + w of the LDS tile is column w
// (odd = 0: x[n], outputs X[2k]; odd = 1: x[n] w_L^n, outputs X[2k+1]).
// The last stage writes its outputs straight from registers: butterfly b = j W2 + l holds
// positions j + r Ns, and ky = 2 pos + odd sits at offset 8 pos + 4 odd + w % 4 of its tile, so
// consecutive lanes write consecutive addresses.  In the inverse the first stage likewise reads
// its inputs (rows j + r M / R of line l) straight from global memory.  (The forward reads its
// input through LDS: both lines of a column take the same x[n], so register loads per butterfly
// would fetch every element twice.)  Columns kx >= hx (the padding of the last tile group) are
// read as zero and not written.
template <class C, int LOGM, int LOGE, int BS, int MINW>
__global__ void __launch_bounds__(BS, MINW)
    k_fft_y_fwd_ct(C* __restrict__ data, int64_t sa, int64_t sb, int hx, const C* __restrict__ twh,
                   const C* __restrict__ twL, FastDiv gxd, FastDiv gyd, int ngroups) {
    constexpr int M = 1 << LOGM, LOGW2 = LOGE - LOGM, W2 = 1 << LOGW2, LOGW = LOGW2 - 1, W = W2 / 2;
    constexpr int n1T = 8 * M; // tile stride; offsets inside a work-group's tiles stay below 2^LOGE
    constexpr int NST = ct_nstages(LOGM);
    using S0 = CtStage<LOGM, LOGW2, 0>;
    using SL = CtStage<LOGM, LOGW2, NST - 1>;
    static_assert(W >= kTile && NST >= 2 && S0::NB % BS == 0 && SL::NB % BS == 0, "tile shape");
    extern __shared__ unsigned char smem_raw[];
    C* buf = reinterpret_cast<C*>(smem_raw);
    C* tws = buf + lds_padded(1 << LOGE);
    for (int t = threadIdx.x; t < M; t += BS) tws[t] = twh[t];
    const int tid = static_cast<int>(threadIdx.x);
    // group g = (plane b * gy + a) * gx + column tile; persistent groups take g, g + gridDim.x, ...
    // and load the next group's columns into registers while the current one is transformed
    const auto base_of = [&](int grp, int& kx0) -> C* {
        const uint32_t q = fdiv(static_cast<uint32_t>(grp), gxd), bx = static_cast<uint32_t>(grp) - q * gxd.d;
        const uint32_t bz = fdiv(q, gyd), by = q - bz * gyd.d;
        kx0 = static_cast<int>(bx) * W;
        return data + static_cast<int64_t>(bz) * sa + static_cast<int64_t>(by) * sb + (kx0 / kTile) * n1T;
    };
    // inputs y < M of the W columns, contiguous: g = (tile * M + y) * 4 + w % 4; both lines of
    // column w are filled (x and x w_L^y) so the even and odd outputs share one pass
    constexpr int IT = M * W / BS;
    C v[IT], t[IT];
    const auto fetch = [&](int grp) {
        int kx0;
        const C* base = base_of(grp, kx0);
        const int ft = opaque_int(tid);
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const int g = ft + i * BS;
            const int tl = g >> (LOGM + 2), y = (g >> 2) & (M - 1), w = (tl << 2) | (g & 3);
            v[i] = kx0 + w < hx ? base[tl * n1T + (g & (4 * M - 1))] : C{0, 0};
            t[i] = twL[y];
        }
    };
    int grp = static_cast<int>(blockIdx.x);
    fetch(grp);
    for (; grp < ngroups; grp += static_cast<int>(gridDim.x)) {
        const int lt = opaque_int(tid);
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const int g = lt + i * BS;
            const int tl = g >> (LOGM + 2), y = (g >> 2) & (M - 1), w = (tl << 2) | (g & 3);
            buf[lds_pad((y << LOGW2) + w)] = v[i];
            buf[lds_pad((y << LOGW2) + W + w)] = cmul(v[i], t[i]);
        }
        __syncthreads();
        const int next = grp + static_cast<int>(gridDim.x);
        if (next < ngroups) fetch(next);
        ct_fft<C, LOGM, LOGW2, BS, false, 0, NST - 1>(buf, tws, lt);
        int kx0;
        C* base = base_of(grp, kx0);
        constexpr int NBL = SL::NB / BS;
#pragma unroll
        for (int i = 0; i < NBL; ++i) {
            const int b = lt + i * BS, j = b >> LOGW2, l = b & (W2 - 1), w = l & (W - 1), odd = l >> LOGW;
            C u[SL::R];
#pragma unroll
            for (int r = 0; r < SL::R; ++r) u[r] = buf[lds_pad(b + r * SL::NB)];
            SL::template bfly<C, false>(u, b, tws);
            if (kx0 + w < hx) {
                C* dst = base + (w >> 2) * n1T + 4 * odd + (w & 3);
#pragma unroll
                for (int r = 0; r < SL::R; ++r) dst[8 * (j + r * SL::NS)] = u[r];
            }
        }
        __syncthreads(); // the next group's columns overwrite buf
    }
}

template <class C> __device__ __forceinline__ C shfl_xor_c(C v, int mask) {
    return C{__shfl_xor(v.x, mask, 64), __shfl_xor(v.y, mask, 64)};
}

// Inverse counterpart.  The last stage leaves IFFT_M(X_even)[n] and IFFT_M(X_odd)[n] of column w
// in lanes b and b ^ W of one wave (same j, lines w and W + w); the pair trades half of its
// outputs through a lane shuffle and each lane writes x[n] = even[n] + w_L^-n odd[n] for half of
// the positions.
template <class C, int LOGM, int LOGE, int BS, int MINW>
__global__ void __launch_bounds__(BS, MINW)
    k_fft_y_inv_ct(C* __restrict__ data, int64_t sa, int64_t sb, int hx, const C* __restrict__ twh,
                   const C* __restrict__ twL, FastDiv gxd, FastDiv gyd, int ngroups) {
    constexpr int M = 1 << LOGM, LOGW2 = LOGE - LOGM, W2 = 1 << LOGW2, LOGW = LOGW2 - 1, W = W2 / 2;
    constexpr int n1T = 8 * M;
    constexpr int NST = ct_nstages(LOGM);
    using S0 = CtStage<LOGM, LOGW2, 0>;
    using SL = CtStage<LOGM, LOGW2, NST - 1>;
    static_assert(W >= kTile && W2 <= 64 && NST >= 2 && S0::NB % BS == 0 && SL::NB % BS == 0 && SL::R >= 2,
                  "tile shape");
    extern __shared__ unsigned char smem_raw[];
    C* buf = reinterpret_cast<C*>(smem_raw);
    C* tws = buf + lds_padded(1 << LOGE);
    for (int t = threadIdx.x; t < M; t += BS) tws[t] = twh[t];
    const int tid = static_cast<int>(threadIdx.x);
    // groups as in k_fft_y_fwd_ct; the next group's first-stage inputs are loaded into registers
    // while the current one is transformed
    const auto base_of = [&](int grp, int& kx0) -> C* {
        const uint32_t q = fdiv(static_cast<uint32_t>(grp), gxd), bx = static_cast<uint32_t>(grp) - q * gxd.d;
        const uint32_t bz = fdiv(q, gyd), by = q - bz * gyd.d;
        kx0 = static_cast<int>(bx) * W;
        return data + static_cast<int64_t>(bz) * sa + static_cast<int64_t>(by) * sb + (kx0 / kTile) * n1T;
    };
    constexpr int NB0 = S0::NB / BS;
    C v0[NB0][S0::R];
    const auto fetch = [&](int grp) {
        int kx0;
        const C* base = base_of(grp, kx0);
        const int ft = opaque_int(tid);
#pragma unroll
        for (int i = 0; i < NB0; ++i) {
            const int b = ft + i * BS, j = b >> LOGW2, l = b & (W2 - 1), w = l & (W - 1), odd = l >> LOGW;
            const C* src = base + (w >> 2) * n1T + 4 * odd + (w & 3);
            const bool ok = kx0 + w < hx;
#pragma unroll
            for (int r = 0; r < S0::R; ++r) v0[i][r] = ok ? src[8 * (j + r * (M / S0::R))] : C{0, 0};
        }
    };
    int grp = static_cast<int>(blockIdx.x);
    fetch(grp);
    for (; grp < ngroups; grp += static_cast<int>(gridDim.x)) {
        const int lt = opaque_int(tid);
#pragma unroll
        for (int i = 0; i < NB0; ++i) {
            const int b = lt + i * BS;
            S0::template bfly<C, true>(v0[i], b, tws);
#pragma unroll
            for (int r = 0; r < S0::R; ++r) buf[lds_pad((S0::opos(b, r) << LOGW2) + (b & (W2 - 1)))] = v0[i][r];
        }
        __syncthreads();
        const int next = grp + static_cast<int>(gridDim.x);
        if (next < ngroups) fetch(next);
        ct_fft<C, LOGM, LOGW2, BS, true, 1, NST - 1>(buf, tws, lt);
        int kx0;
        C* base = base_of(grp, kx0);
        constexpr int NBL = SL::NB / BS, H = SL::R / 2;
#pragma unroll
        for (int i = 0; i < NBL; ++i) {
            const int b = lt + i * BS, j = b >> LOGW2, l = b & (W2 - 1), w = l & (W - 1), odd = l >> LOGW;
            C v[SL::R];
#pragma unroll
            for (int r = 0; r < SL::R; ++r) v[r] = buf[lds_pad(b + r * SL::NB)];
            SL::template bfly<C, true>(v, b, tws);
            const bool ok = kx0 + w < hx;
            C* dst = base + (w >> 2) * n1T + (w & 3);
#pragma unroll
            for (int q = 0; q < H; ++q) {
                // the even lane keeps positions q, the odd lane q + H
                const C recv = shfl_xor_c(odd ? v[q] : v[q + H], W);
                const int n = j + (odd ? q + H : q) * SL::NS;
                const C e = odd ? recv : v[q], o = odd ? v[q + H] : recv;
                if (ok) dst[4 * n] = cadd(e, cmul(conj_if(twL[n], true), o));
            }
        }
        __syncthreads(); // the next group's first stage overwrites buf
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

// ---- FP32 path: the real, mirror-symmetric tensor spectrum ------------------------------------
// The padded tensor [6][n2][n1][n0] holds the pair distance j for j < n/2 and j - n above, so the
// planes x = n0/2, y = n1/2, z = n2/2 (distance -n/2) have no mirror image.  No cell pair of the
// cropped convolution is that far apart, so zeroing them leaves the cropped field unchanged, and
// then each component is even or odd along every axis: xx, yy, zz even; xy odd in x and y; xz odd
// in x and z; yz odd in y and z.  The spectrum is then real, with N^(-ky) = sy N^(ky) and
// N^(-kz) = sz N^(kz), sy = -1 for xy and yz, sz = -1 for xz and yz.  Stored as
// Ts[comp][kz <= n2/2][ky <= n1/2][kx] (row stride hxs): a quarter of the k-points, real.
__global__ void k_zero_unpaired(double* __restrict__ N, int64_t n0, int64_t n1, int64_t n2, int64_t rows) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int64_t y = r % n1, z = (r / n1) % n2;
    double* row = N + r * n0;
    if (y == n1 / 2 || (n2 > 1 && z == n2 / 2)) {
        for (int64_t x = 0; x < n0; ++x) row[x] = 0.0;
    } else {
        row[n0 / 2] = 0.0;
    }
}

// R = T - S: the full FP64 spectrum T minus the symmetric quarter's value at the same k-point,
// in place on T ([6][n2][plane], the spectrum layout); zero in the padding columns kx >= hx.
__global__ void k_tensor_residual(double2* __restrict__ T, const float* __restrict__ Ts, int64_t n1, int64_t n2,
                                  int64_t hx, int64_t hxs, int64_t n1h, int64_t n2h, int64_t total) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const int64_t plane = n1 * hxs, K = n2 * plane;
    const int64_t c = e / K, rem = e - c * K, kz = rem / plane, p = rem - kz * plane;
    const int64_t kt = p / (kTile * n1), r = p - kt * (kTile * n1), ky = r / kTile, kx = kt * kTile + r % kTile;
    if (kx >= hx) {
        T[e] = double2{0.0, 0.0};
        return;
    }
    const bool my = ky >= n1h, mz = kz >= n2h;
    const int64_t kys = my ? n1 - ky : ky, kzs = mz ? n2 - kz : kz;
    const double sy = my ? -1.0 : 1.0, sz = mz ? -1.0 : 1.0;
    const double sg = c == 1 ? sy : c == 2 ? sz : c == 4 ? sy * sz : 1.0;
    const int64_t Q = n1h * hxs;
    const double sv = sg * static_cast<double>(Ts[(c * n2h + kzs) * Q + tcol(kx, kTile * n1h) + kTile * kys]);
    T[e] = double2{T[e].x - sv, T[e].y};
}

// R scaled by the component's power of two, rounded to FP16 pairs
__global__ void k_residual_f16(const double2* __restrict__ R, __half2* __restrict__ out, int64_t K, TensorScale sc,
                               int64_t total) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const double f = sc.s[e / K];
    out[e] = __floats2half2_rn(static_cast<float>(R[e].x * f), static_cast<float>(R[e].y * f));
}

// Ts holds, per component and kz <= n2/2, a plane of n1h = n1/2 + 1 rows ky in column tiles:
// column kx of row kys at tcol(kx, 4 n1h) + 4 kys.  T is the transformed tensor in the spectrum
// layout (planes of n1 rows).
__global__ void k_tensor_sym_f32(const double2* __restrict__ T, float* __restrict__ Ts, int64_t n1, int64_t n2,
                                 int64_t hx, int64_t hxs, int64_t n1h, int64_t n2h, int64_t total) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const int64_t Q = n1h * hxs, p = e % Q, r2 = e / Q, kz = r2 % n2h, c = r2 / n2h;
    const int64_t kt = p / (kTile * n1h), r = p - kt * (kTile * n1h), ky = r / kTile, kx = kt * kTile + r % kTile;
    Ts[e] = kx < hx ? static_cast<float>(T[(c * n2 + kz) * n1 * hxs + tcol(kx, kTile * n1) + kTile * ky].x) : 0.0f;
}

// H^ = N^ M^ with the symmetric tensor, in place on the spectrum [3][n2][plane]; one thread per
// plane position e (column tiles, padding columns kx >= hx skipped).
template <bool RES>
__global__ void __launch_bounds__(256)
    k_kspace_mul_sym(float2* __restrict__ s, const float* __restrict__ t, int n1, int n2, int hx, FastDiv pld,
                     FastDiv n1Td, int hxs, int n1h, int n2h, int total, TensorRes res) {
    const int e = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (e >= total) return;
    const int kz = static_cast<int>(fdiv(e, pld)), p = e - kz * static_cast<int>(pld.d);
    const int kt = static_cast<int>(fdiv(p, n1Td)), r = p - kt * static_cast<int>(n1Td.d);
    const int ky = r >> 2, kx = kt * kTile + (r & (kTile - 1));
    if (kx >= hx) return;
    const bool my = ky >= n1h, mz = kz >= n2h;
    const int kys = my ? n1 - ky : ky, kzs = mz ? n2 - kz : kz;
    const float sy = my ? -1.0f : 1.0f, sz = mz ? -1.0f : 1.0f;
    const int64_t Q = static_cast<int64_t>(n1h) * hxs, tc = static_cast<int64_t>(n2h) * Q;
    const int64_t to = kzs * Q + tcol(kx, kTile * n1h) + kTile * kys;
    const float nxx = t[to], nxy = sy * t[tc + to], nxz = sz * t[2 * tc + to], nyy = t[3 * tc + to],
                nyz = sy * sz * t[4 * tc + to], nzz = t[5 * tc + to];
    const int64_t K = static_cast<int64_t>(total);
    const float2 m0 = s[e], m1 = s[K + e], m2 = s[2 * K + e];
    if constexpr (RES) {
        // the reference spectrum: the symmetric value plus the stored residual (complex)
        float2 T[6];
        const float sv[6] = {nxx, nxy, nxz, nyy, nyz, nzz};
#pragma unroll
        for (int c = 0; c < 6; ++c) T[c] = res_entry(res, c, c * K + e, sv[c]);
        s[e] = cadd(cadd(cmul(T[0], m0), cmul(T[1], m1)), cmul(T[2], m2));
        s[K + e] = cadd(cadd(cmul(T[1], m0), cmul(T[3], m1)), cmul(T[4], m2));
        s[2 * K + e] = cadd(cadd(cmul(T[2], m0), cmul(T[4], m1)), cmul(T[5], m2));
        return;
    }
    s[e] = float2{nxx * m0.x + nxy * m1.x + nxz * m2.x, nxx * m0.y + nxy * m1.y + nxz * m2.y};
    s[K + e] = float2{nxy * m0.x + nyy * m1.x + nyz * m2.x, nxy * m0.y + nyy * m1.y + nyz * m2.y};
    s[2 * K + e] = float2{nxz * m0.x + nyz * m1.x + nzz * m2.x, nxz * m0.y + nyz * m1.y + nzz * m2.y};
}

// Fused z stage with the symmetric tensor.  One thread per (kys <= n1/2, kx < hx), taken in
// (kx mod 4, kys, kx / 4) order so neighbouring lanes touch neighbouring addresses of the column
// tiles: it loads the tensor column once (6 (NZ + 1) reals) and serves both spectrum columns
// ky = kys and ky = n1 - kys.
template <int NZ, bool RES>
__global__ void __launch_bounds__(256)
    k_zfused_sym(float2* __restrict__ s, const float* __restrict__ t, int n1, int hx, int hxs, FastDiv n1hd,
                 ZTw<float2> tw, TensorRes res) {
    constexpr int NZ2 = 2 * NZ, NH = NZ + 1;
    const int e = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    const int r = e >> 2, kt = static_cast<int>(fdiv(r, n1hd)), kys = r - kt * static_cast<int>(n1hd.d);
    const int kx = kt * kTile + (e & (kTile - 1));
    if (kx >= hx) return;
    const int n1h = static_cast<int>(n1hd.d);
    const int64_t plane = static_cast<int64_t>(n1) * hxs;
    const int64_t Q = static_cast<int64_t>(n1h) * hxs, tc = static_cast<int64_t>(NH) * Q;
    const int64_t tcx = tcol(kx, kTile * n1h) + kTile * kys;
    float tv[6][NH];
#pragma unroll
    for (int c = 0; c < 6; ++c)
#pragma unroll
        for (int k = 0; k < NH; ++k) tv[c][k] = t[c * tc + k * Q + tcx];
    const int64_t scx = tcol(kx, kTile * static_cast<int64_t>(n1));
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1 && (kys == 0 || kys == n1h - 1)) break;
        const int ky = pass == 0 ? kys : n1 - kys;
        const float sy = pass == 0 ? 1.0f : -1.0f;
        const int64_t col = scx + kTile * ky;
        float2 m[3][NZ], h[3][NZ];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int z = 0; z < NZ; ++z) {
                m[c][z] = s[(c * NZ + z) * plane + col];
                h[c][z] = float2{0.0f, 0.0f};
            }
#pragma unroll
        for (int k = 0; k < NZ2; ++k) {
            float2 X[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                float2 acc = m[c][0];
#pragma unroll
                for (int z = 1; z < NZ; ++z) acc = cadd(acc, cmul(m[c][z], tw.w[(k * z) % NZ2]));
                X[c] = acc;
            }
            const int kzs = k <= NZ ? k : NZ2 - k;
            const float sz = k <= NZ ? 1.0f : -1.0f;
            const float nxx = tv[0][kzs], nxy = sy * tv[1][kzs], nxz = sz * tv[2][kzs], nyy = tv[3][kzs],
                        nyz = sy * sz * tv[4][kzs], nzz = tv[5][kzs];
            float2 H0, H1, H2;
            if constexpr (RES) {
                // the reference spectrum at (kx, ky, k): symmetric value plus stored residual
                float2 T[6];
                const float sv[6] = {nxx, nxy, nxz, nyy, nyz, nzz};
#pragma unroll
                for (int c = 0; c < 6; ++c) T[c] = res_entry(res, c, (c * NZ2 + k) * plane + col, sv[c]);
                H0 = cadd(cadd(cmul(T[0], X[0]), cmul(T[1], X[1])), cmul(T[2], X[2]));
                H1 = cadd(cadd(cmul(T[1], X[0]), cmul(T[3], X[1])), cmul(T[4], X[2]));
                H2 = cadd(cadd(cmul(T[2], X[0]), cmul(T[4], X[1])), cmul(T[5], X[2]));
            } else {
                H0 = float2{nxx * X[0].x + nxy * X[1].x + nxz * X[2].x, nxx * X[0].y + nxy * X[1].y + nxz * X[2].y};
                H1 = float2{nxy * X[0].x + nyy * X[1].x + nyz * X[2].x, nxy * X[0].y + nyy * X[1].y + nyz * X[2].y};
                H2 = float2{nxz * X[0].x + nyz * X[1].x + nzz * X[2].x, nxz * X[0].y + nyz * X[1].y + nzz * X[2].y};
            }
#pragma unroll
            for (int z = 0; z < NZ; ++z) {
                const float2 w = conj_if(tw.w[(k * z) % NZ2], true);
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
}

// In-place radix-2 DIT transform of length N (power of two, N <= 8) in registers, forward
// or inverse (conjugate twiddles, no 1/N), unscaled: A[j] = sum_n a[n] w_N^(+-n j).  tw holds
// the powers of the order-TW root of unity, tw.w[i] = exp(-2 pi i i / TW), TW a multiple of N.
template <int N, int TW> __device__ __forceinline__ void dft_pow2(float2 (&a)[N], const ZTw<float2>& tw, bool inv) {
    static_assert(N >= 1 && N <= 8 && (N & (N - 1)) == 0 && TW % N == 0, "length 1, 2, 4 or 8 dividing TW");
    constexpr int LOG = N == 8 ? 3 : N == 4 ? 2 : N == 2 ? 1 : 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        int r = 0;
#pragma unroll
        for (int b = 0; b < LOG; ++b) r |= ((i >> b) & 1) << (LOG - 1 - b);
        if (r > i) {
            const float2 t = a[i];
            a[i] = a[r];
            a[r] = t;
        }
    }
#pragma unroll
    for (int len = 2; len <= N; len <<= 1) {
#pragma unroll
        for (int i = 0; i < N; i += len) {
#pragma unroll
            for (int j = 0; j < len / 2; ++j) {
                const float2 u = a[i + j];
                const float2 v = j == 0 ? a[i + j + len / 2] : cmul(a[i + j + len / 2], conj_if(tw.w[j * (TW / len)], inv));
                a[i + j] = cadd(u, v);
                a[i + j + len / 2] = float2{u.x - v.x, u.y - v.y};
            }
        }
    }
}

// Fused z stage for power-of-two NZ as k_zfused_sym, with the length-2 NZ transforms as FFTs done
// in place: the zero-padded forward is X[2j] = DFT_NZ(x)[j], X[2j+1] = DFT_NZ(x w_2NZ^n)[j]; the
// tensor product replaces X[k] by H[k]; the cropped inverse is h[z] = IDFT_NZ(H_even)[z] +
// w_2NZ^-z IDFT_NZ(H_odd)[z], z < NZ.  Only the 3 x 2 NZ spectrum values of a column are live
// (the direct DFT holds the NZ inputs and the NZ output sums at once), and each tensor plane kz is
// read when it is used, so NZ = 8 fits 165 VGPRs (3 waves per SIMD) instead of 256 plus AGPRs.
template <int NZ, bool RES>
__global__ void __launch_bounds__(256, RES ? 2 : 3)
    k_zfused_sym_fft(float2* __restrict__ s, const float* __restrict__ t, int n1, int hx, int hxs, FastDiv n1hd,
                     ZTw<float2> tw, TensorRes res) {
    constexpr int NZ2 = 2 * NZ, NH = NZ + 1;
    static_assert(NZ2 <= 16, "twiddle table of 16");
    const int e = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    const int r = e >> 2, kt = static_cast<int>(fdiv(r, n1hd)), kys = r - kt * static_cast<int>(n1hd.d);
    const int kx = kt * kTile + (e & (kTile - 1));
    if (kx >= hx) return;
    const int n1h = static_cast<int>(n1hd.d);
    const int64_t plane = static_cast<int64_t>(n1) * hxs;
    const int64_t Q = static_cast<int64_t>(n1h) * hxs, tc = static_cast<int64_t>(NH) * Q;
    // every access is (uniform plane base, 32-bit lane byte offset): one offset VGPR per array
    // instead of a 64-bit address per access (a spectrum plane and a tensor plane are below 4 GiB)
    const uint32_t tb = static_cast<uint32_t>(tcol(kx, kTile * n1h) + kTile * kys) * 4u;
    const int64_t scx = tcol(kx, kTile * static_cast<int64_t>(n1));
#pragma unroll 1
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1 && (kys == 0 || kys == n1h - 1)) break;
        const int ky = pass == 0 ? kys : n1 - kys;
        const float sy = pass == 0 ? 1.0f : -1.0f;
        const int64_t col = scx + kTile * ky;
        const uint32_t sb = static_cast<uint32_t>(opaque_int(static_cast<int>(col))) * 8u;
        const auto sp = [&](int pl) { return reinterpret_cast<float2*>(reinterpret_cast<char*>(s + pl * plane) + sb); };
        // X[c][0..NZ) even k-points, X[c][NZ..2NZ) odd ones
        float2 ev[3][NZ], od[3][NZ];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int z = 0; z < NZ; ++z) ev[c][z] = *sp(c * NZ + z);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
#pragma unroll
            for (int z = 0; z < NZ; ++z) od[c][z] = z == 0 ? ev[c][0] : cmul(ev[c][z], tw.w[z]);
            dft_pow2<NZ, NZ2>(ev[c], tw, false);
            dft_pow2<NZ, NZ2>(od[c], tw, false);
        }
        // H = N X at k-points kz and 2 NZ - kz (the tensor's parity in kz), the tensor plane kz read
        // once per column through an offset the compiler cannot hoist out of the column loop
        const uint32_t tbp = static_cast<uint32_t>(opaque_int(static_cast<int>(tb)));
        const auto product = [&](float2& X0, float2& X1, float2& X2, int k, const float (&tv)[6]) {
            const float2 x0 = X0, x1 = X1, x2 = X2;
            const float sz = k <= NZ ? 1.0f : -1.0f;
            const float nxx = tv[0], nxy = sy * tv[1], nxz = sz * tv[2], nyy = tv[3], nyz = sy * sz * tv[4], nzz = tv[5];
            if constexpr (RES) {
                float2 T[6];
                const float sv[6] = {nxx, nxy, nxz, nyy, nyz, nzz};
#pragma unroll
                for (int c = 0; c < 6; ++c) T[c] = res_entry(res, c, (c * NZ2 + k) * plane + col, sv[c]);
                X0 = cadd(cadd(cmul(T[0], x0), cmul(T[1], x1)), cmul(T[2], x2));
                X1 = cadd(cadd(cmul(T[1], x0), cmul(T[3], x1)), cmul(T[4], x2));
                X2 = cadd(cadd(cmul(T[2], x0), cmul(T[4], x1)), cmul(T[5], x2));
            } else {
                X0 = float2{nxx * x0.x + nxy * x1.x + nxz * x2.x, nxx * x0.y + nxy * x1.y + nxz * x2.y};
                X1 = float2{nxy * x0.x + nyy * x1.x + nyz * x2.x, nxy * x0.y + nyy * x1.y + nyz * x2.y};
                X2 = float2{nxz * x0.x + nyz * x1.x + nzz * x2.x, nxz * x0.y + nyz * x1.y + nzz * x2.y};
            }
        };
        const auto at_k = [&](int k, const float (&tv)[6]) {
            if (k & 1)
                product(od[0][k >> 1], od[1][k >> 1], od[2][k >> 1], k, tv);
            else
                product(ev[0][k >> 1], ev[1][k >> 1], ev[2][k >> 1], k, tv);
        };
#pragma unroll
        for (int kz = 0; kz <= NZ; ++kz) {
            float tv[6];
#pragma unroll
            for (int c = 0; c < 6; ++c)
                tv[c] = *reinterpret_cast<const float*>(reinterpret_cast<const char*>(t + c * tc + kz * Q) + tbp);
            at_k(kz, tv);
            if (kz != 0 && kz != NZ) at_k(NZ2 - kz, tv);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            dft_pow2<NZ, NZ2>(ev[c], tw, true);
            dft_pow2<NZ, NZ2>(od[c], tw, true);
#pragma unroll
            for (int z = 0; z < NZ; ++z)
                *sp(c * NZ + z) = z == 0 ? cadd(ev[c][0], od[c][0])
                                                       : cadd(ev[c][z], cmul(od[c][z], conj_if(tw.w[z], true)));
        }
    }
}

int lines_per_block(int L, int elems) {
    int W = elems / L;
    if (W < 1) W = 1;
    if (W > 64) W = 64;
    return W;
}

const char* nm(const char* f32, const char* f64, size_t csz) { return csz == 8 ? f32 : f64; }

// Raises a kernel's dynamic LDS limit above the 64 KiB default once per kernel (gfx950: 160 KiB).
void allow_lds(const void* fn, size_t bytes) {
    static std::mutex mu;
    static std::map<const void*, size_t> granted;
    if (bytes <= 65536) return;
    require(bytes <= 160 * 1024, "kernel LDS request exceeds 160 KiB");
    std::lock_guard<std::mutex> lk(mu);
    size_t& g = granted[fn];
    if (g >= bytes) return;
    MAFX_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
    g = bytes;
}

} // namespace

template <class C>
void fft_x_forward(mafx_ctx* ctx, const double* in, const double* ms, double scale, int64_t Xi, int64_t Yi,
                   int64_t Zi, int ncomp, int64_t Yr, int64_t Zr, C* out, int64_t n2o) {
    const FftPlan& ph = ctx->pxh;
    const int M = ph.L;
    require(M <= lds_elems<C>(), "x transform length exceeds the LDS budget");
    // compile-time shapes for the demag input (rows of nx = M reals, 256 <= M <= 2048): 32 KiB tiles
    // of whole row groups of one plane, 512 lanes under a 6-wave register cap
    constexpr int LOGE = sizeof(C) == sizeof(float2) ? 12 : 11;
    if (ncomp == 3 && Xi == M && Yi == Yr && Zi == Zr && ctx->n0 == 2 * M && M >= 256 && M <= 2048 &&
        (M & (M - 1)) == 0 && Yr % std::max(1, (1 << LOGE) / M) == 0) {
        using KC = void (*)(const double*, const double*, double, int, FastDiv, FastDiv, const C*, const C*, C*, int64_t,
                            int64_t, int, int64_t, int);
        const auto pick = [M](auto ms_tag) -> KC {
            constexpr bool HM = decltype(ms_tag)::value;
            return M == 256 ? k_fft_x_fwd_ct<C, 8, LOGE, 512, 6, HM>
                   : M == 512 ? k_fft_x_fwd_ct<C, 9, LOGE, 512, 6, HM>
                   : M == 1024 ? k_fft_x_fwd_ct<C, 10, LOGE, 512, 6, HM>
                               : k_fft_x_fwd_ct<C, 11, LOGE, 512, 6, HM>;
        };
        const KC fc = ms ? pick(std::true_type{}) : pick(std::false_type{});
        const int W = std::max(1, (1 << LOGE) / M);
        const int64_t groups = 3 * Zr * (Yr / W);
        require(groups < (int64_t{1} << 31) && kTile * ctx->n1 * (M / kTile + 1) < (int64_t{1} << 31),
                "x transform: too many rows");
        const double bytes = static_cast<double>(3 * Zr * Yr) * (Xi * 8.0 + ctx->hx * sizeof(C));
        ProfScope ps(ctx, nm("fft_x_fwd_f32", "fft_x_fwd_f64", sizeof(C)), bytes);
        const size_t lds = (lds_padded(1 << LOGE) + M / 2) * sizeof(C);
        allow_lds(reinterpret_cast<const void*>(fc), lds);
        const bool pf = !(ms && sizeof(C) == sizeof(double2)); // the kernel's PF
        const unsigned grid =
            static_cast<unsigned>(pf ? std::min<int64_t>(groups, int64_t{kXGroupsPerCU} * ctx->num_cus) : groups);
        hipLaunchKernelGGL(fc, dim3(grid), dim3(512), lds, ctx->stream, in, ms, scale, static_cast<int>(Yr),
                           make_fastdiv(static_cast<uint32_t>(Yr / W)), make_fastdiv(static_cast<uint32_t>(Zr)),
                           tw_of<C>(ph), tw_of<C>(ctx->px), out, ctx->n1 * ctx->hxs, n2o,
                           static_cast<int>(kTile * ctx->n1), Xi * Yi * Zi, static_cast<int>(groups));
        MAFX_HIP(hipGetLastError());
        return;
    }
    // Power-of-two rows that fit use the 32 KiB tile under a 6-wave register cap (three resident
    // work-groups per CU), measured faster than the 64 KiB tile at 1024 x 1024 x 4 in both precisions.
    constexpr int E64 = lds_elems<C>();
    const bool small = is_pow2_plan(desc_of(ph)) && M <= E64 / 2;
    const int W = lines_per_block(M, small ? E64 / 2 : E64);
    const int64_t rows = static_cast<int64_t>(ncomp) * Zr * Yr;
    require(rows * (M + 1) < (int64_t{1} << 31) && Xi < (int64_t{1} << 30), "x transform: too many rows");
    const unsigned grid = static_cast<unsigned>((rows + W - 1) / W);
    const double bytes = static_cast<double>(rows) * (Xi * 8.0 + ctx->hx * sizeof(C));
    const FftDesc d = desc_of(ph);
    ProfScope ps(ctx, nm("fft_x_fwd_f32", "fft_x_fwd_f64", sizeof(C)), bytes);
    const C* twL = tw_of<C>(ctx->px);
    auto fn = is_pow2_plan(d) ? k_fft_x_fwd<C, true> : k_fft_x_fwd<C, false>;
    if (small) fn = k_fft_x_fwd<C, true, E64 / 2, 6>;
    const size_t lds = (lds_padded(W * M) + M) * sizeof(C);
    allow_lds(reinterpret_cast<const void*>(fn), lds);
    hipLaunchKernelGGL(fn, dim3(grid), dim3(kAxisBlock), lds, ctx->stream, in, ms, scale, static_cast<int>(Xi), Yi, Zi,
                       make_fastdiv(static_cast<uint32_t>(Yr)), make_fastdiv(static_cast<uint32_t>(Zr)),
                       static_cast<int>(rows), make_fastdiv(W), d, make_fastdiv(M), make_fastdiv(M + 1), tw_of<C>(ph),
                       twL, out, ctx->hxs, ctx->n1, n2o);
    MAFX_HIP(hipGetLastError());
}

template <class C, class T>
void fft_x_inverse(mafx_ctx* ctx, const C* spec, int64_t n2s, T* out, double hscale) {
    const FftPlan& ph = ctx->pxh;
    const int M = ph.L;
    const mafx_mesh& g = ctx->mesh;
    require(M <= lds_elems<C>(), "x transform length exceeds the LDS budget");
    // compile-time shapes, as in fft_x_forward
    constexpr int LOGE = sizeof(C) == sizeof(float2) ? 12 : 11;
    if (g.nx == M && ctx->n0 == 2 * M && M >= 256 && M <= 2048 && (M & (M - 1)) == 0 &&
        g.ny % std::max(1, (1 << LOGE) / M) == 0) {
        using KC = void (*)(const C*, int64_t, int64_t, int, int, FastDiv, FastDiv, const C*, const C*, T*, int64_t,
                            double);
        // FP32 groups (38 KiB of LDS with the half twiddle table) fit four per CU under an 8-wave cap
        constexpr int XW = sizeof(C) == sizeof(float2) ? 8 : 6;
        KC fc = M == 256 ? k_fft_x_inv_ct<C, T, 8, LOGE, 512, XW>
                : M == 512 ? k_fft_x_inv_ct<C, T, 9, LOGE, 512, XW>
                : M == 1024 ? k_fft_x_inv_ct<C, T, 10, LOGE, 512, XW>
                            : k_fft_x_inv_ct<C, T, 11, LOGE, 512, XW>;
        const int W = std::max(1, (1 << LOGE) / M);
        const int64_t groups = 3 * g.nz * (g.ny / W);
        require(groups < (int64_t{1} << 31) && kTile * ctx->n1 * (M / kTile + 1) < (int64_t{1} << 31),
                "x transform: too many rows");
        const double bytes = static_cast<double>(3 * g.nz * g.ny) * (ctx->hx * sizeof(C) + g.nx * sizeof(T));
        ProfScope ps(ctx, nm("fft_x_inv_f32", "fft_x_inv_f64", sizeof(C)), bytes);
        const size_t lds = (lds_padded(1 << LOGE) + M / 2) * sizeof(C);
        allow_lds(reinterpret_cast<const void*>(fc), lds);
        hipLaunchKernelGGL(fc, dim3(static_cast<unsigned>(groups)), dim3(512), lds, ctx->stream, spec, ctx->n1 * ctx->hxs,
                           n2s, static_cast<int>(kTile * ctx->n1), static_cast<int>(g.ny),
                           make_fastdiv(static_cast<uint32_t>(g.ny / W)), make_fastdiv(static_cast<uint32_t>(g.nz)),
                           tw_of<C>(ph), tw_of<C>(ctx->px), out, ctx->N, hscale);
        MAFX_HIP(hipGetLastError());
        return;
    }
    constexpr int E64 = lds_elems<C>();
    const bool small = is_pow2_plan(desc_of(ph)) && M <= E64 / 2; // as in fft_x_forward
    const int W = lines_per_block(M, small ? E64 / 2 : E64);
    const int64_t rows = 3 * g.nz * g.ny;
    require(rows * (M + 1) < (int64_t{1} << 31), "x transform: too many rows");
    const unsigned grid = static_cast<unsigned>((rows + W - 1) / W);
    const double bytes = static_cast<double>(rows) * (ctx->hx * sizeof(C) + g.nx * sizeof(T));
    const FftDesc d = desc_of(ph);
    ProfScope ps(ctx, nm("fft_x_inv_f32", "fft_x_inv_f64", sizeof(C)), bytes);
    auto fn = is_pow2_plan(d) ? k_fft_x_inv<C, T, true> : k_fft_x_inv<C, T, false>;
    if (small) fn = k_fft_x_inv<C, T, true, E64 / 2, 6>;
    const size_t lds = (lds_padded(W * M) + M) * sizeof(C);
    allow_lds(reinterpret_cast<const void*>(fn), lds);
    const int Mh = static_cast<int>((g.nx + 1) / 2);
    hipLaunchKernelGGL(fn, dim3(grid), dim3(kAxisBlock), lds, ctx->stream, spec, ctx->hxs, ctx->n1, n2s,
                       make_fastdiv(static_cast<uint32_t>(g.ny)), make_fastdiv(static_cast<uint32_t>(g.nz)),
                       static_cast<int>(rows), make_fastdiv(W), d, make_fastdiv(M), make_fastdiv(Mh), tw_of<C>(ph),
                       tw_of<C>(ctx->px), out, static_cast<int>(g.nx), ctx->N, hscale);
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
    auto fn = is_pow2_plan(d) ? k_fft_strided<C, true> : k_fft_strided<C, false>;
    const size_t lds = lds_padded(W * L) * sizeof(C);
    allow_lds(reinterpret_cast<const void*>(fn), lds);
    hipLaunchKernelGGL(fn, dim3(gx, nb, na), dim3(kAxisBlock), lds, ctx->stream, data, sa, sb, stride, ctx->hx, n_in,
                       n_out, make_fastdiv(W), d, tw_of<C>(p), inv ? 1 : 0, kTile * ctx->n1);
    MAFX_HIP(hipGetLastError());
}

template <class C>
void fft_y_split(mafx_ctx* ctx, C* data, int64_t sa, int na, int64_t sb, int nb, int64_t stride, bool inv) {
    const FftPlan& ph = ctx->pyh;
    const int M = ph.L;
    require(nb <= 65535 && na <= 65535, "y transform grid too large");
    const FftDesc d = desc_of(ph);
    constexpr int E64 = lds_elems<C>();
    constexpr bool kF32 = sizeof(C) == sizeof(float2);
    constexpr size_t kLdsMax = 160 * 1024;
    // compile-time shapes: power-of-two M of 256..1024 on 8192-element tiles of whole column tiles
    // (1024 lanes; FP32 under an 8-wave register cap, two groups per CU; FP64 one group per CU)
    if (stride == kTile && ctx->n1 == 2 * M && (M == 256 || M == 512 || M == 1024)) {
        constexpr int LOGE = 13, MW = kF32 ? 8 : 4;
        using KC = void (*)(C*, int64_t, int64_t, int, const C*, const C*, FastDiv, FastDiv, int);
        KC fc = nullptr;
        if (M == 1024) fc = inv ? k_fft_y_inv_ct<C, 10, LOGE, 1024, MW> : k_fft_y_fwd_ct<C, 10, LOGE, 1024, MW>;
        if (M == 512) fc = inv ? k_fft_y_inv_ct<C, 9, LOGE, 1024, MW> : k_fft_y_fwd_ct<C, 9, LOGE, 1024, MW>;
        if (M == 256) fc = inv ? k_fft_y_inv_ct<C, 8, LOGE, 1024, MW> : k_fft_y_fwd_ct<C, 8, LOGE, 1024, MW>;
        const int W = (1 << LOGE) / (2 * M);
        const unsigned gx = static_cast<unsigned>((ctx->hx + W - 1) / W);
        require(static_cast<int64_t>(gx) * W <= ctx->hxs, "y transform: column tiles exceed the padded rows");
        const double bytes = static_cast<double>(na) * nb * ctx->hx * (3.0 * M) * sizeof(C);
        ProfScope ps(ctx,
                     inv ? nm("fft_y_inv_f32", "fft_y_inv_f64", sizeof(C)) : nm("fft_y_fwd_f32", "fft_y_fwd_f64", sizeof(C)),
                     bytes);
        const size_t lds = (lds_padded(1 << LOGE) + M) * sizeof(C);
        allow_lds(reinterpret_cast<const void*>(fc), lds);
        // The FP32 forward runs two persistent groups per CU (72 KiB tiles); measured at 1024 x
        // 1024 x 4, the inverse and the FP64 forms are faster with one column tile per group
        const int64_t groups = static_cast<int64_t>(gx) * nb * na;
        require(groups < (int64_t{1} << 31), "y transform: too many column tiles");
        const unsigned grid = static_cast<unsigned>(
            kF32 && !inv ? std::min<int64_t>(groups, int64_t{kYGroupsPerCU} * ctx->num_cus) : groups);
        hipLaunchKernelGGL(fc, dim3(grid), dim3(1024), lds, ctx->stream, data, sa, sb, static_cast<int>(ctx->hx),
                           tw_of<C>(ph), tw_of<C>(ctx->py), make_fastdiv(gx), make_fastdiv(static_cast<uint32_t>(nb)),
                           static_cast<int>(groups));
        MAFX_HIP(hipGetLastError());
        return;
    }
    // Tile shapes, measured at 1024 x 1024 x 4 (one column pair = 2 lines): FP32 power-of-two
    // lines of up to E64/2 run 1024-lane work-groups on a 64 KiB tile under an 8-wave register cap
    // (two groups per CU); FP64 ones 1024-lane groups on a 128 KiB tile; both stage the twiddles in
    // LDS.  Longer or non-power-of-two lines use 512-lane groups, a 128 KiB tile and global twiddles.
    using KT = void (*)(C*, int64_t, int64_t, int64_t, int64_t, FastDiv, FastDiv, FftDesc, const C*, const C*, int64_t);
    KT fn = nullptr;
    int bs = 512, elems = 2 * E64;
    bool twl = false;
    if (is_pow2_plan(d) && M <= E64 / 2) {
        constexpr int EL = kF32 ? E64 : 2 * E64, MW = kF32 ? 8 : 4;
        bs = 1024;
        twl = true;
        elems = EL;
        fn = inv ? k_fft_y_inv2<C, true, 1024, EL, MW, true> : k_fft_y_fwd2<C, true, 1024, EL, MW, true>;
    } else if (is_pow2_plan(d)) {
        fn = inv ? k_fft_y_inv2<C, true, 512, 2 * E64, 1, false> : k_fft_y_fwd2<C, true, 512, 2 * E64, 1, false>;
    } else {
        fn = inv ? k_fft_y_inv2<C, false, 512, 2 * E64, 1, false> : k_fft_y_fwd2<C, false, 512, 2 * E64, 1, false>;
    }
    // lines per work-group: even, within the tile and, with the twiddles, within the LDS
    auto lds_of = [&](int w2) { return (lds_padded(w2 * M) + (twl ? M : 0)) * sizeof(C); };
    int W2 = std::min(elems / M, 64) & ~1;
    while (W2 >= 2 && lds_of(W2) > kLdsMax) W2 -= 2;
    require(W2 >= 2, "y transform length exceeds the LDS budget");
    const int W = W2 / 2;
    const unsigned gx = static_cast<unsigned>((ctx->hx + W - 1) / W);
    const double bytes = static_cast<double>(na) * nb * ctx->hx * (3.0 * M) * sizeof(C);
    ProfScope ps(ctx, inv ? nm("fft_y_inv_f32", "fft_y_inv_f64", sizeof(C)) : nm("fft_y_fwd_f32", "fft_y_fwd_f64", sizeof(C)),
                 bytes);
    const size_t lds = lds_of(W2);
    allow_lds(reinterpret_cast<const void*>(fn), lds);
    hipLaunchKernelGGL(fn, dim3(gx, nb, na), dim3(bs), lds, ctx->stream, data, sa, sb, stride, ctx->hx, make_fastdiv(W),
                       make_fastdiv(W2), d, tw_of<C>(ph), tw_of<C>(ctx->py), kTile * ctx->n1);
    MAFX_HIP(hipGetLastError());
}

template <class C> void kspace_mul(mafx_ctx* ctx, C* spec, const C* tensor, int64_t K) {
    ProfScope ps(ctx, nm("kspace_mul_f32", "kspace_mul_f64", sizeof(C)), static_cast<double>(K) * 12 * sizeof(C));
    hipLaunchKernelGGL(k_kspace_mul<C>, dim3(grid_for(K, 256)), dim3(256), 0, ctx->stream, spec, tensor, K);
    MAFX_HIP(hipGetLastError());
}

template <class C> void zfused(mafx_ctx* ctx, C* spec, const C* tensor, int nz) {
    const int64_t plane = ctx->n1 * ctx->hxs;
    ZTw<C> tw{};
    const int n2 = 2 * nz;
    const long double two_pi = 6.283185307179586476925286766559L;
    for (int t = 0; t < n2; ++t) {
        const long double a = two_pi * t / n2;
        using T = typename real_of<C>::type;
        tw.w[t] = C{static_cast<T>(std::cos(a)), static_cast<T>(-std::sin(a))};
    }
    const double bytes = static_cast<double>(ctx->n1 * ctx->hx) * (2.0 * 3 * nz + 6.0 * n2) * sizeof(C);
    ProfScope ps(ctx, nm("zfused_mul_f32", "zfused_mul_f64", sizeof(C)), bytes);
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

void zero_unpaired_planes(mafx_ctx* ctx, double* N, int ncomp) {
    const int64_t rows = static_cast<int64_t>(ncomp) * ctx->n2 * ctx->n1;
    hipLaunchKernelGGL(k_zero_unpaired, dim3(grid_for(rows, 256)), dim3(256), 0, ctx->stream, N, ctx->n0, ctx->n1,
                       ctx->n2, rows);
    MAFX_HIP(hipGetLastError());
}

void tensor_sym_f32(mafx_ctx* ctx, const double2* T, float* Ts) {
    const int64_t total = 6 * ctx->n2h * ctx->n1h * ctx->hxs;
    hipLaunchKernelGGL(k_tensor_sym_f32, dim3(grid_for(total, 256)), dim3(256), 0, ctx->stream, T, Ts, ctx->n1,
                       ctx->n2, ctx->hx, ctx->hxs, ctx->n1h, ctx->n2h, total);
    MAFX_HIP(hipGetLastError());
}

void tensor_residual_f16(mafx_ctx* ctx, double2* T, const float* Ts, __half2* out, float inv_scale[6]) {
    const int64_t K = ctx->n2 * ctx->n1 * ctx->hxs, total = 6 * K;
    hipLaunchKernelGGL(k_tensor_residual, dim3(grid_for(total, 256)), dim3(256), 0, ctx->stream, T, Ts, ctx->n1,
                       ctx->n2, ctx->hx, ctx->hxs, ctx->n1h, ctx->n2h, total);
    MAFX_HIP(hipGetLastError());
    // per component: the power of two that takes max |R| to within [8192, 16384), far from FP16's
    // overflow (65504), so the largest residuals keep FP16's full 11-bit precision
    TensorScale sc{};
    for (int c = 0; c < 6; ++c) {
        const double mx = max_abs(ctx, reinterpret_cast<const double*>(T + c * K), 2 * K);
        int ex = 0;
        if (mx > 0) {
            std::frexp(mx, &ex); // mx = f 2^ex, f in [0.5, 1)
        }
        sc.s[c] = mx > 0 ? std::ldexp(1.0, 14 - ex) : 1.0;
        inv_scale[c] = static_cast<float>(1.0 / sc.s[c]);
    }
    hipLaunchKernelGGL(k_residual_f16, dim3(grid_for(total, 256)), dim3(256), 0, ctx->stream, T, out, K, sc, total);
    MAFX_HIP(hipGetLastError());
}

void kspace_mul_sym(mafx_ctx* ctx, float2* spec, const float* ts, int64_t n2, const TensorRes* res) {
    const int64_t plane = ctx->n1 * ctx->hxs, total = n2 * plane;
    require(total < (int64_t{1} << 31), "spectrum too large for the symmetric tensor product");
    const double bytes = static_cast<double>(n2 * ctx->n1 * ctx->hx) * (6.0 * sizeof(float2)) +
                         6.0 * sizeof(float) * static_cast<double>(ctx->n2h * ctx->n1h * ctx->hx) +
                         (res ? 6.0 * sizeof(__half2) * static_cast<double>(n2 * ctx->n1 * ctx->hx) : 0.0);
    ProfScope ps(ctx, res ? "kspace_mul_symres_f32" : "kspace_mul_sym_f32", bytes);
    const TensorRes rv = res ? *res : TensorRes{};
    auto fn = res ? k_kspace_mul_sym<true> : k_kspace_mul_sym<false>;
    hipLaunchKernelGGL(fn, dim3(grid_for(total, 256)), dim3(256), 0, ctx->stream, spec, ts,
                       static_cast<int>(ctx->n1), static_cast<int>(n2), static_cast<int>(ctx->hx),
                       make_fastdiv(static_cast<uint32_t>(plane)), make_fastdiv(static_cast<uint32_t>(kTile * ctx->n1)),
                       static_cast<int>(ctx->hxs), static_cast<int>(ctx->n1h), static_cast<int>(ctx->n2h),
                       static_cast<int>(total), rv);
    MAFX_HIP(hipGetLastError());
}

void zfused_sym(mafx_ctx* ctx, float2* spec, const float* ts, int nz, const TensorRes* res) {
    ZTw<float2> tw{};
    const int n2 = 2 * nz;
    const long double two_pi = 6.283185307179586476925286766559L;
    for (int t = 0; t < n2; ++t) {
        const long double a = two_pi * t / n2;
        tw.w[t] = float2{static_cast<float>(std::cos(a)), static_cast<float>(-std::sin(a))};
    }
    // (kx mod 4, kys, kx / 4) over whole column tiles; columns kx >= hx return at once
    const int64_t threads = ctx->n1h * kTile * ((ctx->hx + kTile - 1) / kTile);
    require(threads < (int64_t{1} << 31), "spectrum too large for the fused z stage");
    // spectrum columns read and written once, the tensor quarter read once
    const double bytes = static_cast<double>(ctx->n1 * ctx->hx) * (2.0 * 3 * nz) * sizeof(float2) +
                         static_cast<double>(ctx->n1h * ctx->hx) * 6.0 * (nz + 1) * sizeof(float);
    const double res_bytes = res ? static_cast<double>(ctx->n1 * ctx->hx) * 6.0 * (2 * nz) * sizeof(__half2) : 0.0;
    ProfScope ps(ctx, res ? "zfused_mul_symres_f32" : "zfused_mul_f32", bytes + res_bytes);
    const dim3 g(grid_for(threads, 256)), b(256);
    const int n1 = static_cast<int>(ctx->n1), hxs = static_cast<int>(ctx->hxs), hx = static_cast<int>(ctx->hx);
    const FastDiv n1hd = make_fastdiv(static_cast<uint32_t>(ctx->n1h));
    const TensorRes rv = res ? *res : TensorRes{};
    auto launch = [&](auto nzc) {
        constexpr int NZ = decltype(nzc)::value;
        // NZ = 8 through the in-place FFT form (165 VGPRs, 3 waves per SIMD, against 256 + AGPRs
        // and one wave for the direct DFT: 740 -> 536 us at 2048 x 1024 x 8); at NZ = 4 the direct
        // DFT is faster (92 against 113 us at 1024 x 1024 x 4)
        if constexpr (NZ == 8) {
            if (res)
                hipLaunchKernelGGL((k_zfused_sym_fft<NZ, true>), g, b, 0, ctx->stream, spec, ts, n1, hx, hxs, n1hd, tw, rv);
            else
                hipLaunchKernelGGL((k_zfused_sym_fft<NZ, false>), g, b, 0, ctx->stream, spec, ts, n1, hx, hxs, n1hd, tw, rv);
            return;
        }
        if (res)
            hipLaunchKernelGGL((k_zfused_sym<NZ, true>), g, b, 0, ctx->stream, spec, ts, n1, hx, hxs, n1hd, tw, rv);
        else
            hipLaunchKernelGGL((k_zfused_sym<NZ, false>), g, b, 0, ctx->stream, spec, ts, n1, hx, hxs, n1hd, tw, rv);
    };
    switch (nz) {
    case 2: launch(std::integral_constant<int, 2>{}); break;
    case 3: launch(std::integral_constant<int, 3>{}); break;
    case 4: launch(std::integral_constant<int, 4>{}); break;
    case 5: launch(std::integral_constant<int, 5>{}); break;
    case 6: launch(std::integral_constant<int, 6>{}); break;
    case 7: launch(std::integral_constant<int, 7>{}); break;
    case 8: launch(std::integral_constant<int, 8>{}); break;
    default: throw Error(MAFX_ERR_ARG, "fused z stage supports 2 <= nz <= 8");
    }
    MAFX_HIP(hipGetLastError());
}

template void fft_x_forward<double2>(mafx_ctx*, const double*, const double*, double, int64_t, int64_t, int64_t, int,
                                     int64_t, int64_t, double2*, int64_t);
template void fft_x_forward<float2>(mafx_ctx*, const double*, const double*, double, int64_t, int64_t, int64_t, int,
                                    int64_t, int64_t, float2*, int64_t);
template void fft_x_inverse<double2, double>(mafx_ctx*, const double2*, int64_t, double*, double);
template void fft_x_inverse<float2, float>(mafx_ctx*, const float2*, int64_t, float*, double);
template void fft_strided<double2>(mafx_ctx*, const FftPlan&, double2*, int64_t, int, int64_t, int, int64_t, int,
                                   int, bool, const char*);
template void fft_strided<float2>(mafx_ctx*, const FftPlan&, float2*, int64_t, int, int64_t, int, int64_t, int, int,
                                  bool, const char*);
template void fft_y_split<double2>(mafx_ctx*, double2*, int64_t, int, int64_t, int, int64_t, bool);
template void fft_y_split<float2>(mafx_ctx*, float2*, int64_t, int, int64_t, int, int64_t, bool);
template void kspace_mul<double2>(mafx_ctx*, double2*, const double2*, int64_t);
template void zfused<double2>(mafx_ctx*, double2*, const double2*, int);

} // namespace mafx
