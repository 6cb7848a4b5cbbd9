// Host launchers of the FFT axis kernels (fft.hip).
#pragma once
#include <hip/hip_fp16.h>

#include "fft_device.hpp"
#include "mafx_internal.hpp"

namespace mafx {

FftDesc desc_of(const FftPlan& p);

// twiddle table of a plan in the precision of C
template <class C> inline const C* tw_of(const FftPlan& p) {
    if constexpr (sizeof(C) == sizeof(double2))
        return reinterpret_cast<const C*>(p.tw64.p);
    else
        return reinterpret_cast<const C*>(p.tw32.p);
}

// complex elements of a 64 KiB LDS budget
template <class C> constexpr int lds_elems() { return 65536 / static_cast<int>(sizeof(C)); }

// rows (c, z, y) of the real input (row length Xi <= n0) -> x half spectrum at
// out[(c n2o + z) n1 hxs + 4 y + tcol(k, 4 n1)], k <= n0/2 (column tiles), through one FFT of
// length n0/2 per row
template <class C>
void fft_x_forward(mafx_ctx* ctx, const double* in, const double* ms, double scale, int64_t Xi, int64_t Yi,
                   int64_t Zi, int ncomp, int64_t Yr, int64_t Zr, C* out, int64_t n2o);
// inverse real transform along x of the spectrum rows (c < 3, z < nz, y < ny; n2s z-planes per
// component) cropped to x < nx and scaled: out[c N + (z ny + y) nx + x]
template <class C, class T> void fft_x_inverse(mafx_ctx* ctx, const C* spec, int64_t n2s, T* out, double hscale);
// general strided complex transform (full length, optional zero padding / cropping) of the lines
// data[a sa + b sb + tcol(kx, 4 n1) + pos stride]
template <class C>
void fft_strided(mafx_ctx* ctx, const FftPlan& p, C* data, int64_t sa, int na, int64_t sb, int nb, int64_t stride,
                 int n_in, int n_out, bool inv, const char* name);
// y transform of the zero-padded spectrum (inputs / outputs y < ny) as two length-ny transforms
template <class C>
void fft_y_split(mafx_ctx* ctx, C* data, int64_t sa, int na, int64_t sb, int nb, int64_t stride, bool inv);
template <class C> void kspace_mul(mafx_ctx* ctx, C* spec, const C* tensor, int64_t K);
// z DFT (length 2 nz) + tensor product + inverse z DFT, 2 <= nz <= 8, spectrum planes [3][nz]
template <class C> void zfused(mafx_ctx* ctx, C* spec, const C* tensor, int nz);
// FP32 path, real mirror-symmetric tensor spectrum (fft.hip): zero the unpaired planes of the padded
// real-space tensor [ncomp][n2][n1][n0]; keep the real quarter Ts[6][n2h] of its spectrum, planes of
// n1h rows in column tiles;
// apply it to a spectrum of n2 z-planes; fused z stage for 2 <= nz <= 8
void zero_unpaired_planes(mafx_ctx* ctx, double* N, int ncomp);
void tensor_sym_f32(mafx_ctx* ctx, const double2* T, float* Ts);
// The reference's own spectrum on the FP32 path (MAFX_TERM_FULL_TENSOR): the symmetric quarter plus
// the residual R = T - S over the whole spectrum [6][n2][plane], stored as FP16 pairs scaled per
// component by a power of two (R / S is ~1e-4, so FP16's 2^-11 leaves ~1e-7 of the entry).
struct TensorRes {
    const __half2* r; // [6][n2][plane], spectrum layout
    float inv[6];     // 1 / scale per component
};
struct TensorScale {
    double s[6];
};
// R = T - S in place on T, then out = fp16(R scale); inv_scale receives 1 / scale per component
void tensor_residual_f16(mafx_ctx* ctx, double2* T, const float* Ts, __half2* out, float inv_scale[6]);
void kspace_mul_sym(mafx_ctx* ctx, float2* spec, const float* ts, int64_t n2, const TensorRes* res = nullptr);
void zfused_sym(mafx_ctx* ctx, float2* spec, const float* ts, int nz, const TensorRes* res = nullptr);

} // namespace mafx
