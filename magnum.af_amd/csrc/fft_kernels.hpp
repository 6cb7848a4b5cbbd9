// Host launchers of the FFT axis kernels (fft.hip).
#pragma once
#include "fft_device.hpp"
#include "mafx_internal.hpp"

namespace mafx {

FftDesc desc_of(const FftPlan& p);
// rows (c, z, y) of the real input -> x half spectrum at out[((c n2o + z) n1 + y) hx + k]
template <class C>
void fft_x_forward(mafx_ctx* ctx, const FftPlan& p, const double* in, const double* ms, double scale, int64_t Xi,
                   int64_t Yi, int64_t Zi, int ncomp, int64_t Yr, int64_t Zr, C* out, int64_t n2o);
template <class C>
void fft_strided(mafx_ctx* ctx, const FftPlan& p, C* data, int64_t sa, int na, int64_t sb, int nb, int64_t stride,
                 int n_in, int n_out, bool inv, const char* name);
template <class C> void kspace_mul(mafx_ctx* ctx, C* spec, const C* tensor, int64_t K);
// z DFT (length 2 nz) + tensor product + inverse z DFT, 2 <= nz <= 8, spectrum planes [3][nz]
template <class C> void zfused(mafx_ctx* ctx, C* spec, const C* tensor, int nz);
void convert_tensor_f32(mafx_ctx* ctx, const double2* a, float2* b, int64_t n);

} // namespace mafx
