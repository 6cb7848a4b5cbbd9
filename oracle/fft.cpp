// Oracle FFT: unnormalised mixed-radix decimation-in-time DFT in FP64 (any length; radix 4
// first, then the prime factors).  Stands in for the FFTW transforms ArrayFire's CPU backend
// calls in demag_field.cpp:146,229-257 (af::fftR2C / af::fftC2R).  Test infrastructure only.
//
// X[k] = sum_n x[n] exp(sign * 2 pi i n k / n_total).  Pinned in tests against numpy.fft and a
// brute-force DFT.
#include "oracle_internal.hpp"

#include <cmath>
#include <map>
#include <mutex>
#include <stdexcept>

namespace oracle {
namespace {

struct Twiddles {
    int64_t n{0};
    std::vector<cplx> fwd, bwd; // exp(-+2 pi i t / n)
};

const Twiddles& twiddles(int64_t n) {
    static std::mutex mu;
    static std::map<int64_t, std::unique_ptr<Twiddles>> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto& slot = cache[n];
    if (!slot) {
        slot = std::make_unique<Twiddles>();
        slot->n = n;
        slot->fwd.resize(n);
        slot->bwd.resize(n);
        const long double two_pi = 6.283185307179586476925286766559L;
        for (int64_t t = 0; t < n; ++t) {
            const long double a = two_pi * static_cast<long double>(t) / static_cast<long double>(n);
            const double c = static_cast<double>(std::cos(a));
            const double s = static_cast<double>(std::sin(a));
            slot->fwd[t] = cplx(c, -s);
            slot->bwd[t] = cplx(c, s);
        }
    }
    return *slot;
}

int64_t pick_factor(int64_t n) {
    if (n % 4 == 0) return 4;
    for (int64_t p = 2; p * p <= n; ++p)
        if (n % p == 0) return p;
    return n;
}

// out[k] for k < n receives the DFT of in[0], in[stride], ..., in[(n-1)*stride].
// tw is the table of the full length N0, twstep = N0 / n.
void dft_rec(const cplx* in, int64_t stride, cplx* out, int64_t n, const cplx* tw, int64_t twstep,
             int64_t N0) {
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    const int64_t p = pick_factor(n);
    const int64_t m = n / p;
    for (int64_t q = 0; q < p; ++q) dft_rec(in + q * stride, stride * p, out + q * m, m, tw, twstep * p, N0);
    // combine: X[k + r m] = sum_q w_n^{q (k + r m)} Y_q[k]
    std::vector<cplx> t(p), o(p);
    for (int64_t k = 0; k < m; ++k) {
        for (int64_t q = 0; q < p; ++q) t[q] = out[q * m + k];
        for (int64_t r = 0; r < p; ++r) {
            cplx acc(0.0, 0.0);
            const int64_t kk = k + r * m;
            for (int64_t q = 0; q < p; ++q) {
                const int64_t e = ((q * kk) % n) * twstep; // exponent on the N0 table
                acc += t[q] * tw[e % N0];
            }
            o[r] = acc;
        }
        for (int64_t r = 0; r < p; ++r) out[r * m + k] = o[r];
    }
}

} // namespace

void fft_line(cplx* a, int64_t n, int sign) {
    if (n <= 1) return;
    const Twiddles& T = twiddles(n);
    std::vector<cplx> in(a, a + n);
    dft_rec(in.data(), 1, a, n, sign < 0 ? T.fwd.data() : T.bwd.data(), 1, n);
}

void fft_axis(std::vector<cplx>& a, int64_t n0, int64_t n1, int64_t n2, int axis, int sign) {
    const int64_t len = axis == 0 ? n0 : axis == 1 ? n1 : n2;
    if (len <= 1) return;
    const int64_t lines = n0 * n1 * n2 / len;
    const int64_t stride = axis == 0 ? 1 : axis == 1 ? n0 : n0 * n1;
    twiddles(len); // build the table once outside the parallel region
#pragma omp parallel
    {
        std::vector<cplx> buf(len);
#pragma omp for schedule(static)
        for (int64_t l = 0; l < lines; ++l) {
            int64_t base;
            if (axis == 0) {
                base = l * n0;
            } else if (axis == 1) {
                const int64_t i0 = l % n0, i2 = l / n0;
                base = i0 + n0 * n1 * i2;
            } else {
                base = l; // l = i0 + n0 * i1
            }
            for (int64_t i = 0; i < len; ++i) buf[i] = a[base + i * stride];
            fft_line(buf.data(), len, sign);
            for (int64_t i = 0; i < len; ++i) a[base + i * stride] = buf[i];
        }
    }
}

void fft3(std::vector<cplx>& a, int64_t n0, int64_t n1, int64_t n2, int sign) {
    fft_axis(a, n0, n1, n2, 0, sign);
    fft_axis(a, n0, n1, n2, 1, sign);
    fft_axis(a, n0, n1, n2, 2, sign);
}

} // namespace oracle
