// LDS-staged Stockham autosort FFT for gfx950.
//
// A work-group transforms W lines of length L held in LDS in [pos][w] order (line index fastest,
// so consecutive lanes touch consecutive LDS words and, when the lines are adjacent columns of a
// strided axis, consecutive global addresses).  Each radix-R stage reads all of a thread's
// butterflies into registers, synchronises, and writes the outputs back to the autosorted
// positions, so one LDS buffer suffices.  Twiddles come from an FP64-accurate table
// tw[t] = exp(-2 pi i t / L) built on the host and rounded once to the working precision.
//
// Stage recurrence (Stockham autosort, e.g. Govindaraju et al., SC'08): with Ns the product of
// the radices already applied, butterfly j (< L/R) reads v_r = x[j + r L/R], multiplies v_r by
// w_L^(r k L/(Ns R)) with k = j mod Ns, applies an R-point DFT and stores y[(j/Ns) Ns R + k + r Ns].
#pragma once
#include <hip/hip_runtime.h>

#include "fastdiv.hpp"

namespace mafx {

template <class C> struct real_of;
template <> struct real_of<double2> { using type = double; };
template <> struct real_of<float2> { using type = float; };

template <class C> __device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <class C> __device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <class C> __device__ __forceinline__ C cmul(C a, C b) {
    return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// multiply by -i (forward) or +i (inverse)
template <class C> __device__ __forceinline__ C mul_mi(C a, bool inv) { return inv ? C{-a.y, a.x} : C{a.y, -a.x}; }
template <class C> __device__ __forceinline__ C conj_if(C a, bool inv) { return inv ? C{a.x, -a.y} : a; }

constexpr int kFftBlock = 256;

// Block coordinates of a 3-D grid after xcd_swizzle of the x-fastest linear block id.
__device__ __forceinline__ uint3 swizzled_block() {
    const uint32_t gx = gridDim.x, gy = gridDim.y;
    const uint32_t id = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const uint32_t l = xcd_swizzle(id, gx * gy * gridDim.z);
    return uint3{l % gx, (l / gx) % gy, l / (gx * gy)};
}

// Column tiles of spectra and tensors.  A plane (fixed component and z) of n1 rows ky and hxs
// columns kx stores column kx of row ky at tcol(kx, 4 n1) + 4 ky with
// tcol(kx, n1T) = (kx / 4) n1T + kx % 4: each group of 4 columns is one contiguous block over all
// rows.  A y line tile of 4 columns is then a contiguous 32 n1-byte block (FP32) instead of n1
// separate 32-byte pieces, and 4 consecutive x rows of one column tile form 128 contiguous bytes.
constexpr int kTile = 4;
__host__ __device__ inline int64_t tcol(int64_t kx, int64_t n1T) { return (kx >> 2) * n1T + (kx & 3); }

// LDS index padding: one spare element after every 32.  The autosort writes of a stage move
// groups of W lanes at a stride of R W elements, a multiple of the bank count for the power-of-two
// shapes here; the padding spreads consecutive groups over different banks.
__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 5); }
// LDS elements to allocate for n logical elements
constexpr int lds_padded(int n) { return n + (n >> 5) + 1; }

// R-point DFT in registers; forward uses w = exp(-2 pi i / R), inverse its conjugate.
// Generic odd radices read w_R^p = tw[p L / R] from the length-L table.
template <class C, int R>
__device__ __forceinline__ void dft(C (&v)[R], bool inv, const C* __restrict__ tw, int L) {
    using T = typename real_of<C>::type;
    if constexpr (R == 2) {
        const C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        const C a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        const C b0 = cadd(v[1], v[3]), b1 = mul_mi(csub(v[1], v[3]), inv);
        v[0] = cadd(a0, b0);
        v[2] = csub(a0, b0);
        v[1] = cadd(a1, b1);
        v[3] = csub(a1, b1);
    } else if constexpr (R == 8) {
        C e[4] = {v[0], v[2], v[4], v[6]};
        C o[4] = {v[1], v[3], v[5], v[7]};
        dft<C, 4>(e, inv, tw, L);
        dft<C, 4>(o, inv, tw, L);
        const T h = static_cast<T>(0.70710678118654752440084436210485);
        // o[q] *= w8^q
        const C w1 = inv ? C{h, h} : C{h, -h};
        const C w3 = inv ? C{-h, h} : C{-h, -h};
        o[1] = cmul(o[1], w1);
        o[2] = mul_mi(o[2], inv);
        o[3] = cmul(o[3], w3);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            v[q] = cadd(e[q], o[q]);
            v[q + 4] = csub(e[q], o[q]);
        }
    } else {
        C out[R];
        const int step = L / R;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            C acc = v[0];
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const C w = conj_if(tw[((r * q) % R) * step], inv);
                acc = cadd(acc, cmul(v[r], w));
            }
            out[q] = acc;
        }
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = out[q];
    }
}

// Packs row k of the half spectrum X[0..M] for the half-length inverse transform of a real
// sequence of length 2M: Zp[k] = (X[k] + conj X[M-k]) + i w_{2M}^-k (X[k] - conj X[M-k]); the
// unnormalised inverse FFT_M of Zp then holds x[2n] + i x[2n+1].
// (Xk = X[k], Xm = X[M-k], twk = w_{2M}^k.)
template <class C> __device__ __forceinline__ C c2r_combine(C Xk, C Xm, C twk) {
    const C S{Xk.x + Xm.x, Xk.y - Xm.y};
    const C D{Xk.x - Xm.x, Xk.y + Xm.y};
    const C wD = cmul(conj_if(twk, true), D);
    return C{S.x - wD.y, S.y + wD.x};
}

// Radix schedule handed to kernels by value, with the per-stage constants precomputed on the
// host: ns[s] = Ns (product of the radices before stage s) as a divisor, twstep[s] = L / (Ns R).
struct FftDesc {
    int L;
    int nstages;
    int radix[16];
    int twstep[16];
    FastDiv ns[16];
};

// One Stockham stage over W lines of length L in buf[pos * W + w].
// MAXB = max butterflies per thread = ceil(W L / R / BLOCK).  Butterfly b covers line w = b mod W,
// index j = b / W; its inputs sit at pos j + r L/R, i.e. at b + r (L/R) W, so the reads need no
// division.
template <class C, int R, int MAXB, int BLOCK>
__device__ __forceinline__ void fft_stage(C* buf, int L, const FastDiv& W, const FastDiv& Ns, int twstep,
                                          const C* __restrict__ tw, bool inv) {
    const int m = L / R;
    const int nb = m * static_cast<int>(W.d);
    const int rs = m * static_cast<int>(W.d);
    C v[MAXB][R];
#pragma unroll
    for (int i = 0; i < MAXB; ++i) {
        const int b = threadIdx.x + i * BLOCK;
        if (b < nb) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[i][r] = buf[lds_pad(b + r * rs)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < MAXB; ++i) {
        const int b = threadIdx.x + i * BLOCK;
        if (b < nb) {
            const int j = static_cast<int>(fdiv(b, W)), w = b - j * static_cast<int>(W.d);
            const int q = static_cast<int>(fdiv(j, Ns));
            const int ns = static_cast<int>(Ns.d);
            const int k = j - q * ns;
            if (ns > 1) {
                const int step = k * twstep;
#pragma unroll
                for (int r = 1; r < R; ++r) v[i][r] = cmul(v[i][r], conj_if(tw[r * step], inv));
            }
            dft<C, R>(v[i], inv, tw, L);
            const int base = (q * ns * R + k) * static_cast<int>(W.d) + w;
            const int ostride = ns * static_cast<int>(W.d);
#pragma unroll
            for (int r = 0; r < R; ++r) buf[lds_pad(base + r * ostride)] = v[i][r];
        }
    }
    __syncthreads();
}

// Runs all stages.  LDS_ELEMS bounds W * L (the LDS budget in complex elements).  POW2 plans
// (radices 2, 4, 8 only) get a stage switch without the odd radices, whose register arrays
// would otherwise set the kernel's VGPR count.
template <class C, int LDS_ELEMS, bool POW2, int B = kFftBlock>
__device__ void fft_lds(C* buf, const FftDesc& d, const FastDiv& W, const C* __restrict__ tw, bool inv) {
    for (int s = 0; s < d.nstages; ++s) {
        const int R = d.radix[s];
        const FastDiv& Ns = d.ns[s];
        const int ts = d.twstep[s];
        switch (R) {
        case 2: fft_stage<C, 2, (LDS_ELEMS / 2 + B - 1) / B, B>(buf, d.L, W, Ns, ts, tw, inv); break;
        case 4: fft_stage<C, 4, (LDS_ELEMS / 4 + B - 1) / B, B>(buf, d.L, W, Ns, ts, tw, inv); break;
        case 8: fft_stage<C, 8, (LDS_ELEMS / 8 + B - 1) / B, B>(buf, d.L, W, Ns, ts, tw, inv); break;
        default:
            if constexpr (!POW2) {
                switch (R) {
                case 3: fft_stage<C, 3, (LDS_ELEMS / 3 + B - 1) / B, B>(buf, d.L, W, Ns, ts, tw, inv); break;
                case 5: fft_stage<C, 5, (LDS_ELEMS / 5 + B - 1) / B, B>(buf, d.L, W, Ns, ts, tw, inv); break;
                case 7: fft_stage<C, 7, (LDS_ELEMS / 7 + B - 1) / B, B>(buf, d.L, W, Ns, ts, tw, inv); break;
                case 11: fft_stage<C, 11, (LDS_ELEMS / 11 + B - 1) / B, B>(buf, d.L, W, Ns, ts, tw, inv); break;
                case 13: fft_stage<C, 13, (LDS_ELEMS / 13 + B - 1) / B, B>(buf, d.L, W, Ns, ts, tw, inv); break;
                default: break; // unreachable: FftPlan::build rejects other factors
                }
            }
            break;
        }
    }
}

// ---- Compile-time shaped transforms ---------------------------------------------------------
// Power-of-two lines whose length M = 2^LOGM and line count W2 = 2^LOGW2 are template constants.
// Every index of a stage is then a shift or a mask instead of a multiply-high division, each lane
// owns an exact number of butterflies, and the twiddle step of every stage is a constant.  Same
// Stockham recurrence, radix schedule (radix-8 stages, then one radix-2 or radix-4 stage) and
// twiddle table as fft_lds, so the arithmetic of each butterfly is unchanged.
constexpr int ct_nstages(int logm) { return logm / 3 + (logm % 3 ? 1 : 0); }
constexpr int ct_logr(int logm, int s) { return s < logm / 3 ? 3 : logm % 3; }
constexpr int ct_logns(int logm, int s) { return 3 * (s < logm / 3 ? s : logm / 3); }

// HT: tw holds only w_M^t for t < M / 2; w_M^(t + M/2) = -w_M^t.
template <int LOGM, bool HT, class C> __device__ __forceinline__ C ct_tw(const C* __restrict__ tw, int idx) {
    if constexpr (HT) {
        const C u = tw[idx & ((1 << (LOGM - 1)) - 1)];
        return idx < (1 << (LOGM - 1)) ? u : C{-u.x, -u.y};
    } else {
        return tw[idx];
    }
}

template <class C, int LOGM, int LOGW2, int LOGR, int LOGNS, int BS, bool INV, bool HT = false>
__device__ __forceinline__ void ct_stage(C* buf, const C* __restrict__ tw, const int tid) {
    constexpr int R = 1 << LOGR, W2 = 1 << LOGW2, NS = 1 << LOGNS;
    constexpr int NB = (1 << (LOGM - LOGR)) << LOGW2; // butterflies; also the stride between inputs
    constexpr int NBT = (NB + BS - 1) / BS;
    constexpr int TWS = LOGM - LOGNS - LOGR; // log2 of the twiddle step M / (Ns R)
    C v[NBT][R];
#pragma unroll
    for (int i = 0; i < NBT; ++i) {
        const int b = tid + i * BS;
        if (NB % BS == 0 || b < NB) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[i][r] = buf[lds_pad(b + r * NB)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NBT; ++i) {
        const int b = tid + i * BS;
        if (NB % BS == 0 || b < NB) {
            const int j = b >> LOGW2, l = b & (W2 - 1), k = j & (NS - 1);
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) v[i][r] = cmul(v[i][r], conj_if(ct_tw<LOGM, HT>(tw, (r * k) << TWS), INV));
            }
            dft<C, R>(v[i], INV, tw, 1 << LOGM);
            const int o = ((((j - k) << LOGR) + k) << LOGW2) + l;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[lds_pad(o + (r << (LOGNS + LOGW2)))] = v[i][r];
        }
    }
    __syncthreads();
}

// All stages of W2 lines of length M held in buf[pos * W2 + line]; tw = w_M^t, t < M.  tid is
// the lane's index in the group (threadIdx.x, or an opaque copy of it in persistent loops).
template <class C, int LOGM, int LOGW2, int BS, bool INV, int S = 0, int SEND = ct_nstages(LOGM), bool HT = false>
__device__ __forceinline__ void ct_fft(C* buf, const C* __restrict__ tw, const int tid = static_cast<int>(threadIdx.x)) {
    if constexpr (S < SEND) {
        ct_stage<C, LOGM, LOGW2, ct_logr(LOGM, S), ct_logns(LOGM, S), BS, INV, HT>(buf, tw, tid);
        ct_fft<C, LOGM, LOGW2, BS, INV, S + 1, SEND, HT>(buf, tw, tid);
    }
}

// v unchanged, but unknown to the optimiser: address arithmetic derived from it stays inside a
// persistent loop instead of being hoisted out and held in registers across iterations.
__device__ __forceinline__ int opaque_int(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

// Pieces of stage S for kernels that keep its inputs or outputs in registers.  Butterfly b of
// stage S covers line b % W2 and index j = b / W2 < M / R; its inputs are positions j + r M / R,
// and its outputs go to positions ct_opos(b, r) of the same line.
template <int LOGM, int LOGW2, int S> struct CtStage {
    static constexpr int LOGR = ct_logr(LOGM, S), LOGNS = ct_logns(LOGM, S);
    static constexpr int R = 1 << LOGR, NS = 1 << LOGNS;
    static constexpr int NB = (1 << (LOGM - LOGR)) << LOGW2;
    __device__ static __forceinline__ int opos(int b, int r) {
        const int j = b >> LOGW2, k = j & (NS - 1);
        return ((j - k) << LOGR) + k + (r << LOGNS);
    }
    template <class C, bool INV, bool HT = false>
    __device__ static __forceinline__ void bfly(C (&v)[R], int b, const C* __restrict__ tw) {
        if constexpr (NS > 1) {
            const int k = (b >> LOGW2) & (NS - 1);
#pragma unroll
            for (int r = 1; r < R; ++r)
                v[r] = cmul(v[r], conj_if(ct_tw<LOGM, HT>(tw, (r * k) << (LOGM - LOGNS - LOGR)), INV));
        }
        dft<C, R>(v, INV, tw, 1 << LOGM);
    }
};

inline bool is_pow2_plan(const FftDesc& d) {
    for (int s = 0; s < d.nstages; ++s)
        if (d.radix[s] != 2 && d.radix[s] != 4 && d.radix[s] != 8) return false;
    return true;
}

} // namespace mafx
