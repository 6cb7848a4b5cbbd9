// Learned warm start for the LBFGS minimizer: a 4-level convolutional encoder-decoder (BF16
// weights, FP32 activations and accumulation), its backward pass and an Adam step, all as HIP
// kernels on the library's stream.  There is no counterpart in the reference; the CPU restatement
// that checks these kernels is oracle/warmstart_net.py (same operations, same order).
//
// Layout: activations [B][C][H][W] f32, x fastest, which for B = 1 and nz = 1 is the library's
// SoA [3][ny][nx] order of m.  Parameters: per layer W [Cout][Cin][3][3] then b [Cout], flat.
//
// Kernels:
//   k_conv3x3<TRANS>  16x16 output tile x 16 output channels per 256-lane work-group, 8 input
//                     channels staged in LDS per round (with the 1-cell halo), BF16 weights
//                     widened to f32 in LDS.  Input channel c < Ca comes from a (optionally at half
//                     resolution, read through nearest upsampling), c >= Ca from b (channel
//                     concatenation).  TRANS is the data-gradient form: weight w[i][o][8 - k].
//   k_conv3x3_wgrad   dW, db as per-chunk partials (fixed tile order per chunk), then
//                     k_reduce_chunks sums the chunks in a fixed order: deterministic.
//   pooling, upsampling and ReLU masks as small elementwise kernels; k_norm_loss normalises the
//   3-vector output per cell and forms the loss and its gradient.
#include <cmath>
#include <cstring>
#include <algorithm>
#include <random>

#include "net.hpp"

namespace mafx {

struct NetAct {
    DevBuf pool;
    float *e1, *p1, *e2, *p2, *e3, *p3, *e4, *d3, *d2, *d1, *y, *m;
    float *dy, *dd1, *dupd2, *de1s, *dd2, *dupd3, *de2s, *dd3, *dupe4, *de3s, *de4, *dp3, *de3, *dp2, *de2, *dp1, *de1;
    DevBuf wpart;
    int64_t wpart_elems{0};
    const float* x{nullptr};
};

namespace {

constexpr int kT = 16, kTP = kT + 2; // output tile edge and with the halo
constexpr int kCOB = 16;             // output channels per work-group
constexpr int kCIB = 8;              // input channels per LDS round
constexpr int kMaxChunks = 64;       // weight-gradient partial chunks

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) { // round to nearest even (finite inputs)
    const uint32_t u = __float_as_uint(f);
    return static_cast<uint16_t>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float gather(const float* __restrict__ a, int Ca, int a_up, const float* __restrict__ b,
                                        int Cb, int n, int ci, int y, int x, int H, int W) {
    if (y < 0 || y >= H || x < 0 || x >= W || ci >= Ca + Cb) return 0.0f;
    if (ci < Ca) {
        if (a_up) {
            const int Ha = H / 2, Wa = W / 2;
            return a[((static_cast<int64_t>(n) * Ca + ci) * Ha + y / 2) * Wa + x / 2];
        }
        return a[((static_cast<int64_t>(n) * Ca + ci) * H + y) * W + x];
    }
    return b[((static_cast<int64_t>(n) * Cb + (ci - Ca)) * H + y) * W + x];
}

template <bool TRANS>
__global__ void __launch_bounds__(256)
    k_conv3x3(const float* __restrict__ a, int Ca, int a_up, const float* __restrict__ b, int Cb,
              const uint16_t* __restrict__ w, int Cw_in, const float* __restrict__ bias, float* __restrict__ o1,
              int Co1, float* __restrict__ o2, int Co2, int H, int W, int tiles_x, int relu) {
    __shared__ float s_in[kCIB][kTP][kTP];
    __shared__ float s_w[kCOB][kCIB][9];
    const int Ci = Ca + Cb, Co = Co1 + Co2;
    const int tile = static_cast<int>(blockIdx.x);
    const int ty0 = (tile / tiles_x) * kT, tx0 = (tile % tiles_x) * kT;
    const int co0 = static_cast<int>(blockIdx.y) * kCOB, n = static_cast<int>(blockIdx.z);
    const int tid = static_cast<int>(threadIdx.x), py = tid / kT, px = tid % kT;
    float acc[kCOB];
#pragma unroll
    for (int o = 0; o < kCOB; ++o) acc[o] = 0.0f;
    for (int ci0 = 0; ci0 < Ci; ci0 += kCIB) {
        for (int e = tid; e < kCIB * kTP * kTP; e += 256) {
            const int c = e / (kTP * kTP), r = e - c * kTP * kTP, yy = r / kTP, xx = r - yy * kTP;
            s_in[c][yy][xx] = gather(a, Ca, a_up, b, Cb, n, ci0 + c, ty0 + yy - 1, tx0 + xx - 1, H, W);
        }
        for (int e = tid; e < kCOB * kCIB * 9; e += 256) {
            const int o = e / (kCIB * 9), r = e - o * kCIB * 9, c = r / 9, k = r - c * 9;
            const int co = co0 + o, ci = ci0 + c;
            float v = 0.0f;
            if (co < Co && ci < Ci)
                v = TRANS ? bf2f(w[(static_cast<int64_t>(ci) * Cw_in + co) * 9 + 8 - k])
                          : bf2f(w[(static_cast<int64_t>(co) * Cw_in + ci) * 9 + k]);
            s_w[o][c][k] = v;
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < kCIB; ++c)
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                const float v = s_in[c][py + k / 3][px + k % 3];
#pragma unroll
                for (int o = 0; o < kCOB; ++o) acc[o] += v * s_w[o][c][k];
            }
        __syncthreads();
    }
    const int y = ty0 + py, x = tx0 + px;
    if (y >= H || x >= W) return;
#pragma unroll
    for (int o = 0; o < kCOB; ++o) {
        const int co = co0 + o;
        if (co >= Co) break;
        float v = acc[o] + (bias ? bias[co] : 0.0f);
        if (relu) v = fmaxf(v, 0.0f);
        if (co < Co1)
            o1[((static_cast<int64_t>(n) * Co1 + co) * H + y) * W + x] = v;
        else
            o2[((static_cast<int64_t>(n) * Co2 + (co - Co1)) * H + y) * W + x] = v;
    }
}

// Work-group (co group of 16, ci group of 16, chunk): lane (to, ti) owns the pair (co0+to, ci0+ti).
__global__ void __launch_bounds__(256)
    k_conv3x3_wgrad(const float* __restrict__ d, int Co, const float* __restrict__ a, int Ca, int a_up,
                    const float* __restrict__ b, int Cb, int B, int H, int W, int tiles_x, int tiles_img, int nchunk,
                    float* __restrict__ part_w, float* __restrict__ part_b) {
    __shared__ float s_d[16][kT * kT];
    __shared__ float s_in[16][kTP][kTP];
    const int Ci = Ca + Cb;
    const int co0 = static_cast<int>(blockIdx.x) * 16, ci0 = static_cast<int>(blockIdx.y) * 16;
    const int chunk = static_cast<int>(blockIdx.z);
    const int tid = static_cast<int>(threadIdx.x), to = tid / 16, ti = tid % 16;
    float acc[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) acc[k] = 0.0f;
    float accb = 0.0f;
    const int total = B * tiles_img;
    for (int t = chunk; t < total; t += nchunk) {
        const int n = t / tiles_img, tile = t - n * tiles_img;
        const int ty0 = (tile / tiles_x) * kT, tx0 = (tile % tiles_x) * kT;
        for (int e = tid; e < 16 * kT * kT; e += 256) {
            const int o = e / (kT * kT), p = e - o * kT * kT, y = ty0 + p / kT, x = tx0 + p % kT;
            const int co = co0 + o;
            s_d[o][p] = (co < Co && y < H && x < W) ? d[((static_cast<int64_t>(n) * Co + co) * H + y) * W + x] : 0.0f;
        }
        for (int e = tid; e < 16 * kTP * kTP; e += 256) {
            const int c = e / (kTP * kTP), r = e - c * kTP * kTP, yy = r / kTP, xx = r - yy * kTP;
            s_in[c][yy][xx] = gather(a, Ca, a_up, b, Cb, n, ci0 + c, ty0 + yy - 1, tx0 + xx - 1, H, W);
        }
        __syncthreads();
        for (int p = 0; p < kT * kT; ++p) {
            const float dv = s_d[to][p];
            const int py = p / kT, px = p % kT;
#pragma unroll
            for (int k = 0; k < 9; ++k) acc[k] += dv * s_in[ti][py + k / 3][px + k % 3];
            accb += dv;
        }
        __syncthreads();
    }
    const int co = co0 + to, ci = ci0 + ti;
    if (co < Co && ci < Ci) {
#pragma unroll
        for (int k = 0; k < 9; ++k) part_w[((static_cast<int64_t>(chunk) * Co + co) * Ci + ci) * 9 + k] = acc[k];
    }
    if (blockIdx.y == 0 && ti == 0 && co < Co) part_b[static_cast<int64_t>(chunk) * Co + co] = accb;
}

// out[j] = sum over c < nchunk of part[c * count + j], in chunk order
__global__ void k_reduce_chunks(const float* __restrict__ part, int nchunk, int64_t count, float* __restrict__ out) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= count) return;
    float s = 0.0f;
    for (int c = 0; c < nchunk; ++c) s += part[c * count + j];
    out[j] = s;
}

// 2x2 average: out [BC][h][w] from in [BC][2h][2w]
__global__ void k_pool(const float* __restrict__ in, float* __restrict__ out, int64_t BC, int h, int w) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= BC * h * w) return;
    const int64_t bc = i / (static_cast<int64_t>(h) * w);
    const int r = static_cast<int>(i - bc * h * w), y = r / w, x = r - y * w;
    const float* p = in + (bc * 2 * h + 2 * y) * (2 * w) + 2 * x;
    out[i] = 0.25f * (p[0] + p[1] + p[2 * w] + p[2 * w + 1]);
}

// out [BC][H][W] = (pool_bwd(dpool) + skip) * (act > 0)
__global__ void k_pool_bwd_add_mask(const float* __restrict__ dpool, const float* __restrict__ skip,
                                    const float* __restrict__ act, float* __restrict__ out, int64_t BC, int H,
                                    int W) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= BC * H * W) return;
    const int64_t bc = i / (static_cast<int64_t>(H) * W);
    const int r = static_cast<int>(i - bc * H * W), y = r / W, x = r - y * W;
    float v = 0.25f * dpool[(bc * (H / 2) + y / 2) * (W / 2) + x / 2];
    if (skip) v += skip[i];
    out[i] = act[i] > 0.0f ? v : 0.0f;
}

// out [BC][h][w] = up_bwd(dfull) * (act > 0), dfull [BC][2h][2w]
__global__ void k_up_bwd_mask(const float* __restrict__ dfull, const float* __restrict__ act, float* __restrict__ out,
                              int64_t BC, int h, int w) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= BC * h * w) return;
    const int64_t bc = i / (static_cast<int64_t>(h) * w);
    const int r = static_cast<int>(i - bc * h * w), y = r / w, x = r - y * w;
    const float* p = dfull + (bc * 2 * h + 2 * y) * (2 * w) + 2 * x;
    const float v = (p[0] + p[1]) + (p[2 * w] + p[2 * w + 1]);
    out[i] = act[i] > 0.0f ? v : 0.0f;
}

__global__ void k_mask(float* __restrict__ d, const float* __restrict__ act, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n && !(act[i] > 0.0f)) d[i] = 0.0f;
}

constexpr float kNormEps = 1e-12f;

// m = y / sqrt(|y|^2 + eps) per cell; with a target also loss partials (per work-group) and
// dy = (dm - m (m . dm)) / n with dm = 2 (m - target) scale.
__global__ void __launch_bounds__(256)
    k_norm_loss(const float* __restrict__ y, const float* __restrict__ target, float* __restrict__ m,
                float* __restrict__ dy, int B, int HW, float scale, float* __restrict__ part) {
    __shared__ float red[256];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    float l = 0.0f;
    if (i < static_cast<int64_t>(B) * HW) {
        const int64_t n = i / HW, p = i - n * HW;
        const int64_t o0 = (n * 3) * HW + p, o1 = o0 + HW, o2 = o1 + HW;
        const float y0 = y[o0], y1 = y[o1], y2 = y[o2];
        const float nrm = sqrtf(y0 * y0 + y1 * y1 + y2 * y2 + kNormEps);
        const float m0 = y0 / nrm, m1 = y1 / nrm, m2 = y2 / nrm;
        m[o0] = m0;
        m[o1] = m1;
        m[o2] = m2;
        if (target) {
            const float r0 = m0 - target[o0], r1 = m1 - target[o1], r2 = m2 - target[o2];
            l = r0 * r0 + r1 * r1 + r2 * r2;
            const float g0 = 2.0f * r0 * scale, g1 = 2.0f * r1 * scale, g2 = 2.0f * r2 * scale;
            const float mg = m0 * g0 + m1 * g1 + m2 * g2;
            dy[o0] = (g0 - m0 * mg) / nrm;
            dy[o1] = (g1 - m1 * mg) / nrm;
            dy[o2] = (g2 - m2 * mg) / nrm;
        }
    }
    if (!target) return;
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (static_cast<int>(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_sum_f64(const float* __restrict__ part, int n, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += part[i];
    out[0] = s;
}

// Adam (Kingma & Ba 2015) with bias correction; bc1 = 1 - beta1^t, bc2 = 1 - beta2^t
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ mo,
                       float* __restrict__ ve, int64_t n, float lr, float b1, float b2, float eps, float bc1, float bc2) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float m = b1 * mo[i] + (1.0f - b1) * gi;
    const float v = b2 * ve[i] + (1.0f - b2) * gi * gi;
    mo[i] = m;
    ve[i] = v;
    w[i] -= lr * (m / bc1) / (sqrtf(v / bc2) + eps);
}

__global__ void k_to_bf16(const float* __restrict__ w, uint16_t* __restrict__ h, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) h[i] = f2bf(w[i]);
}

__global__ void k_pack_input(const double* __restrict__ m0, int64_t N, float p0, float p1, float p2, float p3,
                             float p4, float* __restrict__ x) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= N) return;
    x[i] = static_cast<float>(m0[i]);
    x[N + i] = static_cast<float>(m0[N + i]);
    x[2 * N + i] = static_cast<float>(m0[2 * N + i]);
    x[3 * N + i] = p0;
    x[4 * N + i] = p1;
    x[5 * N + i] = p2;
    x[6 * N + i] = p3;
    x[7 * N + i] = p4;
}

// FP64 unit vectors from the network output (renormalised in FP64)
__global__ void k_unpack_output(const float* __restrict__ m, int64_t N, double* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double a = m[i], b = m[N + i], c = m[2 * N + i];
    const double n = sqrt(a * a + b * b + c * c);
    out[i] = a / n;
    out[N + i] = b / n;
    out[2 * N + i] = c / n;
}

unsigned nblk(int64_t n, int b = 256) { return static_cast<unsigned>((n + b - 1) / b); }

} // namespace

Net::Net(hipStream_t s) : stream(s) {
    int64_t off = 0;
    for (int l = 0; l < kNetLayers; ++l) {
        woff[l] = off;
        off += static_cast<int64_t>(kNetCout[l]) * kNetCin[l] * 9;
        boff[l] = off;
        off += kNetCout[l];
    }
    nparams = off;
    master.alloc(sizeof(float) * nparams);
    bf16w.alloc(sizeof(uint16_t) * nparams);
    grad.alloc(sizeof(float) * nparams);
    adam_m.alloc(sizeof(float) * nparams);
    adam_v.alloc(sizeof(float) * nparams);
    MAFX_HIP(hipMemsetAsync(adam_m.p, 0, adam_m.bytes, stream));
    MAFX_HIP(hipMemsetAsync(adam_v.p, 0, adam_v.bytes, stream));
    scal.alloc(sizeof(double) * 4);
    MAFX_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_scal), sizeof(double) * 4, hipHostMallocDefault));
    // He-normal initial weights, zero biases
    std::vector<float> h(nparams, 0.0f);
    std::mt19937_64 rng(12345);
    for (int l = 0; l < kNetLayers; ++l) {
        std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / (9.0 * kNetCin[l])));
        for (int64_t i = woff[l]; i < boff[l]; ++i) h[i] = static_cast<float>(nd(rng));
    }
    set_params(h.data());
}

Net::~Net() {
    delete act;
    if (host_scal) (void)hipHostFree(host_scal);
}

void Net::set_params(const float* host) {
    MAFX_HIP(hipMemcpyAsync(master.p, host, sizeof(float) * nparams, hipMemcpyHostToDevice, stream));
    sync_bf16();
    MAFX_HIP(hipStreamSynchronize(stream));
}

void Net::get_params(float* host) const {
    MAFX_HIP(hipMemcpyAsync(host, master.p, sizeof(float) * nparams, hipMemcpyDeviceToHost, stream));
    MAFX_HIP(hipStreamSynchronize(stream));
}

void Net::sync_bf16() {
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk(nparams)), dim3(256), 0, stream, master.as<float>(),
                       bf16w.as<uint16_t>(), nparams);
    MAFX_HIP(hipGetLastError());
}

void Net::ensure_shape(int b, int h, int w) {
    require(b >= 1 && h >= 8 && w >= 8 && h % 8 == 0 && w % 8 == 0, "network input: H and W must be multiples of 8");
    require(static_cast<int64_t>(b) * 256 * h * w < (int64_t{1} << 31), "network input too large");
    if (act && b == B && h == H && w == W) return;
    delete act;
    act = new NetAct();
    B = b;
    H = h;
    W = w;
    const int64_t f = static_cast<int64_t>(b) * h * w; // cells at full resolution
    // float counts of each buffer at its resolution
    struct Item {
        float** p;
        int64_t n;
    };
    NetAct& A = *act;
    const Item items[] = {
        {&A.e1, 32 * f},      {&A.p1, 32 * f / 4},   {&A.e2, 64 * f / 4},  {&A.p2, 64 * f / 16},
        {&A.e3, 128 * f / 16}, {&A.p3, 128 * f / 64}, {&A.e4, 128 * f / 64}, {&A.d3, 64 * f / 16},
        {&A.d2, 32 * f / 4},  {&A.d1, 32 * f},       {&A.y, 3 * f},        {&A.m, 3 * f},
        {&A.dy, 3 * f},       {&A.dd1, 32 * f},      {&A.dupd2, 32 * f},   {&A.de1s, 32 * f},
        {&A.dd2, 32 * f / 4}, {&A.dupd3, 64 * f / 4}, {&A.de2s, 64 * f / 4}, {&A.dd3, 64 * f / 16},
        {&A.dupe4, 128 * f / 16}, {&A.de3s, 128 * f / 16}, {&A.de4, 128 * f / 64}, {&A.dp3, 128 * f / 64},
        {&A.de3, 128 * f / 16}, {&A.dp2, 64 * f / 16}, {&A.de2, 64 * f / 4}, {&A.dp1, 32 * f / 4},
        {&A.de1, 32 * f},
    };
    int64_t total = 0;
    for (const Item& it : items) total += (it.n + 63) / 64 * 64;
    A.pool.alloc(sizeof(float) * total);
    float* base = A.pool.as<float>();
    int64_t off = 0;
    for (const Item& it : items) {
        *it.p = base + off;
        off += (it.n + 63) / 64 * 64;
    }
    // weight-gradient partials: the largest layer times the chunk count
    int64_t wmax = 0;
    for (int l = 0; l < kNetLayers; ++l)
        wmax = std::max<int64_t>(wmax, static_cast<int64_t>(kNetCout[l]) * (kNetCin[l] * 9 + 1));
    A.wpart_elems = wmax * kMaxChunks;
    A.wpart.alloc(sizeof(float) * A.wpart_elems);
    loss_part.alloc(sizeof(float) * nblk(f));
}

namespace {

struct ConvIn {
    const float* a;
    int Ca;
    int a_up;
    const float* b;
    int Cb;
};

void conv_fwd(Net* net, int layer, ConvIn in, float* out, int h, int w, bool relu) {
    const int tiles_x = (w + kT - 1) / kT, tiles = tiles_x * ((h + kT - 1) / kT);
    const int Co = kNetCout[layer];
    require(in.Ca + in.Cb == kNetCin[layer], "conv: channel mismatch");
    hipLaunchKernelGGL(k_conv3x3<false>, dim3(tiles, (Co + kCOB - 1) / kCOB, net->B), dim3(256), 0, net->stream, in.a,
                       in.Ca, in.a_up, in.b, in.Cb, net->bf16w.as<uint16_t>() + net->woff[layer], kNetCin[layer],
                       net->master.as<float>() + net->boff[layer], out, Co, static_cast<float*>(nullptr), 0, h, w,
                       tiles_x, relu ? 1 : 0);
    MAFX_HIP(hipGetLastError());
}

// data gradient of layer `layer` from d (its output gradient): channels [0, C1) to o1, the rest to o2
void conv_bwd_data(Net* net, int layer, const float* d, float* o1, int C1, float* o2, int h, int w) {
    const int tiles_x = (w + kT - 1) / kT, tiles = tiles_x * ((h + kT - 1) / kT);
    const int Cin = kNetCin[layer], Cout = kNetCout[layer];
    hipLaunchKernelGGL(k_conv3x3<true>, dim3(tiles, (Cin + kCOB - 1) / kCOB, net->B), dim3(256), 0, net->stream, d,
                       Cout, 0, static_cast<const float*>(nullptr), 0, net->bf16w.as<uint16_t>() + net->woff[layer],
                       Cin, static_cast<const float*>(nullptr), o1, C1, o2, Cin - C1, h, w, tiles_x, 0);
    MAFX_HIP(hipGetLastError());
}

void conv_bwd_weight(Net* net, int layer, const float* d, ConvIn in, int h, int w) {
    NetAct& A = *net->act;
    const int tiles_x = (w + kT - 1) / kT, tiles = tiles_x * ((h + kT - 1) / kT);
    const int Cin = kNetCin[layer], Cout = kNetCout[layer];
    const int total = net->B * tiles;
    const int nchunk = std::min(total, kMaxChunks);
    const int64_t nw = static_cast<int64_t>(Cout) * Cin * 9;
    require(nchunk * (nw + Cout) <= A.wpart_elems, "weight-gradient partials too small");
    float* pw = A.wpart.as<float>();
    float* pb = pw + nchunk * nw;
    hipLaunchKernelGGL(k_conv3x3_wgrad, dim3((Cout + 15) / 16, (Cin + 15) / 16, nchunk), dim3(256), 0, net->stream, d,
                       Cout, in.a, in.Ca, in.a_up, in.b, in.Cb, net->B, h, w, tiles_x, tiles, nchunk, pw, pb);
    hipLaunchKernelGGL(k_reduce_chunks, dim3(nblk(nw)), dim3(256), 0, net->stream, pw, nchunk, nw,
                       net->grad.as<float>() + net->woff[layer]);
    hipLaunchKernelGGL(k_reduce_chunks, dim3(nblk(Cout)), dim3(256), 0, net->stream, pb, nchunk,
                       static_cast<int64_t>(Cout), net->grad.as<float>() + net->boff[layer]);
    MAFX_HIP(hipGetLastError());
}

void pool_fwd(Net* net, const float* in, float* out, int C, int h, int w) { // h, w: output size
    const int64_t n = static_cast<int64_t>(net->B) * C * h * w;
    hipLaunchKernelGGL(k_pool, dim3(nblk(n)), dim3(256), 0, net->stream, in, out, static_cast<int64_t>(net->B) * C, h, w);
}

} // namespace

void Net::forward(const float* x) {
    NetAct& A = *act;
    A.x = x;
    const int h1 = H, w1 = W, h2 = H / 2, w2 = W / 2, h3 = H / 4, w3 = W / 4, h4 = H / 8, w4 = W / 8;
    conv_fwd(this, 0, {x, 8, 0, nullptr, 0}, A.e1, h1, w1, true);
    pool_fwd(this, A.e1, A.p1, 32, h2, w2);
    conv_fwd(this, 1, {A.p1, 32, 0, nullptr, 0}, A.e2, h2, w2, true);
    pool_fwd(this, A.e2, A.p2, 64, h3, w3);
    conv_fwd(this, 2, {A.p2, 64, 0, nullptr, 0}, A.e3, h3, w3, true);
    pool_fwd(this, A.e3, A.p3, 128, h4, w4);
    conv_fwd(this, 3, {A.p3, 128, 0, nullptr, 0}, A.e4, h4, w4, true);
    conv_fwd(this, 4, {A.e4, 128, 1, A.e3, 128}, A.d3, h3, w3, true);
    conv_fwd(this, 5, {A.d3, 64, 1, A.e2, 64}, A.d2, h2, w2, true);
    conv_fwd(this, 6, {A.d2, 32, 1, A.e1, 32}, A.d1, h1, w1, true);
    conv_fwd(this, 7, {A.d1, 32, 0, nullptr, 0}, A.y, h1, w1, false);
    hipLaunchKernelGGL(k_norm_loss, dim3(nblk(static_cast<int64_t>(B) * H * W)), dim3(256), 0, stream, A.y,
                       static_cast<const float*>(nullptr), A.m, static_cast<float*>(nullptr), B, H * W, 0.0f,
                       static_cast<float*>(nullptr));
    MAFX_HIP(hipGetLastError());
}

const float* Net::output() const { return act->m; }

double Net::backward(const float* target) {
    NetAct& A = *act;
    const int h1 = H, w1 = W, h2 = H / 2, w2 = W / 2, h3 = H / 4, w3 = W / 4, h4 = H / 8, w4 = W / 8;
    const int64_t f = static_cast<int64_t>(B) * H * W;
    const unsigned nb = nblk(f);
    hipLaunchKernelGGL(k_norm_loss, dim3(nb), dim3(256), 0, stream, A.y, target, A.m, A.dy, B, H * W,
                       static_cast<float>(1.0 / static_cast<double>(f)), loss_part.as<float>());
    hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(64), 0, stream, loss_part.as<float>(), static_cast<int>(nb),
                       scal.as<double>());
    const int64_t Bl = B;
    // layer 8: y = conv(d1)
    conv_bwd_data(this, 7, A.dy, A.dd1, 32, nullptr, h1, w1);
    conv_bwd_weight(this, 7, A.dy, {A.d1, 32, 0, nullptr, 0}, h1, w1);
    hipLaunchKernelGGL(k_mask, dim3(nblk(32 * f)), dim3(256), 0, stream, A.dd1, A.d1, 32 * f);
    // layer 7: d1 = relu(conv(cat(up(d2), e1)))
    conv_bwd_data(this, 6, A.dd1, A.dupd2, 32, A.de1s, h1, w1);
    conv_bwd_weight(this, 6, A.dd1, {A.d2, 32, 1, A.e1, 32}, h1, w1);
    hipLaunchKernelGGL(k_up_bwd_mask, dim3(nblk(32 * f / 4)), dim3(256), 0, stream, A.dupd2, A.d2, A.dd2, Bl * 32, h2,
                       w2);
    // layer 6: d2 = relu(conv(cat(up(d3), e2)))
    conv_bwd_data(this, 5, A.dd2, A.dupd3, 64, A.de2s, h2, w2);
    conv_bwd_weight(this, 5, A.dd2, {A.d3, 64, 1, A.e2, 64}, h2, w2);
    hipLaunchKernelGGL(k_up_bwd_mask, dim3(nblk(64 * f / 16)), dim3(256), 0, stream, A.dupd3, A.d3, A.dd3, Bl * 64,
                       h3, w3);
    // layer 5: d3 = relu(conv(cat(up(e4), e3)))
    conv_bwd_data(this, 4, A.dd3, A.dupe4, 128, A.de3s, h3, w3);
    conv_bwd_weight(this, 4, A.dd3, {A.e4, 128, 1, A.e3, 128}, h3, w3);
    hipLaunchKernelGGL(k_up_bwd_mask, dim3(nblk(128 * f / 64)), dim3(256), 0, stream, A.dupe4, A.e4, A.de4, Bl * 128,
                       h4, w4);
    // layer 4: e4 = relu(conv(p3))
    conv_bwd_data(this, 3, A.de4, A.dp3, 128, nullptr, h4, w4);
    conv_bwd_weight(this, 3, A.de4, {A.p3, 128, 0, nullptr, 0}, h4, w4);
    hipLaunchKernelGGL(k_pool_bwd_add_mask, dim3(nblk(128 * f / 16)), dim3(256), 0, stream, A.dp3, A.de3s, A.e3,
                       A.de3, Bl * 128, h3, w3);
    // layer 3: e3 = relu(conv(p2))
    conv_bwd_data(this, 2, A.de3, A.dp2, 64, nullptr, h3, w3);
    conv_bwd_weight(this, 2, A.de3, {A.p2, 64, 0, nullptr, 0}, h3, w3);
    hipLaunchKernelGGL(k_pool_bwd_add_mask, dim3(nblk(64 * f / 4)), dim3(256), 0, stream, A.dp2, A.de2s, A.e2, A.de2,
                       Bl * 64, h2, w2);
    // layer 2: e2 = relu(conv(p1))
    conv_bwd_data(this, 1, A.de2, A.dp1, 32, nullptr, h2, w2);
    conv_bwd_weight(this, 1, A.de2, {A.p1, 32, 0, nullptr, 0}, h2, w2);
    hipLaunchKernelGGL(k_pool_bwd_add_mask, dim3(nblk(32 * f)), dim3(256), 0, stream, A.dp1, A.de1s, A.e1, A.de1,
                       Bl * 32, h1, w1);
    // layer 1: e1 = relu(conv(x))
    conv_bwd_weight(this, 0, A.de1, {A.x, 8, 0, nullptr, 0}, h1, w1);
    MAFX_HIP(hipGetLastError());
    MAFX_HIP(hipMemcpyAsync(host_scal, scal.p, sizeof(double), hipMemcpyDeviceToHost, stream));
    MAFX_HIP(hipStreamSynchronize(stream));
    // the partials hold sum |m - target|^2; the loss is its mean over the B H W cells
    return host_scal[0] / static_cast<double>(f);
}

void Net::adam(float lr, float beta1, float beta2, float eps) {
    ++adam_t;
    const float bc1 = static_cast<float>(1.0 - std::pow(static_cast<double>(beta1), static_cast<double>(adam_t)));
    const float bc2 = static_cast<float>(1.0 - std::pow(static_cast<double>(beta2), static_cast<double>(adam_t)));
    hipLaunchKernelGGL(k_adam, dim3(nblk(nparams)), dim3(256), 0, stream, master.as<float>(), grad.as<float>(),
                       adam_m.as<float>(), adam_v.as<float>(), nparams, lr, beta1, beta2, eps, bc1, bc2);
    MAFX_HIP(hipGetLastError());
    sync_bf16();
}

void net_pack_input(Net* net, const double* m0, int nx, int ny, const float planes[5], float* x) {
    const int64_t N = static_cast<int64_t>(nx) * ny;
    hipLaunchKernelGGL(k_pack_input, dim3(nblk(N)), dim3(256), 0, net->stream, m0, N, planes[0], planes[1], planes[2],
                       planes[3], planes[4], x);
    MAFX_HIP(hipGetLastError());
}

void net_unpack_output(Net* net, const float* m, int nx, int ny, double* m_out) {
    const int64_t N = static_cast<int64_t>(nx) * ny;
    hipLaunchKernelGGL(k_unpack_output, dim3(nblk(N)), dim3(256), 0, net->stream, m, N, m_out);
    MAFX_HIP(hipGetLastError());
}

} // namespace mafx
