// Learned warm start: the encoder-decoder network (net.hip).  No counterpart in the reference.
#pragma once
#include <cstdint>
#include <vector>

#include "mafx_internal.hpp"

namespace mafx {

// (in, out) channels of the eight 3x3 convolutions in execution order
constexpr int kNetLayers = 8;
constexpr int kNetCin[kNetLayers] = {8, 32, 64, 128, 256, 128, 64, 32};
constexpr int kNetCout[kNetLayers] = {32, 64, 128, 128, 64, 32, 32, 3};
constexpr int kNetInputs = 8; // m0 (3) + Ms, A, D, Ku, Bz planes

struct NetAct; // activations of one (B, H, W) shape

struct Net {
    // parameters: per layer W [Cout][Cin][3][3] then b [Cout], concatenated (FP32 masters)
    int64_t nparams{0};
    int64_t woff[kNetLayers]{}, boff[kNetLayers]{};
    DevBuf master, bf16w, grad, adam_m, adam_v; // bf16w: 16-bit copies of the weights only
    DevBuf loss_part, scal;
    double* host_scal{nullptr};
    int64_t adam_t{0};
    hipStream_t stream{nullptr};
    NetAct* act{nullptr};
    int B{0}, H{0}, W{0};

    explicit Net(hipStream_t s);
    ~Net();
    void set_params(const float* host);
    void get_params(float* host) const;
    void ensure_shape(int b, int h, int w);
    void forward(const float* x);                 // leaves m in the activations
    const float* output() const;                  // [B][3][H][W], unit vectors
    double backward(const float* target);         // gradients into grad, returns the loss
    void adam(float lr, float beta1, float beta2, float eps);
    void sync_bf16();                             // bf16w <- round(master)
};

// x [1][8][ny][nx] from m0 (SoA f64, nz = 1) and the five parameter planes; m_out (SoA f64) from
// the network output.
void net_pack_input(Net* net, const double* m0, int nx, int ny, const float planes[5], float* x);
void net_unpack_output(Net* net, const float* m, int nx, int ny, double* m_out);

} // namespace mafx
