// surrogate_kernels.h - launchers of the gfx950 kernels of the warm-start network (internal).
//
// Tensors are NHWC: [batch][H][W][channels], H along x (the cell index k), W along y (the
// cell index l), channels innermost. Activations, their gradients and the convolution weights
// are bf16 (raw uint16_t bits); accumulation, biases, the head and the optimiser are FP32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dsmc {
namespace net {

// Per-case inputs of the network: the obstacle box and domain (to rasterise the mask and the
// coordinate channels) and the broadcast scalar features.
constexpr int kInChannels = 32;   // input channels as stored (features padded with zeros)
constexpr int kNumFeatures = 9;   // broadcast features, input channels 3..11
struct NetCase {
    float box[4];                 // x0, x1, y0, y1 of the obstacle
    float dom[4];                 // x1, x2, y1, y2 of the domain
    float feat[kNumFeatures];
};

// 3x3 convolution, stride 1, zero padding 1, as an implicit GEMM on bf16 MFMA.
// Input channels [0, c1) come from x1 [B][H][W][c1]; channels [c1, c1+c2) from x2
// [B][H/2][W/2][c2] read through a nearest-neighbour 2x upsample (the decoder's concat).
// Weights w [cout][9][c1+c2] (tap = 3*ky + kx). Output channels [0, csplit) go to y1
// [B][H][W][csplit], the rest to y2 [B][H][W][cout-csplit].
enum ConvMode { kConvBiasRelu = 0, kConvMask = 1, kConvPlain = 2 };
struct ConvArgs {
    const uint16_t* x1;
    const uint16_t* x2;
    const uint16_t* w;
    const float* bias;     // kConvBiasRelu
    const uint16_t* mask;  // kConvMask: y *= (mask > 0), mask [B][H][W][cout]
    uint16_t* y1;
    uint16_t* y2;
    int c1, c2, cout, csplit;
    int B, H, W;
};
void launch_conv3x3(const ConvArgs& a, ConvMode mode, hipStream_t s);

// Weight gradient of the same convolution: part[s][cout][9][cin] = sum over the pixels of
// split s of dz[p][co] * x(p + tap)[ci]; `rows` pixels per split.
struct WgradArgs {
    const uint16_t* dz;
    const uint16_t* x1;
    const uint16_t* x2;
    float* part;
    int cout, c1, c2;
    int B, H, W;
    int rows;
};
void launch_wgrad(const WgradArgs& a, int splits, hipStream_t s);
// Bias gradient partials: part[s][co] = sum over split s of dz[p][co].
void launch_bias_grad(const uint16_t* dz, int cout, int P, int rows, int splits, float* part,
                      hipStream_t s);
// out[i] = sum_{s < splits} part[s * n + i] (fixed order).
void launch_reduce(const float* part, int splits, size_t n, float* out, hipStream_t s);

// 2x2 average pool: y [B][H/2][W/2][C].
void launch_pool2(const uint16_t* x, int B, int H, int W, int C, uint16_t* y, hipStream_t s);
// out = (dskip + 0.25 * upsample(dpool)) * (act > 0), all [B][H][W][C] except dpool (H/2, W/2).
void launch_pool_back_add(const uint16_t* dskip, const uint16_t* dpool, const uint16_t* act,
                          int B, int H, int W, int C, uint16_t* out, hipStream_t s);
// Coarse gradient of the decoder's upsample: out[B][H/2][W/2][C] = (sum of the 4 children of
// dup [B][H][W][C]) * (act > 0), act being the coarse activation.
void launch_up_back(const uint16_t* dup, const uint16_t* act, int B, int H, int W, int C,
                    uint16_t* out, hipStream_t s);

// The 1x1 output head (C -> 4, FP32 weights) with the masked squared-error loss.
// out[p][j] = bh[j] + sum_c wh[j][c] h[p][c]. With targets: per sample `idx[b]` of the batch,
// tgt [.][HW][4] and wgt [.][HW]; loss_part[block] = sum of scale*w*(out-t)^2;
// g[p][j] = 2*scale*w*(out-t); dh[p][c] = sum_j g[p][j] wh[j][c] * (h[p][c] > 0).
struct HeadArgs {
    const uint16_t* h;
    const float* wh;
    const float* bh;
    int C, P, HW;
    const int* idx;
    const float* tgt;
    const float* wgt;
    float scale;
    float* out;
    float* g;
    uint16_t* dh;
    float* loss_part;
};
int launch_head(const HeadArgs& a, hipStream_t s);  // returns the number of blocks
// Head weight partials: part[s][j*C + c] = sum g[p][j] h[p][c], part[s][4*C + j] = sum g[p][j].
void launch_head_wgrad(const uint16_t* h, const float* g, int C, int P, int rows, int splits,
                       float* part, hipStream_t s);

// Network inputs of batch entry b = case cases[idx[b]] (or cases[b] if idx is null).
void launch_encode(const NetCase* cases, const int* idx, int B, int H, int W, uint16_t* x,
                   hipStream_t s);

// Adam on the FP32 master weights (bias corrections bc1 = 1 - b1^t, bc2 = 1 - b2^t).
void launch_adam(float* w, const float* g, float* m, float* v, size_t n, float lr, float b1,
                 float b2, float eps, float bc1, float bc2, hipStream_t s);
// bf16 copies of one layer's weights w [cout][9][cin]: wb in the same layout and wt
// [cin][9][cout] with the taps mirrored (the data-gradient convolution's weights).
void launch_pack_weights(const float* w, int cout, int cin, uint16_t* wb, uint16_t* wt,
                         hipStream_t s);
// float -> bf16 bits (round to nearest even).
void launch_to_bf16(const float* x, size_t n, uint16_t* y, hipStream_t s);
void launch_from_bf16(const uint16_t* x, size_t n, float* y, hipStream_t s);

}  // namespace net
}  // namespace dsmc
