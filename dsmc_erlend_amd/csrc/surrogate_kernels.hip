// surrogate_kernels.hip - gfx950 kernels of the warm-start network.
//
// The network is a 4-level convolutional encoder-decoder. Its 3x3 convolutions (forward and
// data gradient) are implicit GEMMs on v_mfma_f32_16x16x32_bf16: pixels are the M dimension,
// output channels N, and (tap, input channel) K. A lane's A operand is 8 consecutive channels
// of one pixel (16 bytes of an NHWC tensor) and its B operand 8 consecutive K entries of one
// output channel's weights, so both come straight from global memory (L1/L2 serve the reuse).
// The weight gradient contracts over pixels, which are strided in NHWC, so both of its
// operands are staged transposed in LDS. Its split-K partials are reduced in a fixed order,
// which keeps every gradient bitwise reproducible.
#include "surrogate_kernels.h"

namespace dsmc {
namespace net {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Round to nearest even; NaN stays NaN.
__device__ inline uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ inline uint16_t elem16(const u32x4& v, int j)
{
    return (uint16_t)(v[j >> 1] >> (16 * (j & 1)));
}

__device__ inline bf16x8 as_bf16x8(const u32x4& v) { return __builtin_bit_cast(bf16x8, v); }

__device__ inline u32x4 load16(const uint16_t* p) { return *reinterpret_cast<const u32x4*>(p); }

constexpr int kNT = 256;

// ---------------------------------------------------------------------------------------
// Convolution: a block is 4 waves along the pixels; a wave computes 64 pixels x 32 output
// channels as 4 x 2 MFMA tiles of 16 x 16.
template <int MODE>
__global__ void __launch_bounds__(kNT) k_conv3x3(ConvArgs a)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r = lane & 15, q = lane >> 4;
    const int HW = a.H * a.W;
    const int P = a.B * HW;
    const int pb = blockIdx.x * 256 + wv * 64;
    const int cb = blockIdx.y * 32;
    const int cin = a.c1 + a.c2;
    const int H2 = a.H >> 1, W2 = a.W >> 1;
    int pn[4], py[4], px[4];
    bool pv[4];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
        const int p = pb + 16 * mi + r;
        pv[mi] = p < P;
        const int pp = pv[mi] ? p : 0;
        pn[mi] = pp / HW;
        const int rem = pp - pn[mi] * HW;
        py[mi] = rem / a.W;
        px[mi] = rem - py[mi] * a.W;
    }
    f32x4 acc[4][2];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) acc[mi][ni] = f32x4{0.f, 0.f, 0.f, 0.f};
    const u32x4 zero = {0u, 0u, 0u, 0u};
    for (int tap = 0; tap < 9; ++tap) {
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        const uint16_t* s1[4];
        const uint16_t* s2[4];
        bool nv[4];
#pragma unroll
        for (int mi = 0; mi < 4; ++mi) {
            const int yy = py[mi] + dy, xx = px[mi] + dx;
            nv[mi] = pv[mi] && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W;
            const int yc = nv[mi] ? yy : 0, xc = nv[mi] ? xx : 0;
            s1[mi] = a.x1 + ((size_t)(pn[mi] * a.H + yc) * a.W + xc) * a.c1 + 8 * q;
            s2[mi] = a.x2 + ((size_t)(pn[mi] * H2 + (yc >> 1)) * W2 + (xc >> 1)) * a.c2 + 8 * q;
        }
        const uint16_t* wr0 = a.w + ((size_t)(cb + r) * 9 + tap) * cin + 8 * q;
        const uint16_t* wr1 = a.w + ((size_t)(cb + 16 + r) * 9 + tap) * cin + 8 * q;
        for (int c0 = 0; c0 < cin; c0 += 32) {
            const bool first = c0 < a.c1;
            bf16x8 A[4];
#pragma unroll
            for (int mi = 0; mi < 4; ++mi) {
                u32x4 v = zero;
                if (nv[mi]) v = load16(first ? s1[mi] + c0 : s2[mi] + (c0 - a.c1));
                A[mi] = as_bf16x8(v);
            }
            const bf16x8 B0 = as_bf16x8(load16(wr0 + c0));
            const bf16x8 B1 = as_bf16x8(load16(wr1 + c0));
#pragma unroll
            for (int mi = 0; mi < 4; ++mi) {
                acc[mi][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[mi], B0, acc[mi][0], 0, 0, 0);
                acc[mi][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[mi], B1, acc[mi][1], 0, 0, 0);
            }
        }
    }
    // lane (r, q) holds rows (pixels) 4q + j, column (channel) r of each tile
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int p = pb + 16 * mi + 4 * q + j;
                const int co = cb + 16 * ni + r;
                if (p >= P) continue;
                float v = acc[mi][ni][j];
                if (MODE == kConvBiasRelu) v = fmaxf(v + a.bias[co], 0.f);
                if (MODE == kConvMask) v = bf2f(a.mask[(size_t)p * a.cout + co]) > 0.f ? v : 0.f;
                if (co < a.csplit)
                    a.y1[(size_t)p * a.csplit + co] = f2bf(v);
                else
                    a.y2[(size_t)p * (a.cout - a.csplit) + (co - a.csplit)] = f2bf(v);
            }
}

// ---------------------------------------------------------------------------------------
// Weight gradient: a block owns TCO output channels x TCI input channels of one tap over one
// split of the pixels. Per stage, 64 pixels of dz and of the tap-shifted input are written
// transposed to LDS (channel rows, pixels contiguous), so an MFMA operand is one 16-byte LDS
// read. The 4 waves split the tile 2 x 2.
constexpr int kWKP = 64;          // pixels per LDS stage
constexpr int kWLD = kWKP + 8;    // LDS row stride in elements (144 bytes)

template <int TCO, int TCI>
__global__ void __launch_bounds__(kNT) k_wgrad(WgradArgs a)
{
    __shared__ __attribute__((aligned(16))) uint16_t sdz[TCO * kWLD];
    __shared__ __attribute__((aligned(16))) uint16_t sx[TCI * kWLD];
    constexpr int SM = TCO / 32, SN = TCI / 32;
    const int s = blockIdx.x;
    const int cin = a.c1 + a.c2;
    const int nci = cin / TCI;
    const int tap = blockIdx.y / nci, ci0 = (blockIdx.y % nci) * TCI;
    const int co0 = blockIdx.z * TCO;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int HW = a.H * a.W, P = a.B * HW, H2 = a.H >> 1, W2 = a.W >> 1;
    const int p_begin = s * a.rows;
    const int p_end = min(P, p_begin + a.rows);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r = lane & 15, q = lane >> 4;
    const int wm = wv >> 1, wn = wv & 1;
    const u32x4 zero = {0u, 0u, 0u, 0u};
    f32x4 acc[SM][SN];
#pragma unroll
    for (int mi = 0; mi < SM; ++mi)
#pragma unroll
        for (int ni = 0; ni < SN; ++ni) acc[mi][ni] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int pc = p_begin; pc < p_end; pc += kWKP) {
        for (int e = threadIdx.x; e < kWKP * TCO / 8; e += kNT) {
            const int pix = e / (TCO / 8), cg = (e % (TCO / 8)) * 8;
            const int p = pc + pix;
            u32x4 v = zero;
            if (p < p_end) v = load16(a.dz + (size_t)p * a.cout + co0 + cg);
#pragma unroll
            for (int j = 0; j < 8; ++j) sdz[(cg + j) * kWLD + pix] = elem16(v, j);
        }
        for (int e = threadIdx.x; e < kWKP * TCI / 8; e += kNT) {
            const int pix = e / (TCI / 8), cg = (e % (TCI / 8)) * 8;
            const int p = pc + pix;
            u32x4 v = zero;
            if (p < p_end) {
                const int b = p / HW, rem = p - b * HW;
                const int y = rem / a.W, x = rem - y * a.W;
                const int yy = y + dy, xx = x + dx;
                if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) {
                    const int c = ci0 + cg;
                    v = c < a.c1
                            ? load16(a.x1 + ((size_t)(b * a.H + yy) * a.W + xx) * a.c1 + c)
                            : load16(a.x2 + ((size_t)(b * H2 + (yy >> 1)) * W2 + (xx >> 1)) * a.c2
                                     + (c - a.c1));
                }
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) sx[(cg + j) * kWLD + pix] = elem16(v, j);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kWKP; kk += 32) {
            bf16x8 A[SM], Bv[SN];
#pragma unroll
            for (int mi = 0; mi < SM; ++mi)
                A[mi] = as_bf16x8(*reinterpret_cast<const u32x4*>(
                    &sdz[(wm * (TCO / 2) + 16 * mi + r) * kWLD + kk + 8 * q]));
#pragma unroll
            for (int ni = 0; ni < SN; ++ni)
                Bv[ni] = as_bf16x8(*reinterpret_cast<const u32x4*>(
                    &sx[(wn * (TCI / 2) + 16 * ni + r) * kWLD + kk + 8 * q]));
#pragma unroll
            for (int mi = 0; mi < SM; ++mi)
#pragma unroll
                for (int ni = 0; ni < SN; ++ni)
                    acc[mi][ni] =
                        __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[mi], Bv[ni], acc[mi][ni], 0, 0, 0);
        }
        __syncthreads();
    }
    // lane (r, q): rows (output channels) 4q + j, column (input channel) r
#pragma unroll
    for (int mi = 0; mi < SM; ++mi)
#pragma unroll
        for (int ni = 0; ni < SN; ++ni)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int co = co0 + wm * (TCO / 2) + 16 * mi + 4 * q + j;
                const int ci = ci0 + wn * (TCI / 2) + 16 * ni + r;
                a.part[(((size_t)s * a.cout + co) * 9 + tap) * cin + ci] = acc[mi][ni][j];
            }
}

__global__ void __launch_bounds__(kNT) k_bias_grad(const uint16_t* dz, int cout, int P, int rows,
                                                  float* part)
{
    __shared__ float sm[kNT];
    const int s = blockIdx.x;
    const int groups = kNT / cout;
    const int c = threadIdx.x % cout, g = threadIdx.x / cout;
    const int p0 = s * rows, p1 = min(P, p0 + rows);
    float acc = 0.f;
    if (g < groups)
        for (int p = p0 + g; p < p1; p += groups) acc += bf2f(dz[(size_t)p * cout + c]);
    sm[threadIdx.x] = acc;
    __syncthreads();
    if ((int)threadIdx.x < cout) {
        float t = 0.f;
        for (int k = 0; k < groups; ++k) t += sm[k * cout + threadIdx.x];
        part[(size_t)s * cout + threadIdx.x] = t;
    }
}

__global__ void __launch_bounds__(kNT) k_reduce(const float* part, int splits, size_t n, float* out)
{
    for (size_t i = (size_t)blockIdx.x * kNT + threadIdx.x; i < n; i += (size_t)gridDim.x * kNT) {
        float t = 0.f;
        for (int s = 0; s < splits; ++s) t += part[(size_t)s * n + i];
        out[i] = t;
    }
}

// ---------------------------------------------------------------------------------------
// Pooling and the decoder's upsample: one thread per 8 channels of one pixel.
__global__ void __launch_bounds__(kNT) k_pool2(const uint16_t* x, int B, int H, int W, int C,
                                              uint16_t* y)
{
    const int H2 = H >> 1, W2 = W >> 1, G = C / 8;
    const size_t n = (size_t)B * H2 * W2 * G;
    for (size_t i = (size_t)blockIdx.x * kNT + threadIdx.x; i < n; i += (size_t)gridDim.x * kNT) {
        const int g = (int)(i % G);
        const size_t pc = i / G;
        const int xc = (int)(pc % W2), yc = (int)((pc / W2) % H2), b = (int)(pc / ((size_t)W2 * H2));
        const uint16_t* base = x + ((size_t)(b * H + 2 * yc) * W + 2 * xc) * C + 8 * g;
        const u32x4 v00 = load16(base), v01 = load16(base + C);
        const u32x4 v10 = load16(base + (size_t)W * C), v11 = load16(base + (size_t)W * C + C);
        uint16_t o[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const float t = ((bf2f(elem16(v00, j)) + bf2f(elem16(v01, j))) + bf2f(elem16(v10, j)))
                            + bf2f(elem16(v11, j));
            o[j] = f2bf(0.25f * t);
        }
        uint16_t* dst = y + pc * C + 8 * g;
#pragma unroll
        for (int j = 0; j < 8; ++j) dst[j] = o[j];
    }
}

__global__ void __launch_bounds__(kNT) k_pool_back_add(const uint16_t* dskip, const uint16_t* dpool,
                                                      const uint16_t* act, int B, int H, int W,
                                                      int C, uint16_t* out)
{
    const int H2 = H >> 1, W2 = W >> 1;
    const size_t n = (size_t)B * H * W * C;
    for (size_t i = (size_t)blockIdx.x * kNT + threadIdx.x; i < n; i += (size_t)gridDim.x * kNT) {
        const int c = (int)(i % C);
        const size_t p = i / C;
        const int x = (int)(p % W), y = (int)((p / W) % H), b = (int)(p / ((size_t)W * H));
        const float dp = bf2f(dpool[((size_t)(b * H2 + (y >> 1)) * W2 + (x >> 1)) * C + c]);
        const float v = bf2f(dskip[i]) + 0.25f * dp;
        out[i] = f2bf(bf2f(act[i]) > 0.f ? v : 0.f);
    }
}

__global__ void __launch_bounds__(kNT) k_up_back(const uint16_t* dup, const uint16_t* act, int B,
                                                int H, int W, int C, uint16_t* out)
{
    const int H2 = H >> 1, W2 = W >> 1;
    const size_t n = (size_t)B * H2 * W2 * C;
    for (size_t i = (size_t)blockIdx.x * kNT + threadIdx.x; i < n; i += (size_t)gridDim.x * kNT) {
        const int c = (int)(i % C);
        const size_t pc = i / C;
        const int xc = (int)(pc % W2), yc = (int)((pc / W2) % H2), b = (int)(pc / ((size_t)W2 * H2));
        const size_t f = ((size_t)(b * H + 2 * yc) * W + 2 * xc) * C + c;
        const float t = ((bf2f(dup[f]) + bf2f(dup[f + C])) + bf2f(dup[f + (size_t)W * C]))
                        + bf2f(dup[f + (size_t)W * C + C]);
        out[i] = f2bf(bf2f(act[i]) > 0.f ? t : 0.f);
    }
}

// ---------------------------------------------------------------------------------------
// Output head and loss: one thread per pixel, C channels in.
template <int C>
__global__ void __launch_bounds__(kNT) k_head(HeadArgs a)
{
    __shared__ float sm[kNT / 64];
    const int p = blockIdx.x * kNT + threadIdx.x;
    float lp = 0.f;
    if (p < a.P) {
        float h[C];
#pragma unroll
        for (int c = 0; c < C; c += 8) {
            const u32x4 v = load16(a.h + (size_t)p * C + c);
#pragma unroll
            for (int j = 0; j < 8; ++j) h[c + j] = bf2f(elem16(v, j));
        }
        float o[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float t = a.bh[j];
#pragma unroll
            for (int c = 0; c < C; ++c) t += a.wh[j * C + c] * h[c];
            o[j] = t;
        }
        if (a.out)
            for (int j = 0; j < 4; ++j) a.out[(size_t)p * 4 + j] = o[j];
        if (a.tgt) {
            const int b = p / a.HW, pix = p - b * a.HW;
            const size_t sp = (size_t)(a.idx ? a.idx[b] : b) * a.HW + pix;
            const float w = a.wgt[sp] * a.scale;
            float g[4];
            for (int j = 0; j < 4; ++j) {
                const float e = o[j] - a.tgt[sp * 4 + j];
                lp += w * (e * e);
                g[j] = 2.f * w * e;
                a.g[(size_t)p * 4 + j] = g[j];
            }
#pragma unroll
            for (int c = 0; c < C; ++c) {
                float d = 0.f;
#pragma unroll
                for (int j = 0; j < 4; ++j) d += g[j] * a.wh[j * C + c];
                a.dh[(size_t)p * C + c] = f2bf(h[c] > 0.f ? d : 0.f);
            }
        }
    }
    if (a.loss_part) {
        for (int d = 32; d >= 1; d >>= 1) lp += __shfl_xor(lp, d, 64);
        if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = lp;
        __syncthreads();
        if (threadIdx.x == 0) {
            float t = 0.f;
            for (int w = 0; w < kNT / 64; ++w) t += sm[w];
            a.loss_part[blockIdx.x] = t;
        }
    }
}

__global__ void __launch_bounds__(kNT) k_head_wgrad(const uint16_t* h, const float* g, int C, int P,
                                                   int rows, float* part)
{
    const int s = blockIdx.x;
    const int t = threadIdx.x;
    const int nout = 4 * C + 4;
    if (t >= nout) return;
    const int p0 = s * rows, p1 = min(P, p0 + rows);
    float acc = 0.f;
    if (t < 4 * C) {
        const int j = t / C, c = t % C;
        for (int p = p0; p < p1; ++p) acc += g[(size_t)p * 4 + j] * bf2f(h[(size_t)p * C + c]);
    } else {
        const int j = t - 4 * C;
        for (int p = p0; p < p1; ++p) acc += g[(size_t)p * 4 + j];
    }
    part[(size_t)s * nout + t] = acc;
}

// ---------------------------------------------------------------------------------------
// Inputs: channel 0 the obstacle mask at the cell centre, 1-2 the centre's coordinates scaled
// to [-1, 1], 3.. the case's broadcast features, the rest zero.
__global__ void __launch_bounds__(kNT) k_encode(const NetCase* cases, const int* idx, int B, int H,
                                               int W, uint16_t* x)
{
    const size_t n = (size_t)B * H * W;
    for (size_t p = (size_t)blockIdx.x * kNT + threadIdx.x; p < n; p += (size_t)gridDim.x * kNT) {
        const int b = (int)(p / ((size_t)H * W));
        const int rem = (int)(p - (size_t)b * H * W);
        const int k = rem / W, l = rem - k * W;
        const NetCase& cs = cases[idx ? idx[b] : b];
        const float fx = ((float)k + 0.5f) / (float)H, fy = ((float)l + 0.5f) / (float)W;
        const float X = cs.dom[0] + (cs.dom[1] - cs.dom[0]) * fx;
        const float Y = cs.dom[2] + (cs.dom[3] - cs.dom[2]) * fy;
        const bool in = X > cs.box[0] && X < cs.box[1] && Y > cs.box[2] && Y < cs.box[3];
        float v[kInChannels];
#pragma unroll
        for (int c = 0; c < kInChannels; ++c) v[c] = 0.f;
        v[0] = in ? 1.f : 0.f;
        v[1] = 2.f * fx - 1.f;
        v[2] = 2.f * fy - 1.f;
#pragma unroll
        for (int f = 0; f < kNumFeatures; ++f) v[3 + f] = cs.feat[f];
        uint16_t* dst = x + p * kInChannels;
#pragma unroll
        for (int c = 0; c < kInChannels; ++c) dst[c] = f2bf(v[c]);
    }
}

// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kNT) k_adam(float* w, const float* g, float* m, float* v,
                                             size_t n, float lr, float b1, float b2, float eps,
                                             float bc1, float bc2)
{
    for (size_t i = (size_t)blockIdx.x * kNT + threadIdx.x; i < n; i += (size_t)gridDim.x * kNT) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    }
}

__global__ void __launch_bounds__(kNT) k_pack_weights(const float* w, int cout, int cin,
                                                     uint16_t* wb, uint16_t* wt)
{
    const size_t n = (size_t)cout * 9 * cin;
    for (size_t i = (size_t)blockIdx.x * kNT + threadIdx.x; i < n; i += (size_t)gridDim.x * kNT) {
        const int ci = (int)(i % cin);
        const int tap = (int)((i / cin) % 9);
        const int co = (int)(i / ((size_t)9 * cin));
        const uint16_t h = f2bf(w[i]);
        wb[i] = h;
        wt[((size_t)ci * 9 + (8 - tap)) * cout + co] = h;
    }
}

__global__ void __launch_bounds__(kNT) k_to_bf16(const float* x, size_t n, uint16_t* y)
{
    for (size_t i = (size_t)blockIdx.x * kNT + threadIdx.x; i < n; i += (size_t)gridDim.x * kNT)
        y[i] = f2bf(x[i]);
}

__global__ void __launch_bounds__(kNT) k_from_bf16(const uint16_t* x, size_t n, float* y)
{
    for (size_t i = (size_t)blockIdx.x * kNT + threadIdx.x; i < n; i += (size_t)gridDim.x * kNT)
        y[i] = bf2f(x[i]);
}

inline unsigned grid_for(size_t n)
{
    const size_t b = (n + kNT - 1) / kNT;
    return (unsigned)(b < 65536 ? (b ? b : 1) : 65536);
}

}  // namespace

void launch_conv3x3(const ConvArgs& a, ConvMode mode, hipStream_t s)
{
    const int P = a.B * a.H * a.W;
    const dim3 grid((P + 255) / 256, a.cout / 32);
    if (mode == kConvBiasRelu)
        hipLaunchKernelGGL(k_conv3x3<kConvBiasRelu>, grid, dim3(kNT), 0, s, a);
    else if (mode == kConvMask)
        hipLaunchKernelGGL(k_conv3x3<kConvMask>, grid, dim3(kNT), 0, s, a);
    else
        hipLaunchKernelGGL(k_conv3x3<kConvPlain>, grid, dim3(kNT), 0, s, a);
}

void launch_wgrad(const WgradArgs& a, int splits, hipStream_t s)
{
    const int cin = a.c1 + a.c2;
    const int tco = a.cout % 64 == 0 ? 64 : 32;
    const int tci = cin % 64 == 0 && a.c1 % 64 == 0 ? 64 : 32;
    const dim3 grid(splits, 9 * (cin / tci), a.cout / tco);
    if (tco == 64 && tci == 64)
        hipLaunchKernelGGL((k_wgrad<64, 64>), grid, dim3(kNT), 0, s, a);
    else if (tco == 64)
        hipLaunchKernelGGL((k_wgrad<64, 32>), grid, dim3(kNT), 0, s, a);
    else if (tci == 64)
        hipLaunchKernelGGL((k_wgrad<32, 64>), grid, dim3(kNT), 0, s, a);
    else
        hipLaunchKernelGGL((k_wgrad<32, 32>), grid, dim3(kNT), 0, s, a);
}

void launch_bias_grad(const uint16_t* dz, int cout, int P, int rows, int splits, float* part,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_bias_grad, dim3(splits), dim3(kNT), 0, s, dz, cout, P, rows, part);
}

void launch_reduce(const float* part, int splits, size_t n, float* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce, dim3(grid_for(n)), dim3(kNT), 0, s, part, splits, n, out);
}

void launch_pool2(const uint16_t* x, int B, int H, int W, int C, uint16_t* y, hipStream_t s)
{
    const size_t n = (size_t)B * (H / 2) * (W / 2) * (C / 8);
    hipLaunchKernelGGL(k_pool2, dim3(grid_for(n)), dim3(kNT), 0, s, x, B, H, W, C, y);
}

void launch_pool_back_add(const uint16_t* dskip, const uint16_t* dpool, const uint16_t* act,
                          int B, int H, int W, int C, uint16_t* out, hipStream_t s)
{
    const size_t n = (size_t)B * H * W * C;
    hipLaunchKernelGGL(k_pool_back_add, dim3(grid_for(n)), dim3(kNT), 0, s, dskip, dpool, act, B,
                       H, W, C, out);
}

void launch_up_back(const uint16_t* dup, const uint16_t* act, int B, int H, int W, int C,
                    uint16_t* out, hipStream_t s)
{
    const size_t n = (size_t)B * (H / 2) * (W / 2) * C;
    hipLaunchKernelGGL(k_up_back, dim3(grid_for(n)), dim3(kNT), 0, s, dup, act, B, H, W, C, out);
}

int launch_head(const HeadArgs& a, hipStream_t s)
{
    const int blocks = (a.P + kNT - 1) / kNT;
    if (a.C == 32)
        hipLaunchKernelGGL(k_head<32>, dim3(blocks), dim3(kNT), 0, s, a);
    else
        hipLaunchKernelGGL(k_head<64>, dim3(blocks), dim3(kNT), 0, s, a);
    return blocks;
}

void launch_head_wgrad(const uint16_t* h, const float* g, int C, int P, int rows, int splits,
                       float* part, hipStream_t s)
{
    hipLaunchKernelGGL(k_head_wgrad, dim3(splits), dim3(kNT), 0, s, h, g, C, P, rows, part);
}

void launch_encode(const NetCase* cases, const int* idx, int B, int H, int W, uint16_t* x,
                   hipStream_t s)
{
    hipLaunchKernelGGL(k_encode, dim3(grid_for((size_t)B * H * W)), dim3(kNT), 0, s, cases, idx, B,
                       H, W, x);
}

void launch_adam(float* w, const float* g, float* m, float* v, size_t n, float lr, float b1,
                 float b2, float eps, float bc1, float bc2, hipStream_t s)
{
    hipLaunchKernelGGL(k_adam, dim3(grid_for(n)), dim3(kNT), 0, s, w, g, m, v, n, lr, b1, b2, eps,
                       bc1, bc2);
}

void launch_pack_weights(const float* w, int cout, int cin, uint16_t* wb, uint16_t* wt,
                         hipStream_t s)
{
    hipLaunchKernelGGL(k_pack_weights, dim3(grid_for((size_t)cout * 9 * cin)), dim3(kNT), 0, s, w,
                       cout, cin, wb, wt);
}

void launch_to_bf16(const float* x, size_t n, uint16_t* y, hipStream_t s)
{
    hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(n)), dim3(kNT), 0, s, x, n, y);
}

void launch_from_bf16(const uint16_t* x, size_t n, float* y, hipStream_t s)
{
    hipLaunchKernelGGL(k_from_bf16, dim3(grid_for(n)), dim3(kNT), 0, s, x, n, y);
}

}  // namespace net
}  // namespace dsmc
