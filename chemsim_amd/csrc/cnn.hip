// Warm-start network: a 4-level convolutional encoder-decoder that maps an
// obstacle mask, the Reynolds number and position channels to the steady
// (rho, u) field of a channel flow.  Every FLOP runs in the kernels of
// cnn_kernels.hpp and the small element-wise kernels below (BF16 operands,
// FP32 accumulation, FP32 master weights with Adam).
//
// Layers (base channels c, input 8 channels, all convolutions 3x3, pad 1):
//   e1: 8->c->c          pool
//   e2: c->2c->2c        pool
//   e3: 2c->4c->4c       pool
//   b : 4c->8c->8c       up, concat e3
//   d3: 12c->4c->4c      up, concat e2
//   d2: 6c->2c->2c       up, concat e1
//   d1: 3c->c->c
//   head: c->8 (channels 0..2 used), no activation
// ReLU after every convolution except the head; 2x2 average pooling; nearest
// 2x up-sampling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "cnn_kernels.hpp"
#include "cs_internal.hpp"

namespace cs {
namespace cnn {

// ------------------------------------------------------- element-wise ----
__global__ void k_f32_to_bf16(const float* __restrict__ x, uint16_t* __restrict__ y, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = bf16_rne(x[i]);
}

__global__ void k_bf16_to_f32(const uint16_t* __restrict__ x, float* __restrict__ y, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = bf16_f32(x[i]);
}

// Weights [cout][kpad] f32 -> bf16 copy and the transposed copy used by the
// data gradient: wd[ci][tap*cout + co] = w[co][tap*cin + ci].
__global__ void k_prepare_weights(const float* __restrict__ w, uint16_t* __restrict__ wb, uint16_t* __restrict__ wd,
                                  int cin, int cout, int kpad, int kpad_d)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)cout * kpad) return;
    const int co = (int)(i / kpad), k = (int)(i % kpad);
    const uint16_t v = bf16_rne(w[i]);
    wb[i] = v;
    if (k < 9 * cin && wd) {
        const int tap = k / cin, ci = k - tap * cin;
        wd[(int64_t)ci * kpad_d + tap * cout + co] = v;
    }
}

// 2x2 average pooling (NHWC).
__global__ void k_pool(const uint16_t* __restrict__ x, uint16_t* __restrict__ y, int N, int H, int W, int C)
{
    const int Ho = H / 2, Wo = W / 2;
    const int64_t n_out = (int64_t)N * Ho * Wo * C;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_out) return;
    const int c = (int)(i % C);
    int64_t t = i / C;
    const int xo = (int)(t % Wo);
    t /= Wo;
    const int yo = (int)(t % Ho);
    const int64_t n = t / Ho;
    const int64_t b = ((n * H + 2 * yo) * W + 2 * xo) * C + c;
    const float s = (bf16_f32(x[b]) + bf16_f32(x[b + C])) + (bf16_f32(x[b + (int64_t)W * C]) +
                                                               bf16_f32(x[b + (int64_t)W * C + C]));
    y[i] = bf16_rne(0.25f * s);
}

// Pooling backward, fused with the next layer's ReLU mask: the gradient of
// the pooled tensor's input is the skip gradient accumulated so far plus
// dy / 4 from the pooled output; dz = bf16(that * (y > 0)) (relu = 0: no
// mask), y the input tensor (the conv output that was pooled).
__global__ void k_pool_back(const uint16_t* __restrict__ dy, const float* __restrict__ skip,
                            const uint16_t* __restrict__ y, int relu, uint16_t* __restrict__ dz, int N, int H, int W,
                            int C)
{
    const int64_t n_in = (int64_t)N * H * W * C;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_in) return;
    const int c = (int)(i % C);
    int64_t t = i / C;
    const int x = (int)(t % W);
    t /= W;
    const int yy = (int)(t % H);
    const int64_t n = t / H;
    const int Ho = H / 2, Wo = W / 2;
    float g = skip[i] + 0.25f * bf16_f32(dy[((n * Ho + yy / 2) * Wo + x / 2) * C + c]);
    if (relu && !(bf16_f32(y[i]) > 0.f)) g = 0.f;
    dz[i] = bf16_rne(g);
}

// Backward of a virtual concatenation (up-sampled coarse ++ skip) from the
// gradient dy of the concatenation: the coarse tensor's gradient is the sum
// over its 4 children of the first C1 channels, written as
// dz = bf16(sum * (y > 0)) with the coarse tensor y (the ReLU output that was
// up-sampled; relu = 0: no mask); the skip tensor's gradient (f32) is the
// last C2 channels, the first contribution to it.
__global__ void k_up_concat_back(const uint16_t* __restrict__ dy, const uint16_t* __restrict__ y, int relu,
                                 uint16_t* __restrict__ dz, float* __restrict__ ds, int N, int H, int W, int C1,
                                 int C2)
{
    const int Ct = C1 + C2;
    const int Hc = H / 2, Wc = W / 2;
    const int64_t n_c = (int64_t)N * Hc * Wc * C1;
    const int64_t n_s = (int64_t)N * H * W * C2;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_c) {
        const int c = (int)(i % C1);
        int64_t t = i / C1;
        const int xc = (int)(t % Wc);
        t /= Wc;
        const int yc = (int)(t % Hc);
        const int64_t n = t / Hc;
        float g = 0.f;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                g += bf16_f32(dy[((n * H + 2 * yc + a) * W + 2 * xc + b) * Ct + c]);
        if (relu && !(bf16_f32(y[i]) > 0.f)) g = 0.f;
        dz[i] = bf16_rne(g);
    }
    if (i < n_s) {
        const int c = (int)(i % C2);
        const int64_t p = i / C2;
        ds[i] = 0.f + bf16_f32(dy[p * Ct + C1 + c]);
    }
}

// Bias gradient, first stage: block b sums the rows [b*rows, (b+1)*rows) of
// dz [P][C] per channel into part[b][C] (8 <= C, C divides 256; fixed order).
// A thread reads 8 channels (16 bytes) of a pixel per iteration.  The second
// stage is part of k_reduce_grads.
__global__ void k_bias_partial(const uint16_t* __restrict__ dz, float* __restrict__ part, int64_t P, int C,
                               int64_t rows)
{
    __shared__ float sh[256][9];
    const int t = threadIdx.x, nv = C / 8, step = 256 / nv;
    const int v = t % nv, r0 = t / nv;
    const int64_t pb = (int64_t)blockIdx.x * rows, pe = min(P, pb + rows);
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int64_t p = pb + r0; p < pe; p += step) {
        const uint4 q = *(const uint4*)(dz + p * C + 8 * v);
        const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            s[2 * j] += __uint_as_float(w[j] << 16);
            s[2 * j + 1] += __uint_as_float(w[j] & 0xffff0000u);
        }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) sh[t][j] = s[j];
    __syncthreads();
    if (t < C) {
        const int vv = t / 8, jj = t % 8;
        float acc = 0.f;
        for (int k = 0; k < step; ++k) acc += sh[k * nv + vv][jj];
        part[(int64_t)blockIdx.x * C + t] = acc;
    }
}

// Weight and bias gradients from their partials, in a fixed order.  Blocks
// below nbw: the split-K partials of the weights -- a block takes 64
// outputs, 16 thread groups each sum every 16th split (coalesced rows of 64
// floats), and one thread per output adds the 16 group sums in group order.
// The next C blocks: the bias partials bpart[rows][C], one block per column,
// 256 threads striding the rows and a tree over them.
__global__ __launch_bounds__(1024) void k_reduce_grads(const float* __restrict__ part, float* __restrict__ out,
                                                       int64_t n, int splits, int nbw, const float* __restrict__ bpart,
                                                       float* __restrict__ bout, int rows, int C)
{
    __shared__ float sh[16][64];
    if ((int)blockIdx.x >= nbw) {
        float* sb = &sh[0][0];
        const int c = (int)blockIdx.x - nbw, t = threadIdx.x;
        if (t < 256) {
            float v = 0.f;
            for (int r = t; r < rows; r += 256) v += bpart[(int64_t)r * C + c];
            sb[t] = v;
        }
        __syncthreads();
        for (int k = 128; k > 0; k >>= 1) {
            if (t < k) sb[t] += sb[t + k];
            __syncthreads();
        }
        if (t == 0) bout[c] = sb[0];
        return;
    }
    const int o = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int64_t i = (int64_t)blockIdx.x * 64 + o;
    float v = 0.f;
    if (i < n)
        for (int z = g; z < splits; z += 16) v += part[(int64_t)z * n + i];
    sh[g][o] = v;
    __syncthreads();
    if (g == 0 && i < n) {
        float acc = 0.f;
        for (int k = 0; k < 16; ++k) acc += sh[k][o];
        out[i] = acc;
    }
}

// Weighted MSE over the first `used` channels of the f32 head output y
// ([P][8]): loss partials per block and the head's output gradient
// dz = bf16(2 w (y - t) / norm) (the head has no ReLU).
__global__ void k_mse(const float* __restrict__ y, const float* __restrict__ t, const float* __restrict__ w,
                      uint16_t* __restrict__ dz, float* __restrict__ partial, int64_t P, int used, float inv_norm)
{
    __shared__ float sh[256];
    float s = 0.f;
    for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < P; p += (int64_t)gridDim.x * 256) {
        const float wt = w[p];
        for (int c = 0; c < 8; ++c) {
            float g = 0.f;
            if (c < used) {
                const float d = y[p * 8 + c] - t[p * used + c];
                s += wt * d * d;
                g = 2.f * wt * d * inv_norm;
            }
            dz[p * 8 + c] = bf16_rne(g);
        }
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int64_t n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Gathers minibatch samples from the device dataset (inputs bf16 [S][HW][8],
// targets f32 [S][HW][3], weights f32 [S][HW]) by index list.
__global__ void k_gather_batch(const uint16_t* __restrict__ din, const float* __restrict__ dt,
                               const float* __restrict__ dw, const int* __restrict__ idx, int n, int64_t hw,
                               uint16_t* __restrict__ bin, float* __restrict__ bt, float* __restrict__ bw)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n * hw) return;
    const int b = (int)(i / hw);
    const int64_t p = i % hw;
    const int64_t s = (int64_t)idx[b] * hw + p;
    *(u32x4*)(bin + i * 8) = *(const u32x4*)(din + s * 8);
    for (int c = 0; c < 3; ++c) bt[i * 3 + c] = dt[s * 3 + c];
    bw[i] = dw[s];
}

}  // namespace cnn
}  // namespace cs

using namespace cs;
using namespace cs::cnn;

namespace {

inline int pad32(int k) { return (k + 31) / 32 * 32; }
inline unsigned blocks_for(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

struct Conv {
    int cin = 0, cout = 0, kpad = 0, kpad_d = 0;
    int dil = 1;  // tap spacing
    bool relu = true;
    float *w = nullptr, *b = nullptr, *gw = nullptr, *gb = nullptr;
    float *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;
    uint16_t *wb = nullptr, *wd = nullptr;
    int64_t nw() const { return (int64_t)cout * kpad; }
};

// NHWC bf16 activation and its shape.  A virtual concatenation (lo not null)
// is the nearest 2x up-sampling of lo (C1 channels at H/2 x W/2) followed by
// the C - C1 channels of d; the convolutions read it in place.
struct Act {
    uint16_t* d = nullptr;
    int N = 0, H = 0, W = 0, C = 0;
    const uint16_t* lo = nullptr;
    int C1 = 0;
    int64_t P() const { return (int64_t)N * H * W; }
    int64_t n() const { return P() * C; }
};

}  // namespace

struct cs_cnn {
    cs_ctx* ctx = nullptr;
    int H = 0, W = 0, base = 0, batch = 0;
    hipStream_t stream = nullptr;
    std::vector<Conv> convs;          // 19 convolutions in forward order
    std::vector<void*> allocs;
    // activations: inputs of every conv and every conv output
    std::vector<Act> conv_in, conv_out;
    Act x_in, p1, p2, p3, cat3, cat2, cat1;
    float* head_out = nullptr;        // [P][8] f32
    // scratch
    float* part = nullptr;
    int64_t part_elems = 0;
    float* bpart = nullptr;           // bias-gradient partials [<= 1024][cout]
    float* stage32 = nullptr;         // f32 staging of host inputs (act_max elements)
    int64_t act_max = 0;
    float* skip32[3] = {nullptr, nullptr, nullptr};  // accumulated gradients of e1, e2, e3 outputs
    uint16_t* dz = nullptr;           // bf16 gradients of a conv's output and input (largest activation each)
    uint16_t* dz2 = nullptr;
    float* loss_part = nullptr;
    float* targets = nullptr;
    float* weights = nullptr;
    int adam_t = 0;
    // device dataset
    uint16_t* ds_in = nullptr;
    float* ds_t = nullptr;
    float* ds_w = nullptr;
    int ds_count = 0;
    int* d_idx = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double gemm_flops = 0.0;          // FLOPs of the last forward/backward
};

namespace {

template <class T>
T* dalloc(cs_cnn* net, int64_t n)
{
    void* p = nullptr;
    CS_HIP(hipMalloc(&p, sizeof(T) * (size_t)std::max<int64_t>(n, 1)));
    net->allocs.push_back(p);
    return (T*)p;
}

void gemm(cs_cnn* net, int la, int lb, int epi, const uint16_t* A, int lda, const uint16_t* B, int ldb, void* C, int ldc,
          const float* bias, int M, int N, int K, int splits = 1, const Act* gather = nullptr, int sign = 1,
          int dil = 1, const uint16_t* mask = nullptr)
{
    GemmArgs g{};
    CS_REQUIRE(epi != EPI_BF16_MASK || mask, CS_ERR_ARG, "masked epilogue without a mask");
    g.mask = mask;
    if (lb == GATHER_T) {
        CS_REQUIRE(la != GATHER && gather && gather->C % 8 == 0 && (int64_t)K == gather->P() && dil >= 1, CS_ERR_ARG,
                   "bad gathered operand of the weight gradient");
        g.gH = gather->H;
        g.gW = gather->W;
        g.gC = gather->C;
        g.gsign = 1;
        g.gdil = dil;
        g.glo = gather->lo;
        g.gC1 = gather->C1;
    }
    if (la == GATHER) {
        CS_REQUIRE(gather && gather->C % 8 == 0 && (int64_t)M == gather->P() && dil >= 1, CS_ERR_ARG,
                   "bad gather operand");
        g.gH = gather->H;
        g.gW = gather->W;
        g.gC = gather->C;
        g.gsign = sign;
        g.gdil = dil;
        g.glo = gather->lo;
        g.gC1 = gather->C1;
        CS_REQUIRE(gather->lo == nullptr || (gather->C1 % 8 == 0 && gather->H % 2 == 0 && gather->W % 2 == 0),
                   CS_ERR_ARG, "bad virtual concatenation");
    }
    g.A = A;
    g.B = B;
    g.C = C;
    g.bias = bias;
    g.M = M;
    g.N = N;
    g.K = K;
    g.lda = lda;
    g.ldb = ldb;
    g.ldc = ldc;
    CS_REQUIRE(K % 8 == 0, CS_ERR_ARG, "gemm K must be a multiple of 8");
    g.k_per_split = ((K + splits - 1) / splits + BK - 1) / BK * BK;  // splits cover [0, K)
    g.split_stride = (int64_t)M * ldc;
    const int cfg = pick_tile(M, N);
    int bm, bn;
    tile_dims(cfg, bm, bn);
    dim3 grid((unsigned)((M + bm - 1) / bm), (unsigned)((N + bn - 1) / bn), (unsigned)splits);
    net->gemm_flops += 2.0 * M * N * K;
#define CS_GEMM(LA, LB, EPI) launch_gemm<LA, LB, EPI>(cfg, grid, net->stream, g)
    if (la == KMAJOR && lb == KMAJOR && epi == EPI_F32) {
        CS_GEMM(KMAJOR, KMAJOR, EPI_F32);
    } else if (la == PMAJOR_T && lb == GATHER_T && epi == EPI_F32) {
        CS_GEMM(PMAJOR_T, GATHER_T, EPI_F32);
    } else if (la == GATHER && lb == KMAJOR) {
        if (epi == EPI_F32) CS_GEMM(GATHER, KMAJOR, EPI_F32);
        else if (epi == EPI_BF16) CS_GEMM(GATHER, KMAJOR, EPI_BF16);
        else if (epi == EPI_BF16_MASK) CS_GEMM(GATHER, KMAJOR, EPI_BF16_MASK);
        else CS_GEMM(GATHER, KMAJOR, EPI_BF16_RELU);
    } else {
        throw Error{CS_ERR_ARG, "unsupported gemm layout"};
    }
#undef CS_GEMM
    CS_HIP(hipGetLastError());
}

// The halo kernel takes 3x3 convolutions with dilation 1 on images whose
// width is a multiple of 64 and height a multiple of 4.
bool halo_ok(const Act& x, int dil) { return dil == 1 && x.W % 64 == 0 && x.H % 4 == 0 && x.C % 8 == 0; }

void conv_halo(cs_cnn* net, int epi, const Act& x, const uint16_t* w, int ldw, const float* bias, void* out, int ldc,
               int N, int sign, const uint16_t* mask, int kflops)
{
    HaloArgs a{};
    a.in = x.d;
    a.glo = x.lo;
    a.C1 = x.C1;
    a.w = w;
    a.ldw = ldw;
    a.bias = bias;
    a.out = out;
    a.ldc = ldc;
    a.mask = mask;
    a.N = N;
    a.C = x.C;
    a.H = x.H;
    a.W = x.W;
    a.sign = sign;
    CS_REQUIRE(epi != EPI_BF16_MASK || mask, CS_ERR_ARG, "masked epilogue without a mask");
    const int bn = N <= 16 ? 16 : (N <= 32 ? 32 : 64);
    const dim3 grid((unsigned)((int64_t)x.N * (x.H / 4) * (x.W / 64)), (unsigned)((N + bn - 1) / bn));
    net->gemm_flops += 2.0 * (double)x.P() * N * kflops;
    auto go = [&]<int EPI>() {
        if (bn == 16) hipLaunchKernelGGL((k_conv_halo<EPI, 16>), grid, dim3(256), 0, net->stream, a);
        else if (bn == 32) hipLaunchKernelGGL((k_conv_halo<EPI, 32>), grid, dim3(256), 0, net->stream, a);
        else hipLaunchKernelGGL((k_conv_halo<EPI, 64>), grid, dim3(256), 0, net->stream, a);
    };
    switch (epi) {
    case EPI_F32: go.template operator()<EPI_F32>(); break;
    case EPI_BF16: go.template operator()<EPI_BF16>(); break;
    case EPI_BF16_RELU: go.template operator()<EPI_BF16_RELU>(); break;
    default: go.template operator()<EPI_BF16_MASK>(); break;
    }
    CS_HIP(hipGetLastError());
}

void conv_forward(cs_cnn* net, int li, const Act& x, const Act& y, float* y32 = nullptr)
{
    Conv& c = net->convs[(size_t)li];
    if (halo_ok(x, c.dil)) {
        if (y32) conv_halo(net, EPI_F32, x, c.wb, c.kpad, c.b, y32, 8, c.cout, 1, nullptr, c.kpad);
        else conv_halo(net, c.relu ? EPI_BF16_RELU : EPI_BF16, x, c.wb, c.kpad, c.b, y.d, c.cout, c.cout, 1, nullptr,
                       c.kpad);
        return;
    }
    if (y32)
        gemm(net, GATHER, KMAJOR, EPI_F32, x.d, 0, c.wb, c.kpad, y32, 8, c.b, (int)x.P(), c.cout, c.kpad, 1, &x, 1, c.dil);
    else
        gemm(net, GATHER, KMAJOR, c.relu ? EPI_BF16_RELU : EPI_BF16, x.d, 0, c.wb, c.kpad, y.d, c.cout, c.b,
             (int)x.P(), c.cout, c.kpad, 1, &x, 1, c.dil);
}

// Backward of conv li from the bf16 gradient of its output in dz (already
// masked by the layer's ReLU).  Writes dW, dB and, when dz_out is not null,
// the gradient of its input in bf16: masked by the input's own ReLU when
// mask_in (the input is the previous conv's ReLU output), plain otherwise
// (pool and concat outputs).  The bias partials go to bpart; one reduction
// launch then sums them and the weight gradient's split-K partials.
void conv_backward(cs_cnn* net, int li, const Act& x, uint16_t* dz, uint16_t* dz_out, bool mask_in)
{
    Conv& c = net->convs[(size_t)li];
    const int64_t P = x.P();
    CS_REQUIRE(dz != dz_out, CS_ERR_STATE, "gradient buffers alias");
    CS_REQUIRE(c.cout >= 8 && c.cout <= 256 && 256 % c.cout == 0 && P % 8 == 0, CS_ERR_STATE,
               "unsupported layer shape");
    const int64_t bblocks = std::min<int64_t>(1024, (P + 255) / 256);
    {
        const int64_t rows = (P + bblocks - 1) / bblocks;
        hipLaunchKernelGGL(k_bias_partial, dim3((unsigned)bblocks), dim3(256), 0, net->stream, dz, net->bpart, P,
                           c.cout, rows);
        CS_HIP(hipGetLastError());
    }
    const int nbw = (int)((c.nw() + 63) / 64);
    // weight gradient: dW[co][k] = sum_p dz[p][co] x[p + tap offset][ci],
    // k = tap*cin + ci, split over p.  Dilation 1: the LDS-tile kernel;
    // otherwise the GEMM, whose loaders read both operands from their NHWC
    // tensors and transpose them into LDS.
    if (halo_ok(x, c.dil) && c.cout % 8 == 0) {
        const int bm = c.cout <= 16 ? 16 : (c.cout <= 32 ? 32 : 64);
        const int cob = (c.cout + bm - 1) / bm, cib = (x.C + WG_CI - 1) / WG_CI;
        const int tiles = x.N * (x.H / 4) * (x.W / 64);
        int runs = std::max(1, std::min(tiles, 1024 / (cob * cib)));
        while (runs > 1 && (int64_t)runs * c.cout * c.kpad > net->part_elems) --runs;
        WgradArgs wa{dz, x.d, x.lo, x.C1, net->part, c.kpad, c.cout, x.C, x.H, x.W, tiles};
        const dim3 grid((unsigned)cob, (unsigned)cib, (unsigned)runs);
        if (bm == 16) hipLaunchKernelGGL((k_wgrad_halo<16>), grid, dim3(256), 0, net->stream, wa);
        else if (bm == 32) hipLaunchKernelGGL((k_wgrad_halo<32>), grid, dim3(256), 0, net->stream, wa);
        else hipLaunchKernelGGL((k_wgrad_halo<64>), grid, dim3(256), 0, net->stream, wa);
        CS_HIP(hipGetLastError());
        net->gemm_flops += 2.0 * c.cout * c.kpad * (double)P;
        hipLaunchKernelGGL(k_reduce_grads, dim3((unsigned)(nbw + c.cout)), dim3(1024), 0, net->stream, net->part, c.gw,
                           c.nw(), runs, nbw, net->bpart, c.gb, (int)bblocks, c.cout);
        CS_HIP(hipGetLastError());
    } else {
        // enough K ranges for ~1024 workgroups, each at least 1024 pixels long
        int bm, bn;
        tile_dims(pick_tile(c.cout, c.kpad), bm, bn);
        const int64_t tiles = (int64_t)((c.cout + bm - 1) / bm) * ((c.kpad + bn - 1) / bn);
        int splits = (int)std::max<int64_t>(1, std::min<int64_t>({256, P / 1024, (1024 + tiles - 1) / tiles}));
        while (splits > 1 && (int64_t)splits * c.cout * c.kpad > net->part_elems) --splits;
        gemm(net, PMAJOR_T, GATHER_T, EPI_F32, dz, c.cout, x.d, 0, net->part, c.kpad, nullptr, c.cout, c.kpad, (int)P,
             splits, &x, 1, c.dil);
        hipLaunchKernelGGL(k_reduce_grads, dim3((unsigned)(nbw + c.cout)), dim3(1024), 0, net->stream, net->part, c.gw,
                           c.nw(), splits, nbw, net->bpart, c.gb, (int)bblocks, c.cout);
        CS_HIP(hipGetLastError());
    }
    if (!dz_out) return;
    // data gradient: dx[p][ci] = sum_k colT(dz)[p][k] wd[ci][k], stored in
    // bf16 (masked by x's ReLU when mask_in) -- the conversion the next layer's
    // backward would make from f32, done in the epilogue
    Act dzact{dz, x.N, x.H, x.W, c.cout};
    if (halo_ok(dzact, c.dil)) {
        conv_halo(net, mask_in ? EPI_BF16_MASK : EPI_BF16, dzact, c.wd, c.kpad_d, nullptr, dz_out, c.cin, c.cin, -1,
                  mask_in ? x.d : nullptr, c.kpad_d);
        return;
    }
    gemm(net, GATHER, KMAJOR, mask_in ? EPI_BF16_MASK : EPI_BF16, dz, 0, c.wd, c.kpad_d, dz_out, c.cin, nullptr,
         (int)P, c.cin, c.kpad_d, 1, &dzact, -1, c.dil, mask_in ? x.d : nullptr);
}

void forward(cs_cnn* net, int n)
{
    net->gemm_flops = 0;
    auto resize = [&](Act& a, int H, int W) { a.N = n; a.H = H; a.W = W; };
    const int H = net->H, W = net->W;
    for (auto& a : net->conv_in) a.N = n;
    for (auto& a : net->conv_out) a.N = n;
    resize(net->x_in, H, W);
    resize(net->p1, H / 2, W / 2);
    resize(net->p2, H / 4, W / 4);
    resize(net->p3, H / 8, W / 8);
    resize(net->cat3, H / 4, W / 4);
    resize(net->cat2, H / 2, W / 2);
    resize(net->cat1, H, W);
    auto& ci = net->conv_in;
    auto& co = net->conv_out;
    hipStream_t s = net->stream;
    // encoder
    conv_forward(net, 0, ci[0], co[0]);
    conv_forward(net, 1, ci[1], co[1]);
    hipLaunchKernelGGL(k_pool, dim3(blocks_for(net->p1.n())), dim3(256), 0, s, co[1].d, net->p1.d, n, H, W, co[1].C);
    conv_forward(net, 2, ci[2], co[2]);
    conv_forward(net, 3, ci[3], co[3]);
    hipLaunchKernelGGL(k_pool, dim3(blocks_for(net->p2.n())), dim3(256), 0, s, co[3].d, net->p2.d, n, H / 2, W / 2,
                       co[3].C);
    conv_forward(net, 4, ci[4], co[4]);
    conv_forward(net, 5, ci[5], co[5]);
    hipLaunchKernelGGL(k_pool, dim3(blocks_for(net->p3.n())), dim3(256), 0, s, co[5].d, net->p3.d, n, H / 4, W / 4,
                       co[5].C);
    conv_forward(net, 6, ci[6], co[6]);
    conv_forward(net, 7, ci[7], co[7]);
    // decoder
    conv_forward(net, 8, ci[8], co[8]);
    conv_forward(net, 9, ci[9], co[9]);
    conv_forward(net, 10, ci[10], co[10]);
    conv_forward(net, 11, ci[11], co[11]);
    conv_forward(net, 12, ci[12], co[12]);
    conv_forward(net, 13, ci[13], co[13]);
    conv_forward(net, 14, ci[14], co[14], net->head_out);
    CS_HIP(hipGetLastError());
}

// Backward from the bf16 gradient of the head output (net->dz holds it on
// entry, written by k_mse).  Between consecutive convs the gradient stays in
// bf16 (two buffers used in turn); the skip gradients of the encoder outputs
// add up in f32 (the concat's share, then the pool's), and the kernels that
// finish a gradient write it in bf16, masked by the ReLU of the conv whose
// output it is.
void backward(cs_cnn* net, int n)
{
    auto& ci = net->conv_in;
    auto& co = net->conv_out;
    hipStream_t s = net->stream;
    const int H = net->H, W = net->W;
    uint16_t* zb[2] = {net->dz, net->dz2};
    int zc = 0;  // zb[zc]: bf16 gradient of the current conv's output
    // input gradient kinds: 0 none, 1 masked by the previous conv's ReLU, 2 plain
    auto step = [&](int li, int kind) {
        const bool mask = kind == 1 && net->convs[(size_t)li - 1].relu;
        conv_backward(net, li, ci[(size_t)li], zb[zc], kind ? zb[zc ^ 1] : nullptr, mask);
        zc ^= 1;
    };
    // zb[zc]: bf16 gradient of the concat of conv `li`'s up-sampled output and
    // a skip tensor: conv li's masked output gradient -> zb[zc ^ 1], the skip
    // part -> the skip accumulator
    auto concat_back = [&](int li, int C2, int Hc, int Wc, float* skip) {
        const int C1 = co[(size_t)li].C;
        const int64_t nmax = std::max<int64_t>((int64_t)n * (Hc / 2) * (Wc / 2) * C1, (int64_t)n * Hc * Wc * C2);
        hipLaunchKernelGGL(k_up_concat_back, dim3(blocks_for(nmax)), dim3(256), 0, s, zb[zc], co[(size_t)li].d,
                           net->convs[(size_t)li].relu ? 1 : 0, zb[zc ^ 1], skip, n, Hc, Wc, C1, C2);
        CS_HIP(hipGetLastError());
        zc ^= 1;
    };
    // zb[zc]: bf16 gradient of conv `li`'s pooled output; with the skip
    // accumulator, conv li's masked output gradient -> zb[zc ^ 1]
    auto pool_back = [&](int li, int Hc, int Wc, float* skip) {
        const int C = co[(size_t)li].C;
        hipLaunchKernelGGL(k_pool_back, dim3(blocks_for((int64_t)n * Hc * Wc * C)), dim3(256), 0, s, zb[zc], skip,
                           co[(size_t)li].d, net->convs[(size_t)li].relu ? 1 : 0, zb[zc ^ 1], n, Hc, Wc, C);
        CS_HIP(hipGetLastError());
        zc ^= 1;
    };
    // head and d1
    step(14, 1);
    step(13, 1);
    step(12, 2);  // gradient of cat1 (channels co[11].C + co[1].C)
    concat_back(11, co[1].C, H, W, net->skip32[0]);
    step(11, 1);
    step(10, 2);  // gradient of cat2
    concat_back(9, co[3].C, H / 2, W / 2, net->skip32[1]);
    step(9, 1);
    step(8, 2);   // gradient of cat3
    concat_back(7, co[5].C, H / 4, W / 4, net->skip32[2]);
    step(7, 1);
    step(6, 2);   // gradient of p3 -> with the e3 skip gradient, the gradient of co[5]
    pool_back(5, H / 4, W / 4, net->skip32[2]);
    step(5, 1);
    step(4, 2);   // gradient of p2
    pool_back(3, H / 2, W / 2, net->skip32[1]);
    step(3, 1);
    step(2, 2);   // gradient of p1
    pool_back(1, H, W, net->skip32[0]);
    step(1, 1);
    step(0, 0);
}

void prepare_weights(cs_cnn* net)
{
    for (auto& c : net->convs) {
        hipLaunchKernelGGL(k_prepare_weights, dim3(blocks_for(c.nw())), dim3(256), 0, net->stream, c.w, c.wb, c.wd,
                           c.cin, c.cout, c.kpad, c.kpad_d);
        CS_HIP(hipGetLastError());
    }
}

void check_net(const cs_cnn* n) { CS_REQUIRE(n != nullptr, CS_ERR_ARG, "null network"); }

void upload_inputs(cs_cnn* net, int n, const float* input)
{
    const int64_t cnt = (int64_t)n * net->H * net->W * 8;
    float* tmp = net->stage32;  // large enough f32 staging
    CS_HIP(hipMemcpyAsync(tmp, input, sizeof(float) * (size_t)cnt, hipMemcpyHostToDevice, net->stream));
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(blocks_for(cnt)), dim3(256), 0, net->stream, tmp, net->x_in.d, cnt);
    CS_HIP(hipGetLastError());
}

float loss_and_grad(cs_cnn* net, int n, const float* targets_dev, const float* weights_dev)
{
    const int64_t P = (int64_t)n * net->H * net->W;
    // loss = sum w (y - t)^2 / (3 P)
    const int blocks = 256;
    hipLaunchKernelGGL(k_mse, dim3(blocks), dim3(256), 0, net->stream, net->head_out, targets_dev, weights_dev,
                       net->dz, net->loss_part, P, 3, 1.0f / (float)(P * 3));
    CS_HIP(hipGetLastError());
    std::vector<float> part(blocks);
    CS_HIP(hipMemcpyAsync(part.data(), net->loss_part, sizeof(float) * blocks, hipMemcpyDeviceToHost, net->stream));
    backward(net, n);
    CS_HIP(hipStreamSynchronize(net->stream));
    double s = 0;
    for (float v : part) s += v;
    return (float)(s / (double)(P * 3));
}

void adam(cs_cnn* net, float lr, float b1, float b2, float eps)
{
    net->adam_t += 1;
    const float c1 = 1.f - std::pow(b1, (float)net->adam_t), c2 = 1.f - std::pow(b2, (float)net->adam_t);
    for (auto& c : net->convs) {
        hipLaunchKernelGGL(k_adam, dim3(blocks_for(c.nw())), dim3(256), 0, net->stream, c.w, c.gw, c.mw, c.vw, c.nw(),
                           lr, b1, b2, eps, c1, c2);
        hipLaunchKernelGGL(k_adam, dim3(blocks_for(c.cout)), dim3(256), 0, net->stream, c.b, c.gb, c.mb, c.vb,
                           (int64_t)c.cout, lr, b1, b2, eps, c1, c2);
    }
    CS_HIP(hipGetLastError());
    prepare_weights(net);
}

}  // namespace

// ======================================================================= ABI

extern "C" int cs_cnn_create(cs_ctx* ctx, int height, int width, int base, int batch, uint32_t seed, cs_cnn** out)
{
    return guarded([&] {
        CS_REQUIRE(ctx && out, CS_ERR_ARG, "null argument");
        CS_REQUIRE(height > 0 && width > 0 && height % 8 == 0 && width % 8 == 0, CS_ERR_ARG,
                   "height and width must be positive multiples of 8");
        CS_REQUIRE(base >= 8 && base % 8 == 0 && base <= 256, CS_ERR_ARG, "base channels must be a multiple of 8");
        CS_REQUIRE(batch >= 1 && batch <= 256, CS_ERR_ARG, "batch must be 1..256");
        CS_HIP(hipSetDevice(ctx->device));
        auto* net = new cs_cnn;
        try {
            net->ctx = ctx;
            net->H = height;
            net->W = width;
            net->base = base;
            net->batch = batch;
            CS_HIP(hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking));
            CS_HIP(hipEventCreate(&net->ev0));
            CS_HIP(hipEventCreate(&net->ev1));
            const int c = base;
            // (cin, cout, relu) in forward order
            const int spec[15][3] = {{8, c, 1},          {c, c, 1},          {c, 2 * c, 1},     {2 * c, 2 * c, 1},
                                     {2 * c, 4 * c, 1},  {4 * c, 4 * c, 1},  {4 * c, 8 * c, 1}, {8 * c, 8 * c, 1},
                                     {12 * c, 4 * c, 1}, {4 * c, 4 * c, 1},  {6 * c, 2 * c, 1}, {2 * c, 2 * c, 1},
                                     {3 * c, c, 1},      {c, c, 1},          {c, 8, 0}};
            // tap spacing: the last encoder level and the bottleneck reach
            // across the length of the channel
            const int dil[15] = {1, 1, 1, 1, 1, 2, 4, 8, 2, 1, 1, 1, 1, 1, 1};
            std::mt19937 rng(seed);
            for (int li = 0; li < 15; ++li) {
                const int* sp = spec[li];
                Conv cv;
                cv.dil = dil[li];
                cv.cin = sp[0];
                cv.cout = sp[1];
                cv.relu = sp[2] != 0;
                cv.kpad = pad32(9 * cv.cin);
                cv.kpad_d = pad32(9 * cv.cout);
                cv.w = dalloc<float>(net, cv.nw());
                cv.b = dalloc<float>(net, cv.cout);
                cv.gw = dalloc<float>(net, cv.nw());
                cv.gb = dalloc<float>(net, cv.cout);
                cv.mw = dalloc<float>(net, cv.nw());
                cv.vw = dalloc<float>(net, cv.nw());
                cv.mb = dalloc<float>(net, cv.cout);
                cv.vb = dalloc<float>(net, cv.cout);
                cv.wb = dalloc<uint16_t>(net, cv.nw());
                cv.wd = dalloc<uint16_t>(net, (int64_t)cv.cin * cv.kpad_d);
                CS_HIP(hipMemset(cv.wd, 0, sizeof(uint16_t) * (size_t)cv.cin * cv.kpad_d));
                for (float* p : {cv.mw, cv.vw}) CS_HIP(hipMemset(p, 0, sizeof(float) * (size_t)cv.nw()));
                for (float* p : {cv.mb, cv.vb, cv.b}) CS_HIP(hipMemset(p, 0, sizeof(float) * (size_t)cv.cout));
                // He initialisation over the 9*cin fan-in; padding columns zero.
                std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (9.f * cv.cin)));
                std::vector<float> hw((size_t)cv.nw(), 0.f);
                for (int o = 0; o < cv.cout; ++o)
                    for (int k = 0; k < 9 * cv.cin; ++k) hw[(size_t)o * cv.kpad + k] = nd(rng) * (cv.relu ? 1.f : 0.5f);
                CS_HIP(hipMemcpy(cv.w, hw.data(), sizeof(float) * hw.size(), hipMemcpyHostToDevice));
                net->convs.push_back(cv);
            }
            // activations at the maximum batch
            const int n = batch, H = height, W = width;
            auto act = [&](int h, int w, int ch) {
                Act a{dalloc<uint16_t>(net, (int64_t)n * h * w * ch), n, h, w, ch};
                return a;
            };
            net->x_in = act(H, W, 8);
            const int chans[15] = {c, c, 2 * c, 2 * c, 4 * c, 4 * c, 8 * c, 8 * c, 4 * c, 4 * c, 2 * c, 2 * c, c, c, 8};
            const int scale[15] = {1, 1, 2, 2, 4, 4, 8, 8, 4, 4, 2, 2, 1, 1, 1};
            for (int i = 0; i < 15; ++i) net->conv_out.push_back(act(H / scale[i], W / scale[i], chans[i]));
            net->p1 = act(H / 2, W / 2, c);
            net->p2 = act(H / 4, W / 4, 2 * c);
            net->p3 = act(H / 8, W / 8, 4 * c);
            // decoder inputs: up-sampled coarse output ++ encoder skip, read in place
            auto cat = [&](const Act& coarse, const Act& skip) {
                Act a{skip.d, n, skip.H, skip.W, coarse.C + skip.C};
                a.lo = coarse.d;
                a.C1 = coarse.C;
                return a;
            };
            net->cat3 = cat(net->conv_out[7], net->conv_out[5]);
            net->cat2 = cat(net->conv_out[9], net->conv_out[3]);
            net->cat1 = cat(net->conv_out[11], net->conv_out[1]);
            const Act ins[15] = {net->x_in,        net->conv_out[0], net->p1,          net->conv_out[2], net->p2,
                                 net->conv_out[4], net->p3,          net->conv_out[6], net->cat3,        net->conv_out[8],
                                 net->cat2,        net->conv_out[10], net->cat1,       net->conv_out[12], net->conv_out[13]};
            for (int i = 0; i < 15; ++i) net->conv_in.push_back(ins[i]);
            net->head_out = dalloc<float>(net, (int64_t)n * H * W * 8);
            // scratch: split-K partials of the weight gradient
            int64_t part = 0;
            for (int i = 0; i < 15; ++i) {
                const int64_t P = net->conv_in[(size_t)i].P();
                part = std::max(part, (int64_t)64 * net->convs[(size_t)i].nw());
                part = std::max(part, P * net->convs[(size_t)i].cin);
            }
            net->part = dalloc<float>(net, part);
            net->part_elems = part;
            net->bpart = dalloc<float>(net, (int64_t)1024 * 256);
            const int64_t amax = (int64_t)n * H * W * 96;  // >= every activation (3c <= 96 for c <= 32)
            const int64_t amax2 = std::max<int64_t>(amax, (int64_t)n * H * W * 3 * c);
            net->act_max = amax2;
            net->stage32 = dalloc<float>(net, amax2);
            net->dz = dalloc<uint16_t>(net, amax2);
            net->dz2 = dalloc<uint16_t>(net, amax2);
            net->skip32[0] = dalloc<float>(net, net->conv_out[1].n());
            net->skip32[1] = dalloc<float>(net, net->conv_out[3].n());
            net->skip32[2] = dalloc<float>(net, net->conv_out[5].n());
            net->loss_part = dalloc<float>(net, 256);
            net->targets = dalloc<float>(net, (int64_t)n * H * W * 3);
            net->weights = dalloc<float>(net, (int64_t)n * H * W);
            net->d_idx = dalloc<int>(net, n);
            prepare_weights(net);
            CS_HIP(hipStreamSynchronize(net->stream));
        } catch (...) {
            cs_cnn_destroy(net);
            throw;
        }
        *out = net;
    });
}

extern "C" int cs_cnn_destroy(cs_cnn* net)
{
    return guarded([&] {
        if (!net) return;
        if (net->stream) (void)hipStreamSynchronize(net->stream);
        for (void* p : net->allocs) (void)hipFree(p);
        for (void* p : {(void*)net->ds_in, (void*)net->ds_t, (void*)net->ds_w})
            if (p) (void)hipFree(p);
        if (net->ev0) (void)hipEventDestroy(net->ev0);
        if (net->ev1) (void)hipEventDestroy(net->ev1);
        if (net->stream) (void)hipStreamDestroy(net->stream);
        delete net;
    });
}

// Parameters in a fixed order: for each convolution, weights [cout][9*cin]
// (tap-major, ci fastest) then bias [cout].
static int64_t param_count(const cs_cnn* net)
{
    int64_t n = 0;
    for (const auto& c : net->convs) n += (int64_t)c.cout * 9 * c.cin + c.cout;
    return n;
}

extern "C" int cs_cnn_param_count(const cs_cnn* net, int64_t* n)
{
    return guarded([&] {
        check_net(net);
        CS_REQUIRE(n, CS_ERR_ARG, "null output");
        *n = param_count(net);
    });
}

static void copy_params(cs_cnn* net, float* host, bool to_host, bool grads)
{
    int64_t off = 0;
    for (auto& c : net->convs) {
        float* w = grads ? c.gw : c.w;
        float* b = grads ? c.gb : c.b;
        const int k9 = 9 * c.cin;
        if (to_host) {
            std::vector<float> tmp((size_t)c.nw());
            CS_HIP(hipMemcpy(tmp.data(), w, sizeof(float) * tmp.size(), hipMemcpyDeviceToHost));
            for (int o = 0; o < c.cout; ++o)
                std::memcpy(host + off + (int64_t)o * k9, tmp.data() + (size_t)o * c.kpad, sizeof(float) * k9);
            CS_HIP(hipMemcpy(host + off + (int64_t)c.cout * k9, b, sizeof(float) * c.cout, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> tmp((size_t)c.nw(), 0.f);
            for (int o = 0; o < c.cout; ++o)
                std::memcpy(tmp.data() + (size_t)o * c.kpad, host + off + (int64_t)o * k9, sizeof(float) * k9);
            CS_HIP(hipMemcpy(w, tmp.data(), sizeof(float) * tmp.size(), hipMemcpyHostToDevice));
            CS_HIP(hipMemcpy(b, host + off + (int64_t)c.cout * k9, sizeof(float) * c.cout, hipMemcpyHostToDevice));
        }
        off += (int64_t)c.cout * k9 + c.cout;
    }
}

extern "C" int cs_cnn_get_params(cs_cnn* net, float* out)
{
    return guarded([&] {
        check_net(net);
        CS_REQUIRE(out, CS_ERR_ARG, "null output");
        CS_HIP(hipStreamSynchronize(net->stream));
        copy_params(net, out, true, false);
    });
}

extern "C" int cs_cnn_set_params(cs_cnn* net, const float* in)
{
    return guarded([&] {
        check_net(net);
        CS_REQUIRE(in, CS_ERR_ARG, "null input");
        CS_HIP(hipStreamSynchronize(net->stream));
        copy_params(net, const_cast<float*>(in), false, false);
        prepare_weights(net);
        CS_HIP(hipStreamSynchronize(net->stream));
    });
}

extern "C" int cs_cnn_get_grads(cs_cnn* net, float* out)
{
    return guarded([&] {
        check_net(net);
        CS_REQUIRE(out, CS_ERR_ARG, "null output");
        CS_HIP(hipStreamSynchronize(net->stream));
        copy_params(net, out, true, true);
    });
}

// input: [n][H][W][8] f32; output: [n][H][W][3] f32.
extern "C" int cs_cnn_forward(cs_cnn* net, int n, const float* input, float* output)
{
    return guarded([&] {
        check_net(net);
        CS_REQUIRE(n >= 1 && n <= net->batch && input && output, CS_ERR_ARG, "bad argument");
        upload_inputs(net, n, input);
        forward(net, n);
        const int64_t P = (int64_t)n * net->H * net->W;
        std::vector<float> tmp((size_t)(P * 8));
        CS_HIP(hipMemcpyAsync(tmp.data(), net->head_out, sizeof(float) * tmp.size(), hipMemcpyDeviceToHost,
                              net->stream));
        CS_HIP(hipStreamSynchronize(net->stream));
        for (int64_t p = 0; p < P; ++p)
            for (int c = 0; c < 3; ++c) output[p * 3 + c] = tmp[(size_t)(p * 8 + c)];
    });
}

// Forward + backward on host data: loss = sum w (y - t)^2 / (3 n H W);
// gradients are left in the network (cs_cnn_get_grads, cs_cnn_adam_step).
extern "C" int cs_cnn_loss_grad(cs_cnn* net, int n, const float* input, const float* target, const float* weight,
                                float* loss)
{
    return guarded([&] {
        check_net(net);
        CS_REQUIRE(n >= 1 && n <= net->batch && input && target && weight && loss, CS_ERR_ARG, "bad argument");
        upload_inputs(net, n, input);
        const int64_t P = (int64_t)n * net->H * net->W;
        CS_HIP(hipMemcpyAsync(net->targets, target, sizeof(float) * (size_t)(P * 3), hipMemcpyHostToDevice,
                              net->stream));
        CS_HIP(hipMemcpyAsync(net->weights, weight, sizeof(float) * (size_t)P, hipMemcpyHostToDevice, net->stream));
        forward(net, n);
        *loss = loss_and_grad(net, n, net->targets, net->weights);
    });
}

extern "C" int cs_cnn_adam_step(cs_cnn* net, float lr, float beta1, float beta2, float eps)
{
    return guarded([&] {
        check_net(net);
        adam(net, lr, beta1, beta2, eps);
        CS_HIP(hipStreamSynchronize(net->stream));
    });
}

extern "C" int cs_cnn_set_dataset(cs_cnn* net, int count, const float* inputs, const float* targets,
                                  const float* weights)
{
    return guarded([&] {
        check_net(net);
        CS_REQUIRE(count >= 1 && inputs && targets && weights, CS_ERR_ARG, "bad argument");
        const int64_t hw = (int64_t)net->H * net->W;
        for (void* p : {(void*)net->ds_in, (void*)net->ds_t, (void*)net->ds_w})
            if (p) (void)hipFree(p);
        net->ds_in = nullptr;
        net->ds_t = nullptr;
        net->ds_w = nullptr;
        CS_HIP(hipMalloc(&net->ds_in, sizeof(uint16_t) * (size_t)(count * hw * 8)));
        CS_HIP(hipMalloc(&net->ds_t, sizeof(float) * (size_t)(count * hw * 3)));
        CS_HIP(hipMalloc(&net->ds_w, sizeof(float) * (size_t)(count * hw)));
        // inputs converted on the device, one sample at a time through stage32
        for (int s = 0; s < count; ++s) {
            CS_HIP(hipMemcpyAsync(net->stage32, inputs + (int64_t)s * hw * 8, sizeof(float) * (size_t)(hw * 8),
                                  hipMemcpyHostToDevice, net->stream));
            hipLaunchKernelGGL(k_f32_to_bf16, dim3(blocks_for(hw * 8)), dim3(256), 0, net->stream, net->stage32,
                               net->ds_in + (int64_t)s * hw * 8, hw * 8);
            CS_HIP(hipGetLastError());
            CS_HIP(hipStreamSynchronize(net->stream));
        }
        CS_HIP(hipMemcpy(net->ds_t, targets, sizeof(float) * (size_t)(count * hw * 3), hipMemcpyHostToDevice));
        CS_HIP(hipMemcpy(net->ds_w, weights, sizeof(float) * (size_t)(count * hw), hipMemcpyHostToDevice));
        net->ds_count = count;
    });
}

// One pass over `count` dataset samples in the given order, minibatches of
// the network's batch size, Adam after each; returns the mean loss.
extern "C" int cs_cnn_train_epoch(cs_cnn* net, const int32_t* order, int count, float lr, float* mean_loss)
{
    return guarded([&] {
        check_net(net);
        CS_REQUIRE(order && mean_loss && count >= 1 && net->ds_count > 0, CS_ERR_ARG, "bad argument");
        for (int i = 0; i < count; ++i)
            CS_REQUIRE(order[i] >= 0 && order[i] < net->ds_count, CS_ERR_ARG, "sample index out of range");
        const int64_t hw = (int64_t)net->H * net->W;
        double total = 0;
        int batches = 0;
        for (int s = 0; s + net->batch <= count; s += net->batch) {
            const int n = net->batch;
            CS_HIP(hipMemcpyAsync(net->d_idx, order + s, sizeof(int) * n, hipMemcpyHostToDevice, net->stream));
            hipLaunchKernelGGL(k_gather_batch, dim3(blocks_for(n * hw)), dim3(256), 0, net->stream, net->ds_in,
                               net->ds_t, net->ds_w, net->d_idx, n, hw, net->x_in.d, net->targets, net->weights);
            CS_HIP(hipGetLastError());
            forward(net, n);
            total += loss_and_grad(net, n, net->targets, net->weights);
            adam(net, lr, 0.9f, 0.999f, 1e-8f);
            batches += 1;
        }
        CS_HIP(hipStreamSynchronize(net->stream));
        *mean_loss = batches ? (float)(total / batches) : 0.f;
    });
}

// Times `iters` forward+backward passes at the full batch on the current
// input buffers; returns ms per pass and the GEMM FLOPs of one pass.
extern "C" int cs_cnn_time_step(cs_cnn* net, int iters, float* ms, double* flops)
{
    return guarded([&] {
        check_net(net);
        CS_REQUIRE(iters >= 1 && ms && flops, CS_ERR_ARG, "bad argument");
        const int n = net->batch;
        CS_HIP(hipMemsetAsync(net->weights, 0, sizeof(float) * (size_t)n * net->H * net->W, net->stream));
        CS_HIP(hipMemsetAsync(net->targets, 0, sizeof(float) * (size_t)n * net->H * net->W * 3, net->stream));
        CS_HIP(hipEventRecord(net->ev0, net->stream));
        double f = 0;
        for (int i = 0; i < iters; ++i) {
            forward(net, n);
            f = net->gemm_flops;
            hipLaunchKernelGGL(k_mse, dim3(256), dim3(256), 0, net->stream, net->head_out, net->targets, net->weights,
                               net->dz, net->loss_part, (int64_t)n * net->H * net->W, 3, 1.0f);
            net->gemm_flops = 0;
            backward(net, n);
            f += net->gemm_flops;
        }
        CS_HIP(hipEventRecord(net->ev1, net->stream));
        CS_HIP(hipEventSynchronize(net->ev1));
        float t = 0;
        CS_HIP(hipEventElapsedTime(&t, net->ev0, net->ev1));
        *ms = t / iters;
        *flops = f;
    });
}

// Plain BF16 GEMM C[M][N] (f32) = A[M][K] B[N][K]^T, both K-contiguous, K a
// multiple of 32; exposes the MFMA tile kernel for tests and timing.
extern "C" int cs_gemm_bf16(cs_ctx* ctx, int M, int N, int K, const uint16_t* A, const uint16_t* B, float* C,
                            float* ms)
{
    return guarded([&] {
        CS_REQUIRE(ctx && A && B && C && M > 0 && N > 0 && K > 0 && K % 32 == 0, CS_ERR_ARG, "bad argument");
        CS_HIP(hipSetDevice(ctx->device));
        void *da = nullptr, *db = nullptr, *dc = nullptr;
        hipStream_t s;
        CS_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        auto cleanup = [&] {
            for (void* p : {da, db, dc})
                if (p) (void)hipFree(p);
            (void)hipStreamDestroy(s);
        };
        try {
            CS_HIP(hipMalloc(&da, sizeof(uint16_t) * (size_t)M * K));
            CS_HIP(hipMalloc(&db, sizeof(uint16_t) * (size_t)N * K));
            CS_HIP(hipMalloc(&dc, sizeof(float) * (size_t)M * N));
            CS_HIP(hipMemcpy(da, A, sizeof(uint16_t) * (size_t)M * K, hipMemcpyHostToDevice));
            CS_HIP(hipMemcpy(db, B, sizeof(uint16_t) * (size_t)N * K, hipMemcpyHostToDevice));
            GemmArgs g{};
            g.A = (const uint16_t*)da;
            g.B = (const uint16_t*)db;
            g.C = dc;
            g.bias = nullptr;
            g.M = M;
            g.N = N;
            g.K = K;
            g.lda = K;
            g.ldb = K;
            g.ldc = N;
            g.k_per_split = K;
            g.split_stride = 0;
            const int cfg = pick_tile(M, N);
            int bm, bn;
            tile_dims(cfg, bm, bn);
            dim3 grid((unsigned)((M + bm - 1) / bm), (unsigned)((N + bn - 1) / bn), 1);
            hipEvent_t e0, e1;
            CS_HIP(hipEventCreate(&e0));
            CS_HIP(hipEventCreate(&e1));
            launch_gemm<KMAJOR, KMAJOR, EPI_F32>(cfg, grid, s, g);
            CS_HIP(hipEventRecord(e0, s));
            const int iters = 10;
            for (int i = 0; i < iters; ++i) launch_gemm<KMAJOR, KMAJOR, EPI_F32>(cfg, grid, s, g);
            CS_HIP(hipEventRecord(e1, s));
            CS_HIP(hipEventSynchronize(e1));
            float t = 0;
            CS_HIP(hipEventElapsedTime(&t, e0, e1));
            if (ms) *ms = t / iters;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            CS_HIP(hipMemcpy(C, dc, sizeof(float) * (size_t)M * N, hipMemcpyDeviceToHost));
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
    });
}
