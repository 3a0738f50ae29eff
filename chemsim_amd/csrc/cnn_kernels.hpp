// BF16 MFMA kernels of the warm-start network (gfx950).
//
// Activations are NHWC bf16 (channels contiguous), weights are [Cout][K] with
// K = 9*Cin ordered (tap, ci), tap = 3*(dy+1) + (dx+1).  A 3x3 convolution is
// an implicit GEMM; here the gather is explicit (im2col into a [pixels][Kpad]
// bf16 buffer) and one GEMM kernel serves the forward pass, the data gradient
// and the weight gradient:
//   forward  out[p][co]  = sum_k col(in)[p][k]  * W[co][k]            (NT)
//   dgrad    din[p][ci]  = sum_k colT(dout)[p][k] * Wd[ci][k]         (NT)
//   wgrad    dW[co][k]   = sum_p dout[p][co] * col(in)[p][k]          (TN, split over p)
// GEMM tiles are 64x64x32 in LDS, 4 waves of 64 lanes, each wave a 32x32
// quarter as 2x2 v_mfma_f32_16x16x32_bf16 with FP32 accumulation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cs {
namespace cnn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t bf16_rne(float x)
{
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// Operand layouts.  KMAJOR: element (r, k) at base[r*ld + k] (k contiguous).
// RMAJOR: element (r, k) at base[k*ld + r] (the row index contiguous).
enum Layout : int { KMAJOR = 0, RMAJOR = 1 };

enum Epi : int {
    EPI_F32 = 0,        // C (+ bias) stored as f32
    EPI_BF16 = 1,       // C (+ bias) stored as bf16
    EPI_BF16_RELU = 2   // max(C + bias, 0) stored as bf16
};

struct GemmArgs {
    const uint16_t* A;  // rows M
    const uint16_t* B;  // rows N
    void* C;            // [M][ldc] (or [splits][M][ldc] partials for EPI_F32 with split-K)
    const float* bias;  // [N] or null
    int M, N, K;        // K: multiple of 32 (operands zero-padded)
    int lda, ldb, ldc;
    int k_per_split;    // K range of one blockIdx.z (multiple of 32)
    int64_t split_stride;  // elements between the partial outputs of consecutive z
};

constexpr int BM = 64, BN = 64, BK = 32;
constexpr int LDS_PITCH = BK + 8;  // bf16 elements per LDS row (padding against bank conflicts)

// Loads one 64x32 tile of an operand into LDS as [row][k].
template <int L>
__device__ __forceinline__ void load_tile(const uint16_t* __restrict__ src, int ld, int rows, int r0, int k0,
                                          uint16_t (*lds)[LDS_PITCH])
{
    const int t = threadIdx.x;
    if (L == KMAJOR) {
        const int r = t >> 2, kg = t & 3;  // 64 rows x 4 groups of 8 k
        u32x4 v = {0u, 0u, 0u, 0u};
        if (r0 + r < rows) v = *(const u32x4*)(src + (int64_t)(r0 + r) * ld + k0 + 8 * kg);
        *(u32x4*)(&lds[r][8 * kg]) = v;
    } else {
        const int kk = t >> 3, rg = t & 7;  // 32 k x 8 groups of 8 rows
        u32x4 v = {0u, 0u, 0u, 0u};
        const int r = r0 + 8 * rg;
        if (r + 7 < rows) {
            v = *(const u32x4*)(src + (int64_t)(k0 + kk) * ld + r);
        } else {
            uint16_t tmp[8];
            for (int j = 0; j < 8; ++j) tmp[j] = (r + j < rows) ? src[(int64_t)(k0 + kk) * ld + r + j] : (uint16_t)0;
            v = *(const u32x4*)tmp;
        }
        const uint16_t* e = (const uint16_t*)&v;
#pragma unroll
        for (int j = 0; j < 8; ++j) lds[8 * rg + j][kk] = e[j];
    }
}

template <int LA, int LB, int EPI>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g)
{
    __shared__ __align__(16) uint16_t As[BM][LDS_PITCH];
    __shared__ __align__(16) uint16_t Bs[BN][LDS_PITCH];
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int kb = blockIdx.z * g.k_per_split;
    const int ke = min(g.K, kb + g.k_per_split);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int fr = lane & 15, fk = 8 * (lane >> 4);
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int k0 = kb; k0 < ke; k0 += BK) {
        load_tile<LA>(g.A, g.lda, g.M, m0, k0, As);
        load_tile<LB>(g.B, g.ldb, g.N, n0, k0, Bs);
        __syncthreads();
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = *(const bf16x8*)(&As[wm + 16 * i + fr][fk]);
#pragma unroll
        for (int j = 0; j < 2; ++j) b[j] = *(const bf16x8*)(&Bs[wn + 16 * j + fr][fk]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }

    // D layout: lane holds rows 4*(lane/16) + 0..3 of column lane % 16.
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = n0 + wn + 16 * j + fr;
            if (col >= g.N) continue;
            const float bias = g.bias ? g.bias[col] : 0.f;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = m0 + wm + 16 * i + 4 * (lane >> 4) + q;
                if (row >= g.M) continue;
                float v = acc[i][j][q] + bias;
                if (EPI == EPI_F32) {
                    ((float*)g.C)[(int64_t)blockIdx.z * g.split_stride + (int64_t)row * g.ldc + col] = v;
                } else {
                    if (EPI == EPI_BF16_RELU) v = v > 0.f ? v : 0.f;
                    ((uint16_t*)g.C)[(int64_t)row * g.ldc + col] = bf16_rne(v);
                }
            }
        }
}

// im2col of an NHWC bf16 tensor: col[p][k], k = tap*C + c < 9C, zero beyond
// (up to Kpad) and outside the image.  sign = +1 gathers in[y+dy][x+dx]
// (forward), -1 gathers in[y-dy][x-dx] (data gradient).  One thread writes 8
// consecutive k (C is a multiple of 8).
__global__ void k_im2col(const uint16_t* __restrict__ in, uint16_t* __restrict__ col, int N, int H, int W, int C,
                         int Kpad, int sign)
{
    const int64_t groups = (int64_t)N * H * W * (Kpad / 8);
    const int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= groups) return;
    const int kg = (int)(gi % (Kpad / 8));
    const int64_t p = gi / (Kpad / 8);
    const int k = 8 * kg;
    u32x4 v = {0u, 0u, 0u, 0u};
    if (k < 9 * C) {
        const int tap = k / C, c = k - tap * C;
        const int dy = sign * (tap / 3 - 1), dx = sign * (tap % 3 - 1);
        const int x = (int)(p % W);
        const int64_t ny = p / W;
        const int y = (int)(ny % H);
        const int64_t n = ny / H;
        const int yy = y + dy, xx = x + dx;
        if (yy >= 0 && yy < H && xx >= 0 && xx < W)
            v = *(const u32x4*)(in + ((n * H + yy) * W + xx) * C + c);
    }
    *(u32x4*)(col + p * Kpad + k) = v;
}

}  // namespace cnn
}  // namespace cs
