// BF16 MFMA kernels of the warm-start network (gfx950).
//
// Activations are NHWC bf16 (channels contiguous), weights are [Cout][K] with
// K = 9*Cin ordered (tap, ci), tap = 3*(dy+1) + (dx+1).  A 3x3 convolution is
// a GEMM over the gathered input col(in)[p][k]; one GEMM kernel serves the
// forward pass, the data gradient and the weight gradient:
//   forward  out[p][co] = sum_k col(in)[p][k] * W[co][k]    gather inside the GEMM
//   dgrad    din[p][ci] = sum_k colT(dout)[p][k] * Wd[ci][k] gather (flipped taps) inside the GEMM
//   wgrad    dW[co][k]  = sum_p doutT[co][p] * colT[k][p]   doutT transposed in memory, colT read
//                                                          from three column-shifted copies, split over p
// GEMM tiles (BM x BN x 32, chosen by shape) are double-buffered in LDS;
// 4 waves of 64 lanes each own a sub-tile of v_mfma_f32_16x16x32_bf16
// fragments with FP32 accumulation.
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
// GATHER: the A operand of a 3x3 convolution (taps `dil` cells apart) read
// straight from the NHWC input (implicit im2col): row r = pixel,
// k = tap*C + c, zero outside the image and for k >= 9C.
// SHIFT: the B operand of the weight gradient, colT[k][p] with k = tap*C + c,
// read from three column-shifted transposed copies of the input,
// X3[s][c][p] = in[p + (s - 1) dil][c] (zero across the row ends), at the
// row offset dy*W of the tap; zero outside the image and for k >= 9C.
enum Layout : int { KMAJOR = 0, GATHER = 2, SHIFT = 3 };

enum Epi : int {
    EPI_F32 = 0,        // C (+ bias) stored as f32
    EPI_BF16 = 1,       // C (+ bias) stored as bf16
    EPI_BF16_RELU = 2,  // max(C + bias, 0) stored as bf16
    EPI_BF16_MASK = 3   // C (+ bias) where mask > 0, else 0, stored as bf16 (ReLU backward)
};

struct GemmArgs {
    const uint16_t* A;  // rows M
    const uint16_t* B;  // rows N
    void* C;            // [M][ldc] (or [splits][M][ldc] partials for EPI_F32 with split-K)
    const float* bias;  // [N] or null
    int M, N, K;        // K: multiple of 8; a K range's tail beyond K (or the split's end) reads as zero
    int lda, ldb, ldc;
    int k_per_split;    // K range of one blockIdx.z (multiple of 32)
    int64_t split_stride;  // elements between the partial outputs of consecutive z
    int gH, gW, gC, gsign;  // GATHER: input height, width, channels (multiple of 8), tap sign (+1 / -1)
    int gdil;               // GATHER: tap spacing (dilation)
    const uint16_t* mask;   // EPI_BF16_MASK: [M][ldc] bf16 ReLU output whose zeros mask C
};

constexpr int BK = 32;
constexpr int LDS_PITCH = BK + 8;  // bf16 elements per LDS row (padding against bank conflicts)

// One BK-wide slice of an R-row operand tile: global -> registers (NV 16-byte
// vectors per thread, zero outside `rows`), registers -> LDS as [row][k].
template <int L, int R>
struct TileLoader {
    static_assert(L == KMAJOR, "K-contiguous operand");
    static constexpr int VECS = 4 * R;                  // R rows x BK/8 vectors
    static constexpr int NV = (VECS + 255) / 256;
    u32x4 v[NV];

    // k >= kend reads as zero (kend: end of this block's K range, a multiple of 8)
    __device__ __forceinline__ void load(const uint16_t* __restrict__ src, int ld, int rows, int r0, int k0, int kend)
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int e = t + 256 * i;
            v[i] = u32x4{0u, 0u, 0u, 0u};
            if (VECS % 256 != 0 && e >= VECS) continue;
            const int r = e >> 2, kg = e & 3;
            if (r0 + r < rows && k0 + 8 * kg < kend)
                v[i] = *(const u32x4*)(src + (int64_t)(r0 + r) * ld + k0 + 8 * kg);
        }
    }
    __device__ __forceinline__ void store(uint16_t (*lds)[LDS_PITCH]) const
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int e = t + 256 * i;
            if (VECS % 256 != 0 && e >= VECS) continue;
            *(u32x4*)(&lds[e >> 2][8 * (e & 3)]) = v[i];
        }
    }
};

// The GATHER operand: each vector slot's pixel is fixed for the whole K loop,
// so its coordinates are computed once; a K slice then only decodes the tap.
template <int R>
struct GatherLoader {
    static constexpr int VECS = 4 * R;
    static constexpr int NV = (VECS + 255) / 256;
    u32x4 v[NV];
    int y[NV], x[NV];
    int64_t base[NV];  // element offset of channel 0 of the slot's pixel

    __device__ __forceinline__ void init(const GemmArgs& g, int r0)
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int e = t + 256 * i;
            const int p = r0 + (e >> 2);
            y[i] = -(1 << 28);  // marks an unused slot: every tap falls outside
            x[i] = 0;
            base[i] = 0;
            if ((VECS % 256 == 0 || e < VECS) && p < g.M) {
                x[i] = p % g.gW;
                const int ny = p / g.gW;
                y[i] = ny % g.gH;
                base[i] = (int64_t)p * g.gC;
            }
        }
    }
    __device__ __forceinline__ void load(const uint16_t* __restrict__ in, const GemmArgs& g, int k0, int kend)
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int e = t + 256 * i;
            const int k = k0 + 8 * (e & 3);
            v[i] = u32x4{0u, 0u, 0u, 0u};
            if (k < kend && k < 9 * g.gC) {
                const int tap = k / g.gC, c = k - tap * g.gC;
                const int dy = g.gsign * g.gdil * (tap / 3 - 1), dx = g.gsign * g.gdil * (tap % 3 - 1);
                const int yy = y[i] + dy, xx = x[i] + dx;
                if (yy >= 0 && yy < g.gH && xx >= 0 && xx < g.gW)
                    v[i] = *(const u32x4*)(in + base[i] + (int64_t)(dy * g.gW + dx) * g.gC + c);
            }
        }
    }
    __device__ __forceinline__ void store(uint16_t (*lds)[LDS_PITCH]) const
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int e = t + 256 * i;
            if (VECS % 256 != 0 && e >= VECS) continue;
            *(u32x4*)(&lds[e >> 2][8 * (e & 3)]) = v[i];
        }
    }
};

// The SHIFT operand (rows n of colT, K = pixels): a vector of 8 pixels lies in
// one image row when W is a multiple of 8 (one validity test, one 16-byte
// load); otherwise the pixels are tested and loaded one by one.
template <int R>
struct ShiftLoader {
    static constexpr int VECS = 4 * R;
    static constexpr int NV = (VECS + 255) / 256;
    u32x4 v[NV];

    __device__ __forceinline__ void load(const uint16_t* __restrict__ x3, const GemmArgs& g, int n0, int k0, int kend)
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int e = t + 256 * i;
            v[i] = u32x4{0u, 0u, 0u, 0u};
            if (VECS % 256 != 0 && e >= VECS) continue;
            const int n = n0 + (e >> 2), p0 = k0 + 8 * (e & 3);
            if (n >= g.N || n >= 9 * g.gC || p0 >= kend) continue;
            const int tap = n / g.gC, c = n - tap * g.gC;
            const int dy = g.gdil * (tap / 3 - 1);
            const uint16_t* row = x3 + ((int64_t)(tap % 3) * g.gC + c) * g.K;
            if (g.gW % 8 == 0) {
                const int yy = (p0 / g.gW) % g.gH + dy;
                if (yy >= 0 && yy < g.gH) v[i] = *(const u32x4*)(row + p0 + dy * g.gW);
            } else {
                uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int p = p0 + j;
                    const int yy = (p / g.gW) % g.gH + dy;
                    const uint32_t b = (yy >= 0 && yy < g.gH) ? (uint32_t)row[p + dy * g.gW] : 0u;
                    w[j >> 1] |= b << (16 * (j & 1));
                }
                v[i] = u32x4{w[0], w[1], w[2], w[3]};
            }
        }
    }
    __device__ __forceinline__ void store(uint16_t (*lds)[LDS_PITCH]) const
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int e = t + 256 * i;
            if (VECS % 256 != 0 && e >= VECS) continue;
            *(u32x4*)(&lds[e >> 2][8 * (e & 3)]) = v[i];
        }
    }
};

template <int L, int R>
struct LoaderB {
    TileLoader<L, R> l;
    __device__ __forceinline__ void load(const GemmArgs& g, int n0, int k0, int kend) { l.load(g.B, g.ldb, g.N, n0, k0, kend); }
    __device__ __forceinline__ void store(uint16_t (*lds)[LDS_PITCH]) const { l.store(lds); }
};
template <int R>
struct LoaderB<SHIFT, R> {
    ShiftLoader<R> l;
    __device__ __forceinline__ void load(const GemmArgs& g, int n0, int k0, int kend) { l.load(g.B, g, n0, k0, kend); }
    __device__ __forceinline__ void store(uint16_t (*lds)[LDS_PITCH]) const { l.store(lds); }
};

template <int L, int R>
struct LoaderA {
    TileLoader<L, R> l;
    __device__ __forceinline__ void init(const GemmArgs&, int) {}
    __device__ __forceinline__ void load(const GemmArgs& g, int r0, int k0, int kend) { l.load(g.A, g.lda, g.M, r0, k0, kend); }
    __device__ __forceinline__ void store(uint16_t (*lds)[LDS_PITCH]) const { l.store(lds); }
};
template <int R>
struct LoaderA<GATHER, R> {
    GatherLoader<R> l;
    __device__ __forceinline__ void init(const GemmArgs& g, int r0) { l.init(g, r0); }
    __device__ __forceinline__ void load(const GemmArgs& g, int, int k0, int kend) { l.load(g.A, g, k0, kend); }
    __device__ __forceinline__ void store(uint16_t (*lds)[LDS_PITCH]) const { l.store(lds); }
};

// C[M][N] = A B^T over one K range per blockIdx.z.  Block tile BM x BN,
// 4 waves as WM x WN, each wave a (BM/WM) x (BN/WN) tile of 16x16 MFMA
// fragments.  LDS is double-buffered: the next K slice is fetched into
// registers while the MFMAs run on the current one, one barrier per slice.
template <int LA, int LB, int EPI, int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g)
{
    static_assert(WM * WN == 4, "4 waves");
    constexpr int TM = BM / WM / 16, TN = BN / WN / 16;
    static_assert(TM >= 1 && TN >= 1 && BM % 8 == 0 && BN % 8 == 0, "tile shape");
    __shared__ __align__(16) uint16_t As[2][BM][LDS_PITCH];
    __shared__ __align__(16) uint16_t Bs[2][BN][LDS_PITCH];
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int kb = blockIdx.z * g.k_per_split;
    const int ke = min(g.K, kb + g.k_per_split);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wm = (wave / WN) * (BM / WM), wn = (wave % WN) * (BN / WN);
    const int fr = lane & 15, fk = 8 * (lane >> 4);
    f32x4 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    LoaderA<LA, BM> la;
    LoaderB<LB, BN> lb;
    la.init(g, m0);
    if (kb < ke) {
        la.load(g, m0, kb, ke);
        lb.load(g, n0, kb, ke);
        la.store(As[0]);
        lb.store(Bs[0]);
    }
    __syncthreads();
    int cur = 0;
    for (int k0 = kb; k0 < ke; k0 += BK) {
        const bool more = k0 + BK < ke;
        if (more) {
            la.load(g, m0, k0 + BK, ke);
            lb.load(g, n0, k0 + BK, ke);
        }
        bf16x8 a[TM], b[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) a[i] = *(const bf16x8*)(&As[cur][wm + 16 * i + fr][fk]);
#pragma unroll
        for (int j = 0; j < TN; ++j) b[j] = *(const bf16x8*)(&Bs[cur][wn + 16 * j + fr][fk]);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        if (more) {
            la.store(As[cur ^ 1]);
            lb.store(Bs[cur ^ 1]);
        }
        __syncthreads();
        cur ^= 1;
    }

    // D layout: lane holds rows 4*(lane/16) + 0..3 of column lane % 16.
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
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
                    if (EPI == EPI_BF16_MASK && !(bf16_f32(g.mask[(int64_t)row * g.ldc + col]) > 0.f)) v = 0.f;
                    ((uint16_t*)g.C)[(int64_t)row * g.ldc + col] = bf16_rne(v);
                }
            }
        }
}

// Tile shapes by the GEMM's N (columns of C): the network's GEMMs are
// tall (M = pixels) with N = 8..256 output channels, or, for the weight
// gradient, short (M = output channels) and wide (N = 9 x input channels).
enum TileCfg : int { TILE_256x32 = 0, TILE_128x64 = 1, TILE_128x128 = 2, TILE_32x256 = 3, TILE_64x256 = 4 };

inline int pick_tile(int M, int N)
{
    if (M <= 32 && N >= 128) return TILE_32x256;
    if (M <= 64 && N >= 128) return TILE_64x256;
    if (N <= 32) return TILE_256x32;
    if (N <= 64) return TILE_128x64;
    return TILE_128x128;
}

inline void tile_dims(int cfg, int& bm, int& bn)
{
    switch (cfg) {
    case TILE_256x32: bm = 256; bn = 32; break;
    case TILE_128x64: bm = 128; bn = 64; break;
    case TILE_32x256: bm = 32; bn = 256; break;
    case TILE_64x256: bm = 64; bn = 256; break;
    default: bm = 128; bn = 128; break;
    }
}

template <int LA, int LB, int EPI>
inline void launch_gemm(int cfg, dim3 grid, hipStream_t s, const GemmArgs& g)
{
    switch (cfg) {
    case TILE_256x32: hipLaunchKernelGGL((k_gemm<LA, LB, EPI, 256, 32, 4, 1>), grid, dim3(256), 0, s, g); break;
    case TILE_128x64: hipLaunchKernelGGL((k_gemm<LA, LB, EPI, 128, 64, 2, 2>), grid, dim3(256), 0, s, g); break;
    case TILE_32x256: hipLaunchKernelGGL((k_gemm<LA, LB, EPI, 32, 256, 1, 4>), grid, dim3(256), 0, s, g); break;
    case TILE_64x256: hipLaunchKernelGGL((k_gemm<LA, LB, EPI, 64, 256, 1, 4>), grid, dim3(256), 0, s, g); break;
    default: hipLaunchKernelGGL((k_gemm<LA, LB, EPI, 128, 128, 2, 2>), grid, dim3(256), 0, s, g); break;
    }
}

// Column-shifted transposed copies of an NHWC input for the SHIFT operand:
// x3[s][c][p] = in[p + (s - 1) dil][c] when the shifted pixel lies in p's
// image row, else 0 (s = 0, 1, 2; dil <= 8).  A block loads 64 + 16 pixels x
// 64 channels into LDS with 16-byte loads and writes the three shifted
// 64-pixel rows of each channel with 16-byte stores.  P and C are multiples
// of 8.
__global__ __launch_bounds__(256) void k_shift3_t(const uint16_t* __restrict__ in, uint16_t* __restrict__ x3,
                                                  int64_t P, int W, int C, int dil)
{
    __shared__ uint16_t tile[80][65];  // odd pitch: the 8 pixel groups of a channel hit 8 banks
    const int64_t p0 = (int64_t)blockIdx.x * 64;
    const int c0 = blockIdx.y * 64;
    const int t = threadIdx.x;
    for (int e = t; e < 80 * 8; e += 256) {
        const int j = e >> 3, cv = e & 7;  // pixel p0 - 8 + j, channels c0 + 8cv .. +7
        const int64_t q = p0 - 8 + j;
        const int c = c0 + 8 * cv;
        u32x4 v = {0u, 0u, 0u, 0u};
        if (q >= 0 && q < P && c < C) v = *(const u32x4*)(in + q * C + c);
        const uint16_t* h = (const uint16_t*)&v;
#pragma unroll
        for (int i = 0; i < 8; ++i) tile[j][8 * cv + i] = h[i];
    }
    __syncthreads();
    for (int e = t; e < 3 * 64 * 8; e += 256) {
        const int s = e / 512, rem = e - s * 512;
        const int c = rem >> 3, pv = rem & 7;  // channel c0 + c, pixels p0 + 8pv .. +7
        const int64_t p = p0 + 8 * pv;
        if (c0 + c >= C || p >= P) continue;
        const int dx = (s - 1) * dil;
        uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int xx = (int)((p + i) % W) + dx;
            const uint32_t b = (xx >= 0 && xx < W) ? (uint32_t)tile[8 + 8 * pv + i + dx][c] : 0u;
            w[i >> 1] |= b << (16 * (i & 1));
        }
        *(u32x4*)(x3 + ((int64_t)s * C + c0 + c) * P + p) = u32x4{w[0], w[1], w[2], w[3]};
    }
}

}  // namespace cnn
}  // namespace cs
