// BF16 MFMA kernels of the warm-start network (gfx950).
//
// Activations are NHWC bf16 (channels contiguous), weights are [Cout][K] with
// K = 9*Cin ordered (tap, ci), tap = 3*(dy+1) + (dx+1).  A 3x3 convolution is
// a GEMM over the gathered input col(in)[p][k]; one GEMM kernel serves the
// forward pass, the data gradient and the weight gradient:
//   forward  out[p][co] = sum_k col(in)[p][k] * W[co][k]    gather inside the GEMM
//   dgrad    din[p][ci] = sum_k colT(dout)[p][k] * Wd[ci][k] gather (flipped taps) inside the GEMM
//   wgrad    dW[co][k]  = sum_p dout[p][co] * col(in)[p][k]  both read pixel-major and transposed
//                                                          into LDS by the loaders, split over p
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
// PMAJOR_T: the A operand of the weight gradient, rows m = channels of an
// NHWC tensor [P][lda] (the output gradient dz), K = pixels: element (m, p)
// at A[p*lda + m].  GATHER_T: its B operand, rows n = tap*C + c of the input
// x read at the tap's neighbour of pixel p (zero outside the image and for
// n >= 9C), K = pixels.  Both are read as 16-byte vectors of 8 channels of
// one pixel and written transposed into LDS, so neither needs a transposed
// or shifted copy in memory.
enum Layout : int { KMAJOR = 0, GATHER = 2, PMAJOR_T = 4, GATHER_T = 5 };

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
    // GATHER / GATHER_T of a virtual concatenation (glo not null): channels
    // [0, gC1) are the nearest 2x up-sampling of glo ([gH/2][gW/2][gC1] per
    // image), channels [gC1, gC) the operand tensor itself, which then holds
    // gC - gC1 channels per pixel (the skip tensor).  No concatenated copy exists.
    const uint16_t* glo;
    int gC1;
    const uint16_t* mask;   // EPI_BF16_MASK: [M][ldc] bf16 ReLU output whose zeros mask C
};

constexpr int BK = 64;
constexpr int KV = BK / 8;         // 16-byte vectors per BK-wide row
constexpr int LDS_PITCH = BK + 8;  // bf16 elements per LDS row (padding against bank conflicts)

// Channels c..c+7 of pixel q = (img, yy, xx) of the gathered tensor `in`,
// through the virtual concatenation when g.glo is set.
__device__ __forceinline__ u32x4 gathered8(const uint16_t* __restrict__ in, const GemmArgs& g, int64_t q, int img,
                                           int yy, int xx, int c)
{
    if (g.glo == nullptr) return *(const u32x4*)(in + q * g.gC + c);
    if (c < g.gC1) {
        const int64_t qc = ((int64_t)img * (g.gH >> 1) + (yy >> 1)) * (g.gW >> 1) + (xx >> 1);
        return *(const u32x4*)(g.glo + qc * g.gC1 + c);
    }
    return *(const u32x4*)(in + q * (g.gC - g.gC1) + (c - g.gC1));
}

// One BK-wide slice of an R-row operand tile: global -> registers (NV 16-byte
// vectors per thread, zero outside `rows`), registers -> LDS as [row][k].
template <int L, int R>
struct TileLoader {
    static_assert(L == KMAJOR, "K-contiguous operand");
    static constexpr int VECS = KV * R;                 // R rows x BK/8 vectors
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
            const int r = e / KV, kg = e % KV;
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
            *(u32x4*)(&lds[e / KV][8 * (e % KV)]) = v[i];
        }
    }
};

// The GATHER operand: each vector slot's pixel is fixed for the whole K loop,
// so its coordinates are computed once; a K slice then only decodes the tap.
template <int R>
struct GatherLoader {
    static constexpr int VECS = KV * R;
    static constexpr int NV = (VECS + 255) / 256;
    u32x4 v[NV];
    int y[NV], x[NV], img[NV];
    int64_t base[NV];  // the slot's pixel

    __device__ __forceinline__ void init(const GemmArgs& g, int r0)
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int e = t + 256 * i;
            const int p = r0 + e / KV;
            y[i] = -(1 << 28);  // marks an unused slot: every tap falls outside
            x[i] = 0;
            img[i] = 0;
            base[i] = 0;
            if ((VECS % 256 == 0 || e < VECS) && p < g.M) {
                x[i] = p % g.gW;
                const int ny = p / g.gW;
                y[i] = ny % g.gH;
                img[i] = ny / g.gH;
                base[i] = p;
            }
        }
    }
    __device__ __forceinline__ void load(const uint16_t* __restrict__ in, const GemmArgs& g, int k0, int kend)
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int e = t + 256 * i;
            const int k = k0 + 8 * (e % KV);
            v[i] = u32x4{0u, 0u, 0u, 0u};
            if (k < kend && k < 9 * g.gC) {
                const int tap = k / g.gC, c = k - tap * g.gC;
                const int dy = g.gsign * g.gdil * (tap / 3 - 1), dx = g.gsign * g.gdil * (tap % 3 - 1);
                const int yy = y[i] + dy, xx = x[i] + dx;
                if (yy >= 0 && yy < g.gH && xx >= 0 && xx < g.gW) {
                    if (g.glo == nullptr)
                        v[i] = *(const u32x4*)(in + base[i] * g.gC + (int64_t)(dy * g.gW + dx) * g.gC + c);
                    else
                        v[i] = gathered8(in, g, base[i] + (int64_t)(dy * g.gW + dx), img[i], yy, xx, c);
                }
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
            *(u32x4*)(&lds[e / KV][8 * (e % KV)]) = v[i];
        }
    }
};

// A BK-pixel slice of a pixel-major operand: vector e holds channels
// 8*(e / BK) .. +7 of pixel k0 + e % BK; stored transposed, lds[channel][pixel].
template <int R>
struct PixelVecs {
    static_assert(R % 8 == 0, "rows in groups of 8 channels");
    static constexpr int VECS = BK * R / 8;
    static constexpr int NV = (VECS + 255) / 256;
    u32x4 v[NV];
    __device__ __forceinline__ void store(uint16_t (*lds)[LDS_PITCH]) const
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int e = t + 256 * i;
            if (VECS % 256 != 0 && e >= VECS) continue;
            const int j = e % BK, r = 8 * (e / BK);
            const uint32_t w[4] = {v[i].x, v[i].y, v[i].z, v[i].w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                lds[r + 2 * q][j] = (uint16_t)w[q];
                lds[r + 2 * q + 1][j] = (uint16_t)(w[q] >> 16);
            }
        }
    }
};

// PMAJOR_T: rows m0.. of A = channels of dz [P][lda] (lda a multiple of 8).
template <int R>
struct PixelMajorLoader : PixelVecs<R> {
    using PixelVecs<R>::v;
    static constexpr int VECS = PixelVecs<R>::VECS, NV = PixelVecs<R>::NV;
    __device__ __forceinline__ void load(const uint16_t* __restrict__ src, int ld, int rows, int r0, int k0, int kend)
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int e = t + 256 * i;
            v[i] = u32x4{0u, 0u, 0u, 0u};
            if (VECS % 256 != 0 && e >= VECS) continue;
            const int p = k0 + e % BK, c = r0 + 8 * (e / BK);
            if (p < kend && c < rows) v[i] = *(const u32x4*)(src + (int64_t)p * ld + c);
        }
    }
};

// GATHER_T: rows n0.. of B = (tap, channel) pairs of the NHWC input x, read
// at the tap's neighbour (dilation gdil) of each pixel of the slice.
template <int R>
struct GatherTLoader : PixelVecs<R> {
    using PixelVecs<R>::v;
    static constexpr int VECS = PixelVecs<R>::VECS, NV = PixelVecs<R>::NV;
    __device__ __forceinline__ void load(const uint16_t* __restrict__ x, const GemmArgs& g, int n0, int k0, int kend)
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int e = t + 256 * i;
            v[i] = u32x4{0u, 0u, 0u, 0u};
            if (VECS % 256 != 0 && e >= VECS) continue;
            const int p = k0 + e % BK, n = n0 + 8 * (e / BK);
            if (p >= kend || n >= g.N || n >= 9 * g.gC) continue;
            const int tap = n / g.gC, c = n - tap * g.gC;  // gC % 8 == 0: the 8 rows share the tap
            const int dy = g.gdil * (tap / 3 - 1), dx = g.gdil * (tap % 3 - 1);
            const int py = p / g.gW;
            const int xx = p - py * g.gW + dx, yy = py % g.gH + dy;
            if (yy >= 0 && yy < g.gH && xx >= 0 && xx < g.gW)
                v[i] = gathered8(x, g, (int64_t)p + dy * g.gW + dx, py / g.gH, yy, xx, c);
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
struct LoaderB<GATHER_T, R> {
    GatherTLoader<R> l;
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
struct LoaderA<PMAJOR_T, R> {
    PixelMajorLoader<R> l;
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
#pragma unroll
        for (int ks = 0; ks < BK; ks += 32) {
            bf16x8 a[TM], b[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) a[i] = *(const bf16x8*)(&As[cur][wm + 16 * i + fr][ks + fk]);
#pragma unroll
            for (int j = 0; j < TN; ++j) b[j] = *(const bf16x8*)(&Bs[cur][wn + 16 * j + fr][ks + fk]);
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
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

// Direct 3x3 convolution (dilation 1) on MFMA with the input tile and its
// one-pixel halo staged in LDS once per chunk of 32 input channels, so the
// nine taps read LDS only (the GATHER GEMM fetches every tap from memory).
// Forward: out[p][n] = sum over taps and channels of in[p + tap][c] W[n][tap*C + c]
// (+ bias, epilogue); with sign -1 (flipped taps) and the transposed weights
// it is the data gradient.  A block is 4 waves over a tile of 4 image rows x
// 64 columns of one image (W % 64 == 0, H % 4 == 0); wave w computes row w
// (64 pixels) x BN output channels n0 = blockIdx.y * BN.  The input may be a
// virtual concatenation (in.glo set, as in GemmArgs).
struct HaloArgs {
    const uint16_t* in;     // gathered input, NHWC with C channels (or the skip part)
    const uint16_t* glo;    // virtual concatenation: up-sampled part (C1 channels at H/2 x W/2), or null
    int C1;
    const uint16_t* w;      // weights [N][ldw], k = tap*C + c
    int ldw;
    const float* bias;      // [N] or null
    void* out;              // [P][ldc]
    int ldc;
    const uint16_t* mask;   // EPI_BF16_MASK: [P][ldc] ReLU output whose zeros mask the result
    int N, C, H, W, sign;
};

constexpr int HALO_CP = 40;                 // bf16 per halo pixel: 32 channels + pad (conflict-free 16-byte reads)
constexpr int HALO_PIX = 6 * 66;            // (4 + 2) rows x (64 + 2) columns
constexpr int HALO_WP = 9 * 32 + 8;         // bf16 per weight row: 9 taps x 32 channels + pad

template <int EPI, int BN>
__global__ __launch_bounds__(256) void k_conv_halo(HaloArgs a)
{
    constexpr int TN = BN / 16;
    __shared__ __align__(16) uint16_t halo[HALO_PIX][HALO_CP];
    __shared__ __align__(16) uint16_t wts[BN][HALO_WP];
    const int tiles_x = a.W / 64, tiles_y = a.H / 4;
    int t = blockIdx.x;
    const int tx0 = (t % tiles_x) * 64;
    t /= tiles_x;
    const int ty0 = (t % tiles_y) * 4;
    const int img = t / tiles_y;
    const int n0 = blockIdx.y * BN;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int fr = lane & 15, fk = 8 * (lane >> 4);
    const int C2 = a.C - a.C1;
    f32x4 acc[4][TN];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int c0 = 0; c0 < a.C; c0 += 32) {
        // halo: pixel (ty0 - 1 + hy, tx0 - 1 + hx), channels c0 .. c0 + 31
        for (int e = threadIdx.x; e < HALO_PIX * 4; e += 256) {
            const int hp = e >> 2, v = e & 3;
            const int hy = hp / 66, hx = hp - hy * 66;
            const int yy = ty0 - 1 + hy, xx = tx0 - 1 + hx, c = c0 + 8 * v;
            u32x4 val = {0u, 0u, 0u, 0u};
            if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W && c < a.C) {
                const int64_t q = ((int64_t)img * a.H + yy) * a.W + xx;
                if (a.glo == nullptr) {
                    val = *(const u32x4*)(a.in + q * a.C + c);
                } else if (c < a.C1) {
                    const int64_t qc = ((int64_t)img * (a.H / 2) + yy / 2) * (a.W / 2) + xx / 2;
                    val = *(const u32x4*)(a.glo + qc * a.C1 + c);
                } else {
                    val = *(const u32x4*)(a.in + q * C2 + (c - a.C1));
                }
            }
            *(u32x4*)(&halo[hp][8 * v]) = val;
        }
        // weights of the chunk: wts[n][tap*32 + k] = W[n0 + n][tap*C + c0 + k]
        for (int e = threadIdx.x; e < BN * 36; e += 256) {
            const int n = e / 36, r = e - n * 36;
            const int tap = r >> 2, v = r & 3;
            const int c = c0 + 8 * v;
            u32x4 val = {0u, 0u, 0u, 0u};
            if (n0 + n < a.N && c < a.C) val = *(const u32x4*)(a.w + (int64_t)(n0 + n) * a.ldw + tap * a.C + c);
            *(u32x4*)(&wts[n][32 * tap + 8 * v]) = val;
        }
        __syncthreads();
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = a.sign * (tap / 3 - 1), dx = a.sign * (tap % 3 - 1);
            bf16x8 af[4], bf[TN];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                af[i] = *(const bf16x8*)(&halo[(wave + 1 + dy) * 66 + 16 * i + fr + 1 + dx][fk]);
#pragma unroll
            for (int j = 0; j < TN; ++j) bf[j] = *(const bf16x8*)(&wts[16 * j + fr][32 * tap + fk]);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }

    // D layout: lane holds rows 4*(lane/16) + 0..3 (pixels) of column lane % 16 (channel)
    const int64_t prow = ((int64_t)img * a.H + ty0 + wave) * a.W + tx0;
#pragma unroll
    for (int j = 0; j < TN; ++j) {
        const int col = n0 + 16 * j + fr;
        if (col >= a.N) continue;
        const float bias = a.bias ? a.bias[col] : 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t p = prow + 16 * i + 4 * (lane >> 4) + q;
                float v = acc[i][j][q] + bias;
                if (EPI == EPI_F32) {
                    ((float*)a.out)[p * a.ldc + col] = v;
                } else {
                    if (EPI == EPI_BF16_RELU) v = v > 0.f ? v : 0.f;
                    if (EPI == EPI_BF16_MASK && !(bf16_f32(a.mask[p * a.ldc + col]) > 0.f)) v = 0.f;
                    ((uint16_t*)a.out)[p * a.ldc + col] = bf16_rne(v);
                }
            }
    }
}

// Weight gradient of a 3x3 convolution with dilation 1 from LDS tiles:
// dW[co][tap*C + ci] = sum over pixels p of dz[p][co] x[p + tap][ci].  A
// block owns BM output channels (blockIdx.x), 16 input channels (blockIdx.y)
// and a run of 4 x 64 pixel tiles (blockIdx.z of gridDim.z runs); per tile it
// stages dz transposed (dzT[co][pixel]) and the input tile with its halo as
// three copies shifted by one column (xs[dx + 1][ci][row][col]), so every
// MFMA operand is an aligned 16-byte LDS read and all nine taps reuse the
// staged tile.  The sums of a run go to part[blockIdx.z][co][k] (padding
// columns k >= 9C written as 0); a fixed-order reduction over the runs follows.
struct WgradArgs {
    const uint16_t* dz;    // [P][cout]
    const uint16_t* in;    // input, NHWC with C channels (or the skip part of a virtual concatenation)
    const uint16_t* glo;   // virtual concatenation: up-sampled part (C1 channels at H/2 x W/2), or null
    int C1;
    float* part;           // [gridDim.z][cout][ldc]
    int ldc;
    int cout, C, H, W, tiles;
};

constexpr int WG_CI = 16;          // input channels per block
constexpr int WG_DZP = 256 + 8;    // bf16 per dzT row: 256 pixels + pad (conflict-free A reads)
constexpr int WG_XR = 68;          // bf16 per xs row: 64 columns + pad (conflict-free B reads)

template <int BM>
__global__ __launch_bounds__(256) void k_wgrad_halo(WgradArgs a)
{
    constexpr int NF = (BM / 16) * 9;       // 16x16 output fragments: (output-channel block, tap)
    constexpr int FPW = (NF + 3) / 4;       // per wave
    __shared__ __align__(16) uint16_t dzT[BM][WG_DZP];
    __shared__ __align__(16) uint16_t xs[3][WG_CI][6][WG_XR];
    const int co0 = blockIdx.x * BM, ci0 = blockIdx.y * WG_CI;
    const int z = blockIdx.z, runs = gridDim.z;
    const int t0 = (int)((int64_t)a.tiles * z / runs), t1 = (int)((int64_t)a.tiles * (z + 1) / runs);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int fr = lane & 15, fk = 8 * (lane >> 4);
    const int tiles_x = a.W / 64, tiles_y = a.H / 4;
    const int C2 = a.C - a.C1;
    f32x4 acc[FPW];
#pragma unroll
    for (int i = 0; i < FPW; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int t = t0; t < t1; ++t) {
        const int tx0 = (t % tiles_x) * 64;
        const int ty0 = ((t / tiles_x) % tiles_y) * 4;
        const int img = t / (tiles_x * tiles_y);
        // dz of the tile's 256 pixels, transposed: dzT[co][64 r + c]
        for (int e = threadIdx.x; e < 256 * (BM / 8); e += 256) {
            const int px = e & 255, v = e >> 8;
            const int co = co0 + 8 * v;
            u32x4 val = {0u, 0u, 0u, 0u};
            if (co < a.cout) {
                const int64_t q = ((int64_t)img * a.H + ty0 + (px >> 6)) * a.W + tx0 + (px & 63);
                val = *(const u32x4*)(a.dz + q * a.cout + co);
            }
            const uint32_t w[4] = {val.x, val.y, val.z, val.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                dzT[8 * v + 2 * j][px] = (uint16_t)w[j];
                dzT[8 * v + 2 * j + 1][px] = (uint16_t)(w[j] >> 16);
            }
        }
        // input halo (6 rows x 66 columns), channels ci0 .. ci0 + 15, three shifted copies
        for (int e = threadIdx.x; e < 6 * 66 * 2; e += 256) {
            const int hp = e >> 1, v = e & 1;
            const int hy = hp / 66, hx = hp - hy * 66;
            const int yy = ty0 - 1 + hy, xx = tx0 - 1 + hx, c = ci0 + 8 * v;
            u32x4 val = {0u, 0u, 0u, 0u};
            if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W && c < a.C) {
                const int64_t q = ((int64_t)img * a.H + yy) * a.W + xx;
                if (a.glo == nullptr) {
                    val = *(const u32x4*)(a.in + q * a.C + c);
                } else if (c < a.C1) {
                    const int64_t qc = ((int64_t)img * (a.H / 2) + yy / 2) * (a.W / 2) + xx / 2;
                    val = *(const u32x4*)(a.glo + qc * a.C1 + c);
                } else {
                    val = *(const u32x4*)(a.in + q * C2 + (c - a.C1));
                }
            }
            const uint32_t w[4] = {val.x, val.y, val.z, val.w};
#pragma unroll
            for (int sft = 0; sft < 3; ++sft) {
                const int col = hx - sft;
                if (col < 0 || col >= 64) continue;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    xs[sft][8 * v + 2 * j][hy][col] = (uint16_t)w[j];
                    xs[sft][8 * v + 2 * j + 1][hy][col] = (uint16_t)(w[j] >> 16);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) {  // 32-pixel slices: tile row ks / 2, columns 32 (ks % 2) ..
            const int r = ks >> 1, cx = 32 * (ks & 1);
#pragma unroll
            for (int i = 0; i < FPW; ++i) {
                const int f = wave + 4 * i;
                if (f >= NF) break;
                const int cb = f / 9, tap = f - 9 * cb;
                const int dy = tap / 3 - 1, dx = tap % 3 - 1;
                const bf16x8 av = *(const bf16x8*)(&dzT[16 * cb + fr][64 * r + cx + fk]);
                const bf16x8 bv = *(const bf16x8*)(&xs[dx + 1][fr][r + 1 + dy][cx + fk]);
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[i], 0, 0, 0);
            }
        }
        __syncthreads();
    }

    // D layout: lane holds rows 4*(lane/16) + 0..3 (output channels) of column lane % 16 (input channel)
    float* out = a.part + (int64_t)z * a.cout * a.ldc;
    const int ci = ci0 + fr;
#pragma unroll
    for (int i = 0; i < FPW; ++i) {
        const int f = wave + 4 * i;
        if (f >= NF) break;
        const int cb = f / 9, tap = f - 9 * cb;
        if (ci >= a.C) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int co = co0 + 16 * cb + 4 * (lane >> 4) + q;
            if (co < a.cout) out[(int64_t)co * a.ldc + tap * a.C + ci] = acc[i][q];
        }
    }
    if (blockIdx.y == 0) {
        const int pad = a.ldc - 9 * a.C;
        for (int e = threadIdx.x; e < BM * pad; e += 256) {
            const int co = co0 + e / pad;
            if (co < a.cout) out[(int64_t)co * a.ldc + 9 * a.C + e % pad] = 0.f;
        }
    }
}

}  // namespace cnn
}  // namespace cs
