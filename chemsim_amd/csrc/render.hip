// Device-side render pass (render.rs:7-178 as main.rs:157-176 drives it):
// the display mode's field of the lattice's moments, z-scored over the whole
// lattice, through the logistic function into HSV, then RGB bytes, with the
// solid cells painted blue.  The reference computes the fields with
// ArrayFire and then loops over every pixel on the host; here three passes
// run on the GPU and only the RGB bytes travel.
//
//   scalar modes (density, speed; render_scalar_field, render.rs:23-89):
//     h = 0, s = 1, v = logistic((x - mean) / std)
//   vector modes (velocity, momentum density; render_vector_field, render.rs:91-178):
//     m = vx^2 + vy^2, h = (atan2(vy, vx) + pi) / (2 pi), s = 0.8, v = logistic((m - mean) / std)
//   each channel clamped to [0, 1], HSV -> RGB, byte = clamp(round(256 c), 0, 255);
//   solid cells (0, 0, 255) (render_geometry, render.rs:7-21).
// mean and std are the population mean and standard deviation over all cells
// (af::mean_all, af::stdev_all), accumulated in f64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "cs_internal.hpp"
#include "lattice_kernels.hpp"
#include "lbm_device.hpp"

namespace cs {

// Field-pass grid: 1024 blocks of 256 threads (measured faster than 2048 at 4096^2).
enum : int { RENDER_BLOCKS = 1024 };

template <class St>
__global__ __launch_bounds__(256) void k_render_field(const typename St::S* __restrict__ f, int w, int h, int pitch,
                                                      int64_t P, Consts<typename St::T> k, int mode,
                                                      float* __restrict__ sval, float* __restrict__ phase,
                                                      double* __restrict__ partials)
{
    using T = typename St::T;
    __shared__ double sh[2][256];
    const int64_t n = (int64_t)w * h;
    double s1 = 0.0, s2 = 0.0;
    // cell d = r * w + c; (r, c) advance by the grid stride without a division per cell
    const int64_t stride = (int64_t)gridDim.x * 256;
    const int dr = (int)(stride / w), dc = (int)(stride - (int64_t)dr * w);
    int64_t d = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int r = (int)(d / w), c = (int)(d - (int64_t)r * w);
    for (; d < n; d += stride, r += dr, c += dc) {
        if (c >= w) {
            c -= w;
            ++r;
        }
        T B[9];
        load_own<St>(f, P, (int64_t)r * pitch + c, B);
        T rho, mx, my, ux, uy;
        moments<St>(k, B, rho, mx, my, ux, uy);
        float s;
        if (mode == CS_RENDER_DENSITY) {
            s = (float)rho;
        } else if (mode == CS_RENDER_SPEED) {
            const float vx = (float)ux, vy = (float)uy;
            s = sqrtf(vx * vx + vy * vy);  // lbm.rs:151-154
        } else {
            const float vx = mode == CS_RENDER_VELOCITY ? (float)ux : (float)mx;
            const float vy = mode == CS_RENDER_VELOCITY ? (float)uy : (float)my;
            s = vx * vx + vy * vy;         // render.rs:112
            phase[d] = atan2f(vy, vx);     // af::arg(cplx2(vx, vy)), render.rs:113-114
        }
        sval[d] = s;
        s1 += (double)s;
        s2 += (double)s * (double)s;
    }
    sh[0][threadIdx.x] = s1;
    sh[1][threadIdx.x] = s2;
    __syncthreads();
    for (int q = 128; q > 0; q >>= 1) {
        if ((int)threadIdx.x < q) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + q];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + q];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = sh[0][0];
        partials[2 * blockIdx.x + 1] = sh[1][0];
    }
}

// The field pass for 4-byte storage with rows that split into 4-cell vectors:
// a thread takes 4 consecutive cells of a row per iteration, reading each
// population plane and writing the field (and phase) with one 16-byte access.
// Per-cell values are those of k_render_field.
template <class St>
__global__ __launch_bounds__(256) void k_render_field_vec(const typename St::S* __restrict__ f, int w, int h,
                                                          int pitch, int64_t P, Consts<typename St::T> k, int mode,
                                                          float* __restrict__ sval, float* __restrict__ phase,
                                                          double* __restrict__ partials)
{
    using S = typename St::S;
    using T = typename St::T;
    static_assert(sizeof(S) == 4, "4-byte storage");
    __shared__ double sh[2][256];
    const int wq = w >> 2;  // vectors per row
    const int64_t n4 = (int64_t)wq * h;
    double s1 = 0.0, s2 = 0.0;
    const int64_t stride = (int64_t)gridDim.x * 256;
    const int dr = (int)(stride / wq), dq = (int)(stride - (int64_t)dr * wq);
    int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int r = (int)(g / wq), q = (int)(g - (int64_t)r * wq);
    for (; g < n4; g += stride, r += dr, q += dq) {
        if (q >= wq) {
            q -= wq;
            ++r;
        }
        const int64_t own = (int64_t)r * pitch + 4 * q;
        VecN<S, 16> v[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) v[i].raw = __builtin_nontemporal_load((const RawVec<16>*)(f + (int64_t)i * P + own));
        float sv[4], ph[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            T B[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) B[i] = St::wid(v[i].e[j]);
            T rho, mx, my, ux, uy;
            moments<St>(k, B, rho, mx, my, ux, uy);
            if (mode == CS_RENDER_DENSITY) {
                sv[j] = (float)rho;
            } else if (mode == CS_RENDER_SPEED) {
                const float vx = (float)ux, vy = (float)uy;
                sv[j] = sqrtf(vx * vx + vy * vy);  // lbm.rs:151-154
            } else {
                const float vx = mode == CS_RENDER_VELOCITY ? (float)ux : (float)mx;
                const float vy = mode == CS_RENDER_VELOCITY ? (float)uy : (float)my;
                sv[j] = vx * vx + vy * vy;         // render.rs:112
                ph[j] = atan2f(vy, vx);            // af::arg(cplx2(vx, vy)), render.rs:113-114
            }
            s1 += (double)sv[j];
            s2 += (double)sv[j] * (double)sv[j];
        }
        const int64_t d = (int64_t)r * w + 4 * q;
        *reinterpret_cast<float4*>(sval + d) = make_float4(sv[0], sv[1], sv[2], sv[3]);
        if (mode == CS_RENDER_VELOCITY || mode == CS_RENDER_MOMENTUM)
            *reinterpret_cast<float4*>(phase + d) = make_float4(ph[0], ph[1], ph[2], ph[3]);
    }
    sh[0][threadIdx.x] = s1;
    sh[1][threadIdx.x] = s2;
    __syncthreads();
    for (int qq = 128; qq > 0; qq >>= 1) {
        if ((int)threadIdx.x < qq) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + qq];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + qq];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = sh[0][0];
        partials[2 * blockIdx.x + 1] = sh[1][0];
    }
}

// mean and 1/std as the reference uses them: Matrix::new_filled(avg as f32)
// and scale(1.0 / std as f32) (render.rs:40-45).
__global__ __launch_bounds__(256) void k_render_stats(const double* __restrict__ partials, int nblocks, int64_t n,
                                                      float* __restrict__ stats)
{
    __shared__ double sh[2][256];
    double s1 = 0.0, s2 = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += 256) {
        s1 += partials[2 * b];
        s2 += partials[2 * b + 1];
    }
    sh[0][threadIdx.x] = s1;
    sh[1][threadIdx.x] = s2;
    __syncthreads();
    for (int q = 128; q > 0; q >>= 1) {
        if ((int)threadIdx.x < q) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + q];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + q];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double mean = sh[0][0] / (double)n;
        const double var = fmax(sh[1][0] / (double)n - mean * mean, 0.0);
        stats[0] = (float)mean;
        stats[1] = 1.0f / (float)sqrt(var);
    }
}

// HSV -> RGB, hue in [0, 1] (sextants of 60 degrees).
__device__ __forceinline__ void hsv_to_rgb(float hue, float sat, float val, float& r, float& g, float& b)
{
    const float c = val * sat;
    const float h6 = hue * 6.0f;
    // fmod(h6, 2) for h6 in [0, 6]: h6 - 2 floor(h6 / 2) is exact there (the
    // halving and doubling are exact and the subtraction meets Sterbenz's bound)
    const float x = c * (1.0f - fabsf((h6 - 2.0f * floorf(h6 * 0.5f)) - 1.0f));
    const float m = val - c;
    float r1, g1, b1;
    if (h6 < 1.0f) { r1 = c; g1 = x; b1 = 0.0f; }
    else if (h6 < 2.0f) { r1 = x; g1 = c; b1 = 0.0f; }
    else if (h6 < 3.0f) { r1 = 0.0f; g1 = c; b1 = x; }
    else if (h6 < 4.0f) { r1 = 0.0f; g1 = x; b1 = c; }
    else if (h6 < 5.0f) { r1 = x; g1 = 0.0f; b1 = c; }
    else { r1 = c; g1 = 0.0f; b1 = x; }
    r = r1 + m;
    g = g1 + m;
    b = b1 + m;
}

__device__ __forceinline__ uint8_t to_byte(float v)
{
    // (256.0 * v).round().min(255.0).max(0.0) as u8 (render.rs:80-84); round half away from zero
    return (uint8_t)fmaxf(fminf(roundf(256.0f * v), 255.0f), 0.0f);
}

__device__ __forceinline__ float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

// One pixel's colour (render.rs:7-178): z-score, logistic, HSV -> RGB, the
// geometry overlay; packed as R | G << 8 | B << 16 | 255 << 24.  HUE: the
// velocity and momentum modes (hue from the phase, saturation 0.8); the
// scalar modes have hue 0 and saturation 1, where hsv_to_rgb gives exactly
// (val, 0, 0).
template <bool HUE>
__device__ __forceinline__ uint32_t pixel_rgba(float s, float ph, float avg, float inv, bool wall)
{
    if (wall) return 0u | (0u << 8) | (255u << 16) | (255u << 24);   // render_geometry: blue
    const float z = (s - avg) * inv;
    const float val = clamp01(1.0f / (1.0f + expf(-z)));  // af::sigmoid
    if constexpr (!HUE) {
        return (uint32_t)to_byte(val) | (255u << 24);
    } else {
        const float pi = 3.14159265358979323846f;
        const float hue = clamp01((ph + pi) * (0.318309886183790671538f * 0.5f));  // render.rs:121-125
        float r, g, b;
        hsv_to_rgb(hue, 0.8f, val, r, g, b);
        return (uint32_t)to_byte(r) | ((uint32_t)to_byte(g) << 8) | ((uint32_t)to_byte(b) << 16) | (255u << 24);
    }
}

// Pixel pass: 4 pixels per thread.  CH = 4 writes RGBA (RGB::to_rgba,
// display.rs:27-29) as one 16-byte store, CH = 3 packed RGB as three 4-byte
// stores; `rgb`, `sval` and `phase` are 16-byte aligned.
template <int CH, bool HUE>
__global__ __launch_bounds__(256) void k_render_pixels(const float* __restrict__ sval, const float* __restrict__ phase,
                                                       const float* __restrict__ stats,
                                                       const uint8_t* __restrict__ flags, int w, int h, int cpitch,
                                                       int overlay, uint8_t* __restrict__ rgb)
{
    const int64_t n = (int64_t)w * h;
    const int64_t d0 = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 4;
    if (d0 >= n) return;
    const bool full = d0 + 4 <= n;
    constexpr bool vec = HUE;
    float s[4], ph[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (full) {
        const float4 v = *reinterpret_cast<const float4*>(sval + d0);
        s[0] = v.x; s[1] = v.y; s[2] = v.z; s[3] = v.w;
        if (vec) {
            const float4 q = *reinterpret_cast<const float4*>(phase + d0);
            ph[0] = q.x; ph[1] = q.y; ph[2] = q.z; ph[3] = q.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            s[j] = d0 + j < n ? sval[d0 + j] : 0.0f;
            if (vec) ph[j] = d0 + j < n ? phase[d0 + j] : 0.0f;
        }
    }
    bool wall[4] = {false, false, false, false};
    if (overlay && flags) {
        int rr = (int)(d0 / w), cc = (int)(d0 - (int64_t)rr * w);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (d0 + j < n) {
                const int t = flags[(int64_t)rr * cpitch + cc];
                wall[j] = t == CS_WALL || t == CS_LID;
            }
            if (++cc == w) {
                cc = 0;
                ++rr;
            }
        }
    }
    const float avg = stats[0], inv = stats[1];
    uint32_t px[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) px[j] = pixel_rgba<HUE>(s[j], ph[j], avg, inv, wall[j]);
    if constexpr (CH == 4) {
        if (full) {
            *reinterpret_cast<uint4*>(rgb + 4 * d0) = make_uint4(px[0], px[1], px[2], px[3]);
        } else {
            for (int j = 0; j < 4 && d0 + j < n; ++j) reinterpret_cast<uint32_t*>(rgb)[d0 + j] = px[j];
        }
    } else {
        if (full) {
            // bytes R0 G0 B0 R1 | G1 B1 R2 G2 | B2 R3 G3 B3
            uint32_t* o = reinterpret_cast<uint32_t*>(rgb + 3 * d0);
            o[0] = (px[0] & 0xffffffu) | (px[1] << 24);
            o[1] = ((px[1] >> 8) & 0xffffu) | ((px[2] & 0xffffu) << 16);
            o[2] = ((px[2] >> 16) & 0xffu) | ((px[3] & 0xffffffu) << 8);
        } else {
            for (int j = 0; j < 4 && d0 + j < n; ++j) {
                rgb[3 * (d0 + j)] = (uint8_t)px[j];
                rgb[3 * (d0 + j) + 1] = (uint8_t)(px[j] >> 8);
                rgb[3 * (d0 + j) + 2] = (uint8_t)(px[j] >> 16);
            }
        }
    }
}

template <class St>
static const Consts<typename St::T>& render_consts(const cs_lattice* l)
{
    if constexpr (std::is_same_v<typename St::T, double>) return l->k64;
    else return l->k32;
}

// The lattice's render scratch: two float fields, the partial sums, the
// statistics and a 3-byte-per-pixel staging area, allocated on first use.
struct RenderScratch {
    float* sval;
    float* phase;
    double* partials;
    float* stats;
    uint8_t* rgb;
};

static RenderScratch render_scratch(cs_lattice* l)
{
    const size_t n = (size_t)l->w * l->h;
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };  // 256-byte aligned sections
    const size_t o_phase = up(sizeof(float) * n), o_part = o_phase + up(sizeof(float) * n);
    const size_t o_stats = o_part + up(sizeof(double) * 2 * RENDER_BLOCKS), o_rgb = o_stats + 256;
    if (!l->render_buf) CS_HIP(hipMalloc(&l->render_buf, o_rgb + up(3 * n)));
    char* base = (char*)l->render_buf;
    RenderScratch r;
    r.sval = (float*)base;
    r.phase = (float*)(base + o_phase);
    r.partials = (double*)(base + o_part);
    r.stats = (float*)(base + o_stats);
    r.rgb = (uint8_t*)(base + o_rgb);
    return r;
}

void render_enqueue(cs_lattice* l, int mode, int overlay, uint8_t* d_out, int channels)
{
    CS_REQUIRE(d_out != nullptr && (channels == 3 || channels == 4), CS_ERR_ARG, "bad render target");
    CS_REQUIRE(((uintptr_t)d_out & 15) == 0, CS_ERR_ARG, "render target must be 16-byte aligned");
    const int64_t n = (int64_t)l->w * l->h;
    const RenderScratch r = render_scratch(l);
    const int blocks = (int)std::min<int64_t>(RENDER_BLOCKS, (n + 255) / 256);
    auto field = [&]<class St>() {
        if constexpr (sizeof(typename St::S) == 4) {
            if (l->w % 4 == 0) {
                hipLaunchKernelGGL((k_render_field_vec<St>), dim3((unsigned)blocks), dim3(256), 0, l->stream,
                                   (const typename St::S*)l->buf[l->cur], l->w, l->h, l->pitch, l->plane,
                                   render_consts<St>(l), mode, r.sval, r.phase, r.partials);
                return;
            }
        }
        hipLaunchKernelGGL((k_render_field<St>), dim3((unsigned)blocks), dim3(256), 0, l->stream,
                           (const typename St::S*)l->buf[l->cur], l->w, l->h, l->pitch, l->plane,
                           render_consts<St>(l), mode, r.sval, r.phase, r.partials);
    };
    switch (l->storage) {
    case CS_F32: field.template operator()<StF32>(); break;
    case CS_F64: field.template operator()<StF64>(); break;
    case CS_F32DEV: field.template operator()<StF32Dev>(); break;
    case CS_BF16DEV: field.template operator()<StBF16>(); break;
    case CS_F16DEV: field.template operator()<StF16>(); break;
    default: throw Error{CS_ERR_ARG, "bad storage"};
    }
    hipLaunchKernelGGL(k_render_stats, dim3(1), dim3(256), 0, l->stream, r.partials, blocks, n, r.stats);
    const uint8_t* fl = l->has_flags ? (const uint8_t*)l->flags : nullptr;
    const unsigned pblocks = (unsigned)((n + 1023) / 1024);  // 4 pixels per thread
    auto pixels = [&]<int CH, bool HUE>() {
        hipLaunchKernelGGL((k_render_pixels<CH, HUE>), dim3(pblocks), dim3(256), 0, l->stream, r.sval, r.phase,
                           r.stats, fl, l->w, l->h, l->cpitch, overlay, d_out);
    };
    const bool hue = mode == CS_RENDER_VELOCITY || mode == CS_RENDER_MOMENTUM;
    if (channels == 4) {
        if (hue) pixels.template operator()<4, true>();
        else pixels.template operator()<4, false>();
    } else {
        if (hue) pixels.template operator()<3, true>();
        else pixels.template operator()<3, false>();
    }
    CS_HIP(hipGetLastError());
}

}  // namespace cs

using namespace cs;

extern "C" int cs_lattice_render(cs_lattice* l, int mode, int overlay, uint8_t* rgb, float* ms)
{
    return guarded([&] {
        CS_REQUIRE(l != nullptr && rgb != nullptr, CS_ERR_ARG, "null argument");
        CS_REQUIRE(mode >= CS_RENDER_DENSITY && mode <= CS_RENDER_MOMENTUM, CS_ERR_ARG, "bad render mode");
        CS_HIP(hipSetDevice(l->ctx->device));
        const int64_t n = (int64_t)l->w * l->h;
        uint8_t* drgb = render_scratch(l).rgb;
        CS_HIP(hipEventRecord(l->ev0, l->stream));
        render_enqueue(l, mode, overlay, drgb, 3);
        CS_HIP(hipEventRecord(l->ev1, l->stream));
        CS_HIP(hipMemcpyAsync(rgb, drgb, 3 * (size_t)n, hipMemcpyDeviceToHost, l->stream));
        CS_HIP(hipStreamSynchronize(l->stream));
        if (ms) {
            float t = 0;
            CS_HIP(hipEventElapsedTime(&t, l->ev0, l->ev1));
            *ms = t;
        }
    });
}
