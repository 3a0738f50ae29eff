// Kernels shared by the single-lattice (lattice.hip) and batched (batch.hip)
// translation units: geometry expansion, conversions, equilibrium
// initialisation and moments.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "cs_internal.hpp"
#include "lbm_device.hpp"

namespace cs {

// Node info from flags (layout in cs_internal.hpp).
static __global__ void k_node_info(const uint8_t* __restrict__ flags, uint32_t* __restrict__ node, int w, int h,
                            int pitch, int edge)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (c >= w || r >= h) return;
    const int t = flags[(int64_t)r * pitch + c];
    uint32_t info = (uint32_t)t << NODE_TYPE_SHIFT;
    for (int i = 1; i < 9; ++i) {
        int sr = r + CX[i], sc = c - CY[i];
        if (edge == CS_EDGE_PERIODIC) {
            sr = (sr + h) % h;
            sc = (sc + w) % w;
        } else if (sr < 0 || sr >= h || sc < 0 || sc >= w) {
            continue;
        }
        const int st = flags[(int64_t)sr * pitch + sc];
        if (st == CS_WALL || st == CS_LID) info |= 1u << i;
        if (st == CS_LID) info |= 1u << (9 + i);
    }
    node[(int64_t)r * pitch + c] = info;
}

template <class T>
__device__ __forceinline__ void equilibrium(const Consts<T>& k, T rho, T ux, T uy, T (&F)[9])
{
#pragma clang fp contract(off)
    const T v2 = ux * ux + uy * uy;
    const T vc[9] = {T(0), ux, uy, -ux, -uy, ux + uy, -ux + uy, -ux - uy, ux - uy};
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const T s = ((T(1) + vc[i] * k.A) + (vc[i] * vc[i]) * k.B) + v2 * k.C;
        F[i] = (rho * k.w(i)) * s;
    }
}

// Deviation-form equilibrium g_eq_i = w_i (drho + rho s'_i) with drho = rho - rho0.
template <class T>
__device__ __forceinline__ void equilibrium_dev(const Consts<T>& k, T rho, T ux, T uy, T (&G)[9])
{
#pragma clang fp contract(off)
    const T drho = rho - k.rho0;
    const T v2 = ux * ux + uy * uy;
    const T vc[9] = {T(0), ux, uy, -ux, -uy, ux + uy, -ux + uy, -ux - uy, ux - uy};
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const T s = (vc[i] * k.A + (vc[i] * vc[i]) * k.B) + v2 * k.C;
        G[i] = k.w(i) * (drho + rho * s);
    }
}

// Lattice::density / momentum_density / velocity (lbm.rs:117-149) of the
// stored values (deviations add rho0 to the density).
template <class St>
__device__ __forceinline__ void moments(const Consts<typename St::T>& k, const typename St::T (&B)[9],
                                        typename St::T& rho, typename St::T& mx, typename St::T& my,
                                        typename St::T& ux, typename St::T& uy)
{
#pragma clang fp contract(off)
    using T = typename St::T;
    const T sum = (((((((B[0] + B[1]) + B[2]) + B[3]) + B[4]) + B[5]) + B[6]) + B[7]) + B[8];
    rho = St::DEV ? k.rho0 + sum : sum;
    mx = ((((B[1] - B[3]) + B[5]) - B[6]) - B[7]) + B[8];
    my = ((((B[2] - B[4]) + B[5]) + B[6]) - B[7]) - B[8];
    const T inv = T(1) / rho;
    ux = inv * mx;
    uy = inv * my;
}

template <class St>
__device__ __forceinline__ void load_own(const typename St::S* f, int64_t P, int64_t own, typename St::T (&B)[9])
{
#pragma unroll
    for (int i = 0; i < 9; ++i) B[i] = St::ld(f, (int64_t)i * P + own);
}

// Dense f32 (pitched) <-> storage.
template <class St>
__global__ void k_from_f32(const float* __restrict__ src, typename St::S* __restrict__ dst, int w, int h, int pitch,
                           int64_t P, Consts<typename St::T> k)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (c >= w) return;
    const int64_t own = (int64_t)r * pitch + c;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const typename St::T v = (typename St::T)src[i * P + own];
        St::st(dst, i * P + own, St::DEV ? v - k.base(i) : v);
    }
}

template <class St>
__global__ void k_to_f32(const typename St::S* __restrict__ src, float* __restrict__ dst, int w, int h, int pitch,
                         int64_t P, Consts<typename St::T> k)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (c >= w) return;
    const int64_t own = (int64_t)r * pitch + c;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const typename St::T v = St::ld(src, i * P + own);
        dst[i * P + own] = (float)(St::DEV ? v + k.base(i) : v);
    }
}

// f = compute_equilibrium(rho, u) (lbm.rs:43-71); inputs dense h*w planes in T.
template <class St>
__global__ void k_init_eq(const typename St::T* __restrict__ rho, const typename St::T* __restrict__ ux,
                          const typename St::T* __restrict__ uy, typename St::S* __restrict__ f, int w, int h,
                          int pitch, int64_t P, Consts<typename St::T> k)
{
    using T = typename St::T;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (c >= w) return;
    const int64_t d = (int64_t)r * w + c, own = (int64_t)r * pitch + c;
    T F[9];
    if constexpr (St::DEV) equilibrium_dev<T>(k, rho[d], ux[d], uy[d], F);
    else equilibrium<T>(k, rho[d], ux[d], uy[d], F);
#pragma unroll
    for (int i = 0; i < 9; ++i) St::st(f, i * P + own, F[i]);
}

enum : unsigned {
    M_RHO = 1u, M_MX = 2u, M_MY = 4u, M_UX = 8u, M_UY = 16u, M_SPEED = 32u, M_P = 64u
};

// Moments of the stored populations into dense double planes.
template <class St>
__global__ void k_moments(const typename St::S* __restrict__ f, int w, int h, int pitch, int64_t P,
                          Consts<typename St::T> k, unsigned mask, double* __restrict__ out)
{
    using T = typename St::T;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (c >= w) return;
    const int64_t own = (int64_t)r * pitch + c, d = (int64_t)r * w + c, n = (int64_t)w * h;
    T B[9];
    load_own<St>(f, P, own, B);
    T rho, mx, my, ux, uy;
    moments<St>(k, B, rho, mx, my, ux, uy);
    int slot = 0;
    if (mask & M_RHO) out[(slot++) * n + d] = (double)rho;
    if (mask & M_MX) out[(slot++) * n + d] = (double)mx;
    if (mask & M_MY) out[(slot++) * n + d] = (double)my;
    if (mask & M_UX) out[(slot++) * n + d] = (double)ux;
    if (mask & M_UY) out[(slot++) * n + d] = (double)uy;
    if (mask & M_SPEED) out[(slot++) * n + d] = (double)sqrt(ux * ux + uy * uy);  // lbm.rs:153
    if (mask & M_P) out[(slot++) * n + d] = (double)(rho * k.cs2);                // lbm.rs:786
}


}  // namespace cs
