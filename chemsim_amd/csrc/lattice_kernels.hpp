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

// Derivative at a node from its two neighbours along one axis.  Each
// neighbour is absent (NB_NONE: outside a non-periodic edge), a fluid node at
// distance 1 (NB_FLUID, value v) or a wall (NB_WALL, wall velocity v) half-way
// to the solid node -- where both bounce-back forms put it.  The wall cases
// are the derivative of the quadratic through the three points.
enum : int { NB_NONE = 0, NB_FLUID = 1, NB_WALL = 2 };
template <class T>
__host__ __device__ inline T neq_deriv(T u0, int kp, T vp, int km, T vm)
{
    const T third = T(1) / T(3), four_thirds = T(4) / T(3);
    if (kp == NB_FLUID && km == NB_FLUID) return (vp - vm) * T(0.5);
    if (kp == NB_FLUID && km == NB_WALL) return (vp * third + u0) - vm * four_thirds;
    if (kp == NB_WALL && km == NB_FLUID) return (vp * four_thirds - u0) - vm * third;
    if (kp == NB_WALL && km == NB_WALL) return vp - vm;
    if (kp == NB_FLUID) return vp - u0;
    if (km == NB_FLUID) return u0 - vm;
    if (kp == NB_WALL) return (vp - u0) * T(2);
    if (km == NB_WALL) return (u0 - vm) * T(2);
    return T(0);
}

// f = f_eq(rho, u) + f_neq, with the first-order Chapman-Enskog
// non-equilibrium part of the post-collision populations a lattice holds
// between steps: f_neq_i = -3 (tau - 1) w_i rho (Q_i : S) (the pre-collision
// part -3 tau w_i rho Q_i : S times 1 - 1/tau), Q_i = c_i c_i - I/3, S the
// symmetric velocity gradient from neq_deriv (lattice units, dt = 1),
// f_neq = 0 at solid cells.  Lattice-velocity
// coordinates: X = -row, Y = column.  Walls move with the lid velocity at
// CS_LID cells and rest at CS_WALL cells.
template <class T>
struct NeqArgs {
    T tau, lid_ux, lid_uy;
    int periodic;
};

template <class St>
__global__ void k_init_neq(const typename St::T* __restrict__ rho, const typename St::T* __restrict__ ux,
                           const typename St::T* __restrict__ uy, const uint8_t* __restrict__ flags,
                           typename St::S* __restrict__ f, int w, int h, int pitch, int cpitch, int64_t P,
                           Consts<typename St::T> k, NeqArgs<typename St::T> a)
{
    using T = typename St::T;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (c >= w) return;
    const int64_t d = (int64_t)r * w + c, own = (int64_t)r * pitch + c;
    // kind of node (rr, cc); on NB_FLUID/NB_WALL sets the dense index or the wall flag
    auto kind = [&](int rr, int cc, int64_t& idx, int& t) {
        if (a.periodic) {
            rr = rr < 0 ? rr + h : (rr >= h ? rr - h : rr);
            cc = cc < 0 ? cc + w : (cc >= w ? cc - w : cc);
        } else if (rr < 0 || rr >= h || cc < 0 || cc >= w) {
            return (int)NB_NONE;
        }
        idx = (int64_t)rr * w + cc;
        t = flags ? (int)flags[(int64_t)rr * cpitch + cc] : (int)CS_FLUID;
        return (t == CS_WALL || t == CS_LID) ? (int)NB_WALL : (int)NB_FLUID;
    };
    T F[9];
    if constexpr (St::DEV) equilibrium_dev<T>(k, rho[d], ux[d], uy[d], F);
    else equilibrium<T>(k, rho[d], ux[d], uy[d], F);
    int64_t i0;
    int t0;
    if (kind(r, c, i0, t0) == NB_FLUID) {
        // +X is row r - 1, +Y is column c + 1
        int64_t ip = 0, im = 0;
        int tp = 0, tm = 0;
        const int kpx = kind(r - 1, c, ip, tp), kmx = kind(r + 1, c, im, tm);
        auto val = [&](int kk, int64_t idx, int t, const T* u, T lid) -> T {
            if (kk == NB_FLUID) return u[idx];
            if (kk == NB_WALL) return t == CS_LID ? lid : T(0);
            return T(0);
        };
        const T dxux = neq_deriv<T>(ux[d], kpx, val(kpx, ip, tp, ux, a.lid_ux), kmx, val(kmx, im, tm, ux, a.lid_ux));
        const T dxuy = neq_deriv<T>(uy[d], kpx, val(kpx, ip, tp, uy, a.lid_uy), kmx, val(kmx, im, tm, uy, a.lid_uy));
        const int kpy = kind(r, c + 1, ip, tp), kmy = kind(r, c - 1, im, tm);
        const T dyux = neq_deriv<T>(ux[d], kpy, val(kpy, ip, tp, ux, a.lid_ux), kmy, val(kmy, im, tm, ux, a.lid_ux));
        const T dyuy = neq_deriv<T>(uy[d], kpy, val(kpy, ip, tp, uy, a.lid_uy), kmy, val(kmy, im, tm, uy, a.lid_uy));
        const T sxy = T(0.5) * (dyux + dxuy);
        const T third = T(1) / T(3);
        const T pre = (T(-3) * (a.tau - T(1))) * rho[d];
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const T cx = (T)CX[i], cy = (T)CY[i];
            const T qs = ((cx * cx - third) * dxux + (T(2) * cx * cy) * sxy) + (cy * cy - third) * dyuy;
            F[i] = F[i] + (pre * k.w(i)) * qs;
        }
    }
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
