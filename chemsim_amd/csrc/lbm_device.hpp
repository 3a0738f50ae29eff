// Device code of the D2Q9 BGK hot path for gfx950: storage traits, the cell
// update and the fused pull-step kernel.  Shared by the single-lattice
// (lattice.hip) and batched (batch.hip) launches.
//
// One step of the reference (State::step, lbm.rs:694-714) is stream
// (lbm.rs:716-729), bounce-back (lbm.rs:741-751) and BGK collide over every
// cell (lbm.rs:731-739).  The stored state between steps is post-collision,
// so a pull kernel that gathers each cell's nine incoming populations,
// applies the boundary rule, collides and stores reproduces one reference
// step per launch.  Streaming is the rotated shift of the reference: the
// population i of cell (r, c) comes from (r + cx_i, c - cy_i).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "cs_internal.hpp"

namespace cs {

// ----------------------------------------------------------- storage types --
__device__ __forceinline__ float bf16_bits_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// Round to nearest even; NaN stays a quiet NaN.
__device__ __forceinline__ uint16_t f32_to_bf16_bits(float x)
{
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float f16_bits_to_f32(uint16_t b) { return (float)__builtin_bit_cast(_Float16, b); }
// Narrowing of a value that is first rounded to f32.  The empty constraint
// keeps the compiler from fusing the producing multiply into the conversion
// (v_fma_mixlo_f16 rounds once, not twice, and so differs on rare ties).
__device__ __forceinline__ uint16_t f32_to_f16_bits(float x)
{
    asm volatile("" : "+v"(x));
    return __builtin_bit_cast(uint16_t, (_Float16)x);
}

// Stochastic rounding with gfx950's v_cvt_sr_{bf16,f16}_f32: the magnitude
// plus random bits below the kept mantissa (top 16 bits of r for bf16, top 13
// for f16), truncated.
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint16_t f32_to_bf16_sr_bits(float x, uint32_t r)
{
    const bf16x2_t v = __builtin_amdgcn_cvt_sr_bf16_f32(bf16x2_t{}, x, r, false);
    return __builtin_bit_cast(uint16_t, v[0]);
}
__device__ __forceinline__ uint16_t f32_to_f16_sr_bits(float x, uint32_t r)
{
    const f16x2_t v = __builtin_amdgcn_cvt_sr_f16_f32(f16x2_t{}, x, r, false);
    return __builtin_bit_cast(uint16_t, v[0]);
}

// Random bits of the stochastic rounding: h = lowbias32(cell ^ key), then
// xorshift32; each 32-bit word serves two directions (word, word << 16).
__host__ __device__ __forceinline__ uint32_t lowbias32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__device__ __forceinline__ uint32_t xorshift32(uint32_t x)
{
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x;
}
// Per-step key: lowbias32(seed ^ step * phi).
__host__ __device__ __forceinline__ uint32_t sr_key(uint32_t seed, uint64_t step)
{
    return lowbias32(seed ^ ((uint32_t)step * 0x9e3779b9u));
}

// S: stored element type.  T: arithmetic type.  DEV: the stored value is the
// deviation g_i = f_i - w_i*rho0 and the step runs in deviation arithmetic.
// ld() widens a stored value, st() narrows (round to nearest even).
struct StF32 {
    using S = float;
    using T = float;
    static constexpr bool DEV = false;
    __device__ static __forceinline__ T wid(S v) { return v; }
    __device__ static __forceinline__ S nar(T v) { return v; }
    __device__ static __forceinline__ T ld(const S* p, int64_t o) { return p[o]; }
    __device__ static __forceinline__ void st(S* p, int64_t o, T v) { __builtin_nontemporal_store(v, p + o); }
};
struct StF64 {
    using S = double;
    using T = double;
    static constexpr bool DEV = false;
    __device__ static __forceinline__ T wid(S v) { return v; }
    __device__ static __forceinline__ S nar(T v) { return v; }
    __device__ static __forceinline__ T ld(const S* p, int64_t o) { return p[o]; }
    __device__ static __forceinline__ void st(S* p, int64_t o, T v) { __builtin_nontemporal_store(v, p + o); }
};
struct StF32Dev {
    using S = float;
    using T = float;
    static constexpr bool DEV = true;
    __device__ static __forceinline__ T wid(S v) { return v; }
    __device__ static __forceinline__ S nar(T v) { return v; }
    __device__ static __forceinline__ T ld(const S* p, int64_t o) { return p[o]; }
    __device__ static __forceinline__ void st(S* p, int64_t o, T v) { __builtin_nontemporal_store(v, p + o); }
};
struct StBF16 {
    using S = uint16_t;
    using T = float;
    static constexpr bool DEV = true;
    __device__ static __forceinline__ T wid(S v) { return bf16_bits_to_f32(v); }
    __device__ static __forceinline__ S nar(T v) { return f32_to_bf16_bits(v); }
    __device__ static __forceinline__ S nar_sr(T v, uint32_t r) { return f32_to_bf16_sr_bits(v, r); }
    __device__ static __forceinline__ T ld(const S* p, int64_t o) { return bf16_bits_to_f32(p[o]); }
    __device__ static __forceinline__ void st(S* p, int64_t o, T v) { __builtin_nontemporal_store(f32_to_bf16_bits(v), p + o); }
    __device__ static __forceinline__ void st_sr(S* p, int64_t o, T v, uint32_t r)
    {
        __builtin_nontemporal_store(f32_to_bf16_sr_bits(v, r), p + o);
    }
};
struct StF16 {
    using S = uint16_t;
    using T = float;
    static constexpr bool DEV = true;
    __device__ static __forceinline__ T wid(S v) { return f16_bits_to_f32(v); }
    __device__ static __forceinline__ S nar(T v) { return f32_to_f16_bits(v); }
    __device__ static __forceinline__ S nar_sr(T v, uint32_t r) { return f32_to_f16_sr_bits(v, r); }
    __device__ static __forceinline__ T ld(const S* p, int64_t o) { return f16_bits_to_f32(p[o]); }
    __device__ static __forceinline__ void st(S* p, int64_t o, T v) { __builtin_nontemporal_store(f32_to_f16_bits(v), p + o); }
    __device__ static __forceinline__ void st_sr(S* p, int64_t o, T v, uint32_t r)
    {
        __builtin_nontemporal_store(f32_to_f16_sr_bits(v, r), p + o);
    }
};
template <class St>
constexpr bool is_16bit() { return sizeof(typename St::S) == 2; }

// Element e of a buffer addressed by an offset of type Off.  With Off =
// uint32_t the byte offset is formed in 32 bits (the buffer is < 4 GiB), so
// the access is a wave-uniform base plus a 32-bit per-lane offset.
template <class Off, class S>
__device__ __forceinline__ const S* at(const S* base, Off e)
{
    return (const S*)((const char*)base + (Off)(e * (Off)sizeof(S)));
}
template <class Off, class S>
__device__ __forceinline__ S* at(S* base, Off e)
{
    return (S*)((char*)base + (Off)(e * (Off)sizeof(S)));
}

// Boundary handling variants of the step kernel.
enum StepMode : int {
    MODE_FLUID = 0,  // no flags: every cell fluid
    MODE_REF = 1,    // flags, reference on-node swap at solid cells, collide all
    MODE_HALF = 2    // flags, half-way links, lid, Zou-He, solid carried
};

template <class T>
struct StepArgs {
    const void* fin;
    void* fout;
    const uint32_t* node;
    int w, h, pitch;
    int cpitch;          // row pitch of node
    int tiles_x;
    int64_t plane;
    uint32_t sr_key;     // stochastic rounding key of this step
    uint32_t cell_base;  // added to r*w + c to index the random bits
    Consts<T> k;
};

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD one
// contiguous range of tiles so that the halo rows a tile shares with its
// neighbours stay in the same L2.
__device__ __forceinline__ int xcd_tile(int bid, int nblocks)
{
    const int xcd = bid & 7, k = bid >> 3;
    const int per = nblocks >> 3, rem = nblocks & 7;
    return xcd * per + (xcd < rem ? xcd : rem) + k;
}

// ------------------------------------------------------------ reciprocal --
// 1 / b, bit-identical to the IEEE quotient T(1) / b for every input.  For
// |b| in [2^-126, 2^126) (biased exponent 1..252): the hardware reciprocal
// v_rcp_f32 (about 1 ulp) and one Newton-Raphson step with fused
// multiply-adds, y = y0 + y0 (1 - b y0), the FMA reciprocal iteration of
// P. W. Markstein, "Computation of elementary functions on the IBM RISC
// System/6000 processor", IBM J. Res. Dev. 34(1), 1990.  That this equals the
// correctly rounded reciprocal was checked on MI355X for every one of the
// 2 x 252 x 2^23 inputs of that range (cs_check_reciprocal repeats the check
// on the GPU at hand).  Zero, subnormal, huge (|b| >= 2^126: subnormal
// result), infinite and NaN inputs take the IEEE division, on a branch that
// only a wave holding such a value enters.  5 instructions on the common
// path against 11 for the IEEE sequence; the FP32 step kernels with many
// steps per launch are bound by their arithmetic, and this is worth 7-8%
// there.
template <class T>
__device__ __forceinline__ T recip_rn(T b)
{
    return T(1) / b;
}
template <>
__device__ __forceinline__ float recip_rn<float>(float b)
{
    const uint32_t e = (__builtin_bit_cast(uint32_t, b) >> 23) & 0xFFu;
    if (e - 1u < 252u) {
        const float y0 = __builtin_amdgcn_rcpf(b);
        return __builtin_fmaf(__builtin_fmaf(-b, y0, 1.0f), y0, y0);
    }
    return 1.0f / b;
}

// ------------------------------------------------------------ cell update --
// c_i . u for the nine directions, {0, ux, uy, -ux, -uy, ux + uy, -ux + uy,
// -ux - uy, ux - uy} as the reference sums them.  The last two are written as
// the negations of the two diagonals before them: rounding to nearest is
// symmetric, so (-ux) - uy == -(ux + uy) and ux - uy == -((-ux) + uy) exactly
// (an exact zero may change sign, which 1 + vc A and vc vc, the only uses,
// absorb), and the compiler then shares vc A, vc vc and (vc vc) B between
// the opposite diagonals: 8 fewer operations per cell update.
template <class T>
__device__ __forceinline__ void lattice_dots(T ux, T uy, T (&vc)[9])
{
#pragma clang fp contract(off)
    const T d1 = ux + uy, d2 = -ux + uy;
    vc[0] = T(0);
    vc[1] = ux;
    vc[2] = uy;
    vc[3] = -ux;
    vc[4] = -uy;
    vc[5] = d1;
    vc[6] = d2;
    vc[7] = -d1;
    vc[8] = -d2;
}

// Moments, equilibrium and BGK (lbm.rs:117-149, 43-71, 350-364) with the
// reference's operation order; terms that multiply by a zero lattice
// component are dropped, which leaves every value unchanged.
template <class T>
__device__ __forceinline__ void bgk_collide(const Consts<T>& k, T (&B)[9])
{
#pragma clang fp contract(off)
    const T rho = (((((((B[0] + B[1]) + B[2]) + B[3]) + B[4]) + B[5]) + B[6]) + B[7]) + B[8];
    const T mx = ((((B[1] - B[3]) + B[5]) - B[6]) - B[7]) + B[8];
    const T my = ((((B[2] - B[4]) + B[5]) + B[6]) - B[7]) - B[8];
    const T inv = recip_rn(rho);
    const T ux = inv * mx, uy = inv * my;
    const T v2 = ux * ux + uy * uy;
    T vc[9];
    lattice_dots(ux, uy, vc);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const T s = ((T(1) + vc[i] * k.A) + (vc[i] * vc[i]) * k.B) + v2 * k.C;
        const T feq = (rho * k.w(i)) * s;
        B[i] = B[i] + (B[i] - feq) * k.K;
    }
}

// TRT::evaluate (lbm.rs:401-444) of one cell: f_eq from the cell's moments,
// f_sw = swap_populations (the opposite direction; f_0 for i = 0) and
// f_eq_sw = swap_equilibrium, which keeps f_eq_0 and takes the opposite
// *populations* for i = 1..8 (lbm.rs:311-323), then
//   omega_i = ((f+ - f_eq+) wp + (f- - f_eq-) wm) (-dt / 2),  f_i' = f_i + omega_i
// with f+- = f_i +- f_sw_i and f_eq+- = f_eq_i +- f_eq_sw_i, in that order.
template <class T>
__device__ __forceinline__ void trt_collide(const Consts<T>& k, T (&B)[9])
{
#pragma clang fp contract(off)
    const T rho = (((((((B[0] + B[1]) + B[2]) + B[3]) + B[4]) + B[5]) + B[6]) + B[7]) + B[8];
    const T mx = ((((B[1] - B[3]) + B[5]) - B[6]) - B[7]) + B[8];
    const T my = ((((B[2] - B[4]) + B[5]) + B[6]) - B[7]) - B[8];
    const T inv = recip_rn(rho);
    const T ux = inv * mx, uy = inv * my;
    const T v2 = ux * ux + uy * uy;
    T vc[9];
    lattice_dots(ux, uy, vc);
    T feq[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const T s = ((T(1) + vc[i] * k.A) + (vc[i] * vc[i]) * k.B) + v2 * k.C;
        feq[i] = (rho * k.w(i)) * s;
    }
    T o[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const T fsw = B[OPP[i]];
        const T feqsw = i == 0 ? feq[0] : B[OPP[i]];
        const T fp = B[i] + fsw, fm = B[i] - fsw;
        const T feqp = feq[i] + feqsw, feqm = feq[i] - feqsw;
        o[i] = B[i] + ((fp - feqp) * k.wp + (fm - feqm) * k.wm) * k.hdt;
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) B[i] = o[i];
}

// Moments and equilibrium of one cell as Lattice::equilibrium computes them
// (lbm.rs:117-160, 43-71).
template <class T>
__device__ __forceinline__ void cell_equilibrium(const Consts<T>& k, const T (&B)[9], T& rho, T& ux, T& uy,
                                                 T (&feq)[9])
{
#pragma clang fp contract(off)
    rho = (((((((B[0] + B[1]) + B[2]) + B[3]) + B[4]) + B[5]) + B[6]) + B[7]) + B[8];
    const T mx = ((((B[1] - B[3]) + B[5]) - B[6]) - B[7]) + B[8];
    const T my = ((((B[2] - B[4]) + B[5]) + B[6]) - B[7]) - B[8];
    const T inv = recip_rn(rho);
    ux = inv * mx;
    uy = inv * my;
    const T v2 = ux * ux + uy * uy;
    T vc[9];
    lattice_dots(ux, uy, vc);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const T s = ((T(1) + vc[i] * k.A) + (vc[i] * vc[i]) * k.B) + v2 * k.C;
        feq[i] = (rho * k.w(i)) * s;
    }
}

// Regularized::evaluate (lbm.rs:606-661): f_eq plus the projection of the
// non-equilibrium stress on the Hermite tensors, in the reference's order
//   P_ab = sum_i f_neq_i c_ia c_ib,  Q_i,ab = c_ia c_ib - cs^2 delta_ab,
//   reg_i = (((f_eq_i + P_xx (Q_xx sf)) + P_xy (Q_xy sf)) + P_yx (Q_yx sf)) + P_yy (Q_yy sf),
// sf = w_i / (2 cs^4).  The wrapped operator is never called (as in the reference).
template <class T>
__device__ __forceinline__ void reg_collide(const Consts<T>& k, T (&B)[9])
{
#pragma clang fp contract(off)
    T rho, ux, uy, feq[9];
    cell_equilibrium(k, B, rho, ux, uy, feq);
    T pxx = T(0), pxy = T(0), pyx = T(0), pyy = T(0);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const T fneq = B[i] - feq[i];
        const T cx = T(CX[i]), cy = T(CY[i]);
        pxx = pxx + fneq * (cx * cx);
        pxy = pxy + fneq * (cx * cy);
        pyx = pyx + fneq * (cy * cx);
        pyy = pyy + fneq * (cy * cy);
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const T cx = T(CX[i]), cy = T(CY[i]);
        const T sf = k.sfq[Consts<T>::wclass(i)];
        T reg = feq[i];
        reg = reg + pxx * ((cx * cx - k.cs2) * sf);
        reg = reg + pxy * ((cx * cy) * sf);
        reg = reg + pyx * ((cy * cx) * sf);
        reg = reg + pyy * ((cy * cy - k.cs2) * sf);
        B[i] = reg;
    }
}

// KBC::evaluate (lbm.rs:468-572) of one cell, every expression in the
// reference's order (its diagnostic print of an entropy sum is left out):
//   pi~ = sum_i f_i dx^2 - u v,  n~ = v^2 - u^2,
//   ds_0 = ((8uv) pi~ + n~ n~) rho 0.5,
//   ds_1,3 = ((((u dx - n~) + 1) n~ - (v (4dx) + 8uv) pi~) rho) 0.25,
//   ds_2,4 = ((((v (-dx) - n~) - 1) n~ - (u (4dx) + 8uv) pi~) rho) 0.25,
//   ds_5..8 = ((((8uv + u (4dx)) + 2dx^2) pi~ + (n~ - (v - u) dx) n~) rho) 0.125,
//   dh_i = (f_i - f_eq_i) - ds_i,
//   gamma = -(((sum ds dh / f_eq) / (sum dh dh / f_eq)) (2 - 1/beta) + (-1/beta)),
//   f_i' = f_i + (ds_i (2 (-beta)) + (dh_i gamma) (-beta)).
template <class T>
__device__ __forceinline__ void kbc_collide(const Consts<T>& k, T (&B)[9])
{
#pragma clang fp contract(off)
    T rho, u, v, feq[9];
    cell_equilibrium(k, B, rho, u, v, feq);
    const T dx = k.dx;
    const T uv = u * v, u2 = u * u, v2 = v * v;
    const T dx2 = dx * dx;
    T temp = T(0);
#pragma unroll
    for (int i = 0; i < 9; ++i) temp = temp + B[i] * dx2;
    const T pit = temp - uv;
    const T nt = v2 - u2;
    const T uv8 = uv * T(8);
    const T ds0 = (((uv8 * pit) + nt * nt) * rho) * T(0.5);
    const T ds13 = (((((u * dx - nt) + T(1)) * nt) - ((v * (dx * T(4))) + uv8) * pit) * rho) * T(0.25);
    const T ds24 = (((((v * (-dx) - nt) + T(-1)) * nt) - ((u * (dx * T(4))) + uv8) * pit) * rho) * T(0.25);
    const T ds58 = (((((uv8 + u * (T(4) * dx)) + (T(2) * dx) * dx) * pit) + ((nt - (v - u) * dx) * nt)) * rho) *
                   T(0.125);
    const T ds[9] = {ds0, ds13, ds24, ds13, ds24, ds58, ds58, ds58, ds58};
    T dh[9];
    T num = T(0), den = T(0);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        dh[i] = (B[i] - feq[i]) - ds[i];
        num = num + (ds[i] * dh[i]) / feq[i];
        den = den + (dh[i] * dh[i]) / feq[i];
    }
    const T beta = k.beta;
    const T gamma = (((num / den) * (T(2) - T(1) / beta)) + (-T(1) / beta)) * T(-1);
    const T m2b = T(2) * (-beta);
#pragma unroll
    for (int i = 0; i < 9; ++i) B[i] = B[i] + (ds[i] * m2b + (dh[i] * gamma) * (-beta));
}

// Deviation arithmetic: with g_i = f_i - w_i rho0,
//   drho = sum g_i, rho = rho0 + drho, m = sum g_i c_i, u = m / rho,
//   g_eq_i = w_i (drho + rho ((vc A + vc^2 B) + v2 C)),  g'_i = g_i + (g_i - g_eq_i) K,
// the same BGK update as bgk_collide without carrying the O(1) rest
// populations through the rounding.
template <class T>
__device__ __forceinline__ void bgk_collide_dev(const Consts<T>& k, T (&B)[9])
{
#pragma clang fp contract(off)
    const T drho = (((((((B[0] + B[1]) + B[2]) + B[3]) + B[4]) + B[5]) + B[6]) + B[7]) + B[8];
    const T rho = k.rho0 + drho;
    const T mx = ((((B[1] - B[3]) + B[5]) - B[6]) - B[7]) + B[8];
    const T my = ((((B[2] - B[4]) + B[5]) + B[6]) - B[7]) - B[8];
    const T inv = recip_rn(rho);
    const T ux = inv * mx, uy = inv * my;
    const T v2 = ux * ux + uy * uy;
    T vc[9];
    lattice_dots(ux, uy, vc);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const T s = (vc[i] * k.A + (vc[i] * vc[i]) * k.B) + v2 * k.C;
        const T geq = k.w(i) * (drho + rho * s);
        B[i] = B[i] + (B[i] - geq) * k.K;
    }
}

// Zou-He on deviations: the unknown populations follow the same formulas
// (opposite directions share w_i); only the density sums carry rho0.
template <class T>
__device__ __forceinline__ void zou_he_inlet_dev(T (&B)[9], T rho0, T ux, T uy)
{
#pragma clang fp contract(off)
    const T G = ((B[0] + B[1]) + B[3]) + T(2) * ((B[4] + B[7]) + B[8]);
    const T rho = (rho0 + G) / (T(1) - uy);
    const T ru_t = rho * ux, ru_a = rho * uy;
    const T d13 = B[1] - B[3];
    const T two3 = T(2) / T(3), six1 = T(1) / T(6), half = T(0.5);
    B[2] = B[4] + two3 * ru_a;
    B[5] = ((B[7] - half * d13) + half * ru_t) + six1 * ru_a;
    B[6] = ((B[8] + half * d13) - half * ru_t) + six1 * ru_a;
}

template <class T>
__device__ __forceinline__ void zou_he_outlet_dev(T (&B)[9], T rho0, T rho)
{
#pragma clang fp contract(off)
    const T G = ((B[0] + B[1]) + B[3]) + T(2) * ((B[2] + B[5]) + B[6]);
    const T uy = ((rho0 - rho) + G) / rho;
    const T ru_a = rho * uy;
    const T d13 = B[1] - B[3];
    const T two3 = T(2) / T(3), six1 = T(1) / T(6), half = T(0.5);
    B[4] = B[2] - two3 * ru_a;
    B[7] = (B[5] + half * d13) - six1 * ru_a;
    B[8] = (B[6] - half * d13) - six1 * ru_a;
}

// Zou-He velocity inlet (low-column edge; unknown populations 2, 5, 6).
template <class T>
__device__ __forceinline__ void zou_he_inlet(T (&B)[9], T ux, T uy)
{
#pragma clang fp contract(off)
    const T rho = (((B[0] + B[1]) + B[3]) + T(2) * ((B[4] + B[7]) + B[8])) / (T(1) - uy);
    const T ru_t = rho * ux, ru_a = rho * uy;
    const T d13 = B[1] - B[3];
    const T two3 = T(2) / T(3), six1 = T(1) / T(6), half = T(0.5);
    B[2] = B[4] + two3 * ru_a;
    B[5] = ((B[7] - half * d13) + half * ru_t) + six1 * ru_a;
    B[6] = ((B[8] + half * d13) - half * ru_t) + six1 * ru_a;
}

// Zou-He density outlet (high-column edge; unknown populations 4, 7, 8).
template <class T>
__device__ __forceinline__ void zou_he_outlet(T (&B)[9], T rho)
{
#pragma clang fp contract(off)
    const T uy = (((B[0] + B[1]) + B[3]) + T(2) * ((B[2] + B[5]) + B[6])) / rho - T(1);
    const T ru_a = rho * uy;
    const T d13 = B[1] - B[3];
    const T two3 = T(2) / T(3), six1 = T(1) / T(6), half = T(0.5);
    B[4] = B[2] - two3 * ru_a;
    B[7] = (B[5] + half * d13) - six1 * ru_a;
    B[8] = (B[6] - half * d13) - six1 * ru_a;
}

// Gathers the nine post-stream, post-boundary populations of cell (r, c).
// Out-of-range (r, c) lanes read a clamped in-range address and are masked by
// the caller.  Returns the cell's own flag (0 without flags).  Off is the
// offset type (uint32_t when the 9-plane buffer is < 4 GiB, else int64_t).
template <class St, int EDGE, int MODE, class Off = int64_t>
__device__ __forceinline__ int gather(const typename St::S* __restrict__ fin, const uint32_t* __restrict__ node,
                                      int w, int h, int pitch, int cpitch, int64_t P,
                                      const Consts<typename St::T>& k, int r, int c, typename St::T (&B)[9])
{
    using T = typename St::T;
    int rows[3], cols[3];
    bool vr[3], vc[3];
    if (EDGE == CS_EDGE_PERIODIC) {
        rows[0] = r == 0 ? h - 1 : r - 1;
        rows[2] = r == h - 1 ? 0 : r + 1;
        cols[0] = c == 0 ? w - 1 : c - 1;
        cols[2] = c == w - 1 ? 0 : c + 1;
        vr[0] = vr[2] = vc[0] = vc[2] = true;
    } else {
        vr[0] = r > 0;
        vr[2] = r < h - 1;
        vc[0] = c > 0;
        vc[2] = c < w - 1;
        rows[0] = vr[0] ? r - 1 : r;
        rows[2] = vr[2] ? r + 1 : r;
        cols[0] = vc[0] ? c - 1 : c;
        cols[2] = vc[2] ? c + 1 : c;
    }
    rows[1] = r;
    cols[1] = c;
    vr[1] = vc[1] = true;
    const Off Pl = (Off)P;
    const Off own = (Off)r * (Off)pitch + (Off)c;

    uint32_t info = 0;
    int type = CS_FLUID;
    if (MODE != MODE_FLUID) {
        info = node[(Off)r * (Off)cpitch + (Off)c];
        type = (int)(info >> NODE_TYPE_SHIFT);
    }
    const bool solid = (type == CS_WALL || type == CS_LID);

#pragma unroll
    for (int i = 0; i < 9; ++i) {
        // Normal pull source of direction i.
        Off off = (Off)i * Pl + (Off)rows[CX[i] + 1] * (Off)pitch + (Off)cols[1 - CY[i]];
        bool valid = vr[CX[i] + 1] && vc[1 - CY[i]];
        bool add_lid = false;
        if (MODE == MODE_REF) {
            // Solid cell: take the opposite population after streaming
            // (lbm.rs:745), i.e. the normal source of direction opp(i).
            if (solid) {
                const int j = OPP[i];
                off = (Off)j * Pl + (Off)rows[CX[j] + 1] * (Off)pitch + (Off)cols[1 - CY[j]];
                valid = vr[CX[j] + 1] && vc[1 - CY[j]];
            }
        } else if (MODE == MODE_HALF) {
            if (solid) {
                off = (Off)i * Pl + own;  // carried unchanged (copied raw in finish)
                valid = true;
            } else if ((info >> i) & 1u) {
                off = (Off)OPP[i] * Pl + own;  // half-way bounce-back
                valid = true;
                add_lid = (info >> (9 + i)) & 1u;
            }
        }
        // A population from outside a zero-filled edge is f = 0, i.e. g = -w_i rho0.
        T v = St::ld(at<Off>(fin, off), 0);
        v = valid ? v : (St::DEV ? -k.base(i) : T(0));
        if (MODE == MODE_HALF) v = add_lid ? v + k.lid(i) : v;
        B[i] = v;
    }
    if (MODE == MODE_HALF) {
        if constexpr (St::DEV) {
            if (type == CS_INLET) zou_he_inlet_dev<T>(B, k.rho0, k.in_ux, k.in_uy);
            else if (type == CS_OUTLET) zou_he_outlet_dev<T>(B, k.rho0, k.out_rho);
        } else {
            if (type == CS_INLET) zou_he_inlet<T>(B, k.in_ux, k.in_uy);
            else if (type == CS_OUTLET) zou_he_outlet<T>(B, k.out_rho);
        }
    }
    return type;
}

// Collides and stores cell (r, c).  Solid cells of the half-way mode are
// carried unchanged: their stored values are copied bit for bit.  With SR the
// 16-bit narrowing is stochastic, keyed by (sr_key, cell index).
template <class St, int MODE, bool SR, class Off = int64_t, int COLL = CS_COLL_BGK>
__device__ __forceinline__ void finish(const typename St::S* __restrict__ fin, typename St::S* __restrict__ fout,
                                       int pitch, int64_t P, const Consts<typename St::T>& k, int r, int c,
                                       int type, typename St::T (&B)[9], uint32_t sr_key, uint32_t cell)
{
    const Off Pl = (Off)P;
    const Off own = (Off)r * (Off)pitch + (Off)c;
    if (MODE == MODE_HALF && (type == CS_WALL || type == CS_LID)) {
#pragma unroll
        for (int i = 0; i < 9; ++i) *at<Off>(fout, (Off)i * Pl + own) = *at<Off>(fin, (Off)i * Pl + own);
        return;
    }
    if constexpr (COLL != CS_COLL_BGK) static_assert(!St::DEV, "TRT, Regularized and KBC run in the reference arithmetic");
    if constexpr (COLL == CS_COLL_TRT) {
        trt_collide(k, B);
    } else if constexpr (COLL == CS_COLL_REGULARIZED) {
        reg_collide(k, B);
    } else if constexpr (COLL == CS_COLL_KBC) {
        kbc_collide(k, B);
    } else if constexpr (St::DEV) {
        bgk_collide_dev(k, B);
    } else {
        bgk_collide(k, B);
    }
    if constexpr (SR) {
        uint32_t h = lowbias32(cell ^ sr_key);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            St::st_sr(at<Off>(fout, (Off)(2 * j) * Pl + own), 0, B[2 * j], h);
            St::st_sr(at<Off>(fout, (Off)(2 * j + 1) * Pl + own), 0, B[2 * j + 1], h << 16);
            h = xorshift32(h);
        }
        St::st_sr(at<Off>(fout, (Off)8 * Pl + own), 0, B[8], h);
    } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) St::st(at<Off>(fout, (Off)i * Pl + own), 0, B[i]);
    }
}

// Fused step.  A block is 4 waves; wave y handles row 4*ty + y, lane l the
// columns 64*CPT*tx + l + 64*j, j < CPT.  All loads of a thread's CPT cells
// are issued before any arithmetic.
template <class St, int EDGE, int MODE, int CPT, bool SR, class Off, int COLL = CS_COLL_BGK>
__global__ __launch_bounds__(256) void k_step(StepArgs<typename St::T> a)
{
    using S = typename St::S;
    using T = typename St::T;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int ty = tile / a.tiles_x;
    const int tx = tile - ty * a.tiles_x;
    const int r = ty * 4 + (threadIdx.x >> 6);
    if (r >= a.h) return;
    const int c0 = tx * (64 * CPT) + (threadIdx.x & 63);
    const S* __restrict__ fin = (const S*)a.fin;
    S* __restrict__ fout = (S*)a.fout;
    T B[CPT][9];
    int type[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int c = c0 + 64 * j;
        const int cc = c < a.w ? c : a.w - 1;
        type[j] = gather<St, EDGE, MODE, Off>(fin, a.node, a.w, a.h, a.pitch, a.cpitch, a.plane, a.k, r, cc, B[j]);
    }
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int c = c0 + 64 * j;
        if (c < a.w)
            finish<St, MODE, SR, Off, COLL>(fin, fout, a.pitch, a.plane, a.k, r, c, type[j], B[j], a.sr_key,
                                      a.cell_base + (uint32_t)(r * a.w + c));
    }
}

// ----------------------------------------------------------- vector step --
// A lane owns V consecutive cells of a row, so every population plane is read
// and written with one 8- or 16-byte access per lane (non-temporal loads and
// stores: every element is used once).  The populations that stream along the
// row (cy = +-1) come from the lane's own vector shifted by one cell; the cell
// that crosses into the neighbouring lane is exchanged with a lane shuffle,
// and the first and last lane of a wave (and the last active lane of a row)
// read their outside cell directly.  The per-cell arithmetic and the stored
// values are those of gather/finish, so the results equal k_step bit for bit.
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));
template <int VB>
using RawVec = std::conditional_t<VB == 16, u32x4_t, std::conditional_t<VB == 8, u32x2_t, uint32_t>>;
template <class S, int VB>
union VecN {
    RawVec<VB> raw;
    S e[VB / sizeof(S)];
};

// Neighbour-lane exchange of a row vector's edge element by DPP wave shifts
// (one VALU move; no LDS crossbar).  lane_prev: each lane receives lane - 1's
// value, lane_next: lane + 1's; lane 0 / lane 63 receive 0 (the callers
// replace those with a direct load).  The bound control bit writes that 0
// itself (also for a source lane switched off in EXEC), so no register has to
// be cleared before each shift.
__device__ __forceinline__ uint32_t dpp_prev(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x138 /* wave_shr:1 */, 0xF, 0xF, true);
}
__device__ __forceinline__ uint32_t dpp_next(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x130 /* wave_shl:1 */, 0xF, 0xF, true);
}
template <class S, bool NEXT>
__device__ __forceinline__ S lane_shift(S v)
{
    if constexpr (sizeof(S) == 8) {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        const uint32_t lo = NEXT ? dpp_next((uint32_t)u) : dpp_prev((uint32_t)u);
        const uint32_t hi = NEXT ? dpp_next((uint32_t)(u >> 32)) : dpp_prev((uint32_t)(u >> 32));
        return __builtin_bit_cast(S, ((uint64_t)hi << 32) | lo);
    } else if constexpr (sizeof(S) == 4) {
        const uint32_t u = __builtin_bit_cast(uint32_t, v);
        return __builtin_bit_cast(S, NEXT ? dpp_next(u) : dpp_prev(u));
    } else {
        const uint32_t u = (uint32_t)v;
        return (S)(NEXT ? dpp_next(u) : dpp_prev(u));
    }
}

// One row per wave (ROWS rows per block), 64*V columns per wave, VB bytes
// per lane and plane (V = VB / sizeof(S) cells).  w % V == 0.
//
// With geometry (MODE_REF, MODE_HALF) the lane reads its V node words as one
// vector after the fluid gather.  MODE_REF: a solid cell takes the pulled
// value of the opposite direction (lbm.rs:745), i.e. the pulled pairs swap.
// MODE_HALF: a bounce-back link takes the cell's own pre-stream population of
// the opposite direction (+ the lid term), a solid cell is carried unchanged,
// then Zou-He at inlet/outlet cells -- gather()'s order.  The own values of
// the planes that stream within the row (cx = 0) are the lane's loaded
// vectors; the other six planes are read at the own row only by lanes that
// hold a link or a solid cell.
template <class St, int EDGE, bool SR, int VB = 16, int ROWS = 4, int COLL = CS_COLL_BGK, int MODE = MODE_FLUID>
__global__ __launch_bounds__(64 * ROWS) void k_step_vec(StepArgs<typename St::T> a)
{
    using S = typename St::S;
    using T = typename St::T;
    constexpr int V = VB / (int)sizeof(S);
    static_assert(V >= 2, "at least two cells per lane");
    static_assert(MODE == MODE_FLUID || !SR, "geometry in the vector kernel: 32/64-bit storage only");
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int ty = tile / a.tiles_x;
    const int tx = tile - ty * a.tiles_x;
    const int r = ty * ROWS + (threadIdx.x >> 6);
    if (r >= a.h) return;  // uniform over the wave
    const int lane = threadIdx.x & 63;
    const int c0 = (tx * 64 + lane) * V;
    const bool active = c0 < a.w;
    const int cl = active ? c0 : a.w - V;  // in-range address for inactive lanes
    const bool last = active && (lane == 63 || c0 + V >= a.w);
    const S* __restrict__ fin = (const S*)a.fin;
    S* __restrict__ fout = (S*)a.fout;
    const uint32_t P = (uint32_t)a.plane;

    int rows[3];
    bool vr[3];
    if (EDGE == CS_EDGE_PERIODIC) {
        rows[0] = r == 0 ? a.h - 1 : r - 1;
        rows[2] = r == a.h - 1 ? 0 : r + 1;
        vr[0] = vr[2] = true;
    } else {
        vr[0] = r > 0;
        vr[2] = r < a.h - 1;
        rows[0] = vr[0] ? r - 1 : r;
        rows[2] = vr[2] ? r + 1 : r;
    }
    rows[1] = r;
    vr[1] = true;
    // columns of the cells just outside this lane's vector
    const int cleft = c0 - 1, cright = c0 + V;
    const bool vleft = EDGE == CS_EDGE_PERIODIC || cleft >= 0;
    const bool vright = EDGE == CS_EDGE_PERIODIC || cright < a.w;
    const int cleft_w = cleft < 0 ? a.w - 1 : cleft;
    const int cright_w = cright >= a.w ? 0 : cright;

    T B[V][9];
    S ownv[9][V];  // MODE_HALF: pre-stream values of the lane's own cells
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const int rs = rows[CX[i] + 1];
        const bool rvalid = vr[CX[i] + 1];
        const uint32_t rowbase = (uint32_t)i * P + (uint32_t)rs * (uint32_t)a.pitch;
        VecN<S, VB> v;
        v.raw = __builtin_nontemporal_load((const RawVec<VB>*)at<uint32_t>(fin, rowbase + (uint32_t)cl));
        if constexpr (MODE == MODE_HALF) {
            if (CX[i] == 0) {
#pragma unroll
                for (int j = 0; j < V; ++j) ownv[i][j] = v.e[j];
            }
        }
        const T outside = St::DEV ? -a.k.base(i) : T(0);
        if (CY[i] == 0) {
#pragma unroll
            for (int j = 0; j < V; ++j) B[j][i] = rvalid ? St::wid(v.e[j]) : outside;
        } else if (CY[i] == 1) {
            // source column c - 1: the previous cell, across the lane boundary for j = 0
            S left = lane_shift<S, false>(v.e[V - 1]);
            if (lane == 0 && active) left = *at<uint32_t>(fin, rowbase + (uint32_t)cleft_w);
            const bool lvalid = rvalid && (lane != 0 || vleft);
            B[0][i] = lvalid ? St::wid(left) : outside;
#pragma unroll
            for (int j = 1; j < V; ++j) B[j][i] = rvalid ? St::wid(v.e[j - 1]) : outside;
        } else {
            // source column c + 1: the next cell, across the lane boundary for j = V - 1
            S right = lane_shift<S, true>(v.e[0]);
            if (last) right = *at<uint32_t>(fin, rowbase + (uint32_t)cright_w);
            const bool rtvalid = rvalid && (!last || vright);
#pragma unroll
            for (int j = 0; j < V - 1; ++j) B[j][i] = rvalid ? St::wid(v.e[j + 1]) : outside;
            B[V - 1][i] = rtvalid ? St::wid(right) : outside;
        }
    }
    if (!active) return;
    bool carried[V];
#pragma unroll
    for (int j = 0; j < V; ++j) carried[j] = false;
    if constexpr (MODE != MODE_FLUID) {
        // cpitch is a multiple of V and c0 % V == 0: one aligned vector load
        VecN<uint32_t, 4 * V> nv;
        nv.raw = *(const RawVec<4 * V>*)(a.node + (uint32_t)r * (uint32_t)a.cpitch + (uint32_t)c0);
        if constexpr (MODE == MODE_REF) {
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const int type = (int)(nv.e[j] >> NODE_TYPE_SHIFT);
                if (type == CS_WALL || type == CS_LID) {
#pragma unroll
                    for (int i = 1; i < 9; ++i) {
                        if (OPP[i] > i) {
                            const T t = B[j][i];
                            B[j][i] = B[j][OPP[i]];
                            B[j][OPP[i]] = t;
                        }
                    }
                }
            }
        } else {
            bool need = false;
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const int type = (int)(nv.e[j] >> NODE_TYPE_SHIFT);
                need = need || type == CS_WALL || type == CS_LID || (nv.e[j] & 0x1FFu) != 0u;
            }
            if (need) {
#pragma unroll
                for (int i = 0; i < 9; ++i) {
                    if (CX[i] != 0) {
                        VecN<S, VB> v;
                        v.raw = *(const RawVec<VB>*)at<uint32_t>(fin, (uint32_t)i * P + (uint32_t)r * (uint32_t)a.pitch +
                                                                           (uint32_t)c0);
#pragma unroll
                        for (int j = 0; j < V; ++j) ownv[i][j] = v.e[j];
                    }
                }
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    const uint32_t info = nv.e[j];
                    const int type = (int)(info >> NODE_TYPE_SHIFT);
                    if (type == CS_WALL || type == CS_LID) {
                        carried[j] = true;
#pragma unroll
                        for (int i = 0; i < 9; ++i) B[j][i] = St::wid(ownv[i][j]);
                    } else {
#pragma unroll
                        for (int i = 0; i < 9; ++i) {
                            if ((info >> i) & 1u) {
                                const T v = St::wid(ownv[OPP[i]][j]);
                                B[j][i] = ((info >> (9 + i)) & 1u) ? v + a.k.lid(i) : v;
                            }
                        }
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const int type = (int)(nv.e[j] >> NODE_TYPE_SHIFT);
                if constexpr (St::DEV) {
                    if (type == CS_INLET) zou_he_inlet_dev<T>(B[j], a.k.rho0, a.k.in_ux, a.k.in_uy);
                    else if (type == CS_OUTLET) zou_he_outlet_dev<T>(B[j], a.k.rho0, a.k.out_rho);
                } else {
                    if (type == CS_INLET) zou_he_inlet<T>(B[j], a.k.in_ux, a.k.in_uy);
                    else if (type == CS_OUTLET) zou_he_outlet<T>(B[j], a.k.out_rho);
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < V; ++j) {
        if (carried[j]) continue;  // MODE_HALF solid cell: stored as read
        if constexpr (COLL == CS_COLL_TRT) trt_collide(a.k, B[j]);
        else if constexpr (COLL == CS_COLL_REGULARIZED) reg_collide(a.k, B[j]);
        else if constexpr (COLL == CS_COLL_KBC) kbc_collide(a.k, B[j]);
        else if constexpr (St::DEV) bgk_collide_dev(a.k, B[j]);
        else bgk_collide(a.k, B[j]);
    }
    const uint32_t own = (uint32_t)r * (uint32_t)a.pitch + (uint32_t)c0;
    if constexpr (SR) {
        // the bits of finish(): h = lowbias32(cell ^ key), two directions per word
        S o[9][V];
#pragma unroll
        for (int j = 0; j < V; ++j) {
            uint32_t h = lowbias32((a.cell_base + (uint32_t)(r * a.w + c0 + j)) ^ a.sr_key);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                o[2 * q][j] = St::nar_sr(B[j][2 * q], h);
                o[2 * q + 1][j] = St::nar_sr(B[j][2 * q + 1], h << 16);
                h = xorshift32(h);
            }
            o[8][j] = St::nar_sr(B[j][8], h);
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            VecN<S, VB> v;
#pragma unroll
            for (int j = 0; j < V; ++j) v.e[j] = o[i][j];
            __builtin_nontemporal_store(v.raw, (RawVec<VB>*)at<uint32_t>(fout, (uint32_t)i * P + own));
        }
    } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            VecN<S, VB> v;
#pragma unroll
            for (int j = 0; j < V; ++j) v.e[j] = St::nar(B[j][i]);
            __builtin_nontemporal_store(v.raw, (RawVec<VB>*)at<uint32_t>(fout, (uint32_t)i * P + own));
        }
    }
}

}  // namespace cs
