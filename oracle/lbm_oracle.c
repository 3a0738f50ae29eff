/*
 * lbm_oracle.c -- CPU restatement of taktoa/chemsim's D2Q9 BGK hot path
 * (src/lbm.rs), used as the parity reference.  TEST INFRASTRUCTURE ONLY: the
 * product library never links it.  Build: oracle/Makefile (-ffp-contract=off).
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* D2Q9::directions, lbm.rs:221-231: c_i = (cx, cy). */
static const int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
/* D2Q9::swap_populations, lbm.rs:298-309. */
static const int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
/* Exact weights of the extension terms (Ladd moving wall). */
static const double W_EXACT[9] = {4.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9,
                                  1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};

static inline int is_solid(int t) { return t == ORC_WALL || t == ORC_LID; }

/* Constants in f32 exactly as lbm.rs evaluates them in Scalar = f32. */
void orc_consts_f32_make(const orc_params* p, orc_consts_f32* k)
{
    volatile float dx = (float)p->delta_x, dt = (float)p->delta_t, tau = (float)p->tau;
    volatile float three = 3.0f;
    float cs = dx / (sqrtf(three) * dt);           /* lbm.rs:84 */
    float cs2 = cs * cs;                           /* lbm.rs:55 */
    float cs4 = cs2 * cs2;                         /* lbm.rs:56 */
    k->cs2 = cs2;
    k->A = 1.0f / cs2;                             /* lbm.rs:64 */
    k->B = 1.0f / (2.0f * cs4);                    /* lbm.rs:65 */
    k->C = -1.0f / (2.0f * cs2);                   /* lbm.rs:66 */
    k->K = -dt / tau;                              /* lbm.rs:357 */
    volatile float tau_m = (float)p->tau_minus;
    k->wp = 1.0f / tau;                            /* lbm.rs:431-432, tau = tau_plus */
    k->wm = 1.0f / tau_m;
    k->hdt = -dt * 0.5f;                           /* lbm.rs:440 */
    k->dx = dx;                                    /* lbm.rs:478 */
    volatile float nu = (float)p->ks_viscosity;
    float cs_k = dx / (sqrtf(three) * dt);         /* isothermal_speed_of_sound, lbm.rs:548 */
    k->beta = 1.0f / ((2.0f * nu / (cs_k * cs_k)) + 1.0f); /* lbm.rs:549 */
    volatile float w0n = 16.0f, w1n = 4.0f, w5n = 1.0f, d = 36.0f; /* lbm.rs:209-219 */
    float w0 = w0n / d, w1 = w1n / d, w5 = w5n / d;
    const float ws[9] = {w0, w1, w1, w1, w1, w5, w5, w5, w5};
    for (int i = 0; i < 9; ++i) {
        k->w[i] = ws[i];
        k->sf[i] = ws[i] / (2.0f * cs4);           /* lbm.rs:651 */
        k->lid[i] = (float)(6.0 * W_EXACT[i] * (CX[i] * p->lid_ux + CY[i] * p->lid_uy));
        k->base[i] = ws[i] * (float)p->rho0;
    }
}

void orc_consts_f64_make(const orc_params* p, orc_consts_f64* k)
{
    volatile double dx = p->delta_x, dt = p->delta_t, tau = p->tau, three = 3.0;
    double cs = dx / (sqrt(three) * dt);
    double cs2 = cs * cs;
    double cs4 = cs2 * cs2;
    k->cs2 = cs2;
    k->A = 1.0 / cs2;
    k->B = 1.0 / (2.0 * cs4);
    k->C = -1.0 / (2.0 * cs2);
    k->K = -dt / tau;
    volatile double tau_m = p->tau_minus;
    k->wp = 1.0 / tau;
    k->wm = 1.0 / tau_m;
    k->hdt = -dt * 0.5;
    k->dx = dx;
    volatile double nu = p->ks_viscosity;
    double cs_k = dx / (sqrt(three) * dt);
    k->beta = 1.0 / ((2.0 * nu / (cs_k * cs_k)) + 1.0);
    volatile double w0n = 16.0, w1n = 4.0, w5n = 1.0, d = 36.0;
    double w0 = w0n / d, w1 = w1n / d, w5 = w5n / d;
    const double ws[9] = {w0, w1, w1, w1, w1, w5, w5, w5, w5};
    for (int i = 0; i < 9; ++i) {
        k->w[i] = ws[i];
        k->sf[i] = ws[i] / (2.0 * cs4);
        k->lid[i] = 6.0 * W_EXACT[i] * (CX[i] * p->lid_ux + CY[i] * p->lid_uy);
        k->base[i] = ws[i] * p->rho0;
    }
}

void orc_trt_new(float lambda, float ks_viscosity, float delta_x, float delta_t, float* tau_plus,
                 float* tau_minus)
{
    volatile float dx = delta_x, dt = delta_t, three = 3.0f;
    float cs = dx / (sqrtf(three) * dt);
    float tp = dt * ((ks_viscosity / (cs * cs)) + 0.5f);   /* lbm.rs:386 */
    float tm = dt * ((lambda / ((tp / dt) - 0.5f)) + 0.5f); /* lbm.rs:387 */
    *tau_plus = tp;
    *tau_minus = tm;
}

static void orc_consts_make_body_f32(const orc_params* p, orc_consts_f32* k) { orc_consts_f32_make(p, k); }
static void orc_consts_make_body_f64(const orc_params* p, orc_consts_f64* k) { orc_consts_f64_make(p, k); }

#define REAL float
#define SFX f32
#define KT orc_consts_f32
#include "lbm_oracle_impl.inc"
#undef REAL
#undef SFX
#undef KT

#define REAL double
#define SFX f64
#define KT orc_consts_f64
#include "lbm_oracle_impl.inc"
#undef REAL
#undef SFX
#undef KT

/* ---------------------------------------------------------------- runs --- */

int orc_run_f32_omp(const orc_params* p, const uint8_t* flags, float* f, int nsteps, int threads)
{
    const long n = 9L * p->w * p->h;
    float* a = f;
    float* b = (float*)malloc(sizeof(float) * n);
    if (!b) return -1;
    int used = 1;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
    used = threads;
#else
    (void)threads;
#endif
    for (int s = 0; s < nsteps; ++s) {
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(static)
#endif
        for (int r = 0; r < p->h; ++r) orc_step_rows_f32(p, flags, a, b, r, r + 1);
        float* t = a;
        a = b;
        b = t;
    }
    if (a != f) {
        memcpy(f, a, sizeof(float) * n);
        free(a);
    } else {
        free(b);
    }
    return used;
}

int orc_run_unfused_f32(const orc_params* p, const uint8_t* flags, float* f, int nsteps)
{
    const long n = (long)p->w * p->h;
    float* scratch = (float*)malloc(sizeof(float) * 21 * n);
    float* out = (float*)malloc(sizeof(float) * 9 * n);
    if (!scratch || !out) {
        free(scratch);
        free(out);
        return -1;
    }
    for (int s = 0; s < nsteps; ++s) {
        orc_step_unfused_f32(p, flags, f, out, scratch);
        memcpy(f, out, sizeof(float) * 9 * n);
    }
    free(scratch);
    free(out);
    return 0;
}

/* ------------------------------------------------------ 16-bit storage --- */

static inline uint32_t f32_bits(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static inline float bits_f32(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

uint16_t orc_f32_to_bf16(float x)
{
    uint32_t u = f32_bits(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x0040u); /* quiet NaN */
    u += 0x7fffu + ((u >> 16) & 1u);                                          /* nearest even */
    return (uint16_t)(u >> 16);
}

float orc_bf16_to_f32(uint16_t b) { return bits_f32((uint32_t)b << 16); }

/* IEEE binary16, round to nearest even, subnormals kept, overflow to inf. */
uint16_t orc_f32_to_f16(float x)
{
    uint32_t u = f32_bits(x);
    uint16_t sign = (uint16_t)((u >> 16) & 0x8000u);
    uint32_t a = u & 0x7fffffffu;
    if (a > 0x7f800000u) return sign | 0x7e00u;            /* NaN */
    if (a >= 0x477ff000u) return sign | 0x7c00u;           /* >= 65520 rounds to inf */
    if (a >= 0x38800000u) {                                /* normal half (>= 2^-14) */
        uint32_t mant = a & 0x7fffffu;
        int e = (int)(a >> 23) - 127 + 15;
        uint32_t r = ((uint32_t)e << 10) | (mant >> 13);
        uint32_t rem = mant & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (r & 1u))) r += 1u; /* carry may bump exponent */
        return sign | (uint16_t)r;
    }
    if (a < 0x33000000u) return sign;                      /* < 2^-25 rounds to zero */
    /* subnormal half: value = m * 2^-24 with m in [0, 1024] */
    uint32_t mant = (a & 0x7fffffu) | 0x800000u;
    int e = (int)(a >> 23) - 127;                          /* in [-25, -15] */
    int shift = -e - 1;                                    /* 24-bit mantissa -> units of 2^-24 */
    shift = shift + 0;                                     /* m = mant * 2^(e-23) / 2^-24 = mant >> (-1-e) */
    uint32_t m = mant >> shift;
    uint32_t rem = mant & ((1u << shift) - 1u);
    uint32_t half = 1u << (shift - 1);
    if (rem > half || (rem == half && (m & 1u))) m += 1u;
    return sign | (uint16_t)m;
}

float orc_f16_to_f32(uint16_t b)
{
    uint32_t sign = (uint32_t)(b & 0x8000u) << 16;
    uint32_t e = (b >> 10) & 0x1fu, m = b & 0x3ffu;
    if (e == 0x1fu) return bits_f32(sign | 0x7f800000u | (m << 13));
    if (e == 0) {
        float v = (float)m * (1.0f / 16777216.0f);        /* m * 2^-24, exact */
        return sign ? -v : v;
    }
    return bits_f32(sign | ((e + 112u) << 23) | (m << 13));
}

/* Stochastic rounding (extension): the magnitude plus random bits below the
 * kept mantissa, truncated.  BF16 takes the top 16 bits of r; FP16 the top 13
 * (shifted further for half subnormals).  This is the rounding of gfx950's
 * v_cvt_sr_bf16_f32 / v_cvt_sr_f16_f32, established by comparing against the
 * instructions' outputs on 2^20 (x, r) pairs. */
uint16_t orc_f32_to_bf16_sr(float x, uint32_t r)
{
    uint32_t u = f32_bits(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x0040u);
    return (uint16_t)((u + (r >> 16)) >> 16);
}

uint16_t orc_f32_to_f16_sr(float x, uint32_t r)
{
    uint32_t u = f32_bits(x);
    uint16_t sign = (uint16_t)((u >> 16) & 0x8000u);
    uint32_t a = u & 0x7fffffffu;
    uint32_t rnd = r >> 19;
    if (a > 0x7f800000u) return sign | 0x7e00u;
    if (a >= 0x38800000u) {
        uint32_t v = a + rnd;
        int e = (int)(v >> 23) - 112;
        if (e >= 31) return sign | 0x7c00u;
        return sign | (uint16_t)(((uint32_t)e << 10) | ((v >> 13) & 0x3ffu));
    }
    if (a < 0x00800000u) return sign;
    uint64_t mant = (a & 0x7fffffu) | 0x800000u;
    int shift = -((int)(a >> 23) - 127) - 1;  /* >= 14 */
    if (shift > 40) return sign;
    uint64_t m = (mant + ((uint64_t)rnd << (shift - 13))) >> shift;
    return sign | (uint16_t)m;
}

/* Random bits of the stochastic rounding: key = lowbias32(seed ^ step * phi),
 * h = lowbias32(cell ^ key), then xorshift32; each 32-bit word serves two
 * directions (the word and the word shifted left by 16). */
static inline uint32_t lowbias32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
static inline uint32_t xorshift32(uint32_t x)
{
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x;
}
uint32_t orc_sr_key(uint32_t seed, uint32_t step) { return lowbias32(seed ^ (step * 0x9e3779b9u)); }
void orc_sr_bits(uint32_t key, uint32_t cell, uint32_t R[9])
{
    uint32_t h = lowbias32(cell ^ key);
    for (int j = 0; j < 4; ++j) {
        R[2 * j] = h;
        R[2 * j + 1] = h << 16;
        h = xorshift32(h);
    }
    R[8] = h;
}

void orc_f32_to_bf16_n(const float* x, uint16_t* y, long n)
{
    for (long k = 0; k < n; ++k) y[k] = orc_f32_to_bf16(x[k]);
}

void orc_f32_to_f16_n(const float* x, uint16_t* y, long n)
{
    for (long k = 0; k < n; ++k) y[k] = orc_f32_to_f16(x[k]);
}

static inline float dev_load(int kind, uint16_t v) { return kind == ORC_BF16 ? orc_bf16_to_f32(v) : orc_f16_to_f32(v); }
static inline uint16_t dev_store(int kind, float v) { return kind == ORC_BF16 ? orc_f32_to_bf16(v) : orc_f32_to_f16(v); }

void orc_encode_dev16(const orc_params* p, int kind, const float* f, uint16_t* d)
{
    orc_consts_f32 k;
    orc_consts_f32_make(p, &k);
    const long n = (long)p->w * p->h;
    for (int i = 0; i < 9; ++i)
        for (long j = 0; j < n; ++j) d[i * n + j] = dev_store(kind, f[i * n + j] - k.base[i]);
}

void orc_decode_dev16(const orc_params* p, int kind, const uint16_t* d, float* f)
{
    orc_consts_f32 k;
    orc_consts_f32_make(p, &k);
    const long n = (long)p->w * p->h;
    for (int i = 0; i < 9; ++i)
        for (long j = 0; j < n; ++j) f[i * n + j] = dev_load(kind, d[i * n + j]) + k.base[i];
}

/* ---- deviation arithmetic (extension) ------------------------------------
 * With g_i = f_i - w_i rho0 the step is evaluated on g directly:
 *   drho = sum g_i, rho = rho0 + drho, m = sum g_i c_i, u = m / rho,
 *   g_eq_i = w_i (drho + rho ((vc A + vc^2 B) + v2 C)),
 *   g'_i = g_i + (g_i - g_eq_i) K,
 * which is f' - w_i rho0 of the BGK update (lbm.rs:350-364) without carrying
 * the O(1) rest populations through the rounding.  Streaming, bounce-back and
 * the Ladd term act on g unchanged (opposite directions share w_i); a zero-
 * filled population outside the lattice is g = -w_i rho0. */
/* Constants of the deviation path: those of orc_consts_f32_make except A, B, C,
 * which are evaluated in f64 and rounded once (exactly 3, 4.5, -1.5 for
 * dx = dt = 1), so that sum_i g_eq_i = drho and sum_i g_eq_i c_i = rho u hold
 * up to rounding. */
void orc_consts_devf32_make(const orc_params* p, orc_consts_f32* k)
{
    orc_consts_f64 d;
    orc_consts_f32_make(p, k);
    orc_consts_f64_make(p, &d);
    k->A = (float)d.A;
    k->B = (float)d.B;
    k->C = (float)d.C;
    k->cs2 = (float)d.cs2;
}

static inline void collide_dev(const orc_consts_f32* k, float rho0, float* B)
{
    float drho = (((((((B[0] + B[1]) + B[2]) + B[3]) + B[4]) + B[5]) + B[6]) + B[7]) + B[8];
    float rho = rho0 + drho;
    float mx = ((((B[1] - B[3]) + B[5]) - B[6]) - B[7]) + B[8];
    float my = ((((B[2] - B[4]) + B[5]) + B[6]) - B[7]) - B[8];
    float inv = 1.0f / rho;
    float ux = inv * mx, uy = inv * my;
    float v2 = ux * ux + uy * uy;
    const float vc[9] = {0.0f, ux, uy, -ux, -uy, ux + uy, -ux + uy, -ux - uy, ux - uy};
    for (int i = 0; i < 9; ++i) {
        float s = (vc[i] * k->A + (vc[i] * vc[i]) * k->B) + v2 * k->C;
        float geq = k->w[i] * (drho + rho * s);
        B[i] = B[i] + (B[i] - geq) * k->K;
    }
}

static inline void zou_he_inlet_dev(float* B, float rho0, float ux, float uy)
{
    float G = ((B[0] + B[1]) + B[3]) + 2.0f * ((B[4] + B[7]) + B[8]);
    float rho = (rho0 + G) / (1.0f - uy);
    float ru_t = rho * ux, ru_a = rho * uy;
    float d13 = B[1] - B[3];
    B[2] = B[4] + (2.0f / 3.0f) * ru_a;
    B[5] = ((B[7] - 0.5f * d13) + 0.5f * ru_t) + (1.0f / 6.0f) * ru_a;
    B[6] = ((B[8] + 0.5f * d13) - 0.5f * ru_t) + (1.0f / 6.0f) * ru_a;
}

static inline void zou_he_outlet_dev(float* B, float rho0, float rho)
{
    float G = ((B[0] + B[1]) + B[3]) + 2.0f * ((B[2] + B[5]) + B[6]);
    float uy = ((rho0 - rho) + G) / rho;
    float ru_a = rho * uy;
    float d13 = B[1] - B[3];
    B[4] = B[2] - (2.0f / 3.0f) * ru_a;
    B[7] = (B[5] + 0.5f * d13) - (1.0f / 6.0f) * ru_a;
    B[8] = (B[6] - 0.5f * d13) - (1.0f / 6.0f) * ru_a;
}

static inline float fetch_dev(const orc_params* p, const orc_consts_f32* k, const float* G, int i, int j,
                              int r, int c)
{
    const int w = p->w, h = p->h;
    if (p->edge == ORC_EDGE_PERIODIC) {
        r = (r % h + h) % h;
        c = (c % w + w) % w;
    } else if (r < 0 || r >= h || c < 0 || c >= w) {
        return -k->base[i];
    }
    return G[(long)j * w * h + (long)r * w + c];
}

void orc_step_devf32(const orc_params* p, const uint8_t* flags, const float* G, float* out)
{
    orc_consts_f32 k;
    orc_consts_devf32_make(p, &k);
    const int w = p->w, h = p->h;
    const long n = (long)w * h;
    const float rho0 = (float)p->rho0;
    float B[9];
    for (int r = 0; r < h; ++r)
        for (int c = 0; c < w; ++c) {
            const long idx = (long)r * w + c;
            const int t = flags ? flags[idx] : ORC_FLUID;
            const int solid = is_solid(t);
            if (p->boundary == ORC_BB_HALFWAY) {
                if (solid) {
                    for (int i = 0; i < 9; ++i) out[i * n + idx] = G[i * n + idx];
                    continue;
                }
                for (int i = 0; i < 9; ++i) {
                    int sr = r + CX[i], sc = c - CY[i];
                    if (p->edge == ORC_EDGE_PERIODIC) {
                        sr = (sr % h + h) % h;
                        sc = (sc % w + w) % w;
                    }
                    int inside = sr >= 0 && sr < h && sc >= 0 && sc < w;
                    int st = (inside && flags) ? flags[(long)sr * w + sc] : ORC_FLUID;
                    if (is_solid(st)) {
                        float v = G[OPP[i] * n + idx];
                        if (st == ORC_LID) v = v + k.lid[i];
                        B[i] = v;
                    } else {
                        B[i] = inside ? G[i * n + (long)sr * w + sc] : -k.base[i];
                    }
                }
                if (t == ORC_INLET) zou_he_inlet_dev(B, rho0, (float)p->inlet_ux, (float)p->inlet_uy);
                else if (t == ORC_OUTLET) zou_he_outlet_dev(B, rho0, (float)p->outlet_rho);
            } else {
                for (int i = 0; i < 9; ++i) {
                    int j = solid ? OPP[i] : i;
                    B[i] = fetch_dev(p, &k, G, i, j, r + CX[j], c - CY[j]);
                }
            }
            collide_dev(&k, rho0, B);
            for (int i = 0; i < 9; ++i) out[i * n + idx] = B[i];
        }
}

/* Deviation-form equilibrium g_eq_i = w_i (drho + rho s'_i), drho = rho - rho0. */
void orc_equilibrium_devf32(const orc_params* p, int n, const float* rho, const float* ux,
                            const float* uy, float* G)
{
    orc_consts_f32 k;
    orc_consts_devf32_make(p, &k);
    const float rho0 = (float)p->rho0;
    for (long j = 0; j < n; ++j) {
        float r = rho[j], vx = ux[j], vy = uy[j];
        float drho = r - rho0;
        float v2 = vx * vx + vy * vy;
        const float vc[9] = {0.0f, vx, vy, -vx, -vy, vx + vy, -vx + vy, -vx - vy, vx - vy};
        for (int i = 0; i < 9; ++i) {
            float s = (vc[i] * k.A + (vc[i] * vc[i]) * k.B) + v2 * k.C;
            G[(long)i * n + j] = k.w[i] * (drho + r * s);
        }
    }
}

int orc_run_devf32(const orc_params* p, const uint8_t* flags, float* g, int nsteps)
{
    const long n = 9L * p->w * p->h;
    float* tmp = (float*)malloc(sizeof(float) * n);
    if (!tmp) return -1;
    for (int s = 0; s < nsteps; ++s) {
        orc_step_devf32(p, flags, g, tmp);
        memcpy(g, tmp, sizeof(float) * n);
    }
    free(tmp);
    return 0;
}

/* One 16-bit step: widen the stored deviations to FP32 (exact), run the
 * deviation-form FP32 step, narrow with round to nearest even (rounding 0) or
 * stochastic rounding keyed by (seed, step, cell) (rounding 1).  Solid cells
 * of the half-way mode keep their stored bits. */
void orc_step_dev16_ex(const orc_params* p, int kind, int rounding, uint32_t seed, uint32_t step,
                       const uint8_t* flags, const uint16_t* P, uint16_t* out)
{
    const long n = (long)p->w * p->h;
    if (n <= 0) return;
    float* g = (float*)calloc((size_t)(9 * n), sizeof(float));
    float* o = (float*)calloc((size_t)(9 * n), sizeof(float));
    if (!g || !o) {
        free(g);
        free(o);
        return;
    }
    for (long j = 0; j < 9 * n; ++j) g[j] = dev_load(kind, P[j]);
    orc_step_devf32(p, flags, g, o);
    if (rounding == 0) {
        for (long j = 0; j < 9 * n; ++j) out[j] = dev_store(kind, o[j]);
    } else {
        const uint32_t key = orc_sr_key(seed, step);
        uint32_t R[9];
        for (long c = 0; c < n; ++c) {
            orc_sr_bits(key, (uint32_t)c, R);
            for (int i = 0; i < 9; ++i)
                out[i * n + c] = kind == ORC_BF16 ? orc_f32_to_bf16_sr(o[i * n + c], R[i])
                                                  : orc_f32_to_f16_sr(o[i * n + c], R[i]);
        }
    }
    if (p->boundary == ORC_BB_HALFWAY && flags) {
        for (long j = 0; j < n; ++j)
            if (is_solid(flags[j]))
                for (int i = 0; i < 9; ++i) out[i * n + j] = P[i * n + j];
    }
    free(g);
    free(o);
}

void orc_step_dev16(const orc_params* p, int kind, const uint8_t* flags, const uint16_t* P,
                    uint16_t* out)
{
    orc_step_dev16_ex(p, kind, 0, 0, 0, flags, P, out);
}

int orc_run_dev16_ex(const orc_params* p, int kind, int rounding, uint32_t seed, uint32_t step0,
                     const uint8_t* flags, uint16_t* d, int nsteps)
{
    const long n = 9L * p->w * p->h;
    uint16_t* tmp = (uint16_t*)malloc(sizeof(uint16_t) * n);
    if (!tmp) return -1;
    for (int s = 0; s < nsteps; ++s) {
        orc_step_dev16_ex(p, kind, rounding, seed, step0 + (uint32_t)s, flags, d, tmp);
        memcpy(d, tmp, sizeof(uint16_t) * n);
    }
    free(tmp);
    return 0;
}

int orc_run_dev16(const orc_params* p, int kind, const uint8_t* flags, uint16_t* d, int nsteps)
{
    return orc_run_dev16_ex(p, kind, 0, 0, 0, flags, d, nsteps);
}

/* ----------------------------------------------------------- residual --- */

double orc_residual(int w, int h, const uint8_t* flags, const float* ux, const float* uy,
                    const float* ux_prev, const float* uy_prev)
{
    const long n = (long)w * h;
    double num = 0, den = 0;
    for (long k = 0; k < n; ++k) {
        if (flags && is_solid(flags[k])) continue;
        double dx = (double)ux[k] - (double)ux_prev[k];
        double dy = (double)uy[k] - (double)uy_prev[k];
        num += dx * dx + dy * dy;
        den += (double)ux[k] * (double)ux[k] + (double)uy[k] * (double)uy[k];
    }
    return den > 0 ? sqrt(num / den) : (num > 0 ? INFINITY : 0.0);
}
