// C ABI of the single-lattice path: handles, transfers, the fused step,
// the reference's unfused phases, moments and FP64 reductions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "cs_internal.hpp"
#include "lbm_device.hpp"
#include "lattice_kernels.hpp"

namespace cs {

// ------------------------------------------------------------------ errors --
// The message of the calling thread's last failed call.  A plain byte array:
// thread-local objects with destructors are destroyed when their thread ends,
// before static objects, so a std::string here would be written to after its
// buffer was freed by any call made from a static destructor at process exit
// (such as a host layer's device singleton destroying its context) -- that
// write lands in a freed heap chunk and corrupts the allocator's free lists.
static thread_local char g_error[512];
void set_error(const std::string& msg)
{
    const size_t n = std::min(msg.size(), sizeof g_error - 1);
    std::memcpy(g_error, msg.data(), n);
    g_error[n] = '\0';
}
void clear_error() { g_error[0] = '\0'; }

size_t storage_bytes(int storage)
{
    switch (storage) {
    case CS_F32:
    case CS_F32DEV: return 4;
    case CS_F64: return 8;
    case CS_BF16DEV:
    case CS_F16DEV: return 2;
    default: return 0;
    }
}

// ----------------------------------------------------------- constants ----
// Evaluated exactly as lbm.rs evaluates them in Scalar = f32:
// cs = dx / (sqrt(3) dt) (lbm.rs:84), cs2, cs4 (lbm.rs:55-56), A = 1/cs2,
// B = 1/(2 cs4), C = -1/(2 cs2) (lbm.rs:64-66), K = -dt/tau (lbm.rs:357),
// weights 16/36, 4/36, 1/36 (lbm.rs:209-219).
static const double W_EXACT[9] = {4.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9,
                                  1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};

template <class T>
static Consts<T> make_consts(const cs_params& p)
{
    Consts<T> k{};
    volatile T dx = (T)p.delta_x, dt = (T)p.delta_t, tau = (T)p.tau, three = (T)3;
    T cs = dx / (std::sqrt((T)three) * dt);
    T cs2 = cs * cs;
    T cs4 = cs2 * cs2;
    k.cs2 = cs2;
    k.A = (T)1 / cs2;
    k.B = (T)1 / ((T)2 * cs4);
    k.C = (T)-1 / ((T)2 * cs2);
    k.K = -dt / tau;
    volatile T tau_m = (T)(p.collision == CS_COLL_TRT ? p.tau_minus : p.tau);
    k.wp = (T)1 / tau;
    k.wm = (T)1 / tau_m;
    k.hdt = -dt * (T)0.5;
    k.dx = dx;
    volatile T nu = (T)p.ks_viscosity;
    k.beta = (T)1 / (((T)2 * nu / (cs * cs)) + (T)1);
    volatile T n0 = 16, n1 = 4, n5 = 1, d = 36;
    const T w0 = n0 / d, w1 = n1 / d, w5 = n5 / d;
    const T ws[3] = {w0, w1, w5};
    for (int q = 0; q < 3; ++q) {
        k.wq[q] = ws[q];
        k.baseq[q] = ws[q] * (T)p.rho0;
        k.sfq[q] = ws[q] / ((T)2 * cs4);
    }
    // Ladd moving wall: f_i = f*_opp(i) + 2 w_i rho_w (c_i . u_w) / cs^2,
    // rho_w = 1, cs^2 = 1/3 (extension).  Directions 1, 2, 5, 6; the opposite
    // ones carry the exact negatives.
    const int dirs[4] = {1, 2, 5, 6};
    for (int q = 0; q < 4; ++q) {
        const int i = dirs[q];
        k.lidq[q] = (T)(6.0 * W_EXACT[i] * (CX[i] * p.lid_ux + CY[i] * p.lid_uy));
    }
    k.in_ux = (T)p.inlet_ux;
    k.in_uy = (T)p.inlet_uy;
    k.out_rho = (T)p.outlet_rho;
    k.rho0 = (T)p.rho0;
    return k;
}

Consts<float> make_consts_f32(const cs_params& p) { return make_consts<float>(p); }

// The deviation storages have no reference counterpart: they take A, B, C
// (and cs^2) from the f64 evaluation rounded once -- exactly 3, 4.5, -1.5 for
// dx = dt = 1 -- so that the equilibrium's mass and momentum match rho and
// rho u up to rounding.  The reference's f32 values (1 ulp high) add about
// 1e-7 relative momentum per collision.
Consts<float> make_consts_dev_f32(const cs_params& p)
{
    Consts<float> k = make_consts<float>(p);
    const Consts<double> d = make_consts<double>(p);
    k.A = (float)d.A;
    k.B = (float)d.B;
    k.C = (float)d.C;
    k.cs2 = (float)d.cs2;
    return k;
}

bool is_dev_storage(int s) { return s == CS_F32DEV || s == CS_BF16DEV || s == CS_F16DEV; }

static void set_consts(cs_lattice* l, const cs_params& p)
{
    l->k32 = is_dev_storage(l->storage) ? make_consts_dev_f32(p) : make_consts_f32(p);
    l->k64 = make_consts_f64(p);
}
Consts<double> make_consts_f64(const cs_params& p) { return make_consts<double>(p); }

// ------------------------------------------------------------- kernels ----

// Lattice::equilibrium (lbm.rs:156-160) or non_equilibrium (lbm.rs:162-173).
template <class St>
__global__ void k_equilibrium(const typename St::S* __restrict__ f, int w, int h, int pitch, int64_t P,
                              Consts<typename St::T> k, int non_eq, double* __restrict__ out)
{
    using T = typename St::T;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (c >= w) return;
    const int64_t own = (int64_t)r * pitch + c, d = (int64_t)r * w + c, n = (int64_t)w * h;
    T B[9], F[9];
    load_own<St>(f, P, own, B);
    T rho, mx, my, ux, uy;
    moments<St>(k, B, rho, mx, my, ux, uy);
    if (non_eq && St::DEV) equilibrium_dev<T>(k, rho, ux, uy, F);  // g - g_eq = f - f_eq
    else equilibrium<T>(k, rho, ux, uy, F);
#pragma unroll
    for (int i = 0; i < 9; ++i) out[i * n + d] = (double)(non_eq ? B[i] - F[i] : F[i]);
}

// State::is_unstable (lbm.rs:815-818): any f_eq[0] < 0.
template <class St>
__global__ void k_unstable(const typename St::S* __restrict__ f, int w, int h, int pitch, int64_t P,
                           Consts<typename St::T> k, int* __restrict__ flag)
{
    using T = typename St::T;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (c >= w) return;
    const int64_t own = (int64_t)r * pitch + c;
    T B[9], F[9];
    load_own<St>(f, P, own, B);
    T rho, mx, my, ux, uy;
    moments<St>(k, B, rho, mx, my, ux, uy);
    equilibrium<T>(k, rho, ux, uy, F);
    if (F[0] < T(0)) atomicOr(flag, 1);
}

// ---- the reference's phases alone (REFERENCE mode, F32/F64 storage) ----
// State::stream (lbm.rs:716-729): out_i[r][c] = in_i[r + cx_i][c - cy_i], zero
// outside (or wrapped).
template <class St>
__global__ void k_stream(const typename St::S* __restrict__ fin, typename St::S* __restrict__ fout, int w, int h,
                         int pitch, int64_t P, int edge)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (c >= w) return;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        int sr = r + CX[i], sc = c - CY[i];
        bool valid = true;
        if (edge == CS_EDGE_PERIODIC) {
            sr = (sr + h) % h;
            sc = (sc + w) % w;
        } else {
            valid = sr >= 0 && sr < h && sc >= 0 && sc < w;
        }
        fout[i * P + (int64_t)r * pitch + c] = valid ? fin[i * P + (int64_t)sr * pitch + sc] : (typename St::S)0;
    }
}

// State::bounce_back (lbm.rs:741-751): at solid cells f_i <- f_opp(i), in place.
template <class St>
__global__ void k_bounce_back(typename St::S* __restrict__ f, const uint8_t* __restrict__ flags, int w, int h,
                              int pitch, int cpitch, int64_t P)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (c >= w) return;
    const int64_t own = (int64_t)r * pitch + c;
    const int t = flags[(int64_t)r * cpitch + c];
    if (!(t == CS_WALL || t == CS_LID)) return;
    typename St::S v[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) v[i] = f[i * P + own];
#pragma unroll
    for (int i = 0; i < 9; ++i) f[i * P + own] = v[OPP[i]];
}

// State::collide (lbm.rs:731-739) with BGK::evaluate (lbm.rs:350-364) or
// TRT::evaluate (lbm.rs:401-444), every cell, in place.
template <class St, int COLL>
__global__ void k_collide(typename St::S* __restrict__ f, int w, int h, int pitch, int64_t P,
                          Consts<typename St::T> k)
{
    using T = typename St::T;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (c >= w) return;
    const int64_t own = (int64_t)r * pitch + c;
    T B[9];
    load_own<St>(f, P, own, B);
    if (COLL == CS_COLL_TRT) trt_collide<T>(k, B);
    else if (COLL == CS_COLL_REGULARIZED) reg_collide<T>(k, B);
    else if (COLL == CS_COLL_KBC) kbc_collide<T>(k, B);
    else bgk_collide<T>(k, B);
#pragma unroll
    for (int i = 0; i < 9; ++i) St::st(f, i * P + own, B[i]);
}

// ---- FP64 reductions ---------------------------------------------------
constexpr int RED_BLOCKS = 1024;
constexpr int RED_THREADS = 256;

// Fixed-shape tree reduction of NV doubles per thread into block partials.
template <int NV>
__device__ __forceinline__ void block_reduce_store(double (&v)[NV], double* partials)
{
    __shared__ double sh[NV][RED_THREADS];
    for (int j = 0; j < NV; ++j) sh[j][threadIdx.x] = v[j];
    __syncthreads();
    for (int s = RED_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int j = 0; j < NV; ++j) sh[j][threadIdx.x] += sh[j][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int j = 0; j < NV; ++j) partials[(int64_t)blockIdx.x * NV + j] = sh[j][0];
}

template <int NV>
__global__ void k_reduce_final(const double* __restrict__ partials, int nblocks, double* __restrict__ out)
{
    double v[NV];
    for (int j = 0; j < NV; ++j) v[j] = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += RED_THREADS)
        for (int j = 0; j < NV; ++j) v[j] += partials[(int64_t)b * NV + j];
    __shared__ double sh[NV][RED_THREADS];
    for (int j = 0; j < NV; ++j) sh[j][threadIdx.x] = v[j];
    __syncthreads();
    for (int s = RED_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int j = 0; j < NV; ++j) sh[j][threadIdx.x] += sh[j][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int j = 0; j < NV; ++j) out[j] = sh[j][0];
}

template <class St>
__global__ __launch_bounds__(RED_THREADS) void k_sums(const typename St::S* __restrict__ f,
                                                      const uint8_t* __restrict__ flags, int w, int h, int pitch,
                                                      int cpitch, int64_t P, Consts<typename St::T> k,
                                                      double* partials)
{
    using T = typename St::T;
    double v[3] = {0.0, 0.0, 0.0};
    const int64_t n = (int64_t)w * h;
    for (int64_t d = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; d < n; d += (int64_t)RED_BLOCKS * RED_THREADS) {
        const int r = (int)(d / w), c = (int)(d - (int64_t)r * w);
        const int64_t own = (int64_t)r * pitch + c;
        if (flags) {
            const int t = flags[(int64_t)r * cpitch + c];
            if (t == CS_WALL || t == CS_LID) continue;
        }
        T B[9];
        load_own<St>(f, P, own, B);
        T rho, mx, my, ux, uy;
        moments<St>(k, B, rho, mx, my, ux, uy);
        v[0] += (double)rho;
        v[1] += (double)mx;
        v[2] += (double)my;
    }
    block_reduce_store<3>(v, partials);
}

// Velocity snapshot (dense double planes ux, uy).
template <class St>
__global__ void k_snapshot(const typename St::S* __restrict__ f, int w, int h, int pitch, int64_t P,
                           Consts<typename St::T> k, double* __restrict__ snap)
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
    snap[d] = (double)ux;
    snap[n + d] = (double)uy;
}

template <class St>
__global__ __launch_bounds__(RED_THREADS) void k_residual(const typename St::S* __restrict__ f,
                                                          const uint8_t* __restrict__ flags, int w, int h, int pitch,
                                                          int cpitch, int64_t P, Consts<typename St::T> k,
                                                          const double* __restrict__ snap, double* partials)
{
    using T = typename St::T;
    double v[2] = {0.0, 0.0};
    const int64_t n = (int64_t)w * h;
    for (int64_t d = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; d < n; d += (int64_t)RED_BLOCKS * RED_THREADS) {
        const int r = (int)(d / w), c = (int)(d - (int64_t)r * w);
        const int64_t own = (int64_t)r * pitch + c;
        if (flags) {
            const int t = flags[(int64_t)r * cpitch + c];
            if (t == CS_WALL || t == CS_LID) continue;
        }
        T B[9];
        load_own<St>(f, P, own, B);
        T rho, mx, my, ux, uy;
        moments<St>(k, B, rho, mx, my, ux, uy);
        const double dx = (double)ux - snap[d], dy = (double)uy - snap[n + d];
        v[0] += dx * dx + dy * dy;
        v[1] += (double)ux * (double)ux + (double)uy * (double)uy;
    }
    block_reduce_store<2>(v, partials);
}

// Plain copy, the bandwidth reference: one 16-byte vector per thread,
// non-temporal loads and stores (the fastest of the measured variants:
// 1, 2, 4 or 8 vectors per thread, temporal or non-temporal loads).
typedef float f32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy(const f32x4* __restrict__ src, f32x4* __restrict__ dst, int64_t n4)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n4) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

// --------------------------------------------------------- host helpers ---

static void check_lat(const cs_lattice* lat) { CS_REQUIRE(lat != nullptr, CS_ERR_ARG, "null lattice"); }

// Elements in one population buffer (9 interleaved planes of h rows).
static int64_t pop_elems(const cs_lattice* l) { return (int64_t)l->h * l->pitch; }

template <class St>
static const Consts<typename St::T>& consts_of(const cs_lattice* lat);
template <> const Consts<float>& consts_of<StF32>(const cs_lattice* l) { return l->k32; }
template <> const Consts<float>& consts_of<StF32Dev>(const cs_lattice* l) { return l->k32; }
template <> const Consts<float>& consts_of<StBF16>(const cs_lattice* l) { return l->k32; }
template <> const Consts<float>& consts_of<StF16>(const cs_lattice* l) { return l->k32; }
template <> const Consts<double>& consts_of<StF64>(const cs_lattice* l) { return l->k64; }

// Calls f.template operator()<St>() for the lattice's storage type.
template <class F>
static void dispatch_storage(int storage, F&& f)
{
    switch (storage) {
    case CS_F32: f.template operator()<StF32>(); break;
    case CS_F64: f.template operator()<StF64>(); break;
    case CS_F32DEV: f.template operator()<StF32Dev>(); break;
    case CS_BF16DEV: f.template operator()<StBF16>(); break;
    case CS_F16DEV: f.template operator()<StF16>(); break;
    default: throw Error{CS_ERR_ARG, "bad storage"};
    }
}

static dim3 cell_grid(const cs_lattice* l) { return dim3((unsigned)((l->w + 255) / 256), (unsigned)l->h); }

struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(size_t bytes) { CS_HIP(hipMalloc(&p, bytes)); }
    ~DevBuf() { if (p) (void)hipFree(p); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

// Process-wide launch switches for tests and tuning (cs_set_cells_per_thread,
// cs_set_step_graphs).  Atomic: lattices may step on other threads while a
// switch changes; a lattice stepping during the change may launch either way.
static std::atomic<int> g_cpt_override{0};

template <class St, int EDGE, int MODE, int CPT, bool SR, int COLL>
static void launch_step_cpt(cs_lattice* l, const void* fin, void* fout)
{
    StepArgs<typename St::T> a{};
    a.fin = fin;
    a.fout = fout;
    a.node = l->node;
    a.w = l->w;
    a.h = l->h;
    a.pitch = l->pitch;
    a.cpitch = l->cpitch;
    a.plane = l->plane;
    a.tiles_x = (l->w + 64 * CPT - 1) / (64 * CPT);
    a.sr_key = sr_key(l->params.seed, l->steps);
    a.cell_base = 0;
    a.k = consts_of<St>(l);
    const int tiles_y = (l->h + 3) / 4;
    const int nblocks = a.tiles_x * tiles_y;
    if (pop_elems(l) * (int64_t)sizeof(typename St::S) < ((int64_t)1 << 32))
        hipLaunchKernelGGL((k_step<St, EDGE, MODE, CPT, SR, uint32_t, COLL>), dim3(nblocks), dim3(256), 0, l->stream,
                           a);
    else
        hipLaunchKernelGGL((k_step<St, EDGE, MODE, CPT, SR, int64_t, COLL>), dim3(nblocks), dim3(256), 0, l->stream,
                           a);
}

// Lattices with 4- or 8-byte storage, with or without geometry, take the
// vector kernel (16 bytes per lane and plane, 4 rows per block) when a row
// splits into vectors and the buffers allow 32-bit offsets.  16-bit storage
// keeps the scalar kernel with 4 cells per lane, which measured faster: 4096^2
// Taylor-Green with stochastic rounding, FP16 114.9 us against 128.4 us with
// 8-byte and 134.6 us with 16-byte vector lanes (BF16 113.5 / 122.8 / 137.4),
// with the row-wise neighbours exchanged by DPP lane shifts; the vector kernel
// stays available for 16-bit storage through cs_set_cells_per_thread.
template <class St, int EDGE, int MODE, bool SR, int COLL, int VB = 16>
static bool launch_step_vec(cs_lattice* l, const void* fin, void* fout)
{
    if constexpr (sizeof(typename St::S) < 4 && VB == 16 && MODE != MODE_FLUID) {
        return false;  // node words of 8 cells do not fit one vector
    } else if constexpr (SR && MODE != MODE_FLUID) {
        return false;
    } else {
        constexpr int ROWS = 4;
        constexpr int V = VB / (int)sizeof(typename St::S);
        if (l->w % V != 0 || pop_elems(l) * (int64_t)sizeof(typename St::S) >= ((int64_t)1 << 32)) return false;
        StepArgs<typename St::T> a{};
        a.fin = fin;
        a.fout = fout;
        a.node = MODE == MODE_FLUID ? nullptr : l->node;
        a.w = l->w;
        a.h = l->h;
        a.pitch = l->pitch;
        a.cpitch = l->cpitch;
        a.plane = l->plane;
        a.tiles_x = (l->w + 64 * V - 1) / (64 * V);
        a.sr_key = sr_key(l->params.seed, l->steps);
        a.cell_base = 0;
        a.k = consts_of<St>(l);
        const int nblocks = a.tiles_x * ((l->h + ROWS - 1) / ROWS);
        hipLaunchKernelGGL((k_step_vec<St, EDGE, SR, VB, ROWS, COLL, MODE>), dim3(nblocks), dim3(64 * ROWS), 0,
                           l->stream, a);
        return true;
    }
}

// Which of the two bit-identical step kernels a lattice takes, by size
// (FP32, MI355X, microseconds per step, vector / scalar kernel, periodic
// Taylor-Green and half-way cavity):
//   128^2 4.3 / 2.5, 6.7 / 3.1;  512^2 5.2 / 3.9, 6.3 / 5.3
//   1024^2 12.6 / 17.5, 13.5 / 17.4;  2048^2 56.0 / 54.0, 59.5 / 53.9
//   4096^2 203.7 / 204.5, 210.2 / 226.5;  8192^2 784 / 809, 835 / 885;
// FP64: 256^2 4.5 / 3.3, 5.4 / 4.4;  512^2 7.5 / 8.0, 8.1 / 9.5;
//   4096^2 391 / 397, 399 / 421.
// Small lattices are latency-bound and gain from the scalar kernel's two
// (FP64) or four (FP32) times as many lanes.  Where one population buffer is
// 80-300 MB it partly stays in the 256 MB last-level cache between steps,
// which the scalar kernel's cacheable loads use and the vector kernel's
// non-temporal loads do not.
static bool vector_size_class(const cs_lattice* l)
{
    const int64_t cells = (int64_t)l->w * l->h;
    const int64_t esize = (int64_t)storage_bytes(l->storage);
    const int64_t bytes = pop_elems(l) * esize;
    if (cells < ((int64_t)1 << (esize == 8 ? 18 : 19))) return false;
    return bytes < ((int64_t)80 << 20) || bytes > ((int64_t)300 << 20);
}

template <class St, int EDGE, int MODE, bool SR, int COLL>
static void launch_step_sr(cs_lattice* l, const void* fin, void* fout)
{
    const int ov = g_cpt_override.load(std::memory_order_relaxed);
    if constexpr (is_16bit<St>()) {
        // 16-bit storage: the vector kernel only when asked for (8- or 16-byte lanes)
        if (ov == CS_STEP_VECTOR && launch_step_vec<St, EDGE, MODE, SR, COLL, 8>(l, fin, fout)) return;
        if (ov == CS_STEP_VECTOR_WIDE && launch_step_vec<St, EDGE, MODE, SR, COLL, 16>(l, fin, fout)) return;
    } else if ((ov == CS_STEP_VECTOR || ov == CS_STEP_VECTOR_WIDE || (ov == 0 && vector_size_class(l))) &&
               launch_step_vec<St, EDGE, MODE, SR, COLL>(l, fin, fout)) {
        return;
    }
    int cpt = ov > 0 ? ov : 0;
    // 16-bit storage: 4 cells per lane (with stochastic rounding on the
    // Taylor-Green benchmark 143.6 against 140.4 GLUPS for one cell per lane at
    // 4096^2, 147.8 against 141.3 at 8192^2).
    if (cpt == 0) cpt = (l->storage == CS_BF16DEV || l->storage == CS_F16DEV) ? 4 : 1;
    if (cpt >= 4) launch_step_cpt<St, EDGE, MODE, 4, SR, COLL>(l, fin, fout);
    else if (cpt == 2) launch_step_cpt<St, EDGE, MODE, 2, SR, COLL>(l, fin, fout);
    else launch_step_cpt<St, EDGE, MODE, 1, SR, COLL>(l, fin, fout);
}

template <class St, int EDGE, int MODE>
static void launch_step_mode(cs_lattice* l, const void* fin, void* fout)
{
    if constexpr (is_16bit<St>()) {
        if (l->params.rounding == CS_ROUND_STOCHASTIC) {
            launch_step_sr<St, EDGE, MODE, true, CS_COLL_BGK>(l, fin, fout);
            return;
        }
    }
    // TRT, Regularized, KBC: reference arithmetic only (set_params rejects them for the deviation storages)
    if constexpr (std::is_same_v<St, StF32> || std::is_same_v<St, StF64>) {
        switch (l->params.collision) {
        case CS_COLL_TRT: launch_step_sr<St, EDGE, MODE, false, CS_COLL_TRT>(l, fin, fout); return;
        case CS_COLL_REGULARIZED: launch_step_sr<St, EDGE, MODE, false, CS_COLL_REGULARIZED>(l, fin, fout); return;
        case CS_COLL_KBC: launch_step_sr<St, EDGE, MODE, false, CS_COLL_KBC>(l, fin, fout); return;
        default: break;
        }
    }
    launch_step_sr<St, EDGE, MODE, false, CS_COLL_BGK>(l, fin, fout);
}

template <class St>
static void launch_step(cs_lattice* l)
{
    const void* fin = l->buf[l->cur];
    void* fout = l->buf[l->cur ^ 1];
    const bool periodic = l->params.edge == CS_EDGE_PERIODIC;
    int mode = MODE_FLUID;
    if (l->has_flags) mode = l->params.boundary == CS_BB_HALFWAY ? MODE_HALF : MODE_REF;
    if (periodic) {
        if (mode == MODE_FLUID) launch_step_mode<St, CS_EDGE_PERIODIC, MODE_FLUID>(l, fin, fout);
        else if (mode == MODE_REF) launch_step_mode<St, CS_EDGE_PERIODIC, MODE_REF>(l, fin, fout);
        else launch_step_mode<St, CS_EDGE_PERIODIC, MODE_HALF>(l, fin, fout);
    } else {
        if (mode == MODE_FLUID) launch_step_mode<St, CS_EDGE_ZERO, MODE_FLUID>(l, fin, fout);
        else if (mode == MODE_REF) launch_step_mode<St, CS_EDGE_ZERO, MODE_REF>(l, fin, fout);
        else launch_step_mode<St, CS_EDGE_ZERO, MODE_HALF>(l, fin, fout);
    }
    l->cur ^= 1;
    l->steps += 1;
}

// ---- graph replay of fused steps (launch-bound lattices) ----------------
// A lattice of at most GRAPH_MAX_CELLS cells spends most of a step in launch
// overhead; runs of GRAPH_STEPS steps are captured once into a HIP graph and
// replayed.  GRAPH_STEPS is even, so a run ends on the buffer parity it
// started from.  Stochastic rounding keys each step's random bits by the step
// count, which a replay would repeat, so 16-bit storage with stochastic
// rounding always launches step by step.
constexpr int GRAPH_STEPS = 64;
constexpr int64_t GRAPH_MAX_CELLS = (int64_t)1 << 20;
static std::atomic<bool> g_step_graphs{true};

void clear_step_graphs(cs_lattice* l)
{
    for (hipGraphExec_t& g : l->step_graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

static bool graph_eligible(const cs_lattice* l)
{
    const bool sr16 = (l->storage == CS_BF16DEV || l->storage == CS_F16DEV) &&
                      l->params.rounding == CS_ROUND_STOCHASTIC;
    return g_step_graphs.load(std::memory_order_relaxed) && !sr16 && (int64_t)l->w * l->h <= GRAPH_MAX_CELLS;
}

static void step_once(cs_lattice* l)
{
    dispatch_storage(l->storage, [&]<class St>() { launch_step<St>(l); });
    l->time += l->params.delta_t;  // lbm.rs:713
}

static void advance(cs_lattice* l, int n, int k);

// GRAPH_STEPS steps from the current parity, captured (first time) and
// replayed; k: the temporally blocked plan (0: one step per launch), part of
// the graph's key with the launch-shape override.
static void replay_steps(cs_lattice* l, int k)
{
    const int ov = g_cpt_override.load(std::memory_order_relaxed);
    if (l->graph_cpt != ov || l->graph_tb != k) clear_step_graphs(l);
    const int parity = l->cur;
    if (!l->step_graph[parity]) {
        const int cur0 = l->cur;
        const uint64_t steps0 = l->steps;
        const double time0 = l->time;
        CS_HIP(hipStreamSynchronize(l->stream));  // capture starts on an idle stream
        CS_HIP(hipStreamBeginCapture(l->stream, hipStreamCaptureModeThreadLocal));
        hipGraph_t graph = nullptr;
        try {
            advance(l, GRAPH_STEPS, k);
        } catch (...) {
            (void)hipStreamEndCapture(l->stream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            l->cur = cur0;
            l->steps = steps0;
            l->time = time0;
            throw;
        }
        const hipError_t e = hipStreamEndCapture(l->stream, &graph);
        l->cur = cur0;  // nothing ran: the capture only recorded the launches
        l->steps = steps0;
        l->time = time0;
        CS_HIP(e);
        hipGraphExec_t exec = nullptr;
        const hipError_t e2 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        CS_HIP(e2);
        l->step_graph[parity] = exec;
        l->graph_cpt = ov;
        l->graph_tb = k;
    }
    CS_HIP(hipGraphLaunch(l->step_graph[parity], l->stream));
    for (int s = 0; s < GRAPH_STEPS; ++s) l->time += l->params.delta_t;  // as step by step
    l->steps += GRAPH_STEPS;
}

// n steps from the current state: K-step launches while n >= K, one 4-step
// launch when at least 4 remain (BGK), then one step per launch (plan:
// tb_steps's, 0 for one step per launch).
static void advance(cs_lattice* l, int n, int plan)
{
    const int k = tb_plan_steps(plan);
    if (k > 0) {
        for (; n >= k; n -= k) launch_tb(l, plan);
        const int plan4 = tb_remainder_plan(l, plan);
        if (plan4 != 0 && n >= 4) {
            launch_tb(l, plan4);
            n -= 4;
        }
    }
    for (int s = 0; s < n; ++s) step_once(l);
}

void do_steps(cs_lattice* l, int n)
{
    const int k = tb_steps(l, g_cpt_override.load(std::memory_order_relaxed));
    if (n >= GRAPH_STEPS && graph_eligible(l))
        for (; n >= GRAPH_STEPS; n -= GRAPH_STEPS) replay_steps(l, k);
    advance(l, n, k);
    CS_HIP(hipGetLastError());
}

// Staging area of cs_lattice_moments_f64_noalloc: the moment pass writes its
// planes here instead of into a buffer allocated per call.  Each GPU the
// library runs on holds its own copy of this array until the process ends.
// In host code `scratch` names a host-side stand-in, never the device memory:
// only the per-device address from hipGetSymbolAddress is passed to kernels
// and copies.
static __device__ __align__(256) double scratch[CS_NOALLOC_MAX_VALUES];
constexpr int MAX_DEVICES = 64;
static double* g_scratch_ptr[MAX_DEVICES];  // device address of `scratch`, per device
static std::mutex g_scratch_mu;             // the calls that use it are serialised per process

// Looks up this device's address of `scratch` once (the device is current).
static void init_noalloc_scratch(int device)
{
    CS_REQUIRE(device >= 0 && device < MAX_DEVICES, CS_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    double* p = nullptr;
    CS_HIP(hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)));
    CS_REQUIRE(p != nullptr, CS_ERR_HIP, "no device address for the moment scratch");
    g_scratch_ptr[device] = p;
}

// Runs k_moments and copies the requested planes (in mask bit order) to host.
static void moments_to_host(cs_lattice* l, unsigned mask, double* host)
{
    const int nplanes = __builtin_popcount(mask);
    const int64_t n = (int64_t)l->w * l->h;
    DevBuf out(sizeof(double) * n * nplanes);
    dispatch_storage(l->storage, [&]<class St>() {
        hipLaunchKernelGGL((k_moments<St>), cell_grid(l), dim3(256), 0, l->stream,
                           (const typename St::S*)l->buf[l->cur], l->w, l->h, l->pitch, l->plane, consts_of<St>(l),
                           mask, (double*)out.p);
    });
    CS_HIP(hipGetLastError());
    CS_HIP(hipMemcpyAsync(host, out.p, sizeof(double) * n * nplanes, hipMemcpyDeviceToHost, l->stream));
    CS_HIP(hipStreamSynchronize(l->stream));
}

static void moments_f32(cs_lattice* l, unsigned mask, float* const* outs)
{
    const int nplanes = __builtin_popcount(mask);
    const int64_t n = (int64_t)l->w * l->h;
    std::vector<double> tmp((size_t)n * nplanes);
    moments_to_host(l, mask, tmp.data());
    for (int j = 0; j < nplanes; ++j)
        for (int64_t d = 0; d < n; ++d) outs[j][d] = (float)tmp[(size_t)j * n + d];
}

static void reduce(cs_lattice* l, int nv, double* out_host)
{
    if (nv == 3) hipLaunchKernelGGL((k_reduce_final<3>), dim3(1), dim3(RED_THREADS), 0, l->stream, l->partials, RED_BLOCKS, l->dscratch);
    else hipLaunchKernelGGL((k_reduce_final<2>), dim3(1), dim3(RED_THREADS), 0, l->stream, l->partials, RED_BLOCKS, l->dscratch);
    CS_HIP(hipGetLastError());
    CS_HIP(hipMemcpyAsync(out_host, l->dscratch, sizeof(double) * nv, hipMemcpyDeviceToHost, l->stream));
    CS_HIP(hipStreamSynchronize(l->stream));
}

static double residual_now(cs_lattice* l)
{
    CS_REQUIRE(l->has_snap, CS_ERR_STATE, "residual requested before a velocity snapshot");
    dispatch_storage(l->storage, [&]<class St>() {
        hipLaunchKernelGGL((k_residual<St>), dim3(RED_BLOCKS), dim3(RED_THREADS), 0, l->stream,
                           (const typename St::S*)l->buf[l->cur], l->has_flags ? l->flags : nullptr, l->w, l->h,
                           l->pitch, l->cpitch, l->plane, consts_of<St>(l), (const double*)l->snap, l->partials);
    });
    double v[2];
    reduce(l, 2, v);
    if (v[1] > 0) return std::sqrt(v[0] / v[1]);
    return v[0] > 0 ? INFINITY : 0.0;
}

static void snapshot(cs_lattice* l)
{
    dispatch_storage(l->storage, [&]<class St>() {
        hipLaunchKernelGGL((k_snapshot<St>), cell_grid(l), dim3(256), 0, l->stream,
                           (const typename St::S*)l->buf[l->cur], l->w, l->h, l->pitch, l->plane, consts_of<St>(l),
                           (double*)l->snap);
    });
    CS_HIP(hipGetLastError());
    l->has_snap = true;
}

// Host SoA (9 dense planes of h x w) <-> the device's row-interleaved planes:
// one strided 2D copy per plane.
static void upload_pitched(cs_lattice* l, void* dst, const void* src, size_t elem)
{
    const size_t hplane = (size_t)l->w * l->h * elem;
    for (int i = 0; i < 9; ++i)
        CS_HIP(hipMemcpy2DAsync((char*)dst + (size_t)i * l->plane * elem, (size_t)l->pitch * elem,
                                (const char*)src + i * hplane, (size_t)l->w * elem, (size_t)l->w * elem,
                                (size_t)l->h, hipMemcpyHostToDevice, l->stream));
}

static void download_pitched(cs_lattice* l, void* dst, const void* src, size_t elem)
{
    const size_t hplane = (size_t)l->w * l->h * elem;
    for (int i = 0; i < 9; ++i)
        CS_HIP(hipMemcpy2DAsync((char*)dst + i * hplane, (size_t)l->w * elem,
                                (const char*)src + (size_t)i * l->plane * elem, (size_t)l->pitch * elem,
                                (size_t)l->w * elem, (size_t)l->h, hipMemcpyDeviceToHost, l->stream));
}

}  // namespace cs

using namespace cs;

// ======================================================================= ABI

extern "C" const char* cs_last_error(void) { return g_error; }
extern "C" int cs_abi_version(void) { return CS_ABI_VERSION; }

extern "C" int cs_ctx_create(int device, cs_ctx** out)
{
    return guarded([&] {
        CS_REQUIRE(out != nullptr, CS_ERR_ARG, "null output");
        int count = 0;
        CS_HIP(hipGetDeviceCount(&count));
        CS_REQUIRE(device >= 0 && device < count, CS_ERR_ARG, "device index out of range");
        CS_HIP(hipSetDevice(device));
        auto c = std::make_unique<cs_ctx>();
        c->device = device;
        CS_HIP(hipGetDeviceProperties(&c->prop, device));
        init_noalloc_scratch(device);
        *out = c.release();
    });
}

extern "C" int cs_ctx_destroy(cs_ctx* ctx)
{
    return guarded([&] { delete ctx; });
}

extern "C" int cs_ctx_device_name(cs_ctx* ctx, char* buf, size_t len)
{
    return guarded([&] {
        CS_REQUIRE(ctx && buf && len > 0, CS_ERR_ARG, "bad argument");
        std::snprintf(buf, len, "%s (%s, %d CUs)", ctx->prop.name, ctx->prop.gcnArchName,
                      ctx->prop.multiProcessorCount);
    });
}

extern "C" int cs_lattice_create(cs_ctx* ctx, int w, int h, int storage, cs_lattice** out)
{
    return guarded([&] {
        CS_REQUIRE(ctx && out, CS_ERR_ARG, "null argument");
        CS_REQUIRE(w >= 1 && h >= 1 && (int64_t)w * h <= (int64_t)1 << 31, CS_ERR_ARG, "bad lattice size");
        CS_REQUIRE(storage_bytes(storage) != 0, CS_ERR_ARG, "bad storage");
        CS_HIP(hipSetDevice(ctx->device));
        auto* l = new cs_lattice;
        l->ctx = ctx;
        l->w = w;
        l->h = h;
        l->storage = storage;
        const int align = (int)(256 / storage_bytes(storage));  // 256-byte rows
        const int wpad = (w + align - 1) / align * align;
        l->cpitch = wpad;
        l->plane = wpad;       // row-interleaved populations (cs_internal.hpp)
        l->pitch = 9 * wpad;
        try {
            const size_t bytes = (size_t)pop_elems(l) * storage_bytes(storage);
            CS_HIP(hipMalloc(&l->buf[0], bytes));
            CS_HIP(hipMalloc(&l->buf[1], bytes));
            CS_HIP(hipMemset(l->buf[0], 0, bytes));
            CS_HIP(hipMemset(l->buf[1], 0, bytes));
            CS_HIP(hipStreamCreateWithFlags(&l->stream, hipStreamNonBlocking));
            CS_HIP(hipEventCreate(&l->ev0));
            CS_HIP(hipEventCreate(&l->ev1));
            CS_HIP(hipMalloc(&l->partials, sizeof(double) * 3 * RED_BLOCKS));
            CS_HIP(hipMalloc(&l->dscratch, sizeof(double) * 8));
            CS_HIP(hipMalloc(&l->iscratch, sizeof(int) * 4));
            cs_params p{};
            p.tau = 1.0;
            p.delta_x = 1.0;
            p.delta_t = 1.0;
            p.edge = CS_EDGE_ZERO;
            p.boundary = CS_BB_REFERENCE;
            p.outlet_rho = 1.0;
            p.rho0 = 1.0;
            l->params = p;
            set_consts(l, p);
        } catch (...) {
            cs_lattice_destroy(l);
            throw;
        }
        *out = l;
    });
}

extern "C" int cs_lattice_destroy(cs_lattice* l)
{
    return guarded([&] {
        if (!l) return;
        if (l->stream) (void)hipStreamSynchronize(l->stream);
        for (void* p : {l->buf[0], l->buf[1], (void*)l->flags, (void*)l->node, (void*)l->snap, (void*)l->partials,
                        (void*)l->dscratch, (void*)l->iscratch, l->render_buf})
            if (p) (void)hipFree(p);
        clear_step_graphs(l);
        if (l->ev0) (void)hipEventDestroy(l->ev0);
        if (l->ev1) (void)hipEventDestroy(l->ev1);
        if (l->stream) (void)hipStreamDestroy(l->stream);
        delete l;
    });
}

extern "C" int cs_lattice_shape(const cs_lattice* l, int* w, int* h, int* storage)
{
    return guarded([&] {
        check_lat(l);
        if (w) *w = l->w;
        if (h) *h = l->h;
        if (storage) *storage = l->storage;
    });
}

extern "C" int cs_lattice_set_params(cs_lattice* l, const cs_params* p)
{
    return guarded([&] {
        check_lat(l);
        CS_REQUIRE(p != nullptr, CS_ERR_ARG, "null params");
        CS_REQUIRE(p->edge == CS_EDGE_ZERO || p->edge == CS_EDGE_PERIODIC, CS_ERR_ARG, "bad edge mode");
        CS_REQUIRE(p->boundary == CS_BB_REFERENCE || p->boundary == CS_BB_HALFWAY, CS_ERR_ARG, "bad boundary mode");
        CS_REQUIRE(p->tau > 0 && p->delta_x > 0 && p->delta_t > 0, CS_ERR_ARG, "tau, delta_x, delta_t must be > 0");
        CS_REQUIRE(p->outlet_rho > 0, CS_ERR_ARG, "outlet_rho must be > 0");
        CS_REQUIRE(p->rounding == CS_ROUND_NEAREST || p->rounding == CS_ROUND_STOCHASTIC, CS_ERR_ARG,
                   "bad rounding mode");
        CS_REQUIRE(p->collision >= CS_COLL_BGK && p->collision <= CS_COLL_KBC, CS_ERR_ARG, "bad collision operator");
        CS_REQUIRE(p->collision != CS_COLL_TRT || p->tau_minus > 0, CS_ERR_ARG, "tau_minus must be > 0");
        CS_REQUIRE(p->collision != CS_COLL_KBC || p->ks_viscosity > 0, CS_ERR_ARG, "ks_viscosity must be > 0");
        CS_REQUIRE(p->collision == CS_COLL_BGK || l->storage == CS_F32 || l->storage == CS_F64, CS_ERR_UNSUPPORTED,
                   "the TRT, Regularized and KBC operators run on F32 and F64 storage");
        const bool edge_changed = p->edge != l->params.edge;
        // Re-encode 16-bit storage when the deviation base changes.
        if ((l->storage == CS_F32DEV || l->storage == CS_BF16DEV || l->storage == CS_F16DEV) &&
            p->rho0 != l->params.rho0) {
            const int64_t n = (int64_t)9 * l->w * l->h;
            std::vector<float> tmp((size_t)n);
            if (cs_lattice_download_f32(l, tmp.data()) != CS_OK) throw Error{CS_ERR_HIP, std::string(g_error)};
            l->params = *p;
            set_consts(l, *p);
            if (cs_lattice_upload_f32(l, tmp.data()) != CS_OK) throw Error{CS_ERR_HIP, std::string(g_error)};
        }
        l->params = *p;
        set_consts(l, *p);
        clear_step_graphs(l);
        if (edge_changed && l->has_flags) {
            dim3 g((unsigned)((l->w + 255) / 256), (unsigned)l->h);
            hipLaunchKernelGGL(k_node_info, g, dim3(256), 0, l->stream, l->flags, l->node, l->w, l->h, l->cpitch,
                               (int)l->params.edge);
            CS_HIP(hipGetLastError());
        }
    });
}

extern "C" int cs_lattice_set_flags(cs_lattice* l, const uint8_t* flags)
{
    return guarded([&] {
        check_lat(l);
        clear_step_graphs(l);
        if (!flags) {
            l->has_flags = false;
            return;
        }
        const int64_t cells = (int64_t)l->w * l->h;
        for (int64_t d = 0; d < cells; ++d)
            CS_REQUIRE(flags[d] <= CS_OUTLET, CS_ERR_ARG, "flag value out of range");
        if (!l->flags) {
            const size_t cells = (size_t)l->cpitch * l->h;
            CS_HIP(hipMalloc(&l->flags, cells));
            CS_HIP(hipMalloc(&l->node, sizeof(uint32_t) * cells));
            CS_HIP(hipMemsetAsync(l->flags, 0, cells, l->stream));
        }
        CS_HIP(hipMemcpy2DAsync(l->flags, (size_t)l->cpitch, flags, (size_t)l->w, (size_t)l->w, (size_t)l->h,
                                hipMemcpyHostToDevice, l->stream));
        dim3 g((unsigned)((l->w + 255) / 256), (unsigned)l->h);
        hipLaunchKernelGGL(k_node_info, g, dim3(256), 0, l->stream, l->flags, l->node, l->w, l->h, l->cpitch,
                           (int)l->params.edge);
        CS_HIP(hipGetLastError());
        CS_HIP(hipStreamSynchronize(l->stream));
        l->has_flags = true;
    });
}

extern "C" int cs_lattice_upload_f32(cs_lattice* l, const float* soa)
{
    return guarded([&] {
        check_lat(l);
        CS_REQUIRE(soa != nullptr, CS_ERR_ARG, "null data");
        if (l->storage == CS_F32) {
            upload_pitched(l, l->buf[l->cur], soa, 4);
        } else {
            DevBuf tmp(sizeof(float) * (size_t)pop_elems(l));
            upload_pitched(l, tmp.p, soa, 4);
            dispatch_storage(l->storage, [&]<class St>() {
                hipLaunchKernelGGL((k_from_f32<St>), cell_grid(l), dim3(256), 0, l->stream, (const float*)tmp.p,
                                   (typename St::S*)l->buf[l->cur], l->w, l->h, l->pitch, l->plane, consts_of<St>(l));
            });
            CS_HIP(hipGetLastError());
            CS_HIP(hipStreamSynchronize(l->stream));
        }
        CS_HIP(hipStreamSynchronize(l->stream));
    });
}

extern "C" int cs_lattice_download_f32(cs_lattice* l, float* soa)
{
    return guarded([&] {
        check_lat(l);
        CS_REQUIRE(soa != nullptr, CS_ERR_ARG, "null data");
        if (l->storage == CS_F32) {
            download_pitched(l, soa, l->buf[l->cur], 4);
        } else {
            DevBuf tmp(sizeof(float) * (size_t)pop_elems(l));
            dispatch_storage(l->storage, [&]<class St>() {
                hipLaunchKernelGGL((k_to_f32<St>), cell_grid(l), dim3(256), 0, l->stream,
                                   (const typename St::S*)l->buf[l->cur], (float*)tmp.p, l->w, l->h, l->pitch, l->plane,
                                   consts_of<St>(l));
            });
            CS_HIP(hipGetLastError());
            download_pitched(l, soa, tmp.p, 4);
            CS_HIP(hipStreamSynchronize(l->stream));
        }
        CS_HIP(hipStreamSynchronize(l->stream));
    });
}

extern "C" int cs_lattice_upload_f64(cs_lattice* l, const double* soa)
{
    return guarded([&] {
        check_lat(l);
        CS_REQUIRE(soa != nullptr, CS_ERR_ARG, "null data");
        if (l->storage == CS_F64) {
            upload_pitched(l, l->buf[l->cur], soa, 8);
            CS_HIP(hipStreamSynchronize(l->stream));
            return;
        }
        const size_t n = (size_t)9 * l->w * l->h;
        std::vector<float> tmp(n);
        for (size_t j = 0; j < n; ++j) tmp[j] = (float)soa[j];
        if (cs_lattice_upload_f32(l, tmp.data()) != CS_OK) throw Error{CS_ERR_HIP, std::string(g_error)};
    });
}

extern "C" int cs_lattice_download_f64(cs_lattice* l, double* soa)
{
    return guarded([&] {
        check_lat(l);
        CS_REQUIRE(soa != nullptr, CS_ERR_ARG, "null data");
        if (l->storage == CS_F64) {
            download_pitched(l, soa, l->buf[l->cur], 8);
            CS_HIP(hipStreamSynchronize(l->stream));
            return;
        }
        const size_t n = (size_t)9 * l->w * l->h;
        std::vector<float> tmp(n);
        if (cs_lattice_download_f32(l, tmp.data()) != CS_OK) throw Error{CS_ERR_HIP, std::string(g_error)};
        for (size_t j = 0; j < n; ++j) soa[j] = (double)tmp[j];
    });
}

extern "C" int cs_lattice_upload_raw(cs_lattice* l, const void* data)
{
    return guarded([&] {
        check_lat(l);
        CS_REQUIRE(data != nullptr, CS_ERR_ARG, "null data");
        upload_pitched(l, l->buf[l->cur], data, storage_bytes(l->storage));
        CS_HIP(hipStreamSynchronize(l->stream));
    });
}

extern "C" int cs_lattice_download_raw(cs_lattice* l, void* data)
{
    return guarded([&] {
        check_lat(l);
        CS_REQUIRE(data != nullptr, CS_ERR_ARG, "null data");
        download_pitched(l, data, l->buf[l->cur], storage_bytes(l->storage));
        CS_HIP(hipStreamSynchronize(l->stream));
    });
}

template <class T>
static void init_eq_impl(cs_lattice* l, const T* rho, const T* ux, const T* uy)
{
    CS_REQUIRE(rho && ux && uy, CS_ERR_ARG, "null input");
    const int64_t n = (int64_t)l->w * l->h;
    dispatch_storage(l->storage, [&]<class St>() {
        using TT = typename St::T;
        std::vector<TT> host((size_t)(3 * n));
        for (int64_t d = 0; d < n; ++d) {
            host[(size_t)d] = (TT)rho[d];
            host[(size_t)(n + d)] = (TT)ux[d];
            host[(size_t)(2 * n + d)] = (TT)uy[d];
        }
        DevBuf dev(sizeof(TT) * 3 * (size_t)n);
        CS_HIP(hipMemcpyAsync(dev.p, host.data(), sizeof(TT) * 3 * (size_t)n, hipMemcpyHostToDevice, l->stream));
        const TT* p = (const TT*)dev.p;
        hipLaunchKernelGGL((k_init_eq<St>), cell_grid(l), dim3(256), 0, l->stream, p, p + n, p + 2 * n,
                           (typename St::S*)l->buf[l->cur], l->w, l->h, l->pitch, l->plane, consts_of<St>(l));
        CS_HIP(hipGetLastError());
        CS_HIP(hipStreamSynchronize(l->stream));
    });
}

extern "C" int cs_lattice_init_nonequilibrium(cs_lattice* l, const float* rho, const float* ux, const float* uy)
{
    return guarded([&] {
        check_lat(l);
        CS_REQUIRE(rho && ux && uy, CS_ERR_ARG, "null input");
        const int64_t n = (int64_t)l->w * l->h;
        dispatch_storage(l->storage, [&]<class St>() {
            using TT = typename St::T;
            std::vector<TT> host((size_t)(3 * n));
            for (int64_t d = 0; d < n; ++d) {
                host[(size_t)d] = (TT)rho[d];
                host[(size_t)(n + d)] = (TT)ux[d];
                host[(size_t)(2 * n + d)] = (TT)uy[d];
            }
            DevBuf dev(sizeof(TT) * 3 * (size_t)n);
            CS_HIP(hipMemcpyAsync(dev.p, host.data(), sizeof(TT) * 3 * (size_t)n, hipMemcpyHostToDevice, l->stream));
            const TT* p = (const TT*)dev.p;
            hipLaunchKernelGGL((k_init_neq<St>), cell_grid(l), dim3(256), 0, l->stream, p, p + n, p + 2 * n,
                               l->has_flags ? (const uint8_t*)l->flags : nullptr, (typename St::S*)l->buf[l->cur],
                               l->w, l->h, l->pitch, l->cpitch, l->plane, consts_of<St>(l),
                               NeqArgs<TT>{(TT)l->params.tau, (TT)l->params.lid_ux, (TT)l->params.lid_uy,
                                           l->params.edge == CS_EDGE_PERIODIC ? 1 : 0});
            CS_HIP(hipGetLastError());
            CS_HIP(hipStreamSynchronize(l->stream));
        });
    });
}

extern "C" int cs_lattice_init_equilibrium(cs_lattice* l, const float* rho, const float* ux, const float* uy)
{
    return guarded([&] {
        check_lat(l);
        init_eq_impl<float>(l, rho, ux, uy);
    });
}

extern "C" int cs_lattice_init_equilibrium_f64(cs_lattice* l, const double* rho, const double* ux,
                                               const double* uy)
{
    return guarded([&] {
        check_lat(l);
        init_eq_impl<double>(l, rho, ux, uy);
    });
}

extern "C" int cs_lattice_step(cs_lattice* l, int n)
{
    return guarded([&] {
        check_lat(l);
        CS_REQUIRE(n >= 0, CS_ERR_ARG, "negative step count");
        do_steps(l, n);
    });
}

extern "C" int cs_lattice_step_timed(cs_lattice* l, int n, float* ms)
{
    return guarded([&] {
        check_lat(l);
        CS_REQUIRE(n >= 0 && ms, CS_ERR_ARG, "bad argument");
        CS_HIP(hipEventRecord(l->ev0, l->stream));
        do_steps(l, n);
        CS_HIP(hipEventRecord(l->ev1, l->stream));
        CS_HIP(hipEventSynchronize(l->ev1));
        CS_HIP(hipEventElapsedTime(ms, l->ev0, l->ev1));
    });
}

extern "C" int cs_lattice_temporal_steps(const cs_lattice* l, int* k)
{
    return guarded([&] {
        check_lat(l);
        CS_REQUIRE(k, CS_ERR_ARG, "null output");
        const int t = tb_plan_steps(tb_steps(l, g_cpt_override.load(std::memory_order_relaxed)));
        *k = t > 0 ? t : 1;
    });
}

static void require_unfused(const cs_lattice* l)
{
    CS_REQUIRE(l->storage == CS_F32 || l->storage == CS_F64, CS_ERR_UNSUPPORTED,
               "stream/bounce_back/collide alone need F32 or F64 storage");
    CS_REQUIRE(l->params.boundary == CS_BB_REFERENCE, CS_ERR_UNSUPPORTED,
               "stream/bounce_back/collide alone are defined for the reference boundary mode");
}

extern "C" int cs_lattice_stream(cs_lattice* l)
{
    return guarded([&] {
        check_lat(l);
        require_unfused(l);
        dispatch_storage(l->storage, [&]<class St>() {
            if constexpr (std::is_same_v<St, StF32> || std::is_same_v<St, StF64>) {
                hipLaunchKernelGGL((k_stream<St>), cell_grid(l), dim3(256), 0, l->stream,
                                   (const typename St::S*)l->buf[l->cur], (typename St::S*)l->buf[l->cur ^ 1], l->w,
                                   l->h, l->pitch, l->plane, (int)l->params.edge);
            }
        });
        CS_HIP(hipGetLastError());
        l->cur ^= 1;
    });
}

extern "C" int cs_lattice_bounce_back(cs_lattice* l)
{
    return guarded([&] {
        check_lat(l);
        require_unfused(l);
        if (!l->has_flags) return;
        dispatch_storage(l->storage, [&]<class St>() {
            if constexpr (std::is_same_v<St, StF32> || std::is_same_v<St, StF64>) {
                hipLaunchKernelGGL((k_bounce_back<St>), cell_grid(l), dim3(256), 0, l->stream,
                                   (typename St::S*)l->buf[l->cur], (const uint8_t*)l->flags, l->w, l->h, l->pitch,
                                   l->cpitch, l->plane);
            }
        });
        CS_HIP(hipGetLastError());
    });
}

extern "C" int cs_lattice_collide(cs_lattice* l)
{
    return guarded([&] {
        check_lat(l);
        require_unfused(l);
        dispatch_storage(l->storage, [&]<class St>() {
            if constexpr (std::is_same_v<St, StF32> || std::is_same_v<St, StF64>) {
                auto go = [&]<int COLL>() {
                    hipLaunchKernelGGL((k_collide<St, COLL>), cell_grid(l), dim3(256), 0, l->stream,
                                       (typename St::S*)l->buf[l->cur], l->w, l->h, l->pitch, l->plane,
                                       consts_of<St>(l));
                };
                if (l->params.collision == CS_COLL_TRT) go.template operator()<CS_COLL_TRT>();
                else if (l->params.collision == CS_COLL_REGULARIZED) go.template operator()<CS_COLL_REGULARIZED>();
                else if (l->params.collision == CS_COLL_KBC) go.template operator()<CS_COLL_KBC>();
                else
                    hipLaunchKernelGGL((k_collide<St, CS_COLL_BGK>), cell_grid(l), dim3(256), 0, l->stream,
                                       (typename St::S*)l->buf[l->cur], l->w, l->h, l->pitch, l->plane,
                                       consts_of<St>(l));
            }
        });
        CS_HIP(hipGetLastError());
        // State::step advances time after collide (lbm.rs:713); the phases
        // alone leave it to the caller, as the reference's methods do.
    });
}

extern "C" int cs_lattice_time(const cs_lattice* l, double* t)
{
    return guarded([&] {
        check_lat(l);
        CS_REQUIRE(t != nullptr, CS_ERR_ARG, "null output");
        *t = l->time;
    });
}

extern "C" int cs_lattice_sync(cs_lattice* l)
{
    return guarded([&] {
        check_lat(l);
        CS_HIP(hipStreamSynchronize(l->stream));
    });
}

extern "C" int cs_lattice_device_ptr(cs_lattice* l, void** f, int* pitch, int64_t* plane)
{
    return guarded([&] {
        check_lat(l);
        if (f) *f = l->buf[l->cur];
        if (pitch) *pitch = l->pitch;
        if (plane) *plane = l->plane;
    });
}

extern "C" int cs_lattice_stream_handle(cs_lattice* l, void** stream)
{
    return guarded([&] {
        check_lat(l);
        CS_REQUIRE(stream != nullptr, CS_ERR_ARG, "null output");
        *stream = (void*)l->stream;
    });
}

extern "C" int cs_lattice_density(cs_lattice* l, float* rho)
{
    return guarded([&] {
        check_lat(l);
        CS_REQUIRE(rho, CS_ERR_ARG, "null output");
        float* o[] = {rho};
        moments_f32(l, M_RHO, o);
    });
}

extern "C" int cs_lattice_momentum_density(cs_lattice* l, float* mx, float* my)
{
    return guarded([&] {
        check_lat(l);
        CS_REQUIRE(mx && my, CS_ERR_ARG, "null output");
        float* o[] = {mx, my};
        moments_f32(l, M_MX | M_MY, o);
    });
}

extern "C" int cs_lattice_velocity(cs_lattice* l, float* ux, float* uy)
{
    return guarded([&] {
        check_lat(l);
        CS_REQUIRE(ux && uy, CS_ERR_ARG, "null output");
        float* o[] = {ux, uy};
        moments_f32(l, M_UX | M_UY, o);
    });
}

extern "C" int cs_lattice_speed(cs_lattice* l, float* speed)
{
    return guarded([&] {
        check_lat(l);
        CS_REQUIRE(speed, CS_ERR_ARG, "null output");
        float* o[] = {speed};
        moments_f32(l, M_SPEED, o);
    });
}

extern "C" int cs_lattice_pressure(cs_lattice* l, float* p)
{
    return guarded([&] {
        check_lat(l);
        CS_REQUIRE(p, CS_ERR_ARG, "null output");
        float* o[] = {p};
        moments_f32(l, M_P, o);
    });
}

static void equilibrium_out(cs_lattice* l, int non_eq, float* out)
{
    CS_REQUIRE(out, CS_ERR_ARG, "null output");
    const int64_t n = (int64_t)l->w * l->h;
    DevBuf dev(sizeof(double) * 9 * (size_t)n);
    dispatch_storage(l->storage, [&]<class St>() {
        hipLaunchKernelGGL((k_equilibrium<St>), cell_grid(l), dim3(256), 0, l->stream,
                           (const typename St::S*)l->buf[l->cur], l->w, l->h, l->pitch, l->plane, consts_of<St>(l),
                           non_eq, (double*)dev.p);
    });
    CS_HIP(hipGetLastError());
    std::vector<double> tmp((size_t)(9 * n));
    CS_HIP(hipMemcpyAsync(tmp.data(), dev.p, sizeof(double) * 9 * (size_t)n, hipMemcpyDeviceToHost, l->stream));
    CS_HIP(hipStreamSynchronize(l->stream));
    for (size_t j = 0; j < tmp.size(); ++j) out[j] = (float)tmp[j];
}

extern "C" int cs_lattice_equilibrium(cs_lattice* l, float* feq)
{
    return guarded([&] {
        check_lat(l);
        equilibrium_out(l, 0, feq);
    });
}

extern "C" int cs_lattice_non_equilibrium(cs_lattice* l, float* fneq)
{
    return guarded([&] {
        check_lat(l);
        equilibrium_out(l, 1, fneq);
    });
}

extern "C" int cs_lattice_is_unstable(cs_lattice* l, int* unstable)
{
    return guarded([&] {
        check_lat(l);
        CS_REQUIRE(unstable, CS_ERR_ARG, "null output");
        CS_HIP(hipMemsetAsync(l->iscratch, 0, sizeof(int), l->stream));
        dispatch_storage(l->storage, [&]<class St>() {
            hipLaunchKernelGGL((k_unstable<St>), cell_grid(l), dim3(256), 0, l->stream,
                               (const typename St::S*)l->buf[l->cur], l->w, l->h, l->pitch, l->plane,
                               consts_of<St>(l), l->iscratch);
        });
        CS_HIP(hipGetLastError());
        int v = 0;
        CS_HIP(hipMemcpyAsync(&v, l->iscratch, sizeof(int), hipMemcpyDeviceToHost, l->stream));
        CS_HIP(hipStreamSynchronize(l->stream));
        *unstable = v != 0;
    });
}

extern "C" int cs_lattice_moments_f64(cs_lattice* l, double* rho, double* ux, double* uy)
{
    return guarded([&] {
        check_lat(l);
        const int64_t n = (int64_t)l->w * l->h;
        std::vector<double> tmp((size_t)(3 * n));
        moments_to_host(l, M_RHO | M_UX | M_UY, tmp.data());
        if (rho) std::memcpy(rho, tmp.data(), sizeof(double) * n);
        if (ux) std::memcpy(ux, tmp.data() + n, sizeof(double) * n);
        if (uy) std::memcpy(uy, tmp.data() + 2 * n, sizeof(double) * n);
    });
}

extern "C" int cs_lattice_moments_f64_noalloc(cs_lattice* l, double* rho, double* ux, double* uy)
{
    return guarded([&] {
        check_lat(l);
        const int64_t n = (int64_t)l->w * l->h;
        const unsigned mask = (rho ? M_RHO : 0u) | (ux ? M_UX : 0u) | (uy ? M_UY : 0u);
        const int nplanes = __builtin_popcount(mask);
        if (nplanes == 0) return;
        CS_REQUIRE(n * nplanes <= (int64_t)CS_NOALLOC_MAX_VALUES, CS_ERR_ARG,
                   "lattice too large for the no-allocation moment pass: planes x cells must be at most "
                   "CS_NOALLOC_MAX_VALUES (3 x 2^20)");
        const int dev = l->ctx->device;
        CS_HIP(hipSetDevice(dev));
        CS_REQUIRE(dev >= 0 && dev < MAX_DEVICES && g_scratch_ptr[dev] != nullptr, CS_ERR_STATE,
                   "moment scratch not set up for this device");
        double* const ptr = g_scratch_ptr[dev];
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        dispatch_storage(l->storage, [&]<class St>() {
            hipLaunchKernelGGL((k_moments<St>), cell_grid(l), dim3(256), 0, l->stream,
                               (const typename St::S*)l->buf[l->cur], l->w, l->h, l->pitch, l->plane, consts_of<St>(l),
                               mask, ptr);
        });
        CS_HIP(hipGetLastError());
        int slot = 0;
        for (double* o : {rho, ux, uy})
            if (o) {
                CS_HIP(hipMemcpyAsync(o, ptr + (int64_t)(slot++) * n, sizeof(double) * (size_t)n,
                                      hipMemcpyDeviceToHost, l->stream));
            }
        CS_HIP(hipStreamSynchronize(l->stream));
    });
}

extern "C" int cs_lattice_sums_f64(cs_lattice* l, double out[3])
{
    return guarded([&] {
        check_lat(l);
        CS_REQUIRE(out, CS_ERR_ARG, "null output");
        dispatch_storage(l->storage, [&]<class St>() {
            hipLaunchKernelGGL((k_sums<St>), dim3(RED_BLOCKS), dim3(RED_THREADS), 0, l->stream,
                               (const typename St::S*)l->buf[l->cur], l->has_flags ? l->flags : nullptr, l->w, l->h,
                               l->pitch, l->cpitch, l->plane, consts_of<St>(l), l->partials);
        });
        CS_HIP(hipGetLastError());
        reduce(l, 3, out);
    });
}

extern "C" int cs_lattice_snapshot_velocity(cs_lattice* l)
{
    return guarded([&] {
        check_lat(l);
        if (!l->snap) CS_HIP(hipMalloc(&l->snap, sizeof(double) * 2 * (size_t)l->w * l->h));
        snapshot(l);
        CS_HIP(hipStreamSynchronize(l->stream));
    });
}

extern "C" int cs_lattice_residual_f64(cs_lattice* l, double* residual)
{
    return guarded([&] {
        check_lat(l);
        CS_REQUIRE(residual, CS_ERR_ARG, "null output");
        *residual = residual_now(l);
    });
}

extern "C" int cs_lattice_run_to_residual(cs_lattice* l, double tol, int interval, int64_t max_steps,
                                          int64_t* steps, double* residual)
{
    return guarded([&] {
        check_lat(l);
        CS_REQUIRE(interval >= 1 && max_steps >= 0 && steps && residual, CS_ERR_ARG, "bad argument");
        if (!l->snap) CS_HIP(hipMalloc(&l->snap, sizeof(double) * 2 * (size_t)l->w * l->h));
        int64_t done = 0;
        double res = INFINITY;
        while (done < max_steps) {
            snapshot(l);
            const int n = (int)std::min<int64_t>(interval, max_steps - done);
            do_steps(l, n);
            done += n;
            res = residual_now(l);
            if (!(res == res)) break;  // NaN: diverged
            if (res <= tol) break;
        }
        *steps = done;
        *residual = res;
    });
}

__global__ void k_bgk_evaluate(const float* __restrict__ f, const float* __restrict__ feq, float K,
                               float* __restrict__ out, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = f[i] + (f[i] - feq[i]) * K;
}

// BGK::evaluate (lbm.rs:350-364) on caller-supplied populations, for the
// unfused plugin path: out = f + (f - f_eq) * K with K = (-dt) / tau in f32.
extern "C" int cs_bgk_evaluate(cs_ctx* ctx, int64_t n, const float* f, const float* feq, float tau,
                               float delta_t, float* out)
{
    return guarded([&] {
        CS_REQUIRE(ctx && f && feq && out && n >= 0, CS_ERR_ARG, "bad argument");
        CS_REQUIRE(tau > 0.f && delta_t > 0.f, CS_ERR_ARG, "tau and delta_t must be positive");
        if (n == 0) return;
        CS_HIP(hipSetDevice(ctx->device));
        volatile float dt = delta_t, t = tau;
        const float K = (-dt) / t;
        DevBuf d((size_t)n * 3 * sizeof(float));
        float* df = (float*)d.p;
        CS_HIP(hipMemcpy(df, f, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
        CS_HIP(hipMemcpy(df + n, feq, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_bgk_evaluate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, df, df + n, K,
                           df + 2 * n, n);
        CS_HIP(hipGetLastError());
        CS_HIP(hipMemcpy(out, df + 2 * n, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
    });
}

// TRT::new (lbm.rs:380-390), evaluated in f32 like the reference.
extern "C" int cs_trt_new(float lambda, float ks_viscosity, float delta_x, float delta_t, float* tau_plus,
                          float* tau_minus)
{
    return guarded([&] {
        CS_REQUIRE(tau_plus && tau_minus && delta_x > 0.f && delta_t > 0.f, CS_ERR_ARG, "bad argument");
        volatile float dx = delta_x, dt = delta_t, three = 3.0f;
        const float cs = dx / (std::sqrt((float)three) * dt);
        const float tp = dt * ((ks_viscosity / (cs * cs)) + 0.5f);
        const float tm = dt * ((lambda / ((tp / dt) - 0.5f)) + 0.5f);
        *tau_plus = tp;
        *tau_minus = tm;
    });
}

extern "C" int cs_bench_copy(cs_ctx* ctx, int64_t bytes, int iters, float* ms_per_copy)
{
    return guarded([&] {
        CS_REQUIRE(ctx && ms_per_copy && bytes >= 16 && iters >= 1, CS_ERR_ARG, "bad argument");
        CS_HIP(hipSetDevice(ctx->device));
        const int64_t n4 = bytes / 16;
        DevBuf a((size_t)n4 * 16), b((size_t)n4 * 16);
        CS_HIP(hipMemset(a.p, 0, (size_t)n4 * 16));
        hipStream_t s;
        CS_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        hipEvent_t e0, e1;
        CS_HIP(hipEventCreate(&e0));
        CS_HIP(hipEventCreate(&e1));
        const int blocks = (int)((n4 + 255) / 256);
        hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, s, (const f32x4*)a.p, (f32x4*)b.p, n4);
        CS_HIP(hipEventRecord(e0, s));
        for (int i = 0; i < iters; ++i)
            hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, s, (const f32x4*)a.p, (f32x4*)b.p, n4);
        CS_HIP(hipEventRecord(e1, s));
        CS_HIP(hipEventSynchronize(e1));
        float ms = 0;
        CS_HIP(hipEventElapsedTime(&ms, e0, e1));
        *ms_per_copy = ms / iters;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
    });
}

// Every 32-bit pattern b: recip_rn(b) against the IEEE quotient one / b (one
// = 1 from the host, so nothing folds); two NaNs count as equal.  65536 blocks
// x 256 threads x 256 inputs = 2^32.
__global__ void k_check_reciprocal(unsigned long long* __restrict__ bad, float one)
{
    const uint64_t base = ((uint64_t)blockIdx.x * 256u + threadIdx.x) * 256u;
    unsigned int n = 0;
    for (uint32_t k = 0; k < 256u; ++k) {
        const float b = __builtin_bit_cast(float, (uint32_t)(base + k));
        const float ref = one / b, got = recip_rn(b);
        const bool same = __builtin_bit_cast(uint32_t, ref) == __builtin_bit_cast(uint32_t, got) ||
                          (ref != ref && got != got);
        n += same ? 0u : 1u;
    }
    if (n) atomicAdd(bad, (unsigned long long)n);
}

extern "C" int cs_check_reciprocal(cs_ctx* ctx, uint64_t* mismatches)
{
    return guarded([&] {
        CS_REQUIRE(ctx && mismatches, CS_ERR_ARG, "bad argument");
        CS_HIP(hipSetDevice(ctx->device));
        DevBuf d(sizeof(unsigned long long));
        CS_HIP(hipMemset(d.p, 0, sizeof(unsigned long long)));
        hipLaunchKernelGGL(k_check_reciprocal, dim3(65536), dim3(256), 0, nullptr, (unsigned long long*)d.p, 1.0f);
        CS_HIP(hipGetLastError());
        unsigned long long n = 0;
        CS_HIP(hipMemcpy(&n, d.p, sizeof n, hipMemcpyDeviceToHost));
        *mismatches = (uint64_t)n;
    });
}

// Launch-shape override: cells per thread of the scalar step kernel (1, 2 or
// 4), CS_STEP_VECTOR for the vector kernel wherever rows split into vectors;
// 0 restores the choice by lattice size.
// Graph replay of fused steps on launch-bound lattices (on by default).
extern "C" int cs_set_step_graphs(int enable)
{
    return guarded([&] { g_step_graphs.store(enable != 0, std::memory_order_relaxed); });
}

extern "C" int cs_set_cells_per_thread(int cpt)
{
    return guarded([&] {
        CS_REQUIRE(cpt == 0 || cpt == 1 || cpt == 2 || cpt == 4 || cpt == CS_STEP_VECTOR || cpt == CS_STEP_VECTOR_WIDE,
                   CS_ERR_ARG, "cells per thread must be 0, 1, 2, 4, CS_STEP_VECTOR or CS_STEP_VECTOR_WIDE");
        g_cpt_override.store(cpt, std::memory_order_relaxed);
    });
}
