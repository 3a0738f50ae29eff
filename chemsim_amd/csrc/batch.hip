// Batch of independent lattices stepped together.
//
// The reference steps one lattice per State::step (lbm.rs:694-714).  A batch
// holds many lattices of uneven sizes in one device arena, with two launch
// kinds chosen by the tuning:
//  - persistent (k_batch): a grid whose workgroups take whole lattices from a
//    device-side queue (largest first by default) and run every step of the
//    call on them, with a workgroup barrier between steps -- no
//    inter-workgroup synchronisation;
//  - tiles (k_batch_tiles): one launch per step over the flattened list of
//    4 x 64 tiles of every lattice, so a large lattice is spread over the
//    whole GPU;
//  - temporal (k_batch_tb / k_batch_tbr, 32-bit storage): one launch per 8
//    steps over the flattened list of 64 x 64 regions of every lattice, each
//    block running the 8 steps on its region in LDS or registers (lbm_tb.hpp);
//    the remainder of a call's steps (steps mod 8) runs as tile launches;
//  - persistent temporal (k_batch_tb_grid / k_batch_tbr_grid): the same
//    rounds of 8 steps in one cooperative launch, a co-resident grid whose
//    blocks each take a contiguous share of the region list and meet at a
//    grid-wide barrier after every round.
// The per-cell code is the single-lattice gather/finish either way, so the
// results equal per-lattice launches bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <functional>
#include <numeric>
#include <type_traits>
#include <vector>

#include "cs_internal.hpp"
#include "lattice_kernels.hpp"
#include "lbm_device.hpp"
#include "lbm_tb.hpp"

namespace cs {

// Per-lattice descriptor in device memory.
template <class T>
struct BatchLat {
    int w, h, pitch, mode;  // mode bits: 1 periodic, 2 half-way, 4 has flags, 8 stochastic rounding
    int cpitch;             // row pitch of the node words (elements)
    int64_t plane;          // plane stride (elements); populations row-interleaved, pitch = 9 * plane
    int64_t off;            // element offset of buffer 0 in the arena (buffer 1 follows, h * pitch later)
    int64_t node_off;       // element offset in the node arena
    uint32_t seed;
    Consts<T> k;
};

enum : int { BM_PERIODIC = 1, BM_HALFWAY = 2, BM_FLAGS = 4, BM_SR = 8 };

// Per-lattice step state: which buffer holds the current populations and how
// many steps the lattice has taken (the stochastic-rounding key).  Lattices
// taken out of the batch keep theirs.
struct LatState {
    int cur;
    int pad;
    uint64_t steps;
};

// One entry of the flattened tile list: lattice index and tile index within
// it (tiles of 64 columns x TR rows, row-major).
struct TileRef {
    int lat;
    int tile;
};

// One step of one lattice by one workgroup: wave `wave` of `nwaves` takes
// rows wave, wave + nwaves, ...; lane l the columns l + 64 j.
template <class St, int EDGE, int MODE, bool SR, int CPT>
__device__ __forceinline__ void batch_lattice_step(const BatchLat<typename St::T>& L, const typename St::S* fin,
                                                   typename St::S* fout, const uint32_t* node, uint32_t key,
                                                   int wave, int nwaves, int lane)
{
    using T = typename St::T;
    for (int r = wave; r < L.h; r += nwaves) {
        for (int c0 = lane; c0 < L.w; c0 += 64 * CPT) {
            T B[CPT][9];
            int type[CPT];
#pragma unroll
            for (int j = 0; j < CPT; ++j) {
                const int c = c0 + 64 * j;
                const int cc = c < L.w ? c : L.w - 1;
                type[j] = gather<St, EDGE, MODE, uint32_t>(fin, node, L.w, L.h, L.pitch, L.cpitch, L.plane, L.k, r, cc, B[j]);
            }
#pragma unroll
            for (int j = 0; j < CPT; ++j) {
                const int c = c0 + 64 * j;
                if (c < L.w)
                    finish<St, MODE, SR, uint32_t>(fin, fout, L.pitch, L.plane, L.k, r, c, type[j], B[j], key,
                                         (uint32_t)(r * L.w + c));
            }
        }
    }
}

// Persistent grid over the lattices of one boundary-mode group (order[0..n)).
// The queue counter starts at 0; every workgroup leaves the loop once the
// counter passes n, so the grid always drains.
template <class St, int EDGE, int MODE, bool SR, int CPT>
__global__ __launch_bounds__(1024) void k_batch(const BatchLat<typename St::T>* __restrict__ lats,
                                                const int* __restrict__ order, int n, int* __restrict__ queue,
                                                void* __restrict__ arena, const uint32_t* __restrict__ nodes,
                                                const LatState* __restrict__ state, int steps)
{
    using S = typename St::S;
    __shared__ int s_q;
    const int nwaves = (int)(blockDim.x >> 6), wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    for (;;) {
        if (threadIdx.x == 0) s_q = atomicAdd(queue, 1);
        __syncthreads();
        const int q = __builtin_amdgcn_readfirstlane(s_q);
        __syncthreads();
        if (q >= n) return;
        const int li = __builtin_amdgcn_readfirstlane(order[q]);  // uniform: descriptor loads go scalar
        const BatchLat<typename St::T> L = lats[li];
        S* base = (S*)arena + L.off;
        const uint32_t* node = nodes + L.node_off;
        const int cur0 = state[li].cur;
        const uint64_t step0 = state[li].steps;
        for (int s = 0; s < steps; ++s) {
            const int cur = (cur0 + s) & 1;
            const S* fin = base + (int64_t)cur * L.h * L.pitch;
            S* fout = base + (int64_t)(cur ^ 1) * L.h * L.pitch;
            const uint32_t key = sr_key(L.seed, step0 + (uint64_t)s);
            batch_lattice_step<St, EDGE, MODE, SR, CPT>(L, fin, fout, node, key, wave, nwaves, lane);
            __syncthreads();  // step s fully stored before step s + 1 reads it
        }
    }
}

// Per-lattice velocity snapshot (one block per lattice).
template <class St>
__global__ __launch_bounds__(256) void k_batch_snapshot(const BatchLat<typename St::T>* __restrict__ lats, int n,
                                                        const void* __restrict__ arena,
                                                        const LatState* __restrict__ state, float* __restrict__ snap,
                                                        const int64_t* __restrict__ snap_off)
{
    using T = typename St::T;
    const int li = blockIdx.x;
    if (li >= n) return;
    const BatchLat<T> L = lats[li];
    const typename St::S* f = (const typename St::S*)arena + L.off + (int64_t)state[li].cur * L.h * L.pitch;
    const int64_t cells = (int64_t)L.w * L.h;
    float* s = snap + snap_off[li];
    for (int64_t d = threadIdx.x; d < cells; d += 256) {
        const int r = (int)(d / L.w), c = (int)(d - (int64_t)r * L.w);
        const int64_t own = (int64_t)r * L.pitch + c;
        T B[9];
        load_own<St>(f, L.plane, own, B);
        T rho, mx, my, ux, uy;
        moments<St>(L.k, B, rho, mx, my, ux, uy);
        s[d] = (float)ux;
        s[cells + d] = (float)uy;
    }
}

// Per-lattice residual ||u - u_snap|| / ||u|| over non-solid cells, f64.
template <class St>
__global__ __launch_bounds__(256) void k_batch_residual(const BatchLat<typename St::T>* __restrict__ lats, int n,
                                                        const void* __restrict__ arena,
                                                        const LatState* __restrict__ state,
                                                        const uint32_t* __restrict__ nodes,
                                                        const float* __restrict__ snap,
                                                        const int64_t* __restrict__ snap_off, double* __restrict__ out)
{
    using T = typename St::T;
    __shared__ double sh[2][256];
    const int li = blockIdx.x;
    if (li >= n) return;
    const BatchLat<T> L = lats[li];
    const typename St::S* f = (const typename St::S*)arena + L.off + (int64_t)state[li].cur * L.h * L.pitch;
    const int64_t cells = (int64_t)L.w * L.h;
    const float* s = snap + snap_off[li];
    double num = 0.0, den = 0.0;
    for (int64_t d = threadIdx.x; d < cells; d += 256) {
        const int r = (int)(d / L.w), c = (int)(d - (int64_t)r * L.w);
        const int64_t own = (int64_t)r * L.pitch + c;
        if (L.mode & BM_FLAGS) {
            const int t = (int)(nodes[L.node_off + (int64_t)r * L.cpitch + c] >> NODE_TYPE_SHIFT);
            if (t == CS_WALL || t == CS_LID) continue;
        }
        T B[9];
        load_own<St>(f, L.plane, own, B);
        T rho, mx, my, ux, uy;
        moments<St>(L.k, B, rho, mx, my, ux, uy);
        const double dx = (double)(float)ux - (double)s[d], dy = (double)(float)uy - (double)s[cells + d];
        num += dx * dx + dy * dy;
        den += (double)ux * (double)ux + (double)uy * (double)uy;
    }
    sh[0][threadIdx.x] = num;
    sh[1][threadIdx.x] = den;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + k];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[li] = sh[1][0] > 0 ? sqrt(sh[0][0] / sh[1][0]) : (sh[0][0] > 0 ? INFINITY : 0.0);
}

// Tile of block b: blocks go to the 8 XCDs round-robin; this deals the list to
// them in chunks of K consecutive tiles, so neighbouring tiles (shared halo
// rows) run on one XCD and its L2 while every XCD gets a mix of the list (the
// list runs from the largest lattice to the smallest, so one contiguous run
// per XCD would leave the XCD with the small lattices' part-empty tiles short
// of work).  Blocks past the last whole round of 8 chunks keep their index.
// 512 cavities of sides 64..512, 4-row tiles, microseconds per step: K = 1
// (plain round-robin) 861, K = 2 838, K = 4 783, K = 8 808, K = 16 807, K = 64
// 802, one contiguous run per XCD 858.
template <int K>
__device__ __forceinline__ int xcd_chunk_tile(int b, int n)
{
    const int full = n / (8 * K) * (8 * K);
    if (b >= full) return b;
    const int x = b & 7, k = b >> 3;
    return ((k / K) * 8 + x) * K + (k % K);
}

// One step of every lattice in a tile list: each block updates one tile
// (TR rows x 64 columns, one cell per thread) of its lattice with the
// single-lattice gather/finish.  `s` is the step's index within the call,
// `flip` the number of buffer flips of the call before it (s, unless
// temporally blocked launches came first).  Each wave's 64 columns start on
// a 64-column boundary, so its row segment of a plane is whole aligned cache
// lines.
template <class St, int EDGE, int MODE, bool SR, int TR>
__global__ __launch_bounds__(64 * TR) void k_batch_tiles(const BatchLat<typename St::T>* __restrict__ lats,
                                                     const TileRef* __restrict__ tiles, void* __restrict__ arena,
                                                     const uint32_t* __restrict__ nodes,
                                                     const LatState* __restrict__ state, int flip, int s)
{
    using S = typename St::S;
    using T = typename St::T;
    const TileRef tr = tiles[xcd_chunk_tile<4>(blockIdx.x, gridDim.x)];
    const int li = __builtin_amdgcn_readfirstlane(tr.lat);
    const int t = __builtin_amdgcn_readfirstlane(tr.tile);
    const BatchLat<T> L = lats[li];
    const int tiles_x = (L.w + 63) >> 6;
    const int ty = t / tiles_x, tx = t - ty * tiles_x;
    const int r = ty * TR + (int)(threadIdx.x >> 6);
    if (r >= L.h) return;
    const int c = tx * 64 + (int)(threadIdx.x & 63);
    const int cur = (state[li].cur + flip) & 1;
    const uint32_t key = sr_key(L.seed, state[li].steps + (uint64_t)s);
    const S* fin = (const S*)arena + L.off + (int64_t)cur * L.h * L.pitch;
    S* fout = (S*)arena + L.off + (int64_t)(cur ^ 1) * L.h * L.pitch;
    const uint32_t* node = nodes + L.node_off;
    T B[9];
    const int cc = c < L.w ? c : L.w - 1;
    const int type = gather<St, EDGE, MODE, uint32_t>(fin, node, L.w, L.h, L.pitch, L.cpitch, L.plane, L.k, r, cc, B);
    if (c < L.w)
        finish<St, MODE, SR, uint32_t>(fin, fout, L.pitch, L.plane, L.k, r, c, type, B, key, (uint32_t)(r * L.w + c));
}

// K steps of every lattice in a region list: each block runs tb_region on one
// 64 x 64 region of its lattice (region t of its ceil(w / (64-2K)) x
// ceil(h / (64-2K))).  `flip` is the number of buffer flips of the call
// before this launch.  The list is grouped by lattice, so regions that share
// halo cells are neighbours in it and xcd_tile keeps them on one XCD.
template <class St, int EDGE, int MODE, int K>
__device__ __forceinline__ void batch_tb_region(const BatchLat<typename St::T>* __restrict__ lats,
                                                const TileRef* __restrict__ regions, void* __restrict__ arena,
                                                const uint32_t* __restrict__ nodes,
                                                const LatState* __restrict__ state, int flip,
                                                StepArgs<typename St::T>& a, int& ty, int& tx, bool direct = false)
{
    using S = typename St::S;
    using T = typename St::T;
    constexpr int IN = TB_EDGE - 2 * K;
    const TileRef tr = direct ? regions[0] : regions[xcd_tile(blockIdx.x, gridDim.x)];
    const int li = __builtin_amdgcn_readfirstlane(tr.lat);
    const int t = __builtin_amdgcn_readfirstlane(tr.tile);
    const BatchLat<T> D = lats[li];
    const int tiles_x = (D.w + IN - 1) / IN;
    ty = t / tiles_x;
    const int cur = (state[li].cur + flip) & 1;
    a.fin = (const S*)arena + D.off + (int64_t)cur * D.h * D.pitch;
    a.fout = (S*)arena + D.off + (int64_t)(cur ^ 1) * D.h * D.pitch;
    a.node = nodes + D.node_off;
    a.w = D.w;
    a.h = D.h;
    a.pitch = D.pitch;
    a.cpitch = D.cpitch;
    a.plane = D.plane;
    a.tiles_x = tiles_x;
    a.k = D.k;
    tx = t - ty * tiles_x;
}

// The half-way mode runs the region in LDS (tb_region), the others in
// registers (tbr_region), as single lattices do.
template <class St, int EDGE, int MODE, int K>
__global__ __launch_bounds__(TB_THREADS) void k_batch_tb(const BatchLat<typename St::T>* __restrict__ lats,
                                                      const TileRef* __restrict__ regions, void* __restrict__ arena,
                                                      const uint32_t* __restrict__ nodes,
                                                      const LatState* __restrict__ state, int flip)
{
    __shared__ typename St::T L[9][TB_EDGE][TB_EDGE];
    StepArgs<typename St::T> a{};
    int ty, tx;
    batch_tb_region<St, EDGE, MODE, K>(lats, regions, arena, nodes, state, flip, a, ty, tx);
    tb_region<St, EDGE, MODE, K, CS_COLL_BGK>(a, ty, tx, L);
}

template <class St, int EDGE, int MODE, int K>
__global__ __launch_bounds__(TBR_THREADS, 4) void k_batch_tbr(const BatchLat<typename St::T>* __restrict__ lats,
                                                           const TileRef* __restrict__ regions,
                                                           void* __restrict__ arena,
                                                           const uint32_t* __restrict__ nodes,
                                                           const LatState* __restrict__ state, int flip)
{
    __shared__ typename St::T X[2][TBR_WAVES][2][3][TB_EDGE];
    StepArgs<typename St::T> a{};
    int ty, tx;
    batch_tb_region<St, EDGE, MODE, K>(lats, regions, arena, nodes, state, flip, a, ty, tx);
    tbr_region<St, EDGE, MODE, K, CS_COLL_BGK>(a, ty, tx, X);
}

// Grid-wide barrier of the persistent temporal kind: a monotonic arrival
// counter (barrier r waits for (r + 1) x blocks arrivals) with agent-scope
// release/acquire ordering, so every block's stores of a round are visible to
// every block (on any XCD) after it.  The wait is bounded: past 2^24 polls
// (some seconds) it raises `abort`, every block leaves at its next barrier
// and the host reports the timeout -- the grid always drains.  `next` hands out the
// work: draw i is region i % nregions of round i / nregions.  The host clears
// `next` and `count` before each launch and `abort` once per call.
struct GridBarrier {
    unsigned long long next;
    unsigned count;
    unsigned abort;
};
constexpr size_t GRID_BARRIER_PER_LAUNCH = offsetof(GridBarrier, abort);  // bytes cleared per launch
// Rounds per persistent launch (32768 steps): the arrival count stays far
// below 2^32 (4096 rounds x at most 1024 blocks); a longer call takes one
// launch per 4096 rounds, which costs nothing measurable.
constexpr int GRID_MAX_ROUNDS = 4096;

__device__ __forceinline__ bool grid_barrier(GridBarrier* b, unsigned target)
{
    __shared__ int s_ok;
    __threadfence();  // this thread's stores of the round, at agent scope
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(&b->count, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        int ok = 1;
        unsigned it = 0;
        while (__hip_atomic_load(&b->count, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(&b->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u || ++it > (1u << 24)) {
                __hip_atomic_store(&b->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        s_ok = ok;
    }
    __syncthreads();
    __atomic_thread_fence(__ATOMIC_ACQUIRE);  // every wave reads the other blocks' stores afresh
    return s_ok != 0;
}

// Next piece of work of a persistent grid: the block's draw from `next`
// (uniform over the block) as (round, region), after the block has passed
// every barrier below that round -- it has finished its regions of the
// earlier rounds, and barrier r lets nobody into round r + 1 before every
// block has drawn past round r, i.e. before all of round r is done.  A draw
// past the last round still passes the remaining barriers (every block
// arrives at each exactly once) and returns false, as does an aborted
// barrier.  `round` is the block's current round (0 at the start).
__device__ __forceinline__ bool grid_next(GridBarrier* bar, int nregions, int rounds, int& round, int& region)
{
    __shared__ unsigned long long s_draw;
    if (threadIdx.x == 0) s_draw = atomicAdd(&bar->next, 1ull);
    __syncthreads();
    const unsigned long long draw = s_draw;
    __syncthreads();  // every thread has the draw before thread 0 takes the next one
    const long long q = (long long)(draw / (unsigned long long)nregions);
    const int upto = (int)(q < rounds ? q : rounds - 1);
    for (; round < upto; ++round)
        if (!grid_barrier(bar, (unsigned)(round + 1) * gridDim.x)) return false;
    if (q >= rounds) return false;
    region = (int)(draw - (unsigned long long)q * (unsigned long long)nregions);
    return true;
}

// Rounds round0 .. round0 + rounds - 1 of the call (round r flips the
// buffers r times after the state upload), regions drawn one at a time.
template <class St, int EDGE, int MODE, int K>
__global__ __launch_bounds__(TB_THREADS) void k_batch_tb_grid(const BatchLat<typename St::T>* __restrict__ lats,
                                                           const TileRef* __restrict__ regions,
                                                           void* __restrict__ arena,
                                                           const uint32_t* __restrict__ nodes,
                                                           const LatState* __restrict__ state, int nregions,
                                                           int round0, int rounds, GridBarrier* bar)
{
    __shared__ typename St::T L[9][TB_EDGE][TB_EDGE];
    int round = 0, r = 0;
    while (grid_next(bar, nregions, rounds, round, r)) {
        StepArgs<typename St::T> a{};
        int ty, tx;
        batch_tb_region<St, EDGE, MODE, K>(lats, regions + r, arena, nodes, state, round0 + round, a, ty, tx, true);
        tb_region<St, EDGE, MODE, K, CS_COLL_BGK>(a, ty, tx, L);
        __syncthreads();  // the region's interior stores have read the LDS before the next load
    }
}

template <class St, int EDGE, int MODE, int K>
__global__ __launch_bounds__(TBR_THREADS, 4) void k_batch_tbr_grid(const BatchLat<typename St::T>* __restrict__ lats,
                                                                const TileRef* __restrict__ regions,
                                                                void* __restrict__ arena,
                                                                const uint32_t* __restrict__ nodes,
                                                                const LatState* __restrict__ state, int nregions,
                                                                int round0, int rounds, GridBarrier* bar)
{
    __shared__ typename St::T X[2][TBR_WAVES][2][3][TB_EDGE];
    int round = 0, r = 0;
    while (grid_next(bar, nregions, rounds, round, r)) {
        StepArgs<typename St::T> a{};
        int ty, tx;
        batch_tb_region<St, EDGE, MODE, K>(lats, regions + r, arena, nodes, state, round0 + round, a, ty, tx, true);
        tbr_region<St, EDGE, MODE, K, CS_COLL_BGK>(a, ty, tx, X);
        __syncthreads();  // every wave has read the exchange buffers of the region's last step
    }
}

// Steps per launch of the temporal launch kind.
constexpr int BATCH_TB_STEPS = 8;

// Mode group of a lattice: periodic, mode (fluid/ref/half), stochastic rounding.
inline int mode_group(int mode_bits)
{
    const int m = !(mode_bits & BM_FLAGS) ? MODE_FLUID : (mode_bits & BM_HALFWAY) ? MODE_HALF : MODE_REF;
    return ((mode_bits & BM_PERIODIC) ? 6 : 0) + m * 2 + ((mode_bits & BM_SR) ? 1 : 0);
}

}  // namespace cs

using namespace cs;

struct cs_batch {
    cs_ctx* ctx = nullptr;
    int n = 0, storage = CS_F32;
    size_t esize = 4;
    std::vector<int> w, h, pitch, cpitch;
    std::vector<int64_t> plane, off, node_off;
    std::vector<cs_params> params;
    std::vector<uint8_t> has_flags;
    std::vector<uint8_t> active;
    std::vector<int64_t> snap_off;    // offset of each lattice's (ux, uy) snapshot
    int64_t snap_elems = 0;
    float* snap = nullptr;
    double* d_res = nullptr;
    int64_t arena_elems = 0, node_elems = 0;
    void* arena = nullptr;
    uint32_t* nodes = nullptr;
    void* d_lats = nullptr;
    int* d_order = nullptr;
    int* d_queue = nullptr;
    TileRef* d_tiles = nullptr;
    int64_t max_tiles = 0;
    TileRef* d_regions = nullptr;     // region list of the temporal launch kinds
    GridBarrier* d_bar = nullptr;     // grid barrier of the persistent temporal kind
    int64_t max_regions = 0;
    LatState* d_state = nullptr;
    bool lats_dirty = true;
    std::vector<LatState> state;      // per lattice: current buffer, steps taken
    // Host sources of the asynchronous uploads.  A source must stay alive and
    // unchanged until the copy has run; every batch call synchronises its
    // stream before it returns, and each source is rewritten only by the next
    // call's upload.
    std::vector<int> host_order;
    std::vector<TileRef> host_tiles;
    std::vector<TileRef> host_regions;
    std::vector<unsigned char> host_lats;  // BatchLat<T> descriptors
    std::vector<LatState> host_state;     // `state` as of the launch
    cs_batch_tuning tuning{256, 2, 1, CS_ORDER_LARGEST_FIRST, CS_LAUNCH_PERSISTENT};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace cs {

template <class F>
static void dispatch_batch_storage(int storage, F&& f)
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

template <class T>
static Consts<T> batch_consts(const cs_batch* b, int i);
template <>
Consts<float> batch_consts<float>(const cs_batch* b, int i)
{
    return is_dev_storage(b->storage) ? make_consts_dev_f32(b->params[i]) : make_consts_f32(b->params[i]);
}
template <>
Consts<double> batch_consts<double>(const cs_batch* b, int i)
{
    return make_consts_f64(b->params[i]);
}

template <class T>
static BatchLat<T> describe(const cs_batch* b, int i)
{
    BatchLat<T> L{};
    L.w = b->w[i];
    L.h = b->h[i];
    L.pitch = b->pitch[i];
    L.cpitch = b->cpitch[i];
    const cs_params& p = b->params[i];
    L.mode = (p.edge == CS_EDGE_PERIODIC ? BM_PERIODIC : 0) | (p.boundary == CS_BB_HALFWAY ? BM_HALFWAY : 0) |
             (b->has_flags[i] ? BM_FLAGS : 0) | (p.rounding == CS_ROUND_STOCHASTIC ? BM_SR : 0);
    L.plane = b->plane[i];
    L.off = b->off[i];
    L.node_off = b->node_off[i];
    L.seed = p.seed;
    L.k = batch_consts<T>(b, i);
    return L;
}

static void check_batch(const cs_batch* b, int index)
{
    CS_REQUIRE(b != nullptr, CS_ERR_ARG, "null batch");
    CS_REQUIRE(index >= 0 && index < b->n, CS_ERR_ARG, "lattice index out of range");
}

static void upload_descriptors(cs_batch* b)
{
    if (!b->lats_dirty) return;
    dispatch_batch_storage(b->storage, [&]<class St>() {
        using T = typename St::T;
        b->host_lats.assign(sizeof(BatchLat<T>) * (size_t)b->n, 0);
        for (int i = 0; i < b->n; ++i) {
            const BatchLat<T> d = describe<T>(b, i);
            std::memcpy(b->host_lats.data() + sizeof d * (size_t)i, &d, sizeof d);
        }
        if (!b->d_lats) CS_HIP(hipMalloc(&b->d_lats, sizeof(BatchLat<T>) * (size_t)b->n));
        CS_HIP(hipMemcpyAsync(b->d_lats, b->host_lats.data(), sizeof(BatchLat<T>) * (size_t)b->n,
                              hipMemcpyHostToDevice, b->stream));
    });
    b->lats_dirty = false;
}

// Lattice order (by the tuning's queue order), grouped by boundary mode;
// returns (group, first, count) triples into d_order.
static std::vector<std::array<int, 3>> upload_order(cs_batch* b)
{
    std::vector<int> ord;
    for (int i = 0; i < b->n; ++i)
        if (b->active[(size_t)i]) ord.push_back(i);
    auto cells = [&](int i) { return (int64_t)b->w[(size_t)i] * b->h[(size_t)i]; };
    if (b->tuning.order == CS_ORDER_LARGEST_FIRST)
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int c) { return cells(a) > cells(c); });
    else if (b->tuning.order == CS_ORDER_SMALLEST_FIRST)
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int c) { return cells(a) < cells(c); });
    std::vector<int> grp((size_t)b->n);
    for (int i = 0; i < b->n; ++i) {
        const cs_params& p = b->params[(size_t)i];
        const int bits = (p.edge == CS_EDGE_PERIODIC ? BM_PERIODIC : 0) | (p.boundary == CS_BB_HALFWAY ? BM_HALFWAY : 0) |
                         (b->has_flags[(size_t)i] ? BM_FLAGS : 0) |
                         (p.rounding == CS_ROUND_STOCHASTIC && b->esize == 2 ? BM_SR : 0);
        grp[(size_t)i] = mode_group(bits);
    }
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int c) { return grp[(size_t)a] < grp[(size_t)c]; });
    std::vector<std::array<int, 3>> groups;
    for (int k = 0; k < (int)ord.size(); ++k) {
        const int g = grp[(size_t)ord[(size_t)k]];
        if (groups.empty() || groups.back()[0] != g) groups.push_back({g, k, 0});
        groups.back()[2] += 1;
    }
    b->host_order = ord;
    if (!b->host_order.empty())
        CS_HIP(hipMemcpyAsync(b->d_order, b->host_order.data(), sizeof(int) * b->host_order.size(),
                              hipMemcpyHostToDevice, b->stream));
    return groups;
}

template <class St, int CPT>
static void launch_group(cs_batch* b, int group, int first, int count, int steps, int* queue, int grid)
{
    using T = typename St::T;
    const auto* lats = (const BatchLat<T>*)b->d_lats;
    const int* order = b->d_order + first;
    const dim3 block((unsigned)b->tuning.block_threads);
    auto go = [&]<int EDGE, int MODE, bool SR>() {
        hipLaunchKernelGGL((k_batch<St, EDGE, MODE, SR, CPT>), dim3((unsigned)grid), block, 0, b->stream, lats, order,
                           count, queue, b->arena, b->nodes, b->d_state, steps);
    };
    auto go_sr = [&]<int EDGE, int MODE>() {
        if constexpr (is_16bit<St>()) {
            if (group & 1) {
                go.template operator()<EDGE, MODE, true>();
                return;
            }
        }
        go.template operator()<EDGE, MODE, false>();
    };
    auto go_mode = [&]<int EDGE>() {
        const int m = (group % 6) / 2;
        if (m == MODE_FLUID) go_sr.template operator()<EDGE, MODE_FLUID>();
        else if (m == MODE_HALF) go_sr.template operator()<EDGE, MODE_HALF>();
        else go_sr.template operator()<EDGE, MODE_REF>();
    };
    if (group >= 6) go_mode.template operator()<CS_EDGE_PERIODIC>();
    else go_mode.template operator()<CS_EDGE_ZERO>();
}

// The launch's per-lattice state goes up from its own copy: advance_state
// changes `state` right after the launch is enqueued.
static void upload_state(cs_batch* b)
{
    b->host_state = b->state;
    CS_HIP(hipMemcpyAsync(b->d_state, b->host_state.data(), sizeof(LatState) * (size_t)b->n, hipMemcpyHostToDevice,
                          b->stream));
}

static void advance_state(cs_batch* b, int steps, int flips)
{
    for (int i = 0; i < b->n; ++i) {
        if (!b->active[(size_t)i]) continue;
        b->state[(size_t)i].cur = (b->state[(size_t)i].cur + flips) & 1;
        b->state[(size_t)i].steps += (uint64_t)steps;
    }
}

// Rows per tile of the tiles launch: block_threads / 64 (4, 8 or 16).
static int tile_rows(const cs_batch* b)
{
    const int bt = b->tuning.block_threads;
    return bt >= 1024 ? 16 : (bt >= 512 ? 8 : 4);
}

// Tile list of the active lattices, grouped like upload_order; returns
// (group, first tile, tile count) triples into d_tiles.
static std::vector<std::array<int, 3>> upload_tiles(cs_batch* b, const std::vector<std::array<int, 3>>& groups)
{
    std::vector<TileRef>& tiles = b->host_tiles;
    tiles.clear();
    std::vector<std::array<int, 3>> out;
    for (const auto& g : groups) {
        const int first = (int)tiles.size();
        for (int k = g[1]; k < g[1] + g[2]; ++k) {
            const int li = b->host_order[(size_t)k];
            const int tr = tile_rows(b);
            const int nt = ((b->w[(size_t)li] + 63) / 64) * ((b->h[(size_t)li] + tr - 1) / tr);
            for (int t = 0; t < nt; ++t) tiles.push_back(TileRef{li, t});
        }
        out.push_back({g[0], first, (int)tiles.size() - first});
    }
    CS_REQUIRE((int64_t)tiles.size() <= b->max_tiles, CS_ERR_STATE, "tile list larger than allocated");
    if (!tiles.empty())
        CS_HIP(hipMemcpyAsync(b->d_tiles, tiles.data(), sizeof(TileRef) * tiles.size(), hipMemcpyHostToDevice,
                              b->stream));
    return out;
}

template <class St>
static void launch_tiles_group(cs_batch* b, int group, int first, int count, int flip, int s)
{
    using T = typename St::T;
    const auto* lats = (const BatchLat<T>*)b->d_lats;
    const TileRef* tiles = b->d_tiles + first;
    const int tr = tile_rows(b);
    auto go = [&]<int EDGE, int MODE, bool SR>() {
        if (tr == 16)
            hipLaunchKernelGGL((k_batch_tiles<St, EDGE, MODE, SR, 16>), dim3((unsigned)count), dim3(1024), 0, b->stream,
                               lats, tiles, b->arena, b->nodes, b->d_state, flip, s);
        else if (tr == 8)
            hipLaunchKernelGGL((k_batch_tiles<St, EDGE, MODE, SR, 8>), dim3((unsigned)count), dim3(512), 0, b->stream,
                               lats, tiles, b->arena, b->nodes, b->d_state, flip, s);
        else
            hipLaunchKernelGGL((k_batch_tiles<St, EDGE, MODE, SR, 4>), dim3((unsigned)count), dim3(256), 0, b->stream,
                               lats, tiles, b->arena, b->nodes, b->d_state, flip, s);
    };
    auto go_sr = [&]<int EDGE, int MODE>() {
        if constexpr (is_16bit<St>()) {
            if (group & 1) {
                go.template operator()<EDGE, MODE, true>();
                return;
            }
        }
        go.template operator()<EDGE, MODE, false>();
    };
    auto go_mode = [&]<int EDGE>() {
        const int m = (group % 6) / 2;
        if (m == MODE_FLUID) go_sr.template operator()<EDGE, MODE_FLUID>();
        else if (m == MODE_HALF) go_sr.template operator()<EDGE, MODE_HALF>();
        else go_sr.template operator()<EDGE, MODE_REF>();
    };
    if (group >= 6) go_mode.template operator()<CS_EDGE_PERIODIC>();
    else go_mode.template operator()<CS_EDGE_ZERO>();
}

// Region list (64 x 64 regions, interior 64 - 2 * BATCH_TB_STEPS) of the
// active lattices, grouped like upload_order, into d_regions; returns
// (group, first region, region count) triples.
static std::vector<std::array<int, 3>> upload_regions(cs_batch* b, const std::vector<std::array<int, 3>>& groups)
{
    constexpr int IN = TB_EDGE - 2 * BATCH_TB_STEPS;
    std::vector<TileRef>& regions = b->host_regions;
    regions.clear();
    std::vector<std::array<int, 3>> out;
    for (const auto& g : groups) {
        const int first = (int)regions.size();
        for (int k = g[1]; k < g[1] + g[2]; ++k) {
            const int li = b->host_order[(size_t)k];
            const int nr = ((b->w[(size_t)li] + IN - 1) / IN) * ((b->h[(size_t)li] + IN - 1) / IN);
            for (int t = 0; t < nr; ++t) regions.push_back(TileRef{li, t});
        }
        out.push_back({g[0], first, (int)regions.size() - first});
    }
    CS_REQUIRE((int64_t)regions.size() <= b->max_regions, CS_ERR_STATE, "region list larger than allocated");
    if (!regions.empty())
        CS_HIP(hipMemcpyAsync(b->d_regions, regions.data(), sizeof(TileRef) * regions.size(), hipMemcpyHostToDevice,
                              b->stream));
    return out;
}

template <class St>
static void launch_tb_group(cs_batch* b, int group, int first, int count, int flip)
{
    if constexpr (sizeof(typename St::S) != 4) {
        throw Error{CS_ERR_UNSUPPORTED, "the temporal launch kind needs 32-bit storage"};
    } else {
        using T = typename St::T;
        const auto* lats = (const BatchLat<T>*)b->d_lats;
        const TileRef* regions = b->d_regions + first;
        auto go = [&]<int EDGE, int MODE>() {
            if constexpr (MODE == MODE_HALF)
                hipLaunchKernelGGL((k_batch_tb<St, EDGE, MODE, BATCH_TB_STEPS>), dim3((unsigned)count),
                                   dim3(TB_THREADS), 0, b->stream, lats, regions, b->arena, b->nodes, b->d_state, flip);
            else
                hipLaunchKernelGGL((k_batch_tbr<St, EDGE, MODE, BATCH_TB_STEPS>), dim3((unsigned)count),
                                   dim3(TBR_THREADS), 0, b->stream, lats, regions, b->arena, b->nodes, b->d_state, flip);
        };
        auto go_mode = [&]<int EDGE>() {
            const int m = (group % 6) / 2;
            if (m == MODE_FLUID) go.template operator()<EDGE, MODE_FLUID>();
            else if (m == MODE_HALF) go.template operator()<EDGE, MODE_HALF>();
            else go.template operator()<EDGE, MODE_REF>();
        };
        if (group >= 6) go_mode.template operator()<CS_EDGE_PERIODIC>();
        else go_mode.template operator()<CS_EDGE_ZERO>();
    }
}

static bool batch_tb_storage(int storage) { return storage == CS_F32 || storage == CS_F32DEV; }

// All rounds of 8 steps of one mode group in one cooperative launch: the
// grid is as many blocks as can be resident at once (the launch fails rather
// than start a grid that could not meet at its barriers).
template <class St>
static void launch_tb_grid_group(cs_batch* b, int group, int first, int count, int round0, int rounds)
{
    if constexpr (sizeof(typename St::S) != 4) {
        throw Error{CS_ERR_UNSUPPORTED, "the temporal launch kinds need 32-bit storage"};
    } else {
        using T = typename St::T;
        const auto* lats = (const BatchLat<T>*)b->d_lats;
        const TileRef* regions = b->d_regions + first;
        void* arena = b->arena;
        const uint32_t* nodes = b->nodes;
        const LatState* state = b->d_state;
        GridBarrier* bar = b->d_bar;
        int nregions = count;
        auto go = [&]<int EDGE, int MODE>() {
            const void* fn = MODE == MODE_HALF ? (const void*)k_batch_tb_grid<St, EDGE, MODE, BATCH_TB_STEPS>
                                               : (const void*)k_batch_tbr_grid<St, EDGE, MODE, BATCH_TB_STEPS>;
            const int threads = MODE == MODE_HALF ? TB_THREADS : TBR_THREADS;
            int per_cu = 0;
            CS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0));
            CS_REQUIRE(per_cu >= 1, CS_ERR_UNSUPPORTED, "the persistent temporal kernel does not fit a CU");
            const int grid = std::min(count, per_cu * b->ctx->prop.multiProcessorCount);
            void* args[] = {(void*)&lats, (void*)&regions, (void*)&arena, (void*)&nodes, (void*)&state,
                            (void*)&nregions, (void*)&round0, (void*)&rounds, (void*)&bar};
            CS_HIP(hipMemsetAsync(bar, 0, GRID_BARRIER_PER_LAUNCH, b->stream));
            CS_HIP(hipLaunchCooperativeKernel(fn, dim3((unsigned)grid), dim3((unsigned)threads), args, 0, b->stream));
        };
        auto go_mode = [&]<int EDGE>() {
            const int m = (group % 6) / 2;
            if (m == MODE_FLUID) go.template operator()<EDGE, MODE_FLUID>();
            else if (m == MODE_HALF) go.template operator()<EDGE, MODE_HALF>();
            else go.template operator()<EDGE, MODE_REF>();
        };
        if (group >= 6) go_mode.template operator()<CS_EDGE_PERIODIC>();
        else go_mode.template operator()<CS_EDGE_ZERO>();
    }
}

static void launch_batch(cs_batch* b, int steps)
{
    upload_descriptors(b);
    upload_state(b);
    const auto groups = upload_order(b);
    CS_REQUIRE(groups.size() <= 16, CS_ERR_STATE, "too many boundary-mode groups");
    int flips = steps;
    const int kind = b->tuning.launch;
    if (kind == CS_LAUNCH_TILES || kind == CS_LAUNCH_TEMPORAL || kind == CS_LAUNCH_TEMPORAL_PERSISTENT) {
        int s = 0;
        flips = 0;
        if (kind != CS_LAUNCH_TILES && steps >= BATCH_TB_STEPS) {
            CS_REQUIRE(batch_tb_storage(b->storage), CS_ERR_UNSUPPORTED,
                       "the temporal launch kinds need FP32 or FP32-deviation storage");
            const auto rg = upload_regions(b, groups);
            const int rounds = steps / BATCH_TB_STEPS;
            if (kind == CS_LAUNCH_TEMPORAL_PERSISTENT && rg.size() == 1) {
                // every round of every lattice in one launch (one mode group;
                // GRID_MAX_ROUNDS rounds per launch for very long calls)
                const auto& g = rg[0];
                CS_HIP(hipMemsetAsync(b->d_bar, 0, sizeof(GridBarrier), b->stream));
                for (int r0 = 0; r0 < rounds && g[2] > 0; r0 += GRID_MAX_ROUNDS) {
                    const int nr = std::min(GRID_MAX_ROUNDS, rounds - r0);
                    dispatch_batch_storage(
                        b->storage, [&]<class St>() { launch_tb_grid_group<St>(b, g[0], g[1], g[2], r0, nr); });
                }
                s += rounds * BATCH_TB_STEPS;
                flips += rounds;
            } else {
                // one launch per BATCH_TB_STEPS steps and mode group over the region list
                for (; steps - s >= BATCH_TB_STEPS; s += BATCH_TB_STEPS, ++flips)
                    for (const auto& g : rg)
                        if (g[2] > 0)
                            dispatch_batch_storage(
                                b->storage, [&]<class St>() { launch_tb_group<St>(b, g[0], g[1], g[2], flips); });
            }
        }
        if (s < steps) {
            // one launch per step and mode group over the flattened tile list
            const auto tg = upload_tiles(b, groups);
            for (; s < steps; ++s, ++flips)
                for (const auto& g : tg)
                    if (g[2] > 0)
                        dispatch_batch_storage(
                            b->storage, [&]<class St>() { launch_tiles_group<St>(b, g[0], g[1], g[2], flips, s); });
        }
    } else {
        CS_HIP(hipMemsetAsync(b->d_queue, 0, sizeof(int) * 16, b->stream));
        for (size_t gi = 0; gi < groups.size(); ++gi) {
            const auto& g = groups[gi];
            const int grid = std::max(1, std::min(g[2], b->ctx->prop.multiProcessorCount * b->tuning.workgroups_per_cu));
            dispatch_batch_storage(b->storage, [&]<class St>() {
                if (b->tuning.cells_per_thread >= 2) launch_group<St, 2>(b, g[0], g[1], g[2], steps, b->d_queue + gi, grid);
                else launch_group<St, 1>(b, g[0], g[1], g[2], steps, b->d_queue + gi, grid);
            });
        }
    }
    CS_HIP(hipGetLastError());
    advance_state(b, steps, flips);
    if (kind == CS_LAUNCH_TEMPORAL_PERSISTENT) {
        GridBarrier hb{};
        CS_HIP(hipMemcpyAsync(&hb, b->d_bar, sizeof hb, hipMemcpyDeviceToHost, b->stream));
        CS_HIP(hipStreamSynchronize(b->stream));
        CS_REQUIRE(hb.abort == 0u, CS_ERR_STATE, "persistent temporal launch: grid barrier timed out");
    }
}

// One k_step launch per lattice per step (single-lattice code path).
static void launch_per_lattice(cs_batch* b, int steps)
{
    dispatch_batch_storage(b->storage, [&]<class St>() {
        using T = typename St::T;
        std::vector<BatchLat<T>> L((size_t)b->n);
        for (int i = 0; i < b->n; ++i) L[(size_t)i] = describe<T>(b, i);
        for (int s = 0; s < steps; ++s) {
            for (int i = 0; i < b->n; ++i) {
                if (!b->active[(size_t)i]) continue;
                const int cur = (b->state[(size_t)i].cur + s) & 1;
                const BatchLat<T>& d = L[(size_t)i];
                StepArgs<T> a{};
                typename St::S* base = (typename St::S*)b->arena + d.off;
                a.fin = base + (int64_t)cur * d.h * d.pitch;
                a.fout = base + (int64_t)(cur ^ 1) * d.h * d.pitch;
                a.node = b->nodes + d.node_off;
                a.w = d.w;
                a.h = d.h;
                a.pitch = d.pitch;
                a.cpitch = d.cpitch;
                a.plane = d.plane;
                a.tiles_x = (d.w + 63) / 64;
                a.sr_key = sr_key(d.seed, b->state[(size_t)i].steps + (uint64_t)s);
                a.cell_base = 0;
                a.k = d.k;
                const int nblocks = a.tiles_x * ((d.h + 3) / 4);
                const bool periodic = d.mode & BM_PERIODIC;
                const int mode = !(d.mode & BM_FLAGS) ? MODE_FLUID : (d.mode & BM_HALFWAY) ? MODE_HALF : MODE_REF;
                const bool sr = (d.mode & BM_SR) && is_16bit<St>();
                auto go = [&]<int EDGE, int MODE, bool SR>() {
                    hipLaunchKernelGGL((k_step<St, EDGE, MODE, 1, SR, uint32_t>), dim3((unsigned)nblocks), dim3(256), 0,
                                       b->stream, a);
                };
                auto go_sr = [&]<int EDGE, int MODE>() {
                    if constexpr (is_16bit<St>()) {
                        if (sr) {
                            go.template operator()<EDGE, MODE, true>();
                            return;
                        }
                    }
                    go.template operator()<EDGE, MODE, false>();
                };
                auto go_mode = [&]<int EDGE>() {
                    if (mode == MODE_FLUID) go_sr.template operator()<EDGE, MODE_FLUID>();
                    else if (mode == MODE_HALF) go_sr.template operator()<EDGE, MODE_HALF>();
                    else go_sr.template operator()<EDGE, MODE_REF>();
                };
                if (periodic) go_mode.template operator()<CS_EDGE_PERIODIC>();
                else go_mode.template operator()<CS_EDGE_ZERO>();
            }
        }
    });
    CS_HIP(hipGetLastError());
    advance_state(b, steps, steps);
}

static void* lattice_buf(cs_batch* b, int i)
{
    return (char*)b->arena + (size_t)b->off[(size_t)i] * b->esize +
           (size_t)b->state[(size_t)i].cur * b->h[(size_t)i] * b->pitch[(size_t)i] * b->esize;
}

// Host SoA (9 dense planes of h x w) <-> lattice i's row-interleaved planes,
// one strided 2D copy per plane on the batch stream.
static void copy_planes(cs_batch* b, size_t i, void* dst, const void* src, size_t elem, hipMemcpyKind kind)
{
    const size_t hplane = (size_t)b->w[i] * b->h[i] * elem, dplane = (size_t)b->plane[i] * elem;
    const bool up = kind == hipMemcpyHostToDevice;
    for (int q = 0; q < 9; ++q)
        CS_HIP(hipMemcpy2DAsync((char*)dst + q * (up ? dplane : hplane), up ? (size_t)b->pitch[i] * elem : b->w[i] * elem,
                                (const char*)src + q * (up ? hplane : dplane),
                                up ? b->w[i] * elem : (size_t)b->pitch[i] * elem, (size_t)b->w[i] * elem,
                                (size_t)b->h[i], kind, b->stream));
}

}  // namespace cs

extern "C" int cs_batch_create(cs_ctx* ctx, int count, const int32_t* w, const int32_t* h, int storage,
                               cs_batch** out)
{
    return guarded([&] {
        CS_REQUIRE(ctx && w && h && out && count >= 1, CS_ERR_ARG, "bad argument");
        CS_REQUIRE(storage_bytes(storage) != 0, CS_ERR_ARG, "bad storage");
        CS_HIP(hipSetDevice(ctx->device));
        auto* b = new cs_batch;
        b->ctx = ctx;
        b->n = count;
        b->storage = storage;
        b->esize = storage_bytes(storage);
        const int align = (int)(256 / b->esize);
        int64_t acc = 0, nacc = 0;
        for (int i = 0; i < count; ++i) {
            if (!(w[i] >= 1 && h[i] >= 1 && (int64_t)w[i] * h[i] <= ((int64_t)1 << 31))) {
                delete b;
                throw Error{CS_ERR_ARG, "bad lattice size"};
            }
            const int p = (w[i] + align - 1) / align * align;
            // the batch kernels address a lattice's buffer with 32-bit byte offsets
            if ((int64_t)9 * p * h[i] * (int64_t)b->esize >= ((int64_t)1 << 32)) {
                delete b;
                throw Error{CS_ERR_UNSUPPORTED, "lattice too large for a batch (one population buffer >= 4 GiB)"};
            }
            b->w.push_back(w[i]);
            b->h.push_back(h[i]);
            b->pitch.push_back(9 * p);  // row-interleaved populations (as the single lattice)
            b->cpitch.push_back(p);
            b->plane.push_back(p);
            b->off.push_back(acc);
            b->node_off.push_back(nacc);
            acc += 2 * 9 * (int64_t)p * h[i];
            nacc += ((int64_t)p * h[i] + 63) / 64 * 64;
        }
        b->arena_elems = acc;
        b->node_elems = nacc;
        for (int i = 0; i < count; ++i) {
            constexpr int IN = TB_EDGE - 2 * BATCH_TB_STEPS;
            b->max_tiles += (int64_t)((w[i] + 63) / 64) * ((h[i] + 3) / 4);
            b->max_regions += (int64_t)((w[i] + IN - 1) / IN) * ((h[i] + IN - 1) / IN);
        }
        b->state.assign((size_t)count, LatState{0, 0, 0});
        cs_params p{};
        p.tau = 1.0;
        p.delta_x = 1.0;
        p.delta_t = 1.0;
        p.outlet_rho = 1.0;
        p.rho0 = 1.0;
        b->params.assign((size_t)count, p);
        b->has_flags.assign((size_t)count, 0);
        b->active.assign((size_t)count, 1);
        for (int i = 0; i < count; ++i) {
            b->snap_off.push_back(b->snap_elems);
            b->snap_elems += 2 * (int64_t)w[i] * h[i];
        }
        try {
            CS_HIP(hipMalloc(&b->arena, (size_t)acc * b->esize));
            CS_HIP(hipMemset(b->arena, 0, (size_t)acc * b->esize));
            CS_HIP(hipMalloc(&b->nodes, sizeof(uint32_t) * (size_t)nacc));
            CS_HIP(hipMemset(b->nodes, 0, sizeof(uint32_t) * (size_t)nacc));
            CS_HIP(hipMalloc(&b->d_order, sizeof(int) * (size_t)count));
            CS_HIP(hipMalloc(&b->d_queue, sizeof(int) * 16));
            CS_HIP(hipMalloc(&b->d_tiles, sizeof(TileRef) * (size_t)b->max_tiles));
            CS_HIP(hipMalloc(&b->d_regions, sizeof(TileRef) * (size_t)b->max_regions));
            CS_HIP(hipMalloc(&b->d_bar, sizeof(GridBarrier)));
            CS_HIP(hipMemset(b->d_bar, 0, sizeof(GridBarrier)));
            CS_HIP(hipMalloc(&b->d_state, sizeof(LatState) * (size_t)count));
            CS_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
            CS_HIP(hipEventCreate(&b->ev0));
            CS_HIP(hipEventCreate(&b->ev1));
        } catch (...) {
            cs_batch_destroy(b);
            throw;
        }
        *out = b;
    });
}

extern "C" int cs_batch_destroy(cs_batch* b)
{
    return guarded([&] {
        if (!b) return;
        if (b->stream) (void)hipStreamSynchronize(b->stream);
        for (void* p : {b->arena, (void*)b->nodes, b->d_lats, (void*)b->d_order, (void*)b->d_queue, (void*)b->snap,
                        (void*)b->d_res, (void*)b->d_tiles, (void*)b->d_regions, (void*)b->d_bar, (void*)b->d_state})
            if (p) (void)hipFree(p);
        if (b->ev0) (void)hipEventDestroy(b->ev0);
        if (b->ev1) (void)hipEventDestroy(b->ev1);
        if (b->stream) (void)hipStreamDestroy(b->stream);
        delete b;
    });
}

extern "C" int cs_batch_count(const cs_batch* b, int* count)
{
    return guarded([&] {
        CS_REQUIRE(b && count, CS_ERR_ARG, "null argument");
        *count = b->n;
    });
}

extern "C" int cs_batch_set_params(cs_batch* b, int index, const cs_params* p)
{
    return guarded([&] {
        check_batch(b, index);
        CS_REQUIRE(p != nullptr, CS_ERR_ARG, "null params");
        CS_REQUIRE(p->edge == CS_EDGE_ZERO || p->edge == CS_EDGE_PERIODIC, CS_ERR_ARG, "bad edge mode");
        CS_REQUIRE(p->boundary == CS_BB_REFERENCE || p->boundary == CS_BB_HALFWAY, CS_ERR_ARG, "bad boundary mode");
        CS_REQUIRE(p->tau > 0 && p->delta_x > 0 && p->delta_t > 0 && p->outlet_rho > 0, CS_ERR_ARG,
                   "tau, delta_x, delta_t, outlet_rho must be > 0");
        CS_REQUIRE(p->rounding == CS_ROUND_NEAREST || p->rounding == CS_ROUND_STOCHASTIC, CS_ERR_ARG,
                   "bad rounding mode");
        CS_REQUIRE(p->collision == CS_COLL_BGK, CS_ERR_UNSUPPORTED, "batch lattices use the BGK operator");
        CS_REQUIRE(p->rho0 == b->params[(size_t)index].rho0 || !is_dev_storage(b->storage), CS_ERR_UNSUPPORTED,
                   "rho0 of a deviation-storage batch is fixed at creation (1.0)");
        const bool edge_changed = p->edge != b->params[(size_t)index].edge;
        CS_REQUIRE(!(edge_changed && b->has_flags[(size_t)index]), CS_ERR_STATE,
                   "set the edge mode before the flags of a batch lattice");
        b->params[(size_t)index] = *p;
        b->lats_dirty = true;
    });
}

extern "C" int cs_batch_set_flags(cs_batch* b, int index, const uint8_t* flags)
{
    return guarded([&] {
        check_batch(b, index);
        const size_t i = (size_t)index;
        if (!flags) {
            b->has_flags[i] = 0;
            b->lats_dirty = true;
            return;
        }
        const int64_t cells = (int64_t)b->w[i] * b->h[i];
        for (int64_t d = 0; d < cells; ++d) CS_REQUIRE(flags[d] <= CS_OUTLET, CS_ERR_ARG, "flag value out of range");
        uint8_t* tmp = nullptr;
        CS_HIP(hipMalloc(&tmp, (size_t)b->cpitch[i] * b->h[i]));
        CS_HIP(hipMemcpy2DAsync(tmp, (size_t)b->cpitch[i], flags, (size_t)b->w[i], (size_t)b->w[i], (size_t)b->h[i],
                                hipMemcpyHostToDevice, b->stream));
        dim3 g((unsigned)((b->w[i] + 255) / 256), (unsigned)b->h[i]);
        hipLaunchKernelGGL(k_node_info, g, dim3(256), 0, b->stream, tmp, b->nodes + b->node_off[i], b->w[i], b->h[i],
                           b->cpitch[i], (int)b->params[i].edge);
        hipError_t e = hipGetLastError();
        hipError_t e2 = hipStreamSynchronize(b->stream);
        (void)hipFree(tmp);
        CS_HIP(e);
        CS_HIP(e2);
        b->has_flags[i] = 1;
        b->lats_dirty = true;
    });
}

extern "C" int cs_batch_init_equilibrium(cs_batch* b, int index, const float* rho, const float* ux, const float* uy)
{
    return guarded([&] {
        check_batch(b, index);
        CS_REQUIRE(rho && ux && uy, CS_ERR_ARG, "null input");
        const size_t i = (size_t)index;
        const int64_t n = (int64_t)b->w[i] * b->h[i];
        dispatch_batch_storage(b->storage, [&]<class St>() {
            using T = typename St::T;
            std::vector<T> host((size_t)(3 * n));
            for (int64_t d = 0; d < n; ++d) {
                host[(size_t)d] = (T)rho[d];
                host[(size_t)(n + d)] = (T)ux[d];
                host[(size_t)(2 * n + d)] = (T)uy[d];
            }
            void* dev = nullptr;
            CS_HIP(hipMalloc(&dev, sizeof(T) * 3 * (size_t)n));
            CS_HIP(hipMemcpyAsync(dev, host.data(), sizeof(T) * 3 * (size_t)n, hipMemcpyHostToDevice, b->stream));
            const T* p = (const T*)dev;
            dim3 g((unsigned)((b->w[i] + 255) / 256), (unsigned)b->h[i]);
            hipLaunchKernelGGL((k_init_eq<St>), g, dim3(256), 0, b->stream, p, p + n, p + 2 * n,
                               (typename St::S*)lattice_buf(b, index), b->w[i], b->h[i], b->pitch[i], b->plane[i],
                               batch_consts<T>(b, index));
            hipError_t e = hipGetLastError();
            hipError_t e2 = hipStreamSynchronize(b->stream);
            (void)hipFree(dev);
            CS_HIP(e);
            CS_HIP(e2);
        });
    });
}

extern "C" int cs_batch_upload_f32(cs_batch* b, int index, const float* soa)
{
    return guarded([&] {
        check_batch(b, index);
        CS_REQUIRE(soa, CS_ERR_ARG, "null data");
        const size_t i = (size_t)index;
        float* tmp = nullptr;
        CS_HIP(hipMalloc(&tmp, sizeof(float) * (size_t)b->h[i] * b->pitch[i]));
        copy_planes(b, i, tmp, soa, 4, hipMemcpyHostToDevice);
        dispatch_batch_storage(b->storage, [&]<class St>() {
            dim3 g((unsigned)((b->w[i] + 255) / 256), (unsigned)b->h[i]);
            hipLaunchKernelGGL((k_from_f32<St>), g, dim3(256), 0, b->stream, (const float*)tmp,
                               (typename St::S*)lattice_buf(b, index), b->w[i], b->h[i], b->pitch[i], b->plane[i],
                               batch_consts<typename St::T>(b, index));
        });
        hipError_t e = hipGetLastError();
        hipError_t e2 = hipStreamSynchronize(b->stream);
        (void)hipFree(tmp);
        CS_HIP(e);
        CS_HIP(e2);
    });
}

extern "C" int cs_batch_download_f32(cs_batch* b, int index, float* soa)
{
    return guarded([&] {
        check_batch(b, index);
        CS_REQUIRE(soa, CS_ERR_ARG, "null data");
        const size_t i = (size_t)index;
        float* tmp = nullptr;
        CS_HIP(hipMalloc(&tmp, sizeof(float) * (size_t)b->h[i] * b->pitch[i]));
        dispatch_batch_storage(b->storage, [&]<class St>() {
            dim3 g((unsigned)((b->w[i] + 255) / 256), (unsigned)b->h[i]);
            hipLaunchKernelGGL((k_to_f32<St>), g, dim3(256), 0, b->stream,
                               (const typename St::S*)lattice_buf(b, index), tmp, b->w[i], b->h[i], b->pitch[i],
                               b->plane[i], batch_consts<typename St::T>(b, index));
        });
        hipError_t e = hipGetLastError();
        hipError_t e1 = hipSuccess;
        try {
            copy_planes(b, i, soa, tmp, 4, hipMemcpyDeviceToHost);
        } catch (const Error&) {
            e1 = hipErrorUnknown;
        }
        hipError_t e2 = hipStreamSynchronize(b->stream);
        (void)hipFree(tmp);
        CS_HIP(e);
        CS_HIP(e1);
        CS_HIP(e2);
    });
}

extern "C" int cs_batch_download_raw(cs_batch* b, int index, void* data)
{
    return guarded([&] {
        check_batch(b, index);
        CS_REQUIRE(data, CS_ERR_ARG, "null data");
        const size_t i = (size_t)index;
        copy_planes(b, i, data, lattice_buf(b, index), b->esize, hipMemcpyDeviceToHost);
        CS_HIP(hipStreamSynchronize(b->stream));
    });
}

extern "C" int cs_batch_velocity(cs_batch* b, int index, float* ux, float* uy)
{
    return guarded([&] {
        check_batch(b, index);
        CS_REQUIRE(ux && uy, CS_ERR_ARG, "null output");
        const size_t i = (size_t)index;
        const int64_t n = (int64_t)b->w[i] * b->h[i];
        double* dev = nullptr;
        CS_HIP(hipMalloc(&dev, sizeof(double) * 2 * (size_t)n));
        dispatch_batch_storage(b->storage, [&]<class St>() {
            dim3 g((unsigned)((b->w[i] + 255) / 256), (unsigned)b->h[i]);
            hipLaunchKernelGGL((k_moments<St>), g, dim3(256), 0, b->stream,
                               (const typename St::S*)lattice_buf(b, index), b->w[i], b->h[i], b->pitch[i],
                               b->plane[i], batch_consts<typename St::T>(b, index), (unsigned)(M_UX | M_UY), dev);
        });
        std::vector<double> host((size_t)(2 * n));
        hipError_t e = hipGetLastError();
        hipError_t e1 = hipMemcpyAsync(host.data(), dev, sizeof(double) * 2 * (size_t)n, hipMemcpyDeviceToHost,
                                       b->stream);
        hipError_t e2 = hipStreamSynchronize(b->stream);
        (void)hipFree(dev);
        CS_HIP(e);
        CS_HIP(e1);
        CS_HIP(e2);
        for (int64_t d = 0; d < n; ++d) {
            ux[d] = (float)host[(size_t)d];
            uy[d] = (float)host[(size_t)(n + d)];
        }
    });
}

extern "C" int cs_batch_set_tuning(cs_batch* b, const cs_batch_tuning* t)
{
    return guarded([&] {
        CS_REQUIRE(b && t, CS_ERR_ARG, "null argument");
        CS_REQUIRE(t->block_threads == 256 || t->block_threads == 512 || t->block_threads == 1024, CS_ERR_ARG,
                   "block_threads must be 256, 512 or 1024");
        CS_REQUIRE(t->workgroups_per_cu >= 1 && t->workgroups_per_cu <= 16, CS_ERR_ARG,
                   "workgroups_per_cu must be 1..16");
        CS_REQUIRE(t->cells_per_thread == 1 || t->cells_per_thread == 2, CS_ERR_ARG, "cells_per_thread must be 1 or 2");
        CS_REQUIRE(t->order >= CS_ORDER_LARGEST_FIRST && t->order <= CS_ORDER_SMALLEST_FIRST, CS_ERR_ARG, "bad order");
        CS_REQUIRE(t->launch >= CS_LAUNCH_PERSISTENT && t->launch <= CS_LAUNCH_TEMPORAL_PERSISTENT, CS_ERR_ARG,
                   "bad launch kind");
        CS_REQUIRE((t->launch != CS_LAUNCH_TEMPORAL && t->launch != CS_LAUNCH_TEMPORAL_PERSISTENT) ||
                       batch_tb_storage(b->storage),
                   CS_ERR_UNSUPPORTED, "the temporal launch kinds need FP32 or FP32-deviation storage");
        b->tuning = *t;
    });
}

extern "C" int cs_batch_get_tuning(const cs_batch* b, cs_batch_tuning* t)
{
    return guarded([&] {
        CS_REQUIRE(b && t, CS_ERR_ARG, "null argument");
        *t = b->tuning;
    });
}

static void timed(cs_batch* b, float* ms, const std::function<void()>& body)
{
    CS_HIP(hipEventRecord(b->ev0, b->stream));
    body();
    CS_HIP(hipEventRecord(b->ev1, b->stream));
    CS_HIP(hipEventSynchronize(b->ev1));
    float t = 0;
    CS_HIP(hipEventElapsedTime(&t, b->ev0, b->ev1));
    if (ms) *ms = t;
}

extern "C" int cs_batch_step(cs_batch* b, int steps, float* ms)
{
    return guarded([&] {
        CS_REQUIRE(b && steps >= 0, CS_ERR_ARG, "bad argument");
        if (steps == 0) return;
        timed(b, ms, [&] { launch_batch(b, steps); });
    });
}

extern "C" int cs_batch_step_per_lattice(cs_batch* b, int steps, float* ms)
{
    return guarded([&] {
        CS_REQUIRE(b && steps >= 0, CS_ERR_ARG, "bad argument");
        if (steps == 0) return;
        timed(b, ms, [&] { launch_per_lattice(b, steps); });
    });
}

extern "C" int cs_batch_autotune(cs_batch* b, int trial_steps, cs_batch_tuning* best, float* best_ms)
{
    return guarded([&] {
        CS_REQUIRE(b && trial_steps >= 1, CS_ERR_ARG, "bad argument");
        // Keep the state: trials run on the live arena and are undone afterwards.
        void* backup = nullptr;
        const size_t bytes = (size_t)b->arena_elems * b->esize;
        CS_HIP(hipMalloc(&backup, bytes));
        CS_HIP(hipMemcpyAsync(backup, b->arena, bytes, hipMemcpyDeviceToDevice, b->stream));
        const std::vector<LatState> state0 = b->state;
        const cs_batch_tuning keep = b->tuning;
        cs_batch_tuning win = keep;
        float win_ms = INFINITY;
        auto trial = [&](const cs_batch_tuning& t) {
            b->tuning = t;
            float ms = 0;
            timed(b, &ms, [&] { launch_batch(b, trial_steps); });
            if (ms < win_ms) {
                win_ms = ms;
                win = b->tuning;
            }
        };
        try {
            for (int blk : {256, 512, 1024})
                trial(cs_batch_tuning{blk, 1, 1, CS_ORDER_LARGEST_FIRST, CS_LAUNCH_TILES});
            if (batch_tb_storage(b->storage))
                for (int blk : {256, 512, 1024})  // block_threads: the remainder steps' tiles
                    for (int kind : {CS_LAUNCH_TEMPORAL, CS_LAUNCH_TEMPORAL_PERSISTENT})
                        trial(cs_batch_tuning{blk, 1, 1, CS_ORDER_LARGEST_FIRST, kind});
            for (int blk : {256, 512, 1024})
                for (int wpc : {1, 2, 4, 8})
                    for (int cpt : {1, 2})
                        for (int ord : {CS_ORDER_LARGEST_FIRST, CS_ORDER_GIVEN})
                            trial(cs_batch_tuning{blk, wpc, cpt, ord, CS_LAUNCH_PERSISTENT});
        } catch (...) {
            b->tuning = keep;
            b->state = state0;
            (void)hipMemcpyAsync(b->arena, backup, bytes, hipMemcpyDeviceToDevice, b->stream);
            (void)hipStreamSynchronize(b->stream);
            (void)hipFree(backup);
            throw;
        }
        hipError_t e1 = hipMemcpyAsync(b->arena, backup, bytes, hipMemcpyDeviceToDevice, b->stream);
        hipError_t e2 = hipStreamSynchronize(b->stream);
        (void)hipFree(backup);
        CS_HIP(e1);
        CS_HIP(e2);
        b->state = state0;
        b->tuning = win;
        if (best) *best = win;
        if (best_ms) *best_ms = win_ms;
    });
}

extern "C" int cs_batch_set_active(cs_batch* b, int index, int active)
{
    return guarded([&] {
        check_batch(b, index);
        b->active[(size_t)index] = active ? 1 : 0;
    });
}

extern "C" int cs_batch_snapshot_velocity(cs_batch* b)
{
    return guarded([&] {
        CS_REQUIRE(b, CS_ERR_ARG, "null batch");
        if (!b->snap) {
            CS_HIP(hipMalloc(&b->snap, sizeof(float) * (size_t)b->snap_elems));
            CS_HIP(hipMalloc(&b->d_res, sizeof(double) * (size_t)b->n + sizeof(int64_t) * (size_t)b->n));
            CS_HIP(hipMemcpy((char*)b->d_res + sizeof(double) * (size_t)b->n, b->snap_off.data(),
                             sizeof(int64_t) * (size_t)b->n, hipMemcpyHostToDevice));
        }
        upload_descriptors(b);
        upload_state(b);
        const int64_t* soff = (const int64_t*)((char*)b->d_res + sizeof(double) * (size_t)b->n);
        dispatch_batch_storage(b->storage, [&]<class St>() {
            hipLaunchKernelGGL((k_batch_snapshot<St>), dim3((unsigned)b->n), dim3(256), 0, b->stream,
                               (const BatchLat<typename St::T>*)b->d_lats, b->n, b->arena, b->d_state, b->snap, soff);
        });
        CS_HIP(hipGetLastError());
        CS_HIP(hipStreamSynchronize(b->stream));
    });
}

extern "C" int cs_batch_residuals(cs_batch* b, double* residuals)
{
    return guarded([&] {
        CS_REQUIRE(b && residuals, CS_ERR_ARG, "null argument");
        CS_REQUIRE(b->snap != nullptr, CS_ERR_STATE, "residuals requested before a velocity snapshot");
        upload_descriptors(b);
        upload_state(b);
        const int64_t* soff = (const int64_t*)((char*)b->d_res + sizeof(double) * (size_t)b->n);
        dispatch_batch_storage(b->storage, [&]<class St>() {
            hipLaunchKernelGGL((k_batch_residual<St>), dim3((unsigned)b->n), dim3(256), 0, b->stream,
                               (const BatchLat<typename St::T>*)b->d_lats, b->n, b->arena, b->d_state, b->nodes,
                               b->snap, soff, b->d_res);
        });
        CS_HIP(hipGetLastError());
        CS_HIP(hipMemcpyAsync(residuals, b->d_res, sizeof(double) * (size_t)b->n, hipMemcpyDeviceToHost, b->stream));
        CS_HIP(hipStreamSynchronize(b->stream));
    });
}

extern "C" int cs_batch_moments(cs_batch* b, int index, float* rho, float* ux, float* uy)
{
    return guarded([&] {
        check_batch(b, index);
        CS_REQUIRE(rho && ux && uy, CS_ERR_ARG, "null output");
        const size_t i = (size_t)index;
        const int64_t n = (int64_t)b->w[i] * b->h[i];
        double* dev = nullptr;
        CS_HIP(hipMalloc(&dev, sizeof(double) * 3 * (size_t)n));
        dispatch_batch_storage(b->storage, [&]<class St>() {
            dim3 g((unsigned)((b->w[i] + 255) / 256), (unsigned)b->h[i]);
            hipLaunchKernelGGL((k_moments<St>), g, dim3(256), 0, b->stream,
                               (const typename St::S*)lattice_buf(b, index), b->w[i], b->h[i], b->pitch[i],
                               b->plane[i], batch_consts<typename St::T>(b, index), (unsigned)(M_RHO | M_UX | M_UY),
                               dev);
        });
        std::vector<double> host((size_t)(3 * n));
        hipError_t e = hipGetLastError();
        hipError_t e1 = hipMemcpyAsync(host.data(), dev, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost,
                                       b->stream);
        hipError_t e2 = hipStreamSynchronize(b->stream);
        (void)hipFree(dev);
        CS_HIP(e);
        CS_HIP(e1);
        CS_HIP(e2);
        for (int64_t d = 0; d < n; ++d) {
            rho[d] = (float)host[(size_t)d];
            ux[d] = (float)host[(size_t)(n + d)];
            uy[d] = (float)host[(size_t)(2 * n + d)];
        }
    });
}
