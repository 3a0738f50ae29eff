// Temporally blocked fused step: K steps of the D2Q9 update per launch.
//
// The one-step kernels read and write every population once per step, so
// they are bound by HBM at 72 bytes per cell update (FP32).  Here a block
// loads a 64 x 64 region of the lattice (its (64-2K) x (64-2K) interior and a
// halo of K cells on every side) into LDS once, runs K steps on it there and
// stores the interior: step s updates the cells at least s cells inside the
// region, which pull only from cells that step s-1 updated, so the interior
// after K steps is exact.  HBM moves (64^2 + (64-2K)^2) * 36 bytes per K
// steps of (64-2K)^2 cells -- 12.5 bytes per cell update at K = 8 -- and the
// update runs at LDS and VALU speed.
//
// Cells outside a zero-edged lattice hold the outside value (f = 0, or
// g = -w_i rho0 for deviation storage) and are never updated, which is the
// one-step kernels' rule for a population pulled from outside the edge; a
// periodic lattice is unrolled into the region by wrapping the loads (any
// number of times, so lattices smaller than a region work too).  The
// per-cell arithmetic is the one-step kernels' (same pull order, same
// collision functions), so K steps here equal K launches of k_step bit for
// bit.  32-bit storage.
//
// Geometry (MODE_REF, MODE_HALF): a cell's node word (its flag, which of its
// pull sources are solid, which of those are a moving lid) does not change
// between steps, so each lane keeps its cells' words in registers and the
// pull follows gather(): the reference's on-node swap at solid cells, or
// half-way links to the cell's own opposite population (+ the lid term) and
// Zou-He at inlet/outlet cells.  Solid cells of the half-way mode are carried
// unchanged, i.e. never written.
#pragma once

#include "lbm_device.hpp"

namespace cs {

constexpr int TB_EDGE = 64;        // region side (cells); lane l owns region column l
constexpr int TB_THREADS = 1024;   // 16 waves; wave w owns region rows w, w + 16, w + 32, w + 48
constexpr int TB_SLOTS = TB_EDGE * TB_EDGE / TB_THREADS;
// Small lattices: 32 x 32 regions, one cell per thread (a wave holds two
// region rows), so a lattice of a few dozen regions still spreads over as
// many CUs and each step costs a CU 1024 cells instead of 4096.
constexpr int TB_EDGE_SMALL = 32;

// Collision of one fluid cell: the operator finish() applies.
template <class St, int COLL>
__device__ __forceinline__ void collide_cell(const Consts<typename St::T>& k, typename St::T (&B)[9])
{
    if constexpr (COLL != CS_COLL_BGK) static_assert(!St::DEV, "TRT, Regularized and KBC run in the reference arithmetic");
    if constexpr (COLL == CS_COLL_TRT) trt_collide(k, B);
    else if constexpr (COLL == CS_COLL_REGULARIZED) reg_collide(k, B);
    else if constexpr (COLL == CS_COLL_KBC) kbc_collide(k, B);
    else if constexpr (St::DEV) bgk_collide_dev(k, B);
    else bgk_collide(k, B);
}

// Pulls of region cell (rr, lane) for the step, with gather()'s values: the
// normal source of each direction; on solid cells of the reference mode the
// opposite direction's normal pull (lbm.rs:745), i.e. the pulled pairs swap;
// in the half-way mode the cell's own opposite population (+ lid term) over a
// solid link, then Zou-He.  `special` (uniform over the wave) says whether any
// cell of the wave's row is solid (reference mode) or has a link or an
// inlet/outlet flag (half-way mode); rows without one -- most rows of a
// lattice -- take the plain pulls, nine LDS reads per cell.
template <class St, int MODE, int RE>
__device__ __forceinline__ void tb_gather(const typename St::T (&L)[9][RE][RE], int rr, int lane,
                                          uint32_t info, bool special, const Consts<typename St::T>& k,
                                          typename St::T (&B)[9])
{
    using T = typename St::T;
    const int type = MODE == MODE_FLUID ? (int)CS_FLUID : (int)(info >> NODE_TYPE_SHIFT);
    const bool solid = type == CS_WALL || type == CS_LID;
#pragma unroll
    for (int i = 0; i < 9; ++i) B[i] = L[i][rr + CX[i]][(lane - CY[i]) & (RE - 1)];
    if (!special) return;  // no cell of the wave's row needs more than the plain pulls
    if (MODE == MODE_REF) {
#pragma unroll
        for (int i = 1; i < 9; ++i) {
            if (i < OPP[i]) {
                const T a = B[i], b = B[OPP[i]];
                B[i] = solid ? b : a;
                B[OPP[i]] = solid ? a : b;
            }
        }
    } else if (MODE == MODE_HALF) {
        const bool link = !solid && (info & 0x1FEu) != 0;
        if (__any(link)) {  // uniform over the wave
#pragma unroll
            for (int i = 1; i < 9; ++i) {
                if (link && ((info >> i) & 1u)) {
                    T v = L[OPP[i]][rr][lane];
                    if ((info >> (9 + i)) & 1u) v = v + k.lid(i);
                    B[i] = v;
                }
            }
        }
        if constexpr (St::DEV) {
            if (type == CS_INLET) zou_he_inlet_dev<T>(B, k.rho0, k.in_ux, k.in_uy);
            else if (type == CS_OUTLET) zou_he_outlet_dev<T>(B, k.rho0, k.out_rho);
        } else {
            if (type == CS_INLET) zou_he_inlet<T>(B, k.in_ux, k.in_uy);
            else if (type == CS_OUTLET) zou_he_outlet<T>(B, k.out_rho);
        }
    }
}

// K steps on the region of tile (ty, tx) of the lattice `a` describes, a
// region of RE x RE cells (RE = 64 or 32): rows ty*(RE-2K) - K .., columns
// tx*(RE-2K) - K ..; blocks of TB_THREADS threads (RE = 64: one block per CU,
// the region takes 144 KiB of LDS; RE = 32: 36 KiB).  Thread t owns region
// column t % RE of rows (t / RE) + (1024 / RE) j.  A wave's LDS accesses are
// 64 consecutive words (one row, or two at RE = 32; a pull from column -1 or
// RE wraps within its row and belongs to a lane that discards its result), so
// they are free of bank conflicts, and each lane's nine pull addresses per row
// slot stay the same from step to step.  At RE = 64 a row is one wave, so the
// row tests below are uniform over the wave; at RE = 32 they mask lanes.
template <class St, int EDGE, int MODE, int K, int COLL, int RE>
__device__ __forceinline__ void tb_region(const StepArgs<typename St::T>& a, int ty, int tx,
                                          typename St::T (&L)[9][RE][RE])
{
    using S = typename St::S;
    using T = typename St::T;
    static_assert(sizeof(S) == 4 && sizeof(T) == 4, "32-bit storage and arithmetic");
    static_assert(RE == 64 || RE == 32, "region side");
    static_assert(K >= 1 && 2 * K < RE, "halo");
    constexpr int IN = RE - 2 * K;                   // interior side
    constexpr int SLOTS = RE * RE / TB_THREADS;      // rows per thread
    constexpr int RSTRIDE = TB_THREADS / RE;         // region rows between a thread's slots
    const int r0 = ty * IN - K, c0 = tx * IN - K;
    const int lane = (int)(threadIdx.x & (RE - 1)), row0 = (int)(threadIdx.x / RE);
    const S* __restrict__ fin = (const S*)a.fin;
    S* __restrict__ fout = (S*)a.fout;
    const uint32_t P = (uint32_t)a.plane;
    // this lane's lattice column, and whether it lies inside the lattice
    int gc = c0 + lane;
    bool cin = true;
    if (EDGE == CS_EDGE_PERIODIC) {
        gc %= a.w;
        gc += gc < 0 ? a.w : 0;
    } else {
        cin = gc >= 0 && gc < a.w;
        gc = cin ? gc : 0;
    }
    bool upd[SLOTS];       // the cell is inside the lattice and not carried
    bool special[SLOTS];   // the wave's row(s) need the geometry pulls (uniform over the wave)
    uint32_t info[SLOTS];  // node words (geometry modes)
#pragma unroll
    for (int j = 0; j < SLOTS; ++j) {
        const int rr = row0 + RSTRIDE * j;
        int gr = r0 + rr;
        bool rin = true;
        if (EDGE == CS_EDGE_PERIODIC) {
            gr %= a.h;
            gr += gr < 0 ? a.h : 0;
        } else {
            rin = gr >= 0 && gr < a.h;
            gr = rin ? gr : 0;
        }
        const bool inside = cin && rin;
        info[j] = 0;
        if (MODE != MODE_FLUID) info[j] = inside ? a.node[(uint32_t)gr * (uint32_t)a.cpitch + (uint32_t)gc] : 0u;
        const int type = (int)(info[j] >> NODE_TYPE_SHIFT);
        upd[j] = inside && !(MODE == MODE_HALF && (type == CS_WALL || type == CS_LID));
        const bool needs = MODE == MODE_REF ? (type == CS_WALL || type == CS_LID)
                                            : ((info[j] & 0x1FEu) != 0 || type == CS_INLET || type == CS_OUTLET);
        special[j] = MODE != MODE_FLUID && __any(needs);
        const uint32_t o = (uint32_t)gr * (uint32_t)a.pitch + (uint32_t)gc;
        T v[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) v[i] = St::wid(*at<uint32_t>(fin, (uint32_t)i * P + o));
#pragma unroll
        for (int i = 0; i < 9; ++i) L[i][rr][lane] = inside ? v[i] : (St::DEV ? -a.k.base(i) : T(0));
    }

#pragma unroll 1
    for (int s = 1; s <= K; ++s) {
        __syncthreads();  // step s - 1 fully in LDS
        T B[SLOTS][9];
#pragma unroll
        for (int j = 0; j < SLOTS; ++j) {
            const int rr = row0 + RSTRIDE * j;
            if (rr >= s && rr < RE - s) {  // uniform over the wave at RE = 64
                tb_gather<St, MODE, RE>(L, rr, lane, info[j], special[j], a.k, B[j]);
                collide_cell<St, COLL>(a.k, B[j]);
            }
        }
        __syncthreads();  // every read of step s done before any write
        const bool cols = lane >= s && lane < RE - s;
#pragma unroll
        for (int j = 0; j < SLOTS; ++j) {
            const int rr = row0 + RSTRIDE * j;
            if (rr >= s && rr < RE - s && cols && upd[j]) {
#pragma unroll
                for (int i = 0; i < 9; ++i) L[i][rr][lane] = B[j][i];
            }
        }
    }
    __syncthreads();

    // interior: rows ty*IN .. and columns tx*IN .. of the lattice (no wrap)
    const int oc = c0 + lane;
    const bool ocol = lane >= K && lane < RE - K && oc < a.w;
#pragma unroll
    for (int j = 0; j < SLOTS; ++j) {
        const int rr = row0 + RSTRIDE * j;
        const int orow = r0 + rr;
        if (ocol && rr >= K && rr < RE - K && orow < a.h) {
            const uint32_t o = (uint32_t)orow * (uint32_t)a.pitch + (uint32_t)oc;
#pragma unroll
            for (int i = 0; i < 9; ++i) St::st(at<uint32_t>(fout, (uint32_t)i * P + o), 0, L[i][rr][lane]);
        }
    }
}


// ---- register-resident variant ----------------------------------------------
// The same K steps on the same 64 x 64 region with the cells in registers: a
// block of TBR_WAVES waves, wave w owning region rows 8w .. 8w+7 (lane l:
// column l), so a lane holds 8 cells x 9 populations.  A pull within the
// band comes from the lane's own registers (the row above or below) or a
// neighbouring lane's (DPP wave shift by one lane); only the band's edge
// rows exchange through LDS -- the top row's populations 1, 5, 8 (pulled by
// the band above) and the bottom row's 3, 6, 7 (pulled by the band below),
// published before each step into one of two buffers by step parity, so one
// barrier per step orders them.  Rows are updated in increasing order with
// the old populations 3, 6, 7 of the row above kept aside.  A cell takes its
// new values only inside the updated square and inside the lattice (and not
// when it is a carried solid cell), as in the LDS variant.
constexpr int TBR_WAVES = 8;
constexpr int TBR_ROWS = TB_EDGE / TBR_WAVES;  // rows per wave
constexpr int TBR_THREADS = 64 * TBR_WAVES;

template <class T>
__device__ __forceinline__ T tb_prev(T v)  // lane l gets lane l-1's value
{
    return __builtin_bit_cast(T, dpp_prev(__builtin_bit_cast(uint32_t, v)));
}
template <class T>
__device__ __forceinline__ T tb_next(T v)  // lane l gets lane l+1's value
{
    return __builtin_bit_cast(T, dpp_next(__builtin_bit_cast(uint32_t, v)));
}

template <class St, int EDGE, int MODE, int K, int COLL>
__device__ __forceinline__ void tbr_region(const StepArgs<typename St::T>& a, int ty, int tx,
                                           typename St::T (&X)[2][TBR_WAVES][2][3][TB_EDGE])
{
    using S = typename St::S;
    using T = typename St::T;
    static_assert(sizeof(S) == 4 && sizeof(T) == 4, "32-bit storage and arithmetic");
    static_assert(K >= 1 && 2 * K < TB_EDGE, "halo");
    constexpr int IN = TB_EDGE - 2 * K;
    const int r0 = ty * IN - K, c0 = tx * IN - K;
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const S* __restrict__ fin = (const S*)a.fin;
    S* __restrict__ fout = (S*)a.fout;
    const uint32_t P = (uint32_t)a.plane;

    int gc = c0 + lane;
    bool cin = true;
    if (EDGE == CS_EDGE_PERIODIC) {
        gc %= a.w;
        gc += gc < 0 ? a.w : 0;
    } else {
        cin = gc >= 0 && gc < a.w;
        gc = cin ? gc : 0;
    }
    T F[TBR_ROWS][9];
    bool upd[TBR_ROWS];
    uint32_t info[TBR_ROWS];
#pragma unroll
    for (int j = 0; j < TBR_ROWS; ++j) {
        const int rr = wave * TBR_ROWS + j;
        int gr = r0 + rr;
        bool rin = true;
        if (EDGE == CS_EDGE_PERIODIC) {
            gr %= a.h;
            gr += gr < 0 ? a.h : 0;
        } else {
            rin = gr >= 0 && gr < a.h;
            gr = rin ? gr : 0;
        }
        const bool inside = cin && rin;
        info[j] = 0;
        if (MODE != MODE_FLUID) info[j] = inside ? a.node[(uint32_t)gr * (uint32_t)a.cpitch + (uint32_t)gc] : 0u;
        const int type = (int)(info[j] >> NODE_TYPE_SHIFT);
        upd[j] = inside && !(MODE == MODE_HALF && (type == CS_WALL || type == CS_LID));
        const uint32_t o = (uint32_t)gr * (uint32_t)a.pitch + (uint32_t)gc;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const T v = St::wid(*at<uint32_t>(fin, (uint32_t)i * P + o));
            F[j][i] = inside ? v : (St::DEV ? -a.k.base(i) : T(0));
        }
    }
    const int up = wave > 0 ? wave - 1 : 0;                  // band above (wave 0: any; its top row never updates)
    const int dn = wave < TBR_WAVES - 1 ? wave + 1 : wave;   // band below
    const int lm = (lane - 1) & 63, lp = (lane + 1) & 63;

#pragma unroll 1
    for (int s = 1; s <= K; ++s) {
        const int pb = s & 1;
        X[pb][wave][0][0][lane] = F[0][1];
        X[pb][wave][0][1][lane] = F[0][5];
        X[pb][wave][0][2][lane] = F[0][8];
        X[pb][wave][1][0][lane] = F[TBR_ROWS - 1][3];
        X[pb][wave][1][1][lane] = F[TBR_ROWS - 1][6];
        X[pb][wave][1][2][lane] = F[TBR_ROWS - 1][7];
        __syncthreads();
        // old populations 3, 6 (from column c - 1) and 7 (from column c + 1) of the row above
        T p3 = X[pb][up][1][0][lane], p6 = X[pb][up][1][1][lm], p7 = X[pb][up][1][2][lp];
        const bool cols = lane >= s && lane < TB_EDGE - s;
#pragma unroll
        for (int j = 0; j < TBR_ROWS; ++j) {
            const int rr = wave * TBR_ROWS + j;
            T B[9];
            B[0] = F[j][0];
            B[2] = tb_prev(F[j][2]);
            B[4] = tb_next(F[j][4]);
            B[3] = p3;
            B[6] = p6;
            B[7] = p7;
            if (j + 1 < TBR_ROWS) {
                B[1] = F[j + 1][1];
                B[5] = tb_prev(F[j + 1][5]);
                B[8] = tb_next(F[j + 1][8]);
            } else {
                B[1] = X[pb][dn][0][0][lane];
                B[5] = X[pb][dn][0][1][lm];
                B[8] = X[pb][dn][0][2][lp];
            }
            // the next row's pulls from this one, before it changes
            p3 = F[j][3];
            p6 = tb_prev(F[j][6]);
            p7 = tb_next(F[j][7]);
            const int type = MODE == MODE_FLUID ? (int)CS_FLUID : (int)(info[j] >> NODE_TYPE_SHIFT);
            const bool solid = type == CS_WALL || type == CS_LID;
            if (MODE == MODE_REF) {
#pragma unroll
                for (int i = 1; i < 9; ++i) {
                    if (i < OPP[i]) {
                        const T x = B[i], y = B[OPP[i]];
                        B[i] = solid ? y : x;
                        B[OPP[i]] = solid ? x : y;
                    }
                }
            } else if (MODE == MODE_HALF) {
                const bool link = !solid && (info[j] & 0x1FEu) != 0;
                if (__any(link)) {  // uniform over the wave
#pragma unroll
                    for (int i = 1; i < 9; ++i) {
                        if (link && ((info[j] >> i) & 1u)) {
                            T v = F[j][OPP[i]];
                            if ((info[j] >> (9 + i)) & 1u) v = v + a.k.lid(i);
                            B[i] = v;
                        }
                    }
                }
                if constexpr (St::DEV) {
                    if (type == CS_INLET) zou_he_inlet_dev<T>(B, a.k.rho0, a.k.in_ux, a.k.in_uy);
                    else if (type == CS_OUTLET) zou_he_outlet_dev<T>(B, a.k.rho0, a.k.out_rho);
                } else {
                    if (type == CS_INLET) zou_he_inlet<T>(B, a.k.in_ux, a.k.in_uy);
                    else if (type == CS_OUTLET) zou_he_outlet<T>(B, a.k.out_rho);
                }
            }
            collide_cell<St, COLL>(a.k, B);
            const bool take = cols && rr >= s && rr < TB_EDGE - s && upd[j];
#pragma unroll
            for (int i = 0; i < 9; ++i) F[j][i] = take ? B[i] : F[j][i];
        }
    }

    const int oc = c0 + lane;
    const bool ocol = lane >= K && lane < TB_EDGE - K && oc < a.w;
#pragma unroll
    for (int j = 0; j < TBR_ROWS; ++j) {
        const int rr = wave * TBR_ROWS + j;
        const int orow = r0 + rr;
        if (ocol && rr >= K && rr < TB_EDGE - K && orow < a.h) {
            const uint32_t o = (uint32_t)orow * (uint32_t)a.pitch + (uint32_t)oc;
#pragma unroll
            for (int i = 0; i < 9; ++i) St::st(at<uint32_t>(fout, (uint32_t)i * P + o), 0, F[j][i]);
        }
    }
}

template <class St, int EDGE, int MODE, int K, int COLL>
__global__ __launch_bounds__(TBR_THREADS, 4) void k_step_tbr(StepArgs<typename St::T> a)
{
    __shared__ typename St::T X[2][TBR_WAVES][2][3][TB_EDGE];
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int ty = tile / a.tiles_x;
    tbr_region<St, EDGE, MODE, K, COLL>(a, ty, tile - ty * a.tiles_x, X);
}

// One lattice: tile t of the lattice's ceil(w / (RE-2K)) x ceil(h / (RE-2K))
// regions of RE x RE cells (a.tiles_x = ceil(w / (RE-2K))).
template <class St, int EDGE, int MODE, int K, int COLL, int RE>
__global__ __launch_bounds__(TB_THREADS) void k_step_tb(StepArgs<typename St::T> a)
{
    __shared__ typename St::T L[9][RE][RE];
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int ty = tile / a.tiles_x;
    tb_region<St, EDGE, MODE, K, COLL>(a, ty, tile - ty * a.tiles_x, L);
}

}  // namespace cs
