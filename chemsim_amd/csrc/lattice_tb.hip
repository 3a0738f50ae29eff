// Launches of the temporally blocked step kernel (lbm_tb.hpp) for single
// lattices: which lattices take it, how many steps per launch, and the
// template dispatch over storage, edge, boundary mode and collision operator.
// Kept apart from lattice.hip so its instantiations compile in parallel.
//
// FP32 and FP32-deviation lattices take the temporally blocked kernels with
// K = 8 steps per launch (4096^2 FP32 Taylor-Green with the region in LDS,
// microseconds per step: one-step vector kernel 203; K = 2 197, K = 4 137,
// K = 6 125, K = 8 116 -- beyond 8 the halo's recomputation outgrows the
// saved traffic): k_step_tbr (the region in registers, 88 us per step)
// without geometry and in the reference boundary mode, k_step_tb (the region
// in LDS) in the half-way mode, whose link and Zou-He code does not fit the
// register variant's budget (it spills: 297 against 130 us per step on a
// 4096^2 cavity).  They are taken when the lattice has at least
// TB_MIN_REGIONS regions: a block updates one 64 x 64 region at one CU's
// VALU rate (about 3-4 us per step), so a lattice of fewer regions leaves CUs
// idle and the one-step kernels, spread over every CU and replayed from HIP
// graphs, are as fast or faster (512^2, 121 regions: 3.9 against 3.9 us per
// step; 128^2 cavity, 9 regions: 5.0 against 3.1).  KBC, whose one-step
// kernel is already VALU-bound (18 divisions per cell), keeps it: 4096^2, 342
// against 267 us per step.  cs_set_temporal_steps(4 or 8) forces the kernels
// on every FP32 lattice, (0) switches them off; a launch-shape override from
// cs_set_cells_per_thread selects the one-step kernels.
#include <atomic>
#include <type_traits>

#include "cs_internal.hpp"
#include "lbm_tb.hpp"

using namespace cs;

namespace {

std::atomic<int> g_tb_steps{CS_TEMPORAL_AUTO};
constexpr int64_t TB_MIN_REGIONS = 192;  // 2/3 of a region per CU of MI355X's 256

template <class St>
const Consts<float>& consts32(const cs_lattice* l)
{
    return l->k32;
}

template <class St, int EDGE, int MODE, int K, int COLL>
void launch_k(cs_lattice* l)
{
    constexpr int IN = TB_EDGE - 2 * K;
    StepArgs<float> a{};
    a.fin = l->buf[l->cur];
    a.fout = l->buf[l->cur ^ 1];
    a.node = MODE == MODE_FLUID ? nullptr : l->node;
    a.w = l->w;
    a.h = l->h;
    a.pitch = l->pitch;
    a.cpitch = l->cpitch;
    a.plane = l->plane;
    a.tiles_x = (l->w + IN - 1) / IN;
    a.k = consts32<St>(l);
    const int tiles_y = (l->h + IN - 1) / IN;
    if constexpr (MODE == MODE_HALF)
        hipLaunchKernelGGL((k_step_tb<St, EDGE, MODE, K, COLL>), dim3((unsigned)(a.tiles_x * tiles_y)),
                           dim3(TB_THREADS), 0, l->stream, a);
    else
        hipLaunchKernelGGL((k_step_tbr<St, EDGE, MODE, K, COLL>), dim3((unsigned)(a.tiles_x * tiles_y)),
                           dim3(TBR_THREADS), 0, l->stream, a);
    l->cur ^= 1;
    l->steps += K;
    for (int s = 0; s < K; ++s) l->time += l->params.delta_t;  // as step by step (lbm.rs:713)
}

template <class St, int EDGE, int MODE, int K>
void launch_coll(cs_lattice* l)
{
    if constexpr (std::is_same_v<St, StF32> && K == 8) {
        switch (l->params.collision) {
        case CS_COLL_TRT: launch_k<St, EDGE, MODE, K, CS_COLL_TRT>(l); return;
        case CS_COLL_REGULARIZED: launch_k<St, EDGE, MODE, K, CS_COLL_REGULARIZED>(l); return;
        case CS_COLL_KBC: launch_k<St, EDGE, MODE, K, CS_COLL_KBC>(l); return;
        default: break;
        }
    }
    launch_k<St, EDGE, MODE, K, CS_COLL_BGK>(l);
}

template <class St, int EDGE, int MODE>
void launch_steps(cs_lattice* l, int k)
{
    if (k == 4) launch_coll<St, EDGE, MODE, 4>(l);
    else launch_coll<St, EDGE, MODE, 8>(l);
}

template <class St, int EDGE>
void launch_mode(cs_lattice* l, int k)
{
    if (!l->has_flags) launch_steps<St, EDGE, MODE_FLUID>(l, k);
    else if (l->params.boundary == CS_BB_HALFWAY) launch_steps<St, EDGE, MODE_HALF>(l, k);
    else launch_steps<St, EDGE, MODE_REF>(l, k);
}

template <class St>
void launch_edge(cs_lattice* l, int k)
{
    if (l->params.edge == CS_EDGE_PERIODIC) launch_mode<St, CS_EDGE_PERIODIC>(l, k);
    else launch_mode<St, CS_EDGE_ZERO>(l, k);
}

}  // namespace

namespace cs {

int tb_steps(const cs_lattice* l, int cpt_override)
{
    const int k = g_tb_steps.load(std::memory_order_relaxed);
    if (k == 0 || cpt_override != 0) return 0;
    if (l->storage != CS_F32 && l->storage != CS_F32DEV) return 0;
    if ((int64_t)l->h * l->pitch * 4 >= ((int64_t)1 << 32)) return 0;  // 32-bit offsets
    if (k == CS_TEMPORAL_AUTO) {
        constexpr int IN = TB_EDGE - 2 * 8;
        const int64_t regions = (int64_t)((l->w + IN - 1) / IN) * ((l->h + IN - 1) / IN);
        if (regions < TB_MIN_REGIONS || l->params.collision == CS_COLL_KBC) return 0;
        return 8;
    }
    // K = 4 is kept for BGK; the other operators run 8 steps per launch
    if (k == 4 && l->params.collision != CS_COLL_BGK) return 8;
    return k;
}

void launch_tb(cs_lattice* l, int k)
{
    if (l->storage == CS_F32) launch_edge<StF32>(l, k);
    else launch_edge<StF32Dev>(l, k);
}

}  // namespace cs

extern "C" int cs_set_temporal_steps(int k)
{
    return guarded([&] {
        CS_REQUIRE(k == CS_TEMPORAL_AUTO || k == 0 || k == 4 || k == 8, CS_ERR_ARG,
                   "temporal steps must be CS_TEMPORAL_AUTO, 0, 4 or 8");
        g_tb_steps.store(k, std::memory_order_relaxed);
    });
}
