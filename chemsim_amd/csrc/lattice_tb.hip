// Launches of the temporally blocked step kernel (lbm_tb.hpp) for single
// lattices: which lattices take it, how many steps per launch, and the
// template dispatch over storage, edge, boundary mode and collision operator.
// Kept apart from lattice.hip so its instantiations compile in parallel.
//
// FP32 and FP32-deviation lattices take the temporally blocked kernels with
// K = 8 steps per launch (4096^2 FP32 Taylor-Green with the region in LDS,
// microseconds per step: one-step vector kernel 203; K = 2 197, K = 4 137,
// K = 6 125, K = 8 116 -- beyond 8 the halo's recomputation outgrows the
// saved traffic; with the region in registers K = 6 measured slower than 8
// too).  Lattices of more than TB_SMALL_MAX_REGIONS 32 x 32 regions (about
// 512^2) take 64 x 64 regions: k_step_tbr (the region in registers, 73 us
// per step at 4096^2) without geometry and in the reference boundary mode,
// k_step_tb (the region in LDS) in the half-way mode, whose link and Zou-He
// code does not fit the register variant's budget (it spills: 297 against
// 130 us per step on a 4096^2 cavity).  Smaller lattices take k_step_tb on
// 32 x 32 regions (16 x 16 interior): a block updates its region at one CU's
// rate, so 64 x 64 regions would leave most CUs idle on them (128^2 cavity,
// 9 regions: 5.0 us per step), while 32 x 32 regions cost a CU a quarter of
// the cells per step and spread over four times as many CUs.  Measured, us
// per step (cavity with half-way walls / periodic Taylor-Green): 128^2 1.64 /
// 1.15 against 3.02 / 2.42 for the one-step kernels replayed from HIP
// graphs; 512^2 4.00 / 2.81 against 4.68 / 3.48 with 64 x 64 regions and
// 4.91 / 3.81 one-step; 576^2 (1296 small regions) 4.76 / 3.93 against
// 4.85 / 3.56 with 64 x 64 regions, where those take over.  KBC, whose
// one-step kernel is already VALU-bound (18 divisions per cell), keeps it:
// 4096^2, 314 against 299 us per step.  cs_set_temporal_steps(4 or 8)
// forces the 64 x 64 kernels on every FP32 lattice, (0) switches temporal
// blocking off; a launch-shape override from cs_set_cells_per_thread selects
// the one-step kernels.
#include <atomic>
#include <type_traits>

#include "cs_internal.hpp"
#include "lbm_tb.hpp"

using namespace cs;

namespace {

std::atomic<int> g_tb_steps{CS_TEMPORAL_AUTO};
constexpr int64_t TB_SMALL_MAX_REGIONS = 1024;  // 32 x 32 regions: four per CU of MI355X's 256

template <class St>
const Consts<float>& consts32(const cs_lattice* l)
{
    return l->k32;
}

template <class St, int EDGE, int MODE, int K, int COLL, int RE>
void launch_k(cs_lattice* l)
{
    constexpr int IN = RE - 2 * K;
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
    if constexpr (MODE == MODE_HALF || RE == TB_EDGE_SMALL)
        hipLaunchKernelGGL((k_step_tb<St, EDGE, MODE, K, COLL, RE>), dim3((unsigned)(a.tiles_x * tiles_y)),
                           dim3(TB_THREADS), 0, l->stream, a);
    else
        hipLaunchKernelGGL((k_step_tbr<St, EDGE, MODE, K, COLL>), dim3((unsigned)(a.tiles_x * tiles_y)),
                           dim3(TBR_THREADS), 0, l->stream, a);
    l->cur ^= 1;
    l->steps += K;
    for (int s = 0; s < K; ++s) l->time += l->params.delta_t;  // as step by step (lbm.rs:713)
}

template <class St, int EDGE, int MODE, int K, int RE>
void launch_coll(cs_lattice* l)
{
    if constexpr (std::is_same_v<St, StF32> && K == 8) {
        switch (l->params.collision) {
        case CS_COLL_TRT: launch_k<St, EDGE, MODE, K, CS_COLL_TRT, RE>(l); return;
        case CS_COLL_REGULARIZED: launch_k<St, EDGE, MODE, K, CS_COLL_REGULARIZED, RE>(l); return;
        case CS_COLL_KBC:
            if constexpr (RE == TB_EDGE) {
                launch_k<St, EDGE, MODE, K, CS_COLL_KBC, RE>(l);
                return;
            }
            break;
        default: break;
        }
    }
    launch_k<St, EDGE, MODE, K, CS_COLL_BGK, RE>(l);
}

template <class St, int EDGE, int MODE>
void launch_steps(cs_lattice* l, int plan)
{
    const bool four = tb_plan_steps(plan) == 4;  // BGK only (launch_coll takes the operator at K = 8)
    if ((plan >> 8) == TB_EDGE_SMALL) {
        if (four) launch_coll<St, EDGE, MODE, 4, TB_EDGE_SMALL>(l);
        else launch_coll<St, EDGE, MODE, 8, TB_EDGE_SMALL>(l);
    } else if (four) {
        launch_coll<St, EDGE, MODE, 4, TB_EDGE>(l);
    } else {
        launch_coll<St, EDGE, MODE, 8, TB_EDGE>(l);
    }
}

template <class St, int EDGE>
void launch_mode(cs_lattice* l, int plan)
{
    if (!l->has_flags) launch_steps<St, EDGE, MODE_FLUID>(l, plan);
    else if (l->params.boundary == CS_BB_HALFWAY) launch_steps<St, EDGE, MODE_HALF>(l, plan);
    else launch_steps<St, EDGE, MODE_REF>(l, plan);
}

template <class St>
void launch_edge(cs_lattice* l, int plan)
{
    if (l->params.edge == CS_EDGE_PERIODIC) launch_mode<St, CS_EDGE_PERIODIC>(l, plan);
    else launch_mode<St, CS_EDGE_ZERO>(l, plan);
}

constexpr int tb_plan(int k, int edge) { return k | (edge << 8); }

}  // namespace

namespace cs {

int tb_steps(const cs_lattice* l, int cpt_override)
{
    const int k = g_tb_steps.load(std::memory_order_relaxed);
    if (k == 0 || cpt_override != 0) return 0;
    if (l->storage != CS_F32 && l->storage != CS_F32DEV) return 0;
    if ((int64_t)l->h * l->pitch * 4 >= ((int64_t)1 << 32)) return 0;  // 32-bit offsets
    if (k == CS_TEMPORAL_AUTO) {
        if (l->params.collision == CS_COLL_KBC) return 0;
        constexpr int INS = TB_EDGE_SMALL - 2 * 8;
        const int64_t small = (int64_t)((l->w + INS - 1) / INS) * ((l->h + INS - 1) / INS);
        return tb_plan(8, small <= TB_SMALL_MAX_REGIONS ? TB_EDGE_SMALL : TB_EDGE);
    }
    // K = 4 is kept for BGK; the other operators run 8 steps per launch
    if (k == 4 && l->params.collision != CS_COLL_BGK) return tb_plan(8, TB_EDGE);
    return tb_plan(k, TB_EDGE);
}

int tb_remainder_plan(const cs_lattice* l, int plan)
{
    if (tb_plan_steps(plan) != 8 || l->params.collision != CS_COLL_BGK) return 0;
    return (plan & ~0xFF) | 4;
}

void launch_tb(cs_lattice* l, int plan)
{
    if (l->storage == CS_F32) launch_edge<StF32>(l, plan);
    else launch_edge<StF32Dev>(l, plan);
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
