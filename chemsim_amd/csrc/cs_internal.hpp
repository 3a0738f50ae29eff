// Internal definitions shared by the translation units of libchemsim_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/chemsim.h"

namespace cs {

// ---------------------------------------------------------------- errors --
void set_error(const std::string& msg);
void clear_error();

struct Error {
    int code;
    std::string msg;
};

#define CS_HIP(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw ::cs::Error{CS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
    } while (0)

#define CS_REQUIRE(cond, code, msg)                        \
    do {                                                   \
        if (!(cond)) throw ::cs::Error{(code), (msg)};     \
    } while (0)

// Runs f, converting exceptions into a status code + thread-local message.
template <class F>
int guarded(F&& f)
{
    try {
        clear_error();
        f();
        return CS_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return CS_ERR_NOMEM;
    } catch (...) {
        set_error("unknown error");
        return CS_ERR_ARG;
    }
}

// ------------------------------------------------------------- constants --
// D2Q9 directions of lbm.rs:221-231 and the opposite map of lbm.rs:298-309.
constexpr int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
constexpr int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};

// Per-run constants in the working precision T (float or double):
// K = -dt/tau (lbm.rs:357); A, B, C of compute_equilibrium (lbm.rs:54-66);
// weights (lbm.rs:209-219); Ladd lid terms 6 w_i (c_i . u_lid); deviation
// storage bases w_i * rho0 and rho0; Zou-He boundary values.  Only distinct
// values are stored (three weight classes; the lid terms of directions
// 3, 4, 7, 8 are the negatives of 1, 2, 5, 6), so a kernel keeps them all in
// scalar registers.
template <class T>
struct Consts {
    T K, A, B, C, cs2;
    T wp, wm, hdt;  // TRT: 1/tau_plus, 1/tau_minus, -dt * 0.5 (lbm.rs:431-440)
    T dx, beta;     // KBC: delta_x, 1 / (2 nu / cs^2 + 1) (lbm.rs:478, 547-550)
    T sfq[3];       // Regularized: w / (2 cs^4) of the rest, axis, diagonal classes (lbm.rs:651)
    T wq[3];     // w of the rest, axis and diagonal directions
    T baseq[3];  // w * rho0 of the same classes
    T lidq[4];   // lid term of directions 1, 2, 5, 6
    T in_ux, in_uy, out_rho;
    T rho0;

    __host__ __device__ static constexpr int wclass(int i) { return i == 0 ? 0 : (i < 5 ? 1 : 2); }
    __host__ __device__ T w(int i) const { return wq[wclass(i)]; }
    __host__ __device__ T base(int i) const { return baseq[wclass(i)]; }
    __host__ __device__ T lid(int i) const
    {
        switch (i) {
        case 1: return lidq[0];
        case 2: return lidq[1];
        case 3: return -lidq[0];
        case 4: return -lidq[1];
        case 5: return lidq[2];
        case 6: return lidq[3];
        case 7: return -lidq[2];
        case 8: return -lidq[3];
        default: return T(0);
        }
    }
};

Consts<float> make_consts_f32(const cs_params& p);
Consts<double> make_consts_f64(const cs_params& p);
// Constants of the deviation storages (A, B, C rounded once from f64).
Consts<float> make_consts_dev_f32(const cs_params& p);
bool is_dev_storage(int storage);

// Node info word (one per cell, built on the device from the flags):
//   bits 0..8   source cell of direction i is solid (wall or lid)
//   bits 9..17  source cell of direction i is a moving lid
//   bits 24..27 the cell's own flag
constexpr uint32_t NODE_TYPE_SHIFT = 24;

inline bool is_solid_flag(int t) { return t == CS_WALL || t == CS_LID; }

size_t storage_bytes(int storage);

}  // namespace cs

struct cs_lattice;
namespace cs {
// n fused steps on the lattice's stream (cs_lattice_step without the checks).
void do_steps(cs_lattice* l, int n);
// Temporally blocked steps (lattice_tb.hip): the plan this lattice takes --
// steps per launch K in the low byte (tb_plan_steps), the region side above
// it (0: one-step kernels; cpt_override: the cs_set_cells_per_thread value)
// -- and one launch of that plan (advances cur, steps and time by K).
int tb_steps(const cs_lattice* l, int cpt_override);
inline int tb_plan_steps(int plan) { return plan & 0xFF; }
// The plan of 4 steps per launch on the same regions, for the remainder of
// a run of 8-step launches (BGK lattices; 0 when there is none).
int tb_remainder_plan(const cs_lattice* l, int plan);
void launch_tb(cs_lattice* l, int plan);
// The render passes of the current populations into d_out (device, h*w pixels
// of `channels` bytes: 3 = RGB, 4 = RGBA with alpha 255), enqueued on the
// lattice's stream; allocates the lattice's render scratch on first use.
void render_enqueue(cs_lattice* l, int mode, int overlay, uint8_t* d_out, int channels);
}  // namespace cs

// ----------------------------------------------------------------- handles --
struct cs_ctx {
    int device = 0;
    hipDeviceProp_t prop{};
};

struct cs_lattice {
    cs_ctx* ctx = nullptr;
    int w = 0, h = 0, storage = CS_F32;
    // Populations are row-interleaved: row r holds planes 0..8 of that row
    // one after another, so element (i, r, c) is at i * plane + r * pitch + c
    // with plane = the padded row length and pitch = 9 * plane.  A thread
    // block then reads and writes a few contiguous row groups instead of
    // 9 streams 1/9 of the buffer apart.
    int pitch = 0;          // population row stride, elements
    int64_t plane = 0;      // population plane stride, elements
    int cpitch = 0;         // row pitch of the per-cell arrays (flags, node), elements
    void* buf[2] = {nullptr, nullptr};
    int cur = 0;            // buf[cur] holds the current populations
    uint8_t* flags = nullptr;   // h*pitch flags (device) or null
    uint32_t* node = nullptr;   // h*pitch node info (device) or null
    bool has_flags = false;
    cs_params params{};
    cs::Consts<float> k32{};
    cs::Consts<double> k64{};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float* snap = nullptr;      // 2 planes (ux, uy) of h*pitch
    bool has_snap = false;
    double* partials = nullptr; // reduction partials
    int npartials = 0;
    double* dscratch = nullptr; // small device scratch for reduction results
    int* iscratch = nullptr;
    double time = 0.0;
    uint64_t steps = 0;         // fused steps taken (keys the stochastic rounding)
    void* render_buf = nullptr; // render pass scratch (fields, partials, RGB)
    // Launch-bound lattices replay captured runs of fused steps: one graph per
    // starting buffer parity, valid for the parameters, geometry and launch
    // shape they were captured with (cleared when any of them changes).
    hipGraphExec_t step_graph[2] = {nullptr, nullptr};
    int graph_cpt = INT_MIN;    // cells-per-thread override at capture (INT_MIN: none captured)
    int graph_tb = 0;           // temporally blocked plan (tb_steps) at capture
};
