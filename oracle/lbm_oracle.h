/*
 * lbm_oracle.h -- CPU restatement of the D2Q9 lattice-Boltzmann hot path of
 * taktoa/chemsim (src/lbm.rs).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity reference: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * never by the product library (chemsim_amd/csrc), which has no CPU path.
 *
 * Layout (reference host layout, src/matrix.rs:24-30,120-126): a population
 * set is 9 planes, direction order 0..8 (src/lbm.rs:209-231), each plane h*w
 * values, row-major, index r*w + c.  Row r is the array's dim0 (image y),
 * column c is dim1 (image x).
 *
 * Node flags (one byte per cell, same layout as a plane):
 *   0 fluid, 1 wall (solid), 2 moving lid (solid), 3 Zou-He velocity inlet,
 *   4 Zou-He density outlet.
 * In the reference boundary mode (src/lbm.rs:741-751) a cell is solid when its
 * flag is 1 or 2, and only that matters.
 *
 * Which pieces follow the reference and which are extensions is stated on each
 * declaration; lbm_oracle_impl.inc carries the file:line citations per step.
 */
#ifndef CHEMSIM_LBM_ORACLE_H
#define CHEMSIM_LBM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_EDGE_ZERO = 0, ORC_EDGE_PERIODIC = 1 };
enum { ORC_BB_REFERENCE = 0, ORC_BB_HALFWAY = 1 };
enum { ORC_FLUID = 0, ORC_WALL = 1, ORC_LID = 2, ORC_INLET = 3, ORC_OUTLET = 4 };
enum { ORC_BF16 = 0, ORC_F16 = 1 };
enum { ORC_COLL_BGK = 0, ORC_COLL_TRT = 1, ORC_COLL_REGULARIZED = 2, ORC_COLL_KBC = 3 };

typedef struct orc_params {
    int w, h;          /* columns, rows */
    int edge;          /* ORC_EDGE_*: zero fill (af::convolve2 DEFAULT) or periodic */
    int boundary;      /* ORC_BB_*: on-node mask swap (lbm.rs) or half-way links */
    double tau;        /* BGK relaxation time (lbm.rs:345-347) */
    double delta_x;    /* Discretization (lbm.rs:75-86) */
    double delta_t;
    double lid_ux, lid_uy;     /* moving-wall velocity, lattice components */
    double inlet_ux, inlet_uy; /* Zou-He inlet velocity (extension) */
    double outlet_rho;         /* Zou-He outlet density (extension) */
    double rho0;               /* base density of 16-bit deviation storage */
    int collision;             /* ORC_COLL_*: BGK{tau}, TRT{tau_plus = tau, tau_minus},
                                  Regularized{} or KBC{ks_viscosity} */
    double tau_minus;          /* TRT (lbm.rs:374-377) */
    double ks_viscosity;       /* KBC (lbm.rs:458-460) */
} orc_params;

/* Derived per-run constants, in the working precision.
 * K = -dt/tau (lbm.rs:357); A = 1/cs^2, B = 1/(2 cs^4), C = -1/(2 cs^2)
 * (lbm.rs:54-66, 84); w_i (lbm.rs:209-219); lid_i = 6 w_i (c_i . u_lid)
 * (Ladd moving wall, extension); base_i = w_i * rho0 (16-bit storage). */
typedef struct orc_consts_f32 {
    float K, A, B, C, cs2;
    float w[9], lid[9], base[9];
    float wp, wm, hdt; /* TRT: 1/tau_plus, 1/tau_minus, -dt * 0.5 (lbm.rs:431-440) */
    float dx, beta;    /* KBC: delta_x, 1 / (2 nu / cs^2 + 1) (lbm.rs:478, 547-550) */
    float sf[9];       /* Regularized: w_i / (2 cs^4) (lbm.rs:651) */
} orc_consts_f32;
typedef struct orc_consts_f64 {
    double K, A, B, C, cs2;
    double w[9], lid[9], base[9];
    double wp, wm, hdt;
    double dx, beta;
    double sf[9];
} orc_consts_f64;

void orc_consts_f32_make(const orc_params* p, orc_consts_f32* out);
void orc_consts_f64_make(const orc_params* p, orc_consts_f64* out);
/* Deviation-path constants: as orc_consts_f32_make, with A, B, C, cs2 taken
 * from the f64 evaluation rounded once (extension). */
void orc_consts_devf32_make(const orc_params* p, orc_consts_f32* out);
/* TRT::new (lbm.rs:380-390) in f32: tau_plus, tau_minus from lambda and the
 * kinematic shear viscosity. */
void orc_trt_new(float lambda, float ks_viscosity, float delta_x, float delta_t, float* tau_plus,
                 float* tau_minus);
/* TRT::evaluate (lbm.rs:401-444) of one plane set: out = F + omega with
 * F_eq recomputed from F and the swapped equilibrium of Lattice::swap_equilibrium
 * (lbm.rs:311-323: f_eq with directions 1..8 replaced by the opposite
 * populations).  n cells per plane. */
void orc_trt_f32(const orc_params* p, int n, const float* F, float* out);
void orc_trt_f64(const orc_params* p, int n, const double* F, double* out);
/* Regularized::evaluate (lbm.rs:606-661): f_eq + the non-equilibrium stress
 * P_ab = sum_i (f_i - f_eq_i) c_ia c_ib projected on Q_i,ab = c_ia c_ib -
 * cs^2 delta_ab with weight w_i / (2 cs^4).  The wrapped operator is unused,
 * as in the reference. */
void orc_regularized_f32(const orc_params* p, int n, const float* F, float* out);
void orc_regularized_f64(const orc_params* p, int n, const double* F, double* out);
/* KBC::evaluate (lbm.rs:468-585): the reference's shear part delta_s, the
 * remainder delta_h = f - f_eq - delta_s, the stabiliser gamma* and
 * f + (-2 beta) delta_s - beta gamma* delta_h.  The reference's diagnostic
 * print (lbm.rs:574-582) has no output and is not restated.  n cells per plane. */
void orc_kbc_f32(const orc_params* p, int n, const float* F, float* out);
void orc_kbc_f64(const orc_params* p, int n, const double* F, double* out);

/* Consistent start (extension): F (an equilibrium, plain or deviation form)
 * += the first-order Chapman-Enskog non-equilibrium part of post-collision
 * populations at fluid cells, -3 (tau - 1) rho w_i (c_i c_i - I/3) : S, S from
 * neighbour differences with walls half-way to solid nodes (flags dense h*w
 * or NULL). */
void orc_add_neq_f32(const orc_params* p, const uint8_t* flags, const float* rho, const float* ux,
                     const float* uy, float* F);
void orc_add_neq_f64(const orc_params* p, const uint8_t* flags, const double* rho, const double* ux,
                     const double* uy, double* F);

/* ---- Reference operations, one pass each, as lbm.rs runs them ---------- */
/* State::stream (lbm.rs:716-729): Q_i[r][c] = P_i[r+cx_i][c-cy_i], zero outside
 * (edge = ZERO) or wrapped (edge = PERIODIC, extension). */
void orc_stream_f32(int w, int h, int edge, const float* P, float* Q);
void orc_stream_f64(int w, int h, int edge, const double* P, double* Q);
/* State::bounce_back (lbm.rs:741-751): B_i = solid ? Q_opp(i) : Q_i. */
void orc_bounce_back_f32(int w, int h, const uint8_t* flags, const float* Q, float* B);
void orc_bounce_back_f64(int w, int h, const uint8_t* flags, const double* Q, double* B);
/* Lattice::density / momentum_density / velocity (lbm.rs:117-149). */
void orc_density_f32(int w, int h, const float* F, float* rho);
void orc_density_f64(int w, int h, const double* F, double* rho);
void orc_momentum_f32(int w, int h, const float* F, float* mx, float* my);
void orc_momentum_f64(int w, int h, const double* F, double* mx, double* my);
void orc_velocity_f32(int w, int h, const float* F, float* ux, float* uy);
void orc_velocity_f64(int w, int h, const double* F, double* ux, double* uy);
/* compute_equilibrium (lbm.rs:43-71). */
void orc_equilibrium_f32(const orc_params* p, int n, const float* rho, const float* ux,
                         const float* uy, float* Feq);
void orc_equilibrium_f64(const orc_params* p, int n, const double* rho, const double* ux,
                         const double* uy, double* Feq);
/* BGK::evaluate (lbm.rs:350-364) over every cell (State::collide, lbm.rs:731-739). */
void orc_bgk_f32(const orc_params* p, int n, const float* F, const float* Feq, float* out);
void orc_bgk_f64(const orc_params* p, int n, const double* F, const double* Feq, double* out);
/* State::step (lbm.rs:694-714), unfused: stream, bounce_back, collide.
 * Reference boundary mode only.  scratch holds 9n + 9n + 3n values. */
void orc_step_unfused_f32(const orc_params* p, const uint8_t* flags, const float* P,
                          float* out, float* scratch);
void orc_step_unfused_f64(const orc_params* p, const uint8_t* flags, const double* P,
                          double* out, double* scratch);

/* ---- Fused pull step, every mode ------------------------------------------
 * In the reference boundary mode this equals orc_step_unfused bit for bit.
 * The half-way mode, moving lid, periodic edges and Zou-He cells are
 * extensions with no reference code. */
void orc_step_f32(const orc_params* p, const uint8_t* flags, const float* P, float* out);
void orc_step_f64(const orc_params* p, const uint8_t* flags, const double* P, double* out);
/* Rows [r0, r1) only (used by the threaded baseline). */
void orc_step_rows_f32(const orc_params* p, const uint8_t* flags, const float* P, float* out,
                       int r0, int r1);
/* n steps in place on f (internal second buffer). */
int orc_run_f32(const orc_params* p, const uint8_t* flags, float* f, int nsteps);
int orc_run_f64(const orc_params* p, const uint8_t* flags, double* f, int nsteps);
/* Threaded fused run (CPU baseline); returns threads used. */
int orc_run_f32_omp(const orc_params* p, const uint8_t* flags, float* f, int nsteps, int threads);
/* Unfused run, one pass per lbm.rs op, single thread (CPU baseline). */
int orc_run_unfused_f32(const orc_params* p, const uint8_t* flags, float* f, int nsteps);

/* ---- 16-bit deviation storage (extension) ------------------------------
 * Stored value d_i = f_i - w_i*rho0 in BF16 or FP16 (round to nearest even);
 * the step widens to FP32, runs the deviation-form step (orc_step_devf32)
 * and narrows. */
uint16_t orc_f32_to_bf16(float x);
float    orc_bf16_to_f32(uint16_t b);
uint16_t orc_f32_to_f16(float x);
float    orc_f16_to_f32(uint16_t b);
void orc_f32_to_bf16_n(const float* x, uint16_t* y, long n);
void orc_f32_to_f16_n(const float* x, uint16_t* y, long n);
void orc_step_dev16(const orc_params* p, int kind, const uint8_t* flags, const uint16_t* P,
                    uint16_t* out);
int orc_run_dev16(const orc_params* p, int kind, const uint8_t* flags, uint16_t* d, int nsteps);
/* Stochastic rounding (rounding = 1): see lbm_oracle.c for the bit model and
 * the random-bit schedule keyed by (seed, step, cell index r*w + c). */
uint16_t orc_f32_to_bf16_sr(float x, uint32_t r);
uint16_t orc_f32_to_f16_sr(float x, uint32_t r);
uint32_t orc_sr_key(uint32_t seed, uint32_t step);
void orc_sr_bits(uint32_t key, uint32_t cell, uint32_t R[9]);
void orc_step_dev16_ex(const orc_params* p, int kind, int rounding, uint32_t seed, uint32_t step,
                       const uint8_t* flags, const uint16_t* P, uint16_t* out);
int orc_run_dev16_ex(const orc_params* p, int kind, int rounding, uint32_t seed, uint32_t step0,
                     const uint8_t* flags, uint16_t* d, int nsteps);
/* Deviation arithmetic on f32 deviations g_i = f_i - w_i*rho0 (extension). */
void orc_step_devf32(const orc_params* p, const uint8_t* flags, const float* G, float* out);
int orc_run_devf32(const orc_params* p, const uint8_t* flags, float* g, int nsteps);
void orc_equilibrium_devf32(const orc_params* p, int n, const float* rho, const float* ux,
                            const float* uy, float* G);
/* Encode/decode between full populations (f32) and stored deviations. */
void orc_encode_dev16(const orc_params* p, int kind, const float* f, uint16_t* d);
void orc_decode_dev16(const orc_params* p, int kind, const uint16_t* d, float* f);

/* ---- FP64 reductions (extension) ----------------------------------------
 * out[0] = sum rho, out[1] = sum m_x, out[2] = sum m_y over cells whose flag is
 * not solid (all cells when flags is NULL); moments follow lbm.rs order in the
 * input precision, then widen to f64 and accumulate left to right. */
void orc_sums_f32(int w, int h, const uint8_t* flags, const float* f, double out[3]);
void orc_sums_f64(int w, int h, const uint8_t* flags, const double* f, double out[3]);
/* ||u - u_prev||_2 / ||u||_2 in f64 over non-solid cells. */
double orc_residual(int w, int h, const uint8_t* flags, const float* ux, const float* uy,
                    const float* ux_prev, const float* uy_prev);

#ifdef __cplusplus
}
#endif
#endif
