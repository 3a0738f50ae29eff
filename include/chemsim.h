/*
 * chemsim.h -- C ABI of libchemsim_hip.so, the MI355X (gfx950) implementation
 * of taktoa/chemsim's D2Q9 lattice-Boltzmann hot path (src/lbm.rs).
 *
 * In the reference, the path sits behind the Rust API of src/lbm.rs
 * (Lattice trait :111-176, CollisionOperator plugin trait :327-341, State
 * :670-819) and reaches the GPU through the arrayfire crate's C API
 * (af_convolve2 at lbm.rs:722, af_replace at :745, JIT element-wise arithmetic
 * through src/matrix.rs, af_imin_all at :817, af_get_data_ptr via
 * Matrix::get_underlying at matrix.rs:120-126).  This header is what a thin
 * FFI crate binds instead; each entry point names the reference interface it
 * replaces.
 *
 * Conventions
 *  - Every function returns int status (CS_OK = 0).  On failure
 *    cs_last_error() returns a thread-local message (the arrayfire crate
 *    panics on a non-zero af_err; the Rust binding does the same on a non-zero
 *    status).
 *  - Host arrays are in the reference host layout (matrix.rs:24-30,120-126):
 *    one plane is h*w values, row-major, index r*w + c; a population set is 9
 *    such planes in direction order 0..8 (lbm.rs:209-231).
 *  - Host pointers are borrowed only for the call.  Upload/download/moment
 *    calls are synchronous; cs_lattice_step is asynchronous on the lattice's
 *    own HIP stream.  Calls on one handle are not re-entrant; different
 *    handles may be used from different threads.
 *  - Handles own their device memory; *_destroy frees it.
 */
#ifndef CHEMSIM_H
#define CHEMSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CS_ABI_VERSION 1

typedef struct cs_ctx cs_ctx;         /* one GPU (af::init + set_backend, main.rs:331-332) */
typedef struct cs_lattice cs_lattice; /* State<D2Q9> with BGK (lbm.rs:670-676) */
typedef struct cs_batch cs_batch;     /* many independent lattices, one launch (extension) */
typedef struct cs_cnn cs_cnn;         /* warm-start network (extension) */

enum cs_status {
    CS_OK = 0,
    CS_ERR_ARG = 1,         /* bad argument (null, size, enum) */
    CS_ERR_HIP = 2,         /* HIP runtime error */
    CS_ERR_UNSUPPORTED = 3, /* operation not defined for this storage/mode */
    CS_ERR_STATE = 4,       /* call out of order (e.g. residual before snapshot) */
    CS_ERR_NOMEM = 5,
    CS_ERR_CALLBACK = 6     /* a caller's callback asked to stop */
};

/* Storage precision of the populations. */
enum cs_storage {
    CS_F32 = 0,     /* Scalar = f32, lbm.rs:13 */
    CS_F64 = 1,     /* FP64 oracle run (extension) */
    CS_BF16DEV = 2, /* bf16 deviations f_i - w_i*rho0, FP32 arithmetic (extension) */
    CS_F16DEV = 3,  /* fp16 deviations f_i - w_i*rho0, FP32 arithmetic (extension) */
    CS_F32DEV = 4   /* f32 deviations f_i - w_i*rho0, FP32 arithmetic (extension) */
};
/* The deviation storages (BF16DEV, F16DEV, F32DEV) run the step in deviation
 * arithmetic: drho = sum g_i, g_eq_i = w_i (drho + rho s_i), g' = g + (g - g_eq) K,
 * the BGK update of lbm.rs:350-364 written for g = f - w rho0.  F32 and F64
 * keep the reference's operation order on full populations. */

/* Treatment of the lattice edge.  ZERO is the reference: af::convolve2 with
 * ConvMode::DEFAULT pads with zeros (lbm.rs:722-724). */
enum cs_edge { CS_EDGE_ZERO = 0, CS_EDGE_PERIODIC = 1 };

/* REFERENCE: after streaming, solid cells take the opposite population
 * (State::bounce_back, lbm.rs:741-751) and every cell is then collided
 * (State::collide, lbm.rs:731-739).  HALFWAY (extension): links into solid
 * cells reflect half-way, moving lids add the Ladd term, inlet/outlet cells
 * apply Zou-He, solid cells are carried unchanged. */
enum cs_boundary { CS_BB_REFERENCE = 0, CS_BB_HALFWAY = 1 };

/* Node flags, one byte per cell.  In the REFERENCE mode the solid mask
 * (Geometry = af::Array<bool>, lbm.rs:99) is flag == WALL || flag == LID. */
enum cs_flag { CS_FLUID = 0, CS_WALL = 1, CS_LID = 2, CS_INLET = 3, CS_OUTLET = 4 };

/* Collision operator of a lattice: BGK { tau } (lbm.rs:345-370),
 * TRT { tau_plus = tau, tau_minus } (lbm.rs:374-451), Regularized<_>
 * (lbm.rs:596-666; its evaluate never calls the wrapped operator) or
 * KBC { ks_viscosity } (lbm.rs:458-590).  Operators other than BGK run on
 * F32 and F64 storage. */
enum cs_collision { CS_COLL_BGK = 0, CS_COLL_TRT = 1, CS_COLL_REGULARIZED = 2, CS_COLL_KBC = 3 };

typedef struct cs_params {
    double tau;        /* BGK { tau } (lbm.rs:345-347); TRT tau_plus */
    double delta_x;    /* Discretization { delta_x, delta_t } (lbm.rs:75-79) */
    double delta_t;
    int32_t edge;      /* enum cs_edge */
    int32_t boundary;  /* enum cs_boundary */
    double lid_ux;     /* moving-lid velocity, lattice components (extension) */
    double lid_uy;
    double inlet_ux;   /* Zou-He inlet velocity (extension) */
    double inlet_uy;
    double outlet_rho; /* Zou-He outlet density (extension) */
    double rho0;       /* base density of deviation storage (extension) */
    int32_t rounding;  /* 16-bit storage: CS_ROUND_NEAREST or CS_ROUND_STOCHASTIC */
    uint32_t seed;     /* seed of the stochastic rounding */
    int32_t collision; /* enum cs_collision */
    double tau_minus;  /* TRT { tau_minus } (lbm.rs:374-377) */
    double ks_viscosity; /* KBC { ks_viscosity } (lbm.rs:458-460) */
} cs_params;

/* Narrowing of the 16-bit storages.  STOCHASTIC rounds each stored value up
 * with probability equal to its fractional position between the two
 * neighbouring 16-bit values (gfx950 v_cvt_sr_*), using random bits keyed by
 * (seed, step index, cell index r*w + c); it removes the bias of round-to-
 * nearest on fields that change by less than an ulp per step. */
enum cs_rounding { CS_ROUND_NEAREST = 0, CS_ROUND_STOCHASTIC = 1 };

/* ---- errors, context -------------------------------------------------- */
const char* cs_last_error(void);
int cs_abi_version(void);
/* af::init + af::set_backend (main.rs:331-332): bind one HIP device. */
int cs_ctx_create(int device, cs_ctx** out);
int cs_ctx_destroy(cs_ctx* ctx);
/* Name of the bound device (af::device_info, main.rs:340). */
int cs_ctx_device_name(cs_ctx* ctx, char* buf, size_t len);

/* ---- lattice ------------------------------------------------------------ */
/* D2Q9::new + State::initial (lbm.rs:187-200, 679-692): w columns, h rows.
 * Unlike the reference (matrix.rs:40-44 vs 24-30) non-square sizes work. */
int cs_lattice_create(cs_ctx* ctx, int w, int h, int storage, cs_lattice** out);
int cs_lattice_destroy(cs_lattice* lat);
int cs_lattice_shape(const cs_lattice* lat, int* w, int* h, int* storage);
/* Collision operator + Discretization + boundary choices (lbm.rs:670-676). */
int cs_lattice_set_params(cs_lattice* lat, const cs_params* params);
/* State::geometry (lbm.rs:673; built at main.rs:269-312, edited at main.rs:77-89).
 * flags: h*w bytes of enum cs_flag, or NULL for all fluid.  Allowed between
 * any two steps. */
int cs_lattice_set_flags(cs_lattice* lat, const uint8_t* flags);
/* Population upload/download in the reference host layout
 * (Matrix::new / Matrix::get_underlying, matrix.rs:24-30,120-126).
 * The f32 pair works for every storage (16-bit storage converts); the f64
 * pair keeps full precision of an F64 lattice. */
int cs_lattice_upload_f32(cs_lattice* lat, const float* soa);
int cs_lattice_download_f32(cs_lattice* lat, float* soa);
int cs_lattice_upload_f64(cs_lattice* lat, const double* soa);
int cs_lattice_download_f64(cs_lattice* lat, double* soa);
/* Raw stored values (9*h*w elements of the storage type: float, double or the
 * 16-bit deviation bit patterns). */
int cs_lattice_upload_raw(cs_lattice* lat, const void* data);
int cs_lattice_download_raw(cs_lattice* lat, void* data);
/* f = compute_equilibrium(rho, (ux, uy)) (lbm.rs:43-71), as main.rs:257-267
 * initialises a lattice.  Inputs are h*w planes. */
int cs_lattice_init_equilibrium(cs_lattice* lat, const float* rho, const float* ux, const float* uy);
int cs_lattice_init_equilibrium_f64(cs_lattice* lat, const double* rho, const double* ux,
                                    const double* uy);
/* f = f_eq(rho, u) + f_neq, f_neq the first-order Chapman-Enskog
 * non-equilibrium part of the post-collision populations a lattice holds
 * between steps, -3 (tau - 1) w_i rho (c_i c_i - I/3) : S, S the symmetric
 * gradient of u from neighbour differences (walls half-way to solid cells,
 * moving with the lid velocity at CS_LID cells; f_neq = 0 at solid cells).
 * A consistent start from a predicted field (extension; set the flags and
 * params first). */
int cs_lattice_init_nonequilibrium(cs_lattice* lat, const float* rho, const float* ux, const float* uy);

/* ---- the step ------------------------------------------------------------ */
/* n x State::step (lbm.rs:694-714): stream, bounce-back and BGK collide fused
 * into one pull kernel per step, or (cs_lattice_temporal_steps) k steps per
 * launch, a remainder of 4 or more steps in one 4-step launch (BGK), the
 * rest one step at a time.  Asynchronous. */
int cs_lattice_step(cs_lattice* lat, int n);
/* As cs_lattice_step, bracketed by HIP events on the lattice's stream;
 * synchronises and returns the elapsed milliseconds. */
int cs_lattice_step_timed(cs_lattice* lat, int n, float* ms);
/* Steps per launch the lattice's cs_lattice_step takes now: k > 1 for the
 * temporally blocked kernel (cs_set_temporal_steps), 1 for the one-step
 * kernels.  No reference counterpart. */
int cs_lattice_temporal_steps(const cs_lattice* lat, int* k);
/* The three phases alone (State::stream :716-729, State::bounce_back
 * :741-751, State::collide :731-739), REFERENCE mode, F32/F64 storage.
 * stream + bounce_back + collide equals one cs_lattice_step. */
int cs_lattice_stream(cs_lattice* lat);
int cs_lattice_bounce_back(cs_lattice* lat);
int cs_lattice_collide(cs_lattice* lat);
/* State::time (lbm.rs:671, advanced at :713). */
int cs_lattice_time(const cs_lattice* lat, double* t);
int cs_lattice_sync(cs_lattice* lat);
/* Device address of the current populations: element (i, r, c) at
 * i * (*plane) + r * (*pitch) + c.  The planes are row-interleaved: row r of
 * plane i follows row r of plane i - 1 (*plane = padded row length, *pitch =
 * 9 * *plane). */
int cs_lattice_device_ptr(cs_lattice* lat, void** f, int* pitch, int64_t* plane);
/* HIP stream the lattice's kernels run on (hipStream_t). */
int cs_lattice_stream_handle(cs_lattice* lat, void** stream);

/* ---- moments (render-facing API) --------------------------------------- */
int cs_lattice_density(cs_lattice* lat, float* rho);                /* lbm.rs:117-121, 779-781 */
int cs_lattice_momentum_density(cs_lattice* lat, float* mx, float* my); /* :123-131, 789-792 */
int cs_lattice_velocity(cs_lattice* lat, float* ux, float* uy);     /* :133-149, 794-797 */
int cs_lattice_speed(cs_lattice* lat, float* speed);                /* :151-154, 799-802 */
int cs_lattice_pressure(cs_lattice* lat, float* p);                 /* :784-787 */
int cs_lattice_equilibrium(cs_lattice* lat, float* feq);            /* :156-160, 804-807 */
int cs_lattice_non_equilibrium(cs_lattice* lat, float* fneq);       /* :162-173, 809-812 */
int cs_lattice_is_unstable(cs_lattice* lat, int* unstable);         /* :814-818 */

/* ---- render pass (next row: render.rs:7-178 as main.rs:157-176 drives it) ---- */
/* Display modes of main.rs:37-52. */
enum cs_render_mode { CS_RENDER_DENSITY = 0, CS_RENDER_SPEED = 1, CS_RENDER_VELOCITY = 2, CS_RENDER_MOMENTUM = 3 };
/* render_scalar_field (density, speed) or render_vector_field (velocity,
 * momentum density) of the current populations into rgb (3 bytes per
 * pixel, row-major h x w: pixel (x, y) at 3 (y w + x)), then, if overlay,
 * render_geometry (solid cells (0, 0, 255)).  Synchronous; ms (if not null)
 * receives the GPU time of the passes. */
int cs_lattice_render(cs_lattice* lat, int mode, int overlay, uint8_t* rgb, float* ms);
/* ---- frame recording (next row: display.rs:157-185) --------------------- */
/* encoder.write_frame of display.rs:181-182: receives frame `frame` (0-based)
 * as h x w RGBA pixels (4 bytes, alpha 255; pixel (x, y) at 4 (y w + x)),
 * valid only during the call; return 0 to continue, non-zero to stop. */
typedef int (*cs_frame_fn)(void* user, int64_t frame, const uint8_t* rgba, int w, int h);
enum { CS_RECORD_MAX_DEPTH = 16 };
typedef struct cs_record_stats {
    double wall_ms;     /* host time from the first step to the last frame delivered */
    double gpu_ms;      /* lattice stream: first step to last render pass */
    double wait_ms;     /* host time blocked on frame copies */
    double callback_ms; /* host time inside the frame function */
    int64_t frames;     /* frames delivered */
} cs_record_stats;
/* record (display.rs:157-185) of this lattice: `frames` times, steps_per_frame
 * fused steps (the speed-factor loop of main.rs:128-136) then the render pass
 * of cs_lattice_render into RGBA, handed to fn in frame order.  Up to depth - 1
 * frames are in flight: their steps and render passes run on the lattice's
 * stream while earlier frames are copied to pinned host memory on a second
 * stream and passed to fn (depth 1: no overlap, the reference's order).  Every
 * frame equals step + cs_lattice_render.  When fn stops the recording, the
 * call returns CS_ERR_CALLBACK after the frames in flight finish: the lattice
 * has then advanced by up to depth - 1 frames more than were delivered.
 * Allocates 2 * depth frames (device + pinned host) for the call; stats may be null. */
int cs_lattice_record(cs_lattice* lat, int frames, int steps_per_frame, int mode, int overlay, int depth,
                      cs_frame_fn fn, void* user, cs_record_stats* stats);
/* BGK::evaluate (lbm.rs:350-364) on caller-supplied host populations f and
 * f_eq (n floats each, any layout): out = f + (f - f_eq) * K, K = (-dt)/tau in
 * f32.  The unfused CollisionOperator path of a host layer; synchronous. */
int cs_bgk_evaluate(cs_ctx* ctx, int64_t n, const float* f, const float* feq, float tau, float delta_t,
                    float* out);
/* TRT::new (lbm.rs:380-390): tau_plus and tau_minus from lambda and the
 * kinematic shear viscosity, in f32 as the reference evaluates them. */
int cs_trt_new(float lambda, float ks_viscosity, float delta_x, float delta_t, float* tau_plus, float* tau_minus);
/* Moments in the working precision, widened to f64 (any null output is skipped). */
int cs_lattice_moments_f64(cs_lattice* lat, double* rho, double* ux, double* uy);
/* The same moment pass with no device or host allocation: no counterpart in
 * the reference; for callers that cannot allocate during the call (serving
 * loops and harnesses that forbid hipMalloc).  The planes are staged in a
 * static device array of CS_NOALLOC_MAX_VALUES doubles per GPU (24 MiB, held
 * until the process ends) instead of a buffer allocated per call, so the
 * requested planes x w x h must be at most CS_NOALLOC_MAX_VALUES (three planes
 * up to 2^20 cells, one plane up to 3 x 2^20); larger lattices get CS_ERR_ARG.
 * Synchronous; the calls of one process are serialised on the shared array
 * (calls from several processes use separate copies). */
enum { CS_NOALLOC_MAX_VALUES = 3 << 20 };
int cs_lattice_moments_f64_noalloc(cs_lattice* lat, double* rho, double* ux, double* uy);

/* ---- FP64 reductions (extension) -------------------------------------- */
/* out = {sum rho, sum m_x, sum m_y} over non-solid cells, f64 accumulation. */
int cs_lattice_sums_f64(cs_lattice* lat, double out[3]);
/* Keep the current velocity field; cs_lattice_residual_f64 then returns
 * ||u - u_snapshot||_2 / ||u||_2 over non-solid cells in f64. */
int cs_lattice_snapshot_velocity(cs_lattice* lat);
int cs_lattice_residual_f64(cs_lattice* lat, double* residual);
/* Step until the residual taken every `interval` steps is <= tol, or
 * max_steps.  Returns the steps taken and the last residual. */
int cs_lattice_run_to_residual(cs_lattice* lat, double tol, int interval, int64_t max_steps,
                               int64_t* steps, double* residual);

/* ---- batch of independent lattices (extension) --------------------------
 * Many lattices of uneven sizes stepped by ONE persistent launch: each
 * workgroup takes whole lattices from a device-side queue and runs all steps
 * on them with block-level barriers between steps.  Results equal one
 * cs_lattice_step launch per lattice per step, bit for bit.  Each lattice has
 * its own cs_params and flags; all share one storage kind. */
enum cs_batch_order { CS_ORDER_LARGEST_FIRST = 0, CS_ORDER_GIVEN = 1, CS_ORDER_SMALLEST_FIRST = 2 };
/* CS_LAUNCH_PERSISTENT: workgroups take whole lattices from a queue and run
 * all steps of a call on them (block_threads, workgroups_per_cu,
 * cells_per_thread, order apply).  CS_LAUNCH_TILES: one launch per step over
 * the (block_threads / 64) x 64 tiles of every lattice.  CS_LAUNCH_TEMPORAL
 * (FP32 and FP32-deviation storage): one launch per 8 steps over the 64 x 64
 * regions of every lattice, each run 8 steps in registers or LDS (the
 * temporally blocked kernels); the remaining steps of a call (steps mod 8) as
 * tile launches (block_threads applies).  CS_LAUNCH_TEMPORAL_PERSISTENT: the
 * same rounds of 8 steps in one cooperative launch (per 4096 rounds) whose
 * resident blocks draw regions one at a time and meet at a grid-wide barrier
 * after every round (a batch whose lattices fall into more than one
 * boundary-mode group runs as CS_LAUNCH_TEMPORAL). */
enum cs_batch_launch {
    CS_LAUNCH_PERSISTENT = 0,
    CS_LAUNCH_TILES = 1,
    CS_LAUNCH_TEMPORAL = 2,
    CS_LAUNCH_TEMPORAL_PERSISTENT = 3
};
typedef struct cs_batch_tuning {
    int32_t block_threads;     /* 256, 512 or 1024 */
    int32_t workgroups_per_cu; /* persistent grid = CUs x this (capped at the lattice count) */
    int32_t cells_per_thread;  /* 1 or 2 */
    int32_t order;             /* enum cs_batch_order: queue order */
    int32_t launch;            /* enum cs_batch_launch */
} cs_batch_tuning;

int cs_batch_create(cs_ctx* ctx, int count, const int32_t* w, const int32_t* h, int storage, cs_batch** out);
int cs_batch_destroy(cs_batch* b);
int cs_batch_count(const cs_batch* b, int* count);
int cs_batch_set_params(cs_batch* b, int index, const cs_params* params);
int cs_batch_set_flags(cs_batch* b, int index, const uint8_t* flags);
int cs_batch_init_equilibrium(cs_batch* b, int index, const float* rho, const float* ux, const float* uy);
int cs_batch_upload_f32(cs_batch* b, int index, const float* soa);
int cs_batch_download_f32(cs_batch* b, int index, float* soa);
int cs_batch_download_raw(cs_batch* b, int index, void* data);
int cs_batch_velocity(cs_batch* b, int index, float* ux, float* uy);
int cs_batch_set_tuning(cs_batch* b, const cs_batch_tuning* tuning);
int cs_batch_get_tuning(const cs_batch* b, cs_batch_tuning* tuning);
/* `steps` steps of every active lattice (persistent launch or one tile-list
 * launch per step, by the tuning); synchronises and returns the elapsed
 * milliseconds (HIP events) if ms is not null.  Each lattice keeps its own
 * step count and buffer parity, so lattices taken out with
 * cs_batch_set_active keep their state. */
int cs_batch_step(cs_batch* b, int steps, float* ms);
/* The same steps as one fused launch per lattice per step (the comparison
 * path of the batch). */
int cs_batch_step_per_lattice(cs_batch* b, int steps, float* ms);
/* Times every tuning candidate for trial_steps steps (state restored after),
 * keeps the fastest and returns it. */
int cs_batch_autotune(cs_batch* b, int trial_steps, cs_batch_tuning* best, float* best_ms);
/* Inactive lattices are skipped by the steps (their state is kept). */
int cs_batch_set_active(cs_batch* b, int index, int active);
/* Velocity snapshot of every lattice; cs_batch_residuals then returns, per
 * lattice, ||u - u_snapshot||_2 / ||u||_2 over non-solid cells in f64. */
int cs_batch_snapshot_velocity(cs_batch* b);
int cs_batch_residuals(cs_batch* b, double* residuals);
/* Moments of one lattice (h*w planes), working precision rounded to f32. */
int cs_batch_moments(cs_batch* b, int index, float* rho, float* ux, float* uy);

/* ---- learned warm start (extension; no reference counterpart) ----------
 * A 4-level convolutional encoder-decoder (3x3 convolutions, ReLU, 2x2
 * average pooling, nearest up-sampling, skip concatenations; base channel
 * count c) predicting the steady (rho, u) of a channel flow from 8 input
 * channels per cell.  BF16 MFMA implicit-GEMM kernels with FP32 accumulation,
 * FP32 master weights, Adam.  Tensors are [n][height][width][channels] f32 on
 * the host; height and width are multiples of 8.  Parameters are ordered per
 * convolution: weights [cout][3][3][cin] then bias [cout]. */
int cs_cnn_create(cs_ctx* ctx, int height, int width, int base_channels, int batch, uint32_t seed, cs_cnn** out);
int cs_cnn_destroy(cs_cnn* net);
int cs_cnn_param_count(const cs_cnn* net, int64_t* count);
int cs_cnn_get_params(cs_cnn* net, float* params);
int cs_cnn_set_params(cs_cnn* net, const float* params);
/* Gradients of the last cs_cnn_loss_grad / training minibatch. */
int cs_cnn_get_grads(cs_cnn* net, float* grads);
/* input [n][h][w][8] -> output [n][h][w][3]; n <= batch. */
int cs_cnn_forward(cs_cnn* net, int n, const float* input, float* output);
/* loss = sum weight * (output - target)^2 / (3 n h w); forward + backward. */
int cs_cnn_loss_grad(cs_cnn* net, int n, const float* input, const float* target, const float* weight, float* loss);
int cs_cnn_adam_step(cs_cnn* net, float lr, float beta1, float beta2, float eps);
/* Device-resident training set: inputs [count][h][w][8], targets
 * [count][h][w][3], weights [count][h][w]. */
int cs_cnn_set_dataset(cs_cnn* net, int count, const float* inputs, const float* targets, const float* weights);
/* One pass in the given sample order, minibatches of `batch`, Adam per batch. */
int cs_cnn_train_epoch(cs_cnn* net, const int32_t* order, int count, float lr, float* mean_loss);
/* Average ms of `iters` forward+backward passes and the GEMM FLOPs of one. */
int cs_cnn_time_step(cs_cnn* net, int iters, float* ms, double* flops);
/* C[M][N] (f32) = A[M][K] . B[N][K]^T in BF16 on the MFMA tile kernel
 * (K a multiple of 32); returns the average kernel ms. */
int cs_gemm_bf16(cs_ctx* ctx, int M, int N, int K, const uint16_t* A, const uint16_t* B, float* C, float* ms);

/* ---- measurement helpers ---------------------------------------------- */
/* Plain device-to-device copy kernel of `bytes` bytes, `iters` times; returns
 * the average milliseconds per copy (HIP events).  Calibrates bandwidth and
 * the HBM byte counters. */
int cs_bench_copy(cs_ctx* ctx, int64_t bytes, int iters, float* ms_per_copy);
/* No counterpart in the reference.  The FP32 step kernels compute 1 / rho
 * with the hardware reciprocal and one FMA Newton-Raphson step where that
 * equals the IEEE quotient; this runs every one of the 2^32 single-precision
 * inputs through that reciprocal on the context's GPU and returns how many
 * differ from IEEE 1.0f / b (0 expected; two NaNs count as equal). */
int cs_check_reciprocal(cs_ctx* ctx, uint64_t* mismatches);
/* Launch shape of the step kernel.  Both kernels give the same values: the
 * scalar kernel (one row per wave, 1, 2 or 4 cells per lane) and the vector
 * kernel (a lane owns consecutive cells of a row, one 8- or 16-byte access per
 * plane).  0 = the built-in choice by lattice size; 1, 2 or 4 = the scalar
 * kernel with that many cells per lane; CS_STEP_VECTOR = the vector kernel
 * wherever the row splits into vectors (16-byte lanes for 4- and 8-byte
 * storage, 8-byte lanes for 16-bit storage), else the built-in choice;
 * CS_STEP_VECTOR_WIDE = the same with 16-byte lanes for 16-bit storage too
 * (no geometry).
 * Process-wide; for tests and tuning.  No reference counterpart (ArrayFire
 * picks its own launch shapes). */
enum { CS_STEP_VECTOR = -1, CS_STEP_VECTOR_WIDE = -2 };
int cs_set_cells_per_thread(int cells_per_thread);
/* Replay of captured HIP graphs of 64 fused steps for lattices of at most
 * 2^20 cells, where launch overhead dominates a step (1 = on, the default;
 * 0 = launch step by step).  Results are identical either way; 16-bit storage
 * with stochastic rounding always launches step by step.  Process-wide; for
 * tests and tuning.  No reference counterpart. */
int cs_set_step_graphs(int enable);
/* Steps per launch of the temporally blocked step kernels: CS_TEMPORAL_AUTO
 * (the default: 8 steps for every FP32 and FP32-deviation lattice except
 * with the KBC operator -- on 32 x 32 regions for lattices of at most 1024
 * of them, about 512^2, on 64 x 64 regions above), 4 or 8 (every FP32 and
 * FP32-deviation lattice, on 64 x 64 regions), or 0 for one step per launch.
 * Results are identical either way.  A launch-shape override of
 * cs_set_cells_per_thread selects the one-step kernels instead.
 * Process-wide; for tests and tuning.  No reference counterpart. */
enum { CS_TEMPORAL_AUTO = -1 };
int cs_set_temporal_steps(int k);

#ifdef __cplusplus
}
#endif
#endif
