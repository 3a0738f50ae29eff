//! Raw bindings to `libchemsim_hip.so`: every declaration of
//! `include/chemsim.h`, one to one (names, argument order and C types), for
//! the `lbm` module of chemsim to call instead of the arrayfire crate on its
//! hot path (`State::step`, `lbm.rs:694-714`, and the moment accessors,
//! `lbm.rs:117-160, 779-818`).
//!
//! Conventions of the C ABI: every function returns a status (`CS_OK` = 0);
//! on failure `cs_last_error()` holds a thread-local message.  Host arrays are
//! in the reference host layout (`matrix.rs:24-30,120-126`): a plane is h*w
//! values, row-major; a population set is 9 planes in direction order 0..8.
//! Host pointers are borrowed for the call only.  Handles own their device
//! memory and are freed by their `*_destroy`.  Calls on one handle are not
//! re-entrant; different handles may be used from different threads.
#![no_std]
#![allow(non_camel_case_types)]

use core::ffi::{c_char, c_int, c_void};

pub const CS_ABI_VERSION: c_int = 1;

// ---- opaque handles ------------------------------------------------------
#[repr(C)]
pub struct cs_ctx {
    _private: [u8; 0],
}
#[repr(C)]
pub struct cs_lattice {
    _private: [u8; 0],
}
#[repr(C)]
pub struct cs_batch {
    _private: [u8; 0],
}
#[repr(C)]
pub struct cs_cnn {
    _private: [u8; 0],
}

// ---- enum cs_status --------------------------------------------------------
pub const CS_OK: c_int = 0;
pub const CS_ERR_ARG: c_int = 1;
pub const CS_ERR_HIP: c_int = 2;
pub const CS_ERR_UNSUPPORTED: c_int = 3;
pub const CS_ERR_STATE: c_int = 4;
pub const CS_ERR_NOMEM: c_int = 5;
pub const CS_ERR_CALLBACK: c_int = 6;

// ---- enum cs_storage -------------------------------------------------------
pub const CS_F32: c_int = 0;
pub const CS_F64: c_int = 1;
pub const CS_BF16DEV: c_int = 2;
pub const CS_F16DEV: c_int = 3;
pub const CS_F32DEV: c_int = 4;

// ---- enum cs_edge, cs_boundary, cs_flag, cs_collision, cs_rounding ---------
pub const CS_EDGE_ZERO: i32 = 0;
pub const CS_EDGE_PERIODIC: i32 = 1;
pub const CS_BB_REFERENCE: i32 = 0;
pub const CS_BB_HALFWAY: i32 = 1;
pub const CS_FLUID: u8 = 0;
pub const CS_WALL: u8 = 1;
pub const CS_LID: u8 = 2;
pub const CS_INLET: u8 = 3;
pub const CS_OUTLET: u8 = 4;
pub const CS_COLL_BGK: i32 = 0;
pub const CS_COLL_TRT: i32 = 1;
pub const CS_COLL_REGULARIZED: i32 = 2;
pub const CS_COLL_KBC: i32 = 3;
pub const CS_ROUND_NEAREST: i32 = 0;
pub const CS_ROUND_STOCHASTIC: i32 = 1;

/// `cs_params`: the collision operator (`BGK { tau }`, `TRT { tau_plus = tau,
/// tau_minus }`, `Regularized`, `KBC { ks_viscosity }`), the `Discretization`
/// and the boundary choices of a lattice (`lbm.rs:670-676`).  Field order and
/// types are the C struct's.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct cs_params {
    pub tau: f64,
    pub delta_x: f64,
    pub delta_t: f64,
    pub edge: i32,
    pub boundary: i32,
    pub lid_ux: f64,
    pub lid_uy: f64,
    pub inlet_ux: f64,
    pub inlet_uy: f64,
    pub outlet_rho: f64,
    pub rho0: f64,
    pub rounding: i32,
    pub seed: u32,
    pub collision: i32,
    pub tau_minus: f64,
    pub ks_viscosity: f64,
}

// ---- render and recording ------------------------------------------------
pub const CS_RENDER_DENSITY: c_int = 0;
pub const CS_RENDER_SPEED: c_int = 1;
pub const CS_RENDER_VELOCITY: c_int = 2;
pub const CS_RENDER_MOMENTUM: c_int = 3;
pub const CS_RECORD_MAX_DEPTH: c_int = 16;

/// `encoder.write_frame` of `display.rs:181-182`: frame index and h x w RGBA
/// pixels, valid during the call only; return 0 to continue.
pub type cs_frame_fn =
    Option<unsafe extern "C" fn(user: *mut c_void, frame: i64, rgba: *const u8, w: c_int, h: c_int) -> c_int>;

#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct cs_record_stats {
    pub wall_ms: f64,
    pub gpu_ms: f64,
    pub wait_ms: f64,
    pub callback_ms: f64,
    pub frames: i64,
}

pub const CS_NOALLOC_MAX_VALUES: c_int = 3 << 20;

// ---- batch -------------------------------------------------------------------
pub const CS_ORDER_LARGEST_FIRST: i32 = 0;
pub const CS_ORDER_GIVEN: i32 = 1;
pub const CS_ORDER_SMALLEST_FIRST: i32 = 2;
pub const CS_LAUNCH_PERSISTENT: i32 = 0;
pub const CS_LAUNCH_TILES: i32 = 1;
pub const CS_LAUNCH_TEMPORAL: i32 = 2;
pub const CS_LAUNCH_TEMPORAL_PERSISTENT: i32 = 3;

#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct cs_batch_tuning {
    pub block_threads: i32,
    pub workgroups_per_cu: i32,
    pub cells_per_thread: i32,
    pub order: i32,
    pub launch: i32,
}

pub const CS_STEP_VECTOR: c_int = -1;
pub const CS_STEP_VECTOR_WIDE: c_int = -2;
pub const CS_TEMPORAL_AUTO: c_int = -1;

extern "C" {
    // errors, context
    pub fn cs_last_error() -> *const c_char;
    pub fn cs_abi_version() -> c_int;
    pub fn cs_ctx_create(device: c_int, out: *mut *mut cs_ctx) -> c_int;
    pub fn cs_ctx_destroy(ctx: *mut cs_ctx) -> c_int;
    pub fn cs_ctx_device_name(ctx: *mut cs_ctx, buf: *mut c_char, len: usize) -> c_int;

    // lattice
    pub fn cs_lattice_create(ctx: *mut cs_ctx, w: c_int, h: c_int, storage: c_int, out: *mut *mut cs_lattice)
        -> c_int;
    pub fn cs_lattice_destroy(lat: *mut cs_lattice) -> c_int;
    pub fn cs_lattice_shape(lat: *const cs_lattice, w: *mut c_int, h: *mut c_int, storage: *mut c_int) -> c_int;
    pub fn cs_lattice_set_params(lat: *mut cs_lattice, params: *const cs_params) -> c_int;
    pub fn cs_lattice_set_flags(lat: *mut cs_lattice, flags: *const u8) -> c_int;
    pub fn cs_lattice_upload_f32(lat: *mut cs_lattice, soa: *const f32) -> c_int;
    pub fn cs_lattice_download_f32(lat: *mut cs_lattice, soa: *mut f32) -> c_int;
    pub fn cs_lattice_upload_f64(lat: *mut cs_lattice, soa: *const f64) -> c_int;
    pub fn cs_lattice_download_f64(lat: *mut cs_lattice, soa: *mut f64) -> c_int;
    pub fn cs_lattice_upload_raw(lat: *mut cs_lattice, data: *const c_void) -> c_int;
    pub fn cs_lattice_download_raw(lat: *mut cs_lattice, data: *mut c_void) -> c_int;
    pub fn cs_lattice_init_equilibrium(lat: *mut cs_lattice, rho: *const f32, ux: *const f32, uy: *const f32)
        -> c_int;
    pub fn cs_lattice_init_equilibrium_f64(lat: *mut cs_lattice, rho: *const f64, ux: *const f64, uy: *const f64)
        -> c_int;
    pub fn cs_lattice_init_nonequilibrium(lat: *mut cs_lattice, rho: *const f32, ux: *const f32, uy: *const f32)
        -> c_int;

    // the step
    pub fn cs_lattice_step(lat: *mut cs_lattice, n: c_int) -> c_int;
    pub fn cs_lattice_step_timed(lat: *mut cs_lattice, n: c_int, ms: *mut f32) -> c_int;
    pub fn cs_lattice_temporal_steps(lat: *const cs_lattice, k: *mut c_int) -> c_int;
    pub fn cs_lattice_stream(lat: *mut cs_lattice) -> c_int;
    pub fn cs_lattice_bounce_back(lat: *mut cs_lattice) -> c_int;
    pub fn cs_lattice_collide(lat: *mut cs_lattice) -> c_int;
    pub fn cs_lattice_time(lat: *const cs_lattice, t: *mut f64) -> c_int;
    pub fn cs_lattice_sync(lat: *mut cs_lattice) -> c_int;
    pub fn cs_lattice_device_ptr(lat: *mut cs_lattice, f: *mut *mut c_void, pitch: *mut c_int, plane: *mut i64)
        -> c_int;
    pub fn cs_lattice_stream_handle(lat: *mut cs_lattice, stream: *mut *mut c_void) -> c_int;

    // moments
    pub fn cs_lattice_density(lat: *mut cs_lattice, rho: *mut f32) -> c_int;
    pub fn cs_lattice_momentum_density(lat: *mut cs_lattice, mx: *mut f32, my: *mut f32) -> c_int;
    pub fn cs_lattice_velocity(lat: *mut cs_lattice, ux: *mut f32, uy: *mut f32) -> c_int;
    pub fn cs_lattice_speed(lat: *mut cs_lattice, speed: *mut f32) -> c_int;
    pub fn cs_lattice_pressure(lat: *mut cs_lattice, p: *mut f32) -> c_int;
    pub fn cs_lattice_equilibrium(lat: *mut cs_lattice, feq: *mut f32) -> c_int;
    pub fn cs_lattice_non_equilibrium(lat: *mut cs_lattice, fneq: *mut f32) -> c_int;
    pub fn cs_lattice_is_unstable(lat: *mut cs_lattice, unstable: *mut c_int) -> c_int;

    // render pass and frame recording
    pub fn cs_lattice_render(lat: *mut cs_lattice, mode: c_int, overlay: c_int, rgb: *mut u8, ms: *mut f32)
        -> c_int;
    pub fn cs_lattice_record(lat: *mut cs_lattice, frames: c_int, steps_per_frame: c_int, mode: c_int,
                             overlay: c_int, depth: c_int, fn_: cs_frame_fn, user: *mut c_void,
                             stats: *mut cs_record_stats) -> c_int;
    pub fn cs_bgk_evaluate(ctx: *mut cs_ctx, n: i64, f: *const f32, feq: *const f32, tau: f32, delta_t: f32,
                           out: *mut f32) -> c_int;
    pub fn cs_trt_new(lambda: f32, ks_viscosity: f32, delta_x: f32, delta_t: f32, tau_plus: *mut f32,
                      tau_minus: *mut f32) -> c_int;
    pub fn cs_lattice_moments_f64(lat: *mut cs_lattice, rho: *mut f64, ux: *mut f64, uy: *mut f64) -> c_int;
    pub fn cs_lattice_moments_f64_noalloc(lat: *mut cs_lattice, rho: *mut f64, ux: *mut f64, uy: *mut f64)
        -> c_int;

    // FP64 reductions
    pub fn cs_lattice_sums_f64(lat: *mut cs_lattice, out: *mut f64) -> c_int;
    pub fn cs_lattice_snapshot_velocity(lat: *mut cs_lattice) -> c_int;
    pub fn cs_lattice_residual_f64(lat: *mut cs_lattice, residual: *mut f64) -> c_int;
    pub fn cs_lattice_run_to_residual(lat: *mut cs_lattice, tol: f64, interval: c_int, max_steps: i64,
                                      steps: *mut i64, residual: *mut f64) -> c_int;

    // batch of independent lattices
    pub fn cs_batch_create(ctx: *mut cs_ctx, count: c_int, w: *const i32, h: *const i32, storage: c_int,
                           out: *mut *mut cs_batch) -> c_int;
    pub fn cs_batch_destroy(b: *mut cs_batch) -> c_int;
    pub fn cs_batch_count(b: *const cs_batch, count: *mut c_int) -> c_int;
    pub fn cs_batch_set_params(b: *mut cs_batch, index: c_int, params: *const cs_params) -> c_int;
    pub fn cs_batch_set_flags(b: *mut cs_batch, index: c_int, flags: *const u8) -> c_int;
    pub fn cs_batch_init_equilibrium(b: *mut cs_batch, index: c_int, rho: *const f32, ux: *const f32,
                                     uy: *const f32) -> c_int;
    pub fn cs_batch_upload_f32(b: *mut cs_batch, index: c_int, soa: *const f32) -> c_int;
    pub fn cs_batch_download_f32(b: *mut cs_batch, index: c_int, soa: *mut f32) -> c_int;
    pub fn cs_batch_download_raw(b: *mut cs_batch, index: c_int, data: *mut c_void) -> c_int;
    pub fn cs_batch_velocity(b: *mut cs_batch, index: c_int, ux: *mut f32, uy: *mut f32) -> c_int;
    pub fn cs_batch_set_tuning(b: *mut cs_batch, tuning: *const cs_batch_tuning) -> c_int;
    pub fn cs_batch_get_tuning(b: *const cs_batch, tuning: *mut cs_batch_tuning) -> c_int;
    pub fn cs_batch_step(b: *mut cs_batch, steps: c_int, ms: *mut f32) -> c_int;
    pub fn cs_batch_step_per_lattice(b: *mut cs_batch, steps: c_int, ms: *mut f32) -> c_int;
    pub fn cs_batch_autotune(b: *mut cs_batch, trial_steps: c_int, best: *mut cs_batch_tuning, best_ms: *mut f32)
        -> c_int;
    pub fn cs_batch_set_active(b: *mut cs_batch, index: c_int, active: c_int) -> c_int;
    pub fn cs_batch_snapshot_velocity(b: *mut cs_batch) -> c_int;
    pub fn cs_batch_residuals(b: *mut cs_batch, residuals: *mut f64) -> c_int;
    pub fn cs_batch_moments(b: *mut cs_batch, index: c_int, rho: *mut f32, ux: *mut f32, uy: *mut f32) -> c_int;

    // learned warm start
    pub fn cs_cnn_create(ctx: *mut cs_ctx, height: c_int, width: c_int, base_channels: c_int, batch: c_int,
                         seed: u32, out: *mut *mut cs_cnn) -> c_int;
    pub fn cs_cnn_destroy(net: *mut cs_cnn) -> c_int;
    pub fn cs_cnn_param_count(net: *const cs_cnn, count: *mut i64) -> c_int;
    pub fn cs_cnn_get_params(net: *mut cs_cnn, params: *mut f32) -> c_int;
    pub fn cs_cnn_set_params(net: *mut cs_cnn, params: *const f32) -> c_int;
    pub fn cs_cnn_get_grads(net: *mut cs_cnn, grads: *mut f32) -> c_int;
    pub fn cs_cnn_forward(net: *mut cs_cnn, n: c_int, input: *const f32, output: *mut f32) -> c_int;
    pub fn cs_cnn_loss_grad(net: *mut cs_cnn, n: c_int, input: *const f32, target: *const f32,
                            weight: *const f32, loss: *mut f32) -> c_int;
    pub fn cs_cnn_adam_step(net: *mut cs_cnn, lr: f32, beta1: f32, beta2: f32, eps: f32) -> c_int;
    pub fn cs_cnn_set_dataset(net: *mut cs_cnn, count: c_int, inputs: *const f32, targets: *const f32,
                              weights: *const f32) -> c_int;
    pub fn cs_cnn_train_epoch(net: *mut cs_cnn, order: *const i32, count: c_int, lr: f32, mean_loss: *mut f32)
        -> c_int;
    pub fn cs_cnn_time_step(net: *mut cs_cnn, iters: c_int, ms: *mut f32, flops: *mut f64) -> c_int;
    pub fn cs_gemm_bf16(ctx: *mut cs_ctx, m: c_int, n: c_int, k: c_int, a: *const u16, b: *const u16, c: *mut f32,
                        ms: *mut f32) -> c_int;

    // measurement helpers
    pub fn cs_bench_copy(ctx: *mut cs_ctx, bytes: i64, iters: c_int, ms_per_copy: *mut f32) -> c_int;
    pub fn cs_check_reciprocal(ctx: *mut cs_ctx, mismatches: *mut u64) -> c_int;
    pub fn cs_set_cells_per_thread(cells_per_thread: c_int) -> c_int;
    pub fn cs_set_step_graphs(enable: c_int) -> c_int;
    pub fn cs_set_temporal_steps(k: c_int) -> c_int;
}
