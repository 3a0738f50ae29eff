//! Render and frame recording of an `lbm::State` on the GPU: what `main.rs`'s
//! `LBMSim::render` (`main.rs:138-177`, through `render.rs:7-178`) and
//! `display::record` (`display.rs:157-185`) do with the lattice, through
//! `cs_lattice_render` / `cs_lattice_record`.  Only the image bytes leave the
//! device.

use std::os::raw::{c_int, c_void};
use std::panic::{catch_unwind, resume_unwind, AssertUnwindSafe};

use chemsim_hip_sys as sys;

use crate::lbm::{Scalar, State, D2Q9};

/// `main.rs:36-42`: the field shown.
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub enum DisplayMode {
    Density,
    Speed,
    Velocity,
    MomentumDensity,
}

impl DisplayMode {
    /// `DisplayMode::next` (`main.rs:44-53`).
    pub fn next(&self) -> DisplayMode {
        match self {
            DisplayMode::Density => DisplayMode::Speed,
            DisplayMode::Speed => DisplayMode::Velocity,
            DisplayMode::Velocity => DisplayMode::MomentumDensity,
            DisplayMode::MomentumDensity => DisplayMode::Density,
        }
    }
    fn code(&self) -> c_int {
        match self {
            DisplayMode::Density => sys::CS_RENDER_DENSITY,
            DisplayMode::Speed => sys::CS_RENDER_SPEED,
            DisplayMode::Velocity => sys::CS_RENDER_VELOCITY,
            DisplayMode::MomentumDensity => sys::CS_RENDER_MOMENTUM,
        }
    }
}

fn check(rc: c_int) {
    if rc != sys::CS_OK {
        // SAFETY: cs_last_error returns a NUL-terminated thread-local buffer.
        let msg = unsafe { std::ffi::CStr::from_ptr(sys::cs_last_error()) };
        panic!("chemsim: {}", msg.to_string_lossy());
    }
}

/// `LBMSim::render` (`main.rs:138-177`): `render_scalar_field` of density or
/// speed, or `render_vector_field` of velocity or momentum density, then (if
/// `geometry`) `render_geometry`.  h x w RGB pixels, pixel (x, y) at
/// 3 (y w + x).
pub fn render(state: &State<D2Q9>, mode: DisplayMode, geometry: bool) -> Vec<u8> {
    let (w, h) = state.size();
    let mut rgb = vec![0u8; 3 * w * h];
    let lat = state.device_view();
    // SAFETY: rgb holds 3*h*w bytes; lat is the state's live lattice.
    check(unsafe { sys::cs_lattice_render(lat, mode.code(), geometry as c_int, rgb.as_mut_ptr(), std::ptr::null_mut()) });
    rgb
}

struct Sink<'a> {
    write_frame: &'a mut dyn FnMut(i64, &[u8], usize, usize) -> std::io::Result<()>,
    error: Option<std::io::Error>,
    panic: Option<Box<dyn std::any::Any + Send>>,
}

unsafe extern "C" fn deliver(user: *mut c_void, frame: i64, rgba: *const u8, w: c_int, h: c_int) -> c_int {
    // SAFETY: user is the &mut Sink passed to cs_lattice_record below; rgba
    // holds 4*w*h bytes during this call.
    let sink = &mut *(user as *mut Sink);
    let pixels = std::slice::from_raw_parts(rgba, 4 * (w as usize) * (h as usize));
    // no unwinding across the C ABI
    match catch_unwind(AssertUnwindSafe(|| (sink.write_frame)(frame, pixels, w as usize, h as usize))) {
        Ok(Ok(())) => 0,
        Ok(Err(e)) => {
            sink.error = Some(e);
            1
        }
        Err(p) => {
            sink.panic = Some(p);
            1
        }
    }
}

/// `record` (`display.rs:157-185`): `frames` times, `LBMSim::step`'s
/// speed-factor loop of `steps_per_frame` steps (`main.rs:128-136`) then the
/// render of `mode` as RGBA, handed to `write_frame` (the encoder) in frame
/// order.  A ring of `depth` frames overlaps the copies and `write_frame` with
/// the next frames' steps (1: the reference's order).  The state's time
/// advances by the steps taken; an error of `write_frame` stops the recording
/// and is returned once the frames in flight are done.
pub fn record(
    state: &mut State<D2Q9>,
    frames: usize,
    steps_per_frame: usize,
    mode: DisplayMode,
    geometry: bool,
    depth: usize,
    write_frame: &mut dyn FnMut(i64, &[u8], usize, usize) -> std::io::Result<()>,
) -> std::io::Result<usize> {
    let lat = state.device_lattice();
    let (mut t0, mut t1) = (0.0f64, 0.0f64);
    let mut sink = Sink { write_frame, error: None, panic: None };
    let mut stats = sys::cs_record_stats::default();
    // SAFETY: lat is the state's live lattice; sink outlives the call.
    let rc = unsafe {
        check(sys::cs_lattice_time(lat, &mut t0));
        let rc = sys::cs_lattice_record(lat, frames as c_int, steps_per_frame as c_int, mode.code(),
                                        geometry as c_int, depth as c_int, Some(deliver),
                                        &mut sink as *mut Sink as *mut c_void, &mut stats);
        check(sys::cs_lattice_time(lat, &mut t1));
        rc
    };
    state.time += (t1 - t0) as Scalar;
    if let Some(p) = sink.panic.take() {
        resume_unwind(p);
    }
    if let Some(e) = sink.error.take() {
        return Err(e);
    }
    check(rc);
    Ok(stats.frames as usize)
}
