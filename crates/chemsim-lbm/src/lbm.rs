//! Drop-in replacement of chemsim's `src/lbm.rs` with the populations on an
//! MI355X: the same public names and signatures -- `Scalar`, `Vector`,
//! `compute_equilibrium`, `Discretization`, `Direction`, `Geometry`,
//! `Population(s)`, the `Lattice` trait, `D2Q9`, the `CollisionOperator`
//! plugin trait, `BGK`, `TRT`, `KBC`, `Regularized` and `State` with
//! `initial`, `step`, `stream`, `bounce_back`, `collide` and every moment
//! accessor -- over `libchemsim_hip.so` (crate `chemsim-hip-sys`) instead of
//! the arrayfire crate.
//!
//! `State::step` runs one fused HIP pull kernel per step (stream, bounce-back
//! and collision, `lbm.rs:694-714`) when the operator is one the device
//! implements (`BGK`, `TRT`, `Regularized`, `KBC`); with a user operator it
//! runs the three phases, the collision evaluated on host populations.  The
//! populations stay on the device; `populations()` downloads them on demand
//! and `populations_mut()` edits reach the device before the next device pass.
//!
//! Errors: arrayfire-rust panics when an ArrayFire call fails; so does every
//! failing C-ABI call here, with `cs_last_error()`.
//!
//! Deviations from the reference, all additive or forced by the device:
//! `Geometry` is a host mask type (solid cells, `main.rs:269-312`) instead of
//! `af::Array<bool>`; `Matrix` holds host data; `State` is implemented for
//! `D2Q9` (the only `Lattice` of the reference); `CollisionOperator` has a
//! provided `device_params` method through which an operator selects the
//! device step; the per-phase `println!` timers of `State::step` are not made.

use std::cell::{Cell, OnceCell};
use std::ffi::CStr;
use std::os::raw::c_int;
use std::sync::atomic::{AtomicI32, Ordering};
use std::sync::OnceLock;

use chemsim_hip_sys as sys;

pub use crate::matrix::Matrix;

/// `lbm.rs:13`.
pub type Scalar = f32;

// ---- errors and the device -------------------------------------------------

fn last_error() -> String {
    // SAFETY: cs_last_error returns a NUL-terminated thread-local buffer.
    unsafe { CStr::from_ptr(sys::cs_last_error()).to_string_lossy().into_owned() }
}

fn check(rc: c_int) {
    if rc != sys::CS_OK {
        panic!("chemsim: {}", last_error());
    }
}

struct Ctx(*mut sys::cs_ctx);
// SAFETY: the context handle is immutable after creation and the library
// allows its use from any thread.
unsafe impl Send for Ctx {}
unsafe impl Sync for Ctx {}

static DEVICE_INDEX: AtomicI32 = AtomicI32::new(0);
static CONTEXT: OnceLock<Ctx> = OnceLock::new();

/// Picks the GPU (`af::set_device`, `main.rs:331`); call before the first
/// lattice is created.  The context lives until the process ends.
pub fn set_device(index: i32) {
    DEVICE_INDEX.store(index, Ordering::SeqCst);
}

fn context() -> *mut sys::cs_ctx {
    CONTEXT
        .get_or_init(|| {
            let mut c = std::ptr::null_mut();
            // SAFETY: c is a valid out-pointer.
            check(unsafe { sys::cs_ctx_create(DEVICE_INDEX.load(Ordering::SeqCst), &mut c) });
            Ctx(c)
        })
        .0
}

/// One device lattice, destroyed with its owner.
struct Handle(*mut sys::cs_lattice);

impl Handle {
    fn new(size: (usize, usize), storage: c_int) -> Self {
        let mut p = std::ptr::null_mut();
        // SAFETY: p is a valid out-pointer; the context outlives every lattice.
        check(unsafe { sys::cs_lattice_create(context(), size.0 as c_int, size.1 as c_int, storage, &mut p) });
        Handle(p)
    }
}

impl Drop for Handle {
    fn drop(&mut self) {
        // SAFETY: created by cs_lattice_create, destroyed once.
        unsafe {
            sys::cs_lattice_destroy(self.0);
        }
    }
}

fn params_for(disc: &Discretization, tau: Scalar) -> sys::cs_params {
    sys::cs_params {
        tau: tau as f64,
        delta_x: disc.delta_x as f64,
        delta_t: disc.delta_t as f64,
        edge: sys::CS_EDGE_ZERO,
        boundary: sys::CS_BB_REFERENCE,
        outlet_rho: 1.0,
        rho0: 1.0,
        collision: sys::CS_COLL_BGK,
        tau_minus: tau as f64,
        ..Default::default()
    }
}

// ---- small types (lbm.rs:17-95) --------------------------------------------

#[derive(PartialEq, PartialOrd, Debug, Clone, Copy)]
pub struct Vector(pub Scalar, pub Scalar);

impl Vector {
    #[inline(always)]
    pub fn to_pair(&self) -> (Scalar, Scalar) {
        (self.0, self.1)
    }
}

impl std::ops::Add for Vector {
    type Output = Vector;
    #[inline(always)]
    fn add(self, rhs: Vector) -> Vector {
        Vector(self.0 + rhs.0, self.1 + rhs.1)
    }
}

#[derive(PartialEq, PartialOrd, Debug, Clone, Copy)]
pub struct Discretization {
    pub delta_x: Scalar,
    pub delta_t: Scalar,
}

impl Discretization {
    /// `lbm.rs:81-86`: dx / (sqrt(3) dt) in f32.
    #[inline(always)]
    pub fn isothermal_speed_of_sound(&self) -> Scalar {
        self.delta_x / (Scalar::sqrt(3.0) * self.delta_t)
    }
}

/// A lattice direction: weight, velocity and the 3x3 one-hot streaming
/// stencil (`lbm.rs:90-95, 202-282`).  The stencil's 1 sits at row 1 + cy,
/// column 1 - cx (row-major), which fixes the streaming orientation
/// f_i(r, c) <- f_i(r + cx, c - cy).
#[derive(Clone, Debug)]
pub struct Direction {
    pub w_scalar: Scalar,
    pub c_vector: Vector,
    pub stencil: Matrix,
}

const CX: [i32; 9] = [0, 1, 0, -1, 0, 1, -1, -1, 1];
const CY: [i32; 9] = [0, 0, 1, 0, -1, 1, 1, -1, -1];
const OPP: [usize; 9] = [0, 3, 4, 1, 2, 7, 8, 5, 6];

/// Solid mask (`lbm.rs:99`; built as in `main.rs:269-312`, edited between
/// steps as in `main.rs:77-89`), h x w, element (x, y) at `[y * w + x]`.
/// Every edit bumps a version, so a `State` uploads it before its next step.
#[derive(Clone, Debug)]
pub struct Geometry {
    w: usize,
    h: usize,
    cells: Vec<u8>,
    version: u64,
}

impl Geometry {
    pub fn new(size: (usize, usize)) -> Self {
        Geometry { w: size.0, h: size.1, cells: vec![0; size.0 * size.1], version: 1 }
    }
    pub fn from_row_major(solid: &[bool], size: (usize, usize)) -> Self {
        assert_eq!(solid.len(), size.0 * size.1, "lbm: geometry size");
        let mut g = Geometry::new(size);
        for (c, &s) in g.cells.iter_mut().zip(solid) {
            *c = if s { sys::CS_WALL } else { sys::CS_FLUID };
        }
        g
    }
    pub fn get(&self, x: usize, y: usize) -> bool {
        self.cells[y * self.w + x] != 0
    }
    pub fn set(&mut self, x: usize, y: usize, solid: bool) {
        self.cells[y * self.w + x] = if solid { sys::CS_WALL } else { sys::CS_FLUID };
        self.version += 1;
    }
    pub fn size(&self) -> (usize, usize) {
        (self.w, self.h)
    }
    pub fn version(&self) -> u64 {
        self.version
    }
    fn flags(&self) -> &[u8] {
        &self.cells
    }
}

pub type Population = Matrix;
pub type Populations = Vec<(Direction, Population)>;

fn directions() -> Vec<Direction> {
    // RN(16/36), RN(4/36), RN(1/36) in f32, as lbm.rs:209-219 evaluates them
    let den: Scalar = 36.0;
    (0..9)
        .map(|i| {
            let mut st = vec![0.0f32; 9];
            st[(3 * (1 + CY[i]) + (1 - CX[i])) as usize] = 1.0;
            Direction {
                w_scalar: if i == 0 { 16.0 / den } else if i < 5 { 4.0 / den } else { 1.0 / den },
                c_vector: Vector(CX[i] as Scalar, CY[i] as Scalar),
                stencil: Matrix::new(&st, (3, 3)).expect("3x3 stencil"),
            }
        })
        .collect()
}

fn flatten(p: &Populations, size: (usize, usize)) -> Vec<f32> {
    assert_eq!(p.len(), 9, "lbm: D2Q9 needs 9 populations");
    let mut soa = Vec::with_capacity(9 * size.0 * size.1);
    for (_, pop) in p {
        assert_eq!(pop.get_shape(), size, "lbm: population shape");
        soa.extend_from_slice(pop.as_slice());
    }
    soa
}

fn wrap(soa: Vec<f32>, size: (usize, usize)) -> Populations {
    let n = size.0 * size.1;
    directions()
        .into_iter()
        .enumerate()
        .map(|(i, d)| (d, Matrix::from_vec(soa[i * n..(i + 1) * n].to_vec(), size)))
        .collect()
}

/// `compute_equilibrium` (`lbm.rs:43-71`) on the GPU in the reference's f32
/// operation order; `directions` must be the D2Q9 directions.
pub fn compute_equilibrium(
    density: Matrix,
    velocity: (Matrix, Matrix),
    directions: &[Direction],
    discretization: Discretization,
) -> Populations {
    let size = density.get_shape();
    let (vx, vy) = velocity;
    assert_eq!(size, vx.get_shape());
    assert_eq!(size, vy.get_shape());
    assert_eq!(directions.len(), 9, "lbm: compute_equilibrium needs the D2Q9 directions");
    let lat = Handle::new(size, sys::CS_F32);
    let p = params_for(&discretization, 1.0);
    let mut soa = vec![0.0f32; 9 * size.0 * size.1];
    // SAFETY: every host array holds h*w (or 9*h*w) floats.
    unsafe {
        check(sys::cs_lattice_set_params(lat.0, &p));
        check(sys::cs_lattice_init_equilibrium(lat.0, density.as_slice().as_ptr(), vx.as_slice().as_ptr(),
                                               vy.as_slice().as_ptr()));
        check(sys::cs_lattice_download_f32(lat.0, soa.as_mut_ptr()));
    }
    wrap(soa, size)
}

// ---- Lattice (lbm.rs:111-176) ----------------------------------------------

pub trait Lattice {
    fn size(&self) -> (usize, usize);
    fn populations(&self) -> &Populations;
    fn populations_mut(&mut self) -> &mut Populations;
    fn swap_populations(&self) -> Populations;

    /// `lbm.rs:117-121` on host populations (D2Q9 computes it on the device).
    fn density(&self) -> Matrix {
        let mut result = Matrix::new_filled(0.0, self.size());
        for (_, pop) in self.populations() {
            result += pop.clone();
        }
        result
    }
    fn momentum_density(&self) -> (Matrix, Matrix) {
        let mut md_x = Matrix::new_filled(0.0, self.size());
        let mut md_y = Matrix::new_filled(0.0, self.size());
        for (dir, f_i) in self.populations() {
            md_x = md_x + f_i.scale(dir.c_vector.0);
            md_y = md_y + f_i.scale(dir.c_vector.1);
        }
        (md_x, md_y)
    }
    fn velocity(&self) -> (Matrix, Matrix) {
        let inverse_density = self.density().recip();
        let (md_x, md_y) = self.momentum_density();
        (inverse_density.hadamard(&md_x), inverse_density.hadamard(&md_y))
    }
    fn speed(&self) -> Matrix {
        let (v_x, v_y) = self.velocity();
        (v_x.hadamard(&v_x) + v_y.hadamard(&v_y)).sqrt()
    }
    fn equilibrium(&self, disc: &Discretization) -> Populations {
        let dirs: Vec<Direction> = self.populations().iter().map(|(d, _)| d.clone()).collect();
        compute_equilibrium(self.density(), self.velocity(), &dirs, *disc)
    }
    fn non_equilibrium(&self, disc: &Discretization) -> Populations {
        let f_eq = self.equilibrium(disc);
        self.populations().iter().zip(f_eq).map(|((d, f), (_, e))| (d.clone(), f - &e)).collect()
    }
    fn swap_equilibrium(&self, disc: &Discretization) -> Populations;
}

/// D2Q9 with device-resident populations (`lbm.rs:180-325`).
pub struct D2Q9 {
    size: (usize, usize),
    lat: Handle,
    host: OnceCell<Populations>, // host copy, downloaded on demand
    host_dirty: Cell<bool>,      // edited through populations_mut since the last upload
    params: Cell<Option<sys::cs_params>>,
}

impl D2Q9 {
    /// `D2Q9::new` (`lbm.rs:187-200`): nine equally shaped populations in
    /// direction order, FP32 storage (the reference's).
    pub fn new(populations: &[Population]) -> Self {
        Self::with_storage(populations, sys::CS_F32)
    }

    /// The same with another device storage (`sys::CS_F64`, the deviation
    /// storages `sys::CS_F32DEV`, `sys::CS_BF16DEV`, `sys::CS_F16DEV`).
    pub fn with_storage(populations: &[Population], storage: c_int) -> Self {
        assert!(populations.len() == 9);
        let size = populations[0].get_shape();
        for pop in populations {
            assert_eq!(size, pop.get_shape());
        }
        let host: Populations = directions().into_iter().zip(populations.iter().cloned()).collect();
        let cell = OnceCell::new();
        let _ = cell.set(host);
        D2Q9 {
            size,
            lat: Handle::new(size, storage),
            host: cell,
            host_dirty: Cell::new(true),
            params: Cell::new(None),
        }
    }

    pub fn directions() -> Vec<Direction> {
        directions()
    }

    /// The device lattice with the host edits uploaded first.
    pub fn handle(&self) -> *mut sys::cs_lattice {
        if self.host_dirty.get() {
            let soa = flatten(self.host.get().expect("dirty host copy"), self.size);
            // SAFETY: soa holds 9*h*w floats.
            check(unsafe { sys::cs_lattice_upload_f32(self.lat.0, soa.as_ptr()) });
            self.host_dirty.set(false);
        }
        self.lat.0
    }

    /// The device is about to change the populations: the host copy goes stale.
    pub fn invalidate_host(&mut self) {
        debug_assert!(!self.host_dirty.get(), "host edits must be uploaded (handle()) first");
        self.host = OnceCell::new();
    }

    pub fn set_params(&self, p: &sys::cs_params) {
        if self.params.get() != Some(*p) {
            // SAFETY: p points to a valid cs_params.
            check(unsafe { sys::cs_lattice_set_params(self.lat.0, p) });
            self.params.set(Some(*p));
        }
    }

    fn field(&self, f: unsafe extern "C" fn(*mut sys::cs_lattice, *mut f32) -> c_int) -> Matrix {
        let mut m = Matrix::new_filled(0.0, self.size);
        // SAFETY: m holds h*w floats.
        check(unsafe { f(self.handle(), m.as_mut_slice().as_mut_ptr()) });
        m
    }

    fn fields(&self, f: unsafe extern "C" fn(*mut sys::cs_lattice, *mut f32, *mut f32) -> c_int)
              -> (Matrix, Matrix) {
        let mut a = Matrix::new_filled(0.0, self.size);
        let mut b = Matrix::new_filled(0.0, self.size);
        // SAFETY: a and b hold h*w floats each.
        check(unsafe { f(self.handle(), a.as_mut_slice().as_mut_ptr(), b.as_mut_slice().as_mut_ptr()) });
        (a, b)
    }

    fn pops_of(&self, disc: &Discretization,
               f: unsafe extern "C" fn(*mut sys::cs_lattice, *mut f32) -> c_int) -> Populations {
        let mut p = self.params.get().unwrap_or_else(|| params_for(disc, 1.0));
        p.delta_x = disc.delta_x as f64;
        p.delta_t = disc.delta_t as f64;
        self.set_params(&p);
        let mut soa = vec![0.0f32; 9 * self.size.0 * self.size.1];
        // SAFETY: soa holds 9*h*w floats.
        check(unsafe { f(self.handle(), soa.as_mut_ptr()) });
        wrap(soa, self.size)
    }

    /// `State::pressure` (`lbm.rs:784-787`): rho cs^2 with this discretization.
    pub fn pressure(&self, disc: &Discretization) -> Matrix {
        let mut p = self.params.get().unwrap_or_else(|| params_for(disc, 1.0));
        p.delta_x = disc.delta_x as f64;
        p.delta_t = disc.delta_t as f64;
        self.set_params(&p);
        self.field(sys::cs_lattice_pressure)
    }
}

impl Lattice for D2Q9 {
    fn size(&self) -> (usize, usize) {
        self.size
    }
    fn populations(&self) -> &Populations {
        self.host.get_or_init(|| {
            let mut soa = vec![0.0f32; 9 * self.size.0 * self.size.1];
            // SAFETY: soa holds 9*h*w floats.
            check(unsafe { sys::cs_lattice_download_f32(self.lat.0, soa.as_mut_ptr()) });
            wrap(soa, self.size)
        })
    }
    fn populations_mut(&mut self) -> &mut Populations {
        let _ = self.populations();
        self.host_dirty.set(true);
        self.host.get_mut().expect("host copy")
    }
    /// `lbm.rs:298-309`: direction i carries population opp(i).
    fn swap_populations(&self) -> Populations {
        let p = self.populations();
        (0..9).map(|i| (p[i].0.clone(), p[OPP[i]].1.clone())).collect()
    }
    fn density(&self) -> Matrix {
        self.field(sys::cs_lattice_density)
    }
    fn momentum_density(&self) -> (Matrix, Matrix) {
        self.fields(sys::cs_lattice_momentum_density)
    }
    fn velocity(&self) -> (Matrix, Matrix) {
        self.fields(sys::cs_lattice_velocity)
    }
    fn speed(&self) -> Matrix {
        self.field(sys::cs_lattice_speed)
    }
    fn equilibrium(&self, disc: &Discretization) -> Populations {
        self.pops_of(disc, sys::cs_lattice_equilibrium)
    }
    fn non_equilibrium(&self, disc: &Discretization) -> Populations {
        self.pops_of(disc, sys::cs_lattice_non_equilibrium)
    }
    /// `lbm.rs:311-323`: the equilibrium with directions 1..8 replaced by the
    /// opposite populations (as the reference's TRT uses it).
    fn swap_equilibrium(&self, disc: &Discretization) -> Populations {
        let mut out = self.equilibrium(disc);
        let p = self.populations();
        for i in 1..9 {
            out[i].1 = p[OPP[i]].1.clone();
        }
        out
    }
}

// ---- CollisionOperator (lbm.rs:327-341) --------------------------------------

pub trait CollisionOperator<L> {
    fn evaluate(&self, lattice: &L, equilibrium: &Populations, discretization: &Discretization) -> Populations;

    fn kinematic_shear_viscosity(&self, disc: &Discretization) -> Scalar;

    #[inline(always)]
    fn kinematic_bulk_viscosity(&self, disc: &Discretization) -> Scalar {
        2.0 * self.kinematic_shear_viscosity(disc) / 3.0
    }

    /// Operators the device step implements fill in their parameters and
    /// return true; `State` then runs the fused device step.
    fn device_params(&self, _params: &mut sys::cs_params) -> bool {
        false
    }
}

/// One collision of `lattice`'s populations on a scratch device lattice with
/// the operator's parameters (the `evaluate` of the device operators).
fn collide_on_device(params: sys::cs_params, lattice: &D2Q9) -> Populations {
    let size = lattice.size();
    let tmp = Handle::new(size, sys::CS_F32);
    let f = flatten(lattice.populations(), size);
    let mut out = vec![0.0f32; f.len()];
    // SAFETY: f and out hold 9*h*w floats.
    unsafe {
        check(sys::cs_lattice_set_params(tmp.0, &params));
        check(sys::cs_lattice_upload_f32(tmp.0, f.as_ptr()));
        check(sys::cs_lattice_collide(tmp.0));
        check(sys::cs_lattice_download_f32(tmp.0, out.as_mut_ptr()));
    }
    wrap(out, size)
}

/// `BGK` (`lbm.rs:345-370`).
pub struct BGK {
    pub tau: Scalar,
}

impl<L: Lattice> CollisionOperator<L> for BGK {
    /// f + (f - f_eq) * (-dt / tau) on the GPU over the given populations.
    fn evaluate(&self, lattice: &L, equilibrium: &Populations, discretization: &Discretization) -> Populations {
        let size = lattice.size();
        let f = flatten(lattice.populations(), size);
        let fe = flatten(equilibrium, size);
        let mut out = vec![0.0f32; f.len()];
        // SAFETY: f, fe and out hold the same number of floats.
        check(unsafe {
            sys::cs_bgk_evaluate(context(), f.len() as i64, f.as_ptr(), fe.as_ptr(), self.tau,
                                 discretization.delta_t, out.as_mut_ptr())
        });
        wrap(out, size)
    }
    /// `lbm.rs:366-369`.
    fn kinematic_shear_viscosity(&self, disc: &Discretization) -> Scalar {
        let (dx, dt) = (disc.delta_x, disc.delta_t);
        (dx * dx / (3.0 * dt * dt)) * (self.tau - dt / 2.0)
    }
    fn device_params(&self, p: &mut sys::cs_params) -> bool {
        p.collision = sys::CS_COLL_BGK;
        p.tau = self.tau as f64;
        p.tau_minus = self.tau as f64;
        true
    }
}

/// `TRT` (`lbm.rs:374-451`), run by the device step with the reference's
/// arithmetic (its `swap_equilibrium` included).
pub struct TRT {
    pub tau_minus: Scalar,
    pub tau_plus: Scalar,
}

impl TRT {
    /// `TRT::new` (`lbm.rs:380-390`) in f32.
    pub fn new(lambda: Scalar, ks_viscosity: Scalar, disc: &Discretization) -> Self {
        let (mut tp, mut tm) = (0.0f32, 0.0f32);
        // SAFETY: valid out-pointers.
        check(unsafe { sys::cs_trt_new(lambda, ks_viscosity, disc.delta_x, disc.delta_t, &mut tp, &mut tm) });
        TRT { tau_minus: tm, tau_plus: tp }
    }
    /// `lbm.rs:392-399`.
    pub fn lambda(&self, disc: &Discretization) -> Scalar {
        let dt = disc.delta_t;
        let mut result = 1.0;
        result *= (self.tau_plus / dt) - 0.5;
        result *= (self.tau_minus / dt) - 0.5;
        result
    }
}

impl CollisionOperator<D2Q9> for TRT {
    /// The TRT collide kernel on the lattice's populations (the passed
    /// equilibrium is unused, as in `lbm.rs:411-412`).
    fn evaluate(&self, lattice: &D2Q9, _eq: &Populations, disc: &Discretization) -> Populations {
        let mut p = params_for(disc, 1.0);
        CollisionOperator::<D2Q9>::device_params(self, &mut p);
        collide_on_device(p, lattice)
    }
    /// `lbm.rs:446-450`.
    fn kinematic_shear_viscosity(&self, disc: &Discretization) -> Scalar {
        let cs = disc.isothermal_speed_of_sound();
        cs * cs * (self.tau_plus / disc.delta_t - 0.5)
    }
    fn device_params(&self, p: &mut sys::cs_params) -> bool {
        p.collision = sys::CS_COLL_TRT;
        p.tau = self.tau_plus as f64;
        p.tau_minus = self.tau_minus as f64;
        true
    }
}

/// `KBC` (`lbm.rs:455-590`): entropic multi-relaxation; its diagnostic print
/// of an entropy sum (`lbm.rs:574-582`) is not made.
pub struct KBC {
    ks_viscosity: Scalar,
}

impl KBC {
    /// `KBC::new` (`lbm.rs:462-466`).
    pub fn new(ks_viscosity: Scalar) -> Self {
        KBC { ks_viscosity }
    }
}

impl CollisionOperator<D2Q9> for KBC {
    fn evaluate(&self, lattice: &D2Q9, _eq: &Populations, disc: &Discretization) -> Populations {
        let mut p = params_for(disc, 1.0);
        CollisionOperator::<D2Q9>::device_params(self, &mut p);
        collide_on_device(p, lattice)
    }
    fn kinematic_shear_viscosity(&self, _disc: &Discretization) -> Scalar {
        self.ks_viscosity
    }
    fn device_params(&self, p: &mut sys::cs_params) -> bool {
        p.collision = sys::CS_COLL_KBC;
        p.ks_viscosity = self.ks_viscosity as f64;
        true
    }
}

/// `Regularized<C>` (`lbm.rs:592-666`): as in the reference, `evaluate`
/// never calls the wrapped operator, which only supplies the viscosity.
pub struct Regularized<C> {
    underlying: C,
}

impl<C> Regularized<C> {
    /// `Regularized::new` (`lbm.rs:600-604`).
    pub fn new(underlying: C) -> Self {
        Regularized { underlying }
    }
}

impl<C: CollisionOperator<D2Q9>> CollisionOperator<D2Q9> for Regularized<C> {
    fn evaluate(&self, lattice: &D2Q9, _eq: &Populations, disc: &Discretization) -> Populations {
        let mut p = params_for(disc, 1.0);
        CollisionOperator::<D2Q9>::device_params(self, &mut p);
        collide_on_device(p, lattice)
    }
    fn kinematic_shear_viscosity(&self, disc: &Discretization) -> Scalar {
        self.underlying.kinematic_shear_viscosity(disc)
    }
    fn device_params(&self, p: &mut sys::cs_params) -> bool {
        p.collision = sys::CS_COLL_REGULARIZED;
        true
    }
}

// ---- State (lbm.rs:670-819) --------------------------------------------------

pub struct State<L> {
    pub time: Scalar,
    pub lattice: Box<L>,
    pub geometry: Geometry,
    pub collision: Box<dyn CollisionOperator<L>>,
    pub discretization: Discretization,
    uploaded_geometry: Cell<u64>,
}

impl State<D2Q9> {
    /// `State::initial` (`lbm.rs:679-692`).
    pub fn initial(
        lattice: Box<D2Q9>,
        geometry: Geometry,
        collision: Box<dyn CollisionOperator<D2Q9>>,
        discretization: Discretization,
    ) -> Self {
        assert_eq!(geometry.size(), lattice.size(), "lbm: geometry and lattice sizes differ");
        State { time: 0.0, lattice, geometry, collision, discretization, uploaded_geometry: Cell::new(0) }
    }

    /// stream -> bounce_back -> collide, time += dt (`lbm.rs:694-714`); one
    /// fused device pass when the device implements the operator.
    pub fn step(&mut self) {
        if self.device_operator() {
            let h = self.device_lattice();
            // SAFETY: h is this state's live lattice.
            check(unsafe { sys::cs_lattice_step(h, 1) });
        } else {
            self.stream();
            self.bounce_back();
            self.collide();
        }
        self.time += self.discretization.delta_t;
    }

    /// `lbm.rs:716-729`: f_i(r, c) <- f_i(r + cx, c - cy), zero fill.
    pub fn stream(&mut self) {
        let h = self.device_lattice();
        // SAFETY: h is this state's live lattice.
        check(unsafe { sys::cs_lattice_stream(h) });
    }

    /// `lbm.rs:731-739`: the operator over every cell with the lattice's equilibrium.
    pub fn collide(&mut self) {
        if self.device_operator() {
            let h = self.device_lattice();
            // SAFETY: h is this state's live lattice.
            check(unsafe { sys::cs_lattice_collide(h) });
        } else {
            let eq = self.equilibrium();
            let f_star = self.collision.evaluate(&*self.lattice, &eq, &self.discretization);
            *self.lattice.populations_mut() = f_star;
        }
    }

    /// `lbm.rs:741-751`: f_i <- f_opp(i) at solid cells.
    pub fn bounce_back(&mut self) {
        let h = self.device_lattice();
        // SAFETY: h is this state's live lattice.
        check(unsafe { sys::cs_lattice_bounce_back(h) });
    }

    #[inline(always)]
    pub fn size(&self) -> (usize, usize) {
        self.lattice.size()
    }
    #[inline(always)]
    pub fn delta_x(&self) -> Scalar {
        self.discretization.delta_x
    }
    #[inline(always)]
    pub fn delta_t(&self) -> Scalar {
        self.discretization.delta_t
    }
    #[inline(always)]
    pub fn populations(&self) -> &Populations {
        self.lattice.populations()
    }
    #[inline(always)]
    pub fn isothermal_speed_of_sound(&self) -> Scalar {
        self.discretization.isothermal_speed_of_sound()
    }
    pub fn density(&self) -> Matrix {
        self.lattice.density()
    }
    /// `lbm.rs:784-787`: density scaled by cs^2 (cs in f32, squared in f32).
    pub fn pressure(&self) -> Matrix {
        self.lattice.pressure(&self.discretization)
    }
    pub fn momentum_density(&self) -> (Matrix, Matrix) {
        self.lattice.momentum_density()
    }
    pub fn velocity(&self) -> (Matrix, Matrix) {
        self.lattice.velocity()
    }
    pub fn speed(&self) -> Matrix {
        self.lattice.speed()
    }
    pub fn equilibrium(&self) -> Populations {
        self.lattice.equilibrium(&self.discretization)
    }
    pub fn non_equilibrium(&self) -> Populations {
        self.lattice.non_equilibrium(&self.discretization)
    }
    /// `lbm.rs:815-818`: min over cells of f_eq[0] < 0.
    pub fn is_unstable(&self) -> bool {
        let h = self.device_view();
        let mut u: c_int = 0;
        // SAFETY: valid lattice and out-pointer.
        check(unsafe { sys::cs_lattice_is_unstable(h, &mut u) });
        u != 0
    }

    /// The device lattice with this state's parameters, geometry and host
    /// edits uploaded, for device work that only reads the populations (the
    /// render pass).  The host copy of the populations stays valid.
    pub fn device_view(&self) -> *mut sys::cs_lattice {
        self.prepare();
        self.lattice.handle()
    }

    /// The same for device work that changes the populations: the host
    /// copy is downloaded again on the next host read.
    pub fn device_lattice(&mut self) -> *mut sys::cs_lattice {
        let h = self.device_view();
        self.lattice.invalidate_host();
        h
    }

    fn device_operator(&self) -> bool {
        let mut p = sys::cs_params::default();
        self.collision.device_params(&mut p)
    }

    /// Parameters (the operator's when the device implements it, else BGK
    /// with tau = 1 for the streaming and bounce-back passes) and geometry
    /// reach the device before any device pass.
    fn prepare(&self) {
        let mut p = params_for(&self.discretization, 1.0);
        self.collision.device_params(&mut p);
        self.lattice.set_params(&p);
        if self.geometry.version() != self.uploaded_geometry.get() {
            // SAFETY: the flags hold h*w bytes.
            check(unsafe { sys::cs_lattice_set_flags(self.lattice.handle(), self.geometry.flags().as_ptr()) });
            self.uploaded_geometry.set(self.geometry.version());
        }
    }
}
