//! Host matrix of the drop-in layer: the operations of `src/matrix.rs` that
//! the lattice's callers (`main.rs`, `render.rs`) use on the fields they read
//! back, over a row-major h x w `f32` buffer (the host layout of
//! `matrix.rs:120-126`: element (x, y) at `[y * w + x]`).  The lattice never
//! computes through this type: its populations and moments live on the GPU
//! behind the C ABI, so no ArrayFire array is held here.

use std::ops::{Add, AddAssign, Sub};

/// `matrix::Error` (`matrix.rs:16-21`).
#[derive(Debug, Clone, Copy, PartialEq, Eq)]
pub enum Error {
    InvalidSliceSize,
}

pub type Result<T> = std::result::Result<T, Error>;

#[derive(Clone, Debug, PartialEq)]
pub struct Matrix {
    w: usize,
    h: usize,
    v: Vec<f32>,
}

impl Matrix {
    /// `Matrix::new` (`matrix.rs:24-30`): `slice` row-major, `dims = (w, h)`.
    pub fn new(slice: &[f32], dims: (usize, usize)) -> Result<Self> {
        if slice.len() != dims.0 * dims.1 {
            return Err(Error::InvalidSliceSize);
        }
        Ok(Matrix { w: dims.0, h: dims.1, v: slice.to_vec() })
    }

    /// `Matrix::new_filled` (`matrix.rs:40-44`), `dims = (w, h)`.  The
    /// reference builds this one with transposed dimensions (so a non-square
    /// lattice fails in `Lattice::density`); here every constructor takes
    /// `(w, h)`.
    pub fn new_filled(value: f32, dims: (usize, usize)) -> Self {
        Matrix { w: dims.0, h: dims.1, v: vec![value; dims.0 * dims.1] }
    }

    pub(crate) fn from_vec(v: Vec<f32>, dims: (usize, usize)) -> Self {
        debug_assert_eq!(v.len(), dims.0 * dims.1);
        Matrix { w: dims.0, h: dims.1, v }
    }

    pub fn get_width(&self) -> usize {
        self.w
    }
    pub fn get_height(&self) -> usize {
        self.h
    }
    /// `(w, h)` (`matrix.rs:67-71`).
    pub fn get_shape(&self) -> (usize, usize) {
        (self.w, self.h)
    }
    /// Row-major copy of the elements (`matrix.rs:120-126`).
    pub fn get_underlying(&self) -> Vec<f32> {
        self.v.clone()
    }
    pub fn as_slice(&self) -> &[f32] {
        &self.v
    }
    pub fn as_mut_slice(&mut self) -> &mut [f32] {
        &mut self.v
    }
    pub fn at(&self, x: usize, y: usize) -> f32 {
        self.v[y * self.w + x]
    }
    pub fn set(&mut self, x: usize, y: usize, value: f32) {
        self.v[y * self.w + x] = value;
    }

    pub fn transpose(&self) -> Self {
        let mut t = Vec::with_capacity(self.v.len());
        for x in 0..self.w {
            for y in 0..self.h {
                t.push(self.v[y * self.w + x]);
            }
        }
        Matrix { w: self.h, h: self.w, v: t }
    }
    pub fn recip(&self) -> Self {
        self.map(|a| 1.0 / a)
    }
    pub fn sqrt(&self) -> Self {
        self.map(f32::sqrt)
    }
    pub fn logistic(&self) -> Self {
        self.map(|a| 1.0 / (1.0 + (-a).exp()))
    }
    pub fn shift(&self, shifter: f32) -> Self {
        self.map(|a| a + shifter)
    }
    pub fn scale(&self, scalar: f32) -> Self {
        self.map(|a| a * scalar)
    }
    pub fn clamp(&self, min: f32, max: f32) -> Self {
        self.map(|a| a.max(min).min(max))
    }
    pub fn hadamard(&self, other: &Matrix) -> Self {
        self.zip(other, |a, b| a * b)
    }
    pub fn divide(&self, other: &Matrix) -> Self {
        self.zip(other, |a, b| a / b)
    }
    pub fn sum(&self) -> f64 {
        self.v.iter().map(|&a| a as f64).sum()
    }
    pub fn maximum_real(&self) -> f64 {
        self.v.iter().fold(f32::NEG_INFINITY, |m, &a| m.max(a)) as f64
    }

    fn map(&self, f: impl Fn(f32) -> f32) -> Self {
        Matrix { w: self.w, h: self.h, v: self.v.iter().map(|&a| f(a)).collect() }
    }
    fn zip(&self, other: &Matrix, f: impl Fn(f32, f32) -> f32) -> Self {
        assert_eq!(self.get_shape(), other.get_shape(), "matrix: shape mismatch");
        Matrix { w: self.w, h: self.h, v: self.v.iter().zip(&other.v).map(|(&a, &b)| f(a, b)).collect() }
    }
}

impl<'a> Add<&'a Matrix> for &'a Matrix {
    type Output = Matrix;
    fn add(self, rhs: &'a Matrix) -> Matrix {
        self.zip(rhs, |a, b| a + b)
    }
}
impl Add for Matrix {
    type Output = Matrix;
    fn add(self, rhs: Matrix) -> Matrix {
        &self + &rhs
    }
}
impl<'a> Sub<&'a Matrix> for &'a Matrix {
    type Output = Matrix;
    fn sub(self, rhs: &'a Matrix) -> Matrix {
        self.zip(rhs, |a, b| a - b)
    }
}
impl Sub for Matrix {
    type Output = Matrix;
    fn sub(self, rhs: Matrix) -> Matrix {
        &self - &rhs
    }
}
impl AddAssign for Matrix {
    fn add_assign(&mut self, rhs: Matrix) {
        *self = &*self + &rhs;
    }
}
