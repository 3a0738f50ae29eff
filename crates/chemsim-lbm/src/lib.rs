//! chemsim's lattice-Boltzmann hot path on an MI355X: drop-in replacements of
//! the reference crate's `lbm` and `matrix` modules (`src/lbm.rs`,
//! `src/matrix.rs`) and the render/record surface a display loop uses
//! (`display`), all over `libchemsim_hip.so` through `chemsim-hip-sys`.
pub mod display;
pub mod lbm;
pub mod matrix;
