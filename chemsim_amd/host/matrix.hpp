// Host matrix of the C++ drop-in layer: the operations of matrix.rs that the
// callers of the lattice (main.rs, render.rs) use on the fields they read
// back -- element-wise arithmetic, reductions and shape queries -- over a
// row-major h x w float buffer (the host layout of matrix.rs:120-126,
// element (x, y) at [y * w + x]).  The lattice itself never computes through
// this type: its populations and moments live on the GPU behind the C ABI.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <stdexcept>
#include <utility>
#include <vector>

namespace chemsim {
namespace matrix {

// matrix::Error (matrix.rs:17-21): the slice given to Matrix::new_ had the
// wrong size.
struct Error : std::invalid_argument {
    Error() : std::invalid_argument("matrix: slice size does not match the dimensions (InvalidSliceSize)") {}
};

class Matrix {
public:
    Matrix() = default;

    // Matrix::new (matrix.rs:24-30): `slice` is row-major with dims = (w, h).
    static Matrix new_(const std::vector<float>& slice, std::pair<size_t, size_t> dims)
    {
        if (slice.size() != dims.first * dims.second) throw Error();
        return Matrix(dims.first, dims.second, slice);
    }
    // Matrix::new_filled (matrix.rs:40-44), dims = (w, h).  The reference
    // builds this one with transposed dimensions (a non-square lattice then
    // fails in Lattice::density); here every constructor takes (w, h).
    static Matrix new_filled(float value, std::pair<size_t, size_t> dims)
    {
        return Matrix(dims.first, dims.second, std::vector<float>(dims.first * dims.second, value));
    }

    size_t get_width() const { return w_; }
    size_t get_height() const { return h_; }
    std::pair<size_t, size_t> get_shape() const { return {w_, h_}; }
    // matrix.rs:120-126: row-major copy of the elements.
    std::vector<float> get_underlying() const { return v_; }
    const std::vector<float>& data() const { return v_; }
    std::vector<float>& data() { return v_; }
    float at(size_t x, size_t y) const { return v_[y * w_ + x]; }
    float& at(size_t x, size_t y) { return v_[y * w_ + x]; }

    Matrix recip() const { return map([](float a) { return 1.0f / a; }); }
    Matrix sqrt() const { return map([](float a) { return std::sqrt(a); }); }
    Matrix scale(float s) const { return map([s](float a) { return a * s; }); }
    Matrix shift(float s) const { return map([s](float a) { return a + s; }); }
    Matrix clamp(float lo, float hi) const { return map([lo, hi](float a) { return std::min(std::max(a, lo), hi); }); }
    Matrix logistic() const { return map([](float a) { return 1.0f / (1.0f + std::exp(-a)); }); }
    Matrix hadamard(const Matrix& r) const { return zip(r, [](float a, float b) { return a * b; }); }
    Matrix divide(const Matrix& r) const { return zip(r, [](float a, float b) { return a / b; }); }
    double sum() const
    {
        double s = 0;
        for (float a : v_) s += a;
        return s;
    }
    double maximum_real() const { return v_.empty() ? 0.0 : *std::max_element(v_.begin(), v_.end()); }

    Matrix operator+(const Matrix& r) const { return zip(r, [](float a, float b) { return a + b; }); }
    Matrix operator-(const Matrix& r) const { return zip(r, [](float a, float b) { return a - b; }); }
    Matrix& operator+=(const Matrix& r) { return *this = *this + r; }

private:
    Matrix(size_t w, size_t h, std::vector<float> v) : w_(w), h_(h), v_(std::move(v)) {}
    template <class F>
    Matrix map(F f) const
    {
        Matrix m = *this;
        for (float& a : m.v_) a = f(a);
        return m;
    }
    template <class F>
    Matrix zip(const Matrix& r, F f) const
    {
        if (r.w_ != w_ || r.h_ != h_) throw std::invalid_argument("matrix: shape mismatch");
        Matrix m = *this;
        for (size_t i = 0; i < v_.size(); ++i) m.v_[i] = f(v_[i], r.v_[i]);
        return m;
    }
    size_t w_ = 0, h_ = 0;
    std::vector<float> v_;
};

}  // namespace matrix
}  // namespace chemsim
