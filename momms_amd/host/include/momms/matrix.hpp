// Matrix<T>: a general-stride matrix with a stack of views, host- or
// device-resident.
//
// Behaviour of src/matrix/general_stride.rs:7-273 and src/matrix/view.rs:5-22
// (the Mat<T> and RoCM<T> traits of src/matrix/matrix.rs:62-266): column-major
// by default (rs = 1, cs = h), transpose() swaps the strides, partitioning
// nodes push a view and slide it, and a view past the edge reports padding so
// height()/width() shrink. Storage is 4096-byte aligned (src/util.rs:114-118).
//
// Additions for the GPU engine: Residence::Device keeps the buffer in HBM
// (hipMalloc), so a GPU node can run on it without host copies; get/set on a
// device matrix throw, since the host cannot dereference it.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "momms/thread_comm.hpp"

namespace momms {

// Element types the GEMM engine computes in (src/matrix/matrix.rs:15-59
// implements Scalar for f64 and f32).
template <class T>
struct is_scalar : std::false_type {};
template <>
struct is_scalar<double> : std::true_type {};
template <>
struct is_scalar<float> : std::true_type {};

// One level of the view stack (src/matrix/view.rs:5-22).
struct MatrixView {
  size_t offset = 0;
  size_t padding = 0;
  size_t iter_size = 0;

  size_t physical_size() const { return iter_size > padding ? iter_size - padding : 0; }

  // (iter_size, padding) of a blksz-wide window starting at index.
  std::pair<size_t, size_t> zoomed_size_and_padding(size_t index, size_t blksz) const {
    const size_t zoomed = std::min(blksz, iter_size - index);
    const size_t phys = physical_size();
    const long long pad = (long long)(index + zoomed) - (long long)std::max(index, phys);
    return {zoomed, (size_t)std::max(pad, 0LL)};
  }
};

enum class Residence { Host, Device };

template <class T>
class Matrix {
  static_assert(is_scalar<T>::value, "Matrix<T> holds double or float");

 public:
  Matrix() = default;
  Matrix(size_t h, size_t w, Residence where = Residence::Host) : where_(where) {
    rs_ = 1;
    cs_ = h;
    capacity_ = h * w;
    y_views_.push_back({0, 0, h});
    x_views_.push_back({0, 0, w});
    if (where_ == Residence::Host) {
      const size_t bytes = ((capacity_ * sizeof(T) + 4095) / 4096) * 4096;
      buf_ = static_cast<T*>(std::aligned_alloc(4096, bytes ? bytes : 4096));
      if (!buf_) throw std::bad_alloc();
    } else {
      if (hipMalloc(&buf_, std::max<size_t>(capacity_ * sizeof(T), 256)) != hipSuccess)
        throw std::runtime_error("Matrix: hipMalloc failed");
    }
  }
  static Matrix device(size_t h, size_t w) { return Matrix(h, w, Residence::Device); }

  Matrix(const Matrix&) = delete;
  Matrix& operator=(const Matrix&) = delete;
  Matrix(Matrix&& o) noexcept { *this = std::move(o); }
  Matrix& operator=(Matrix&& o) noexcept {
    if (this != &o) {
      release();
      alpha_ = o.alpha_;
      y_views_ = std::move(o.y_views_);
      x_views_ = std::move(o.x_views_);
      rs_ = o.rs_;
      cs_ = o.cs_;
      buf_ = o.buf_;
      capacity_ = o.capacity_;
      alias_ = o.alias_;
      where_ = o.where_;
      o.buf_ = nullptr;
      o.alias_ = true;
    }
    return *this;
  }
  ~Matrix() { release(); }

  // ---- Mat<T> --------------------------------------------------------------
  T get(size_t y, size_t x) const { return buf_[host_index(y, x)]; }
  void set(size_t y, size_t x, T v) { buf_[host_index(y, x)] = v; }

  size_t iter_height() const { return y_views_.back().iter_size; }
  size_t iter_width() const { return x_views_.back().iter_size; }
  size_t logical_h_padding() const { return y_views_.back().padding; }
  size_t logical_w_padding() const { return x_views_.back().padding; }
  size_t height() const {
    return iter_height() < logical_h_padding() ? 0 : iter_height() - logical_h_padding();
  }
  size_t width() const {
    return iter_width() < logical_w_padding() ? 0 : iter_width() - logical_w_padding();
  }

  void set_scalar(T a) { alpha_ = a; }
  T get_scalar() const { return alpha_; }

  size_t push_y_view(size_t blksz) { return push_view(y_views_, blksz); }
  size_t push_x_view(size_t blksz) { return push_view(x_views_, blksz); }
  void pop_y_view() { pop(y_views_); }
  void pop_x_view() { pop(x_views_); }
  void slide_y_view_to(size_t y, size_t blksz) { slide(y_views_, y, blksz); }
  void slide_x_view_to(size_t x, size_t blksz) { slide(x_views_, x, blksz); }

  void push_y_split(size_t start, size_t end) { push_split(y_views_, height(), start, end); }
  void push_x_split(size_t start, size_t end) { push_split(x_views_, width(), start, end); }
  void pop_y_split() { pop(y_views_); }
  void pop_x_split() { pop(x_views_); }

  // A second handle on the same storage at the current view
  // (general_stride.rs:196-213); with padding 0 rather than the reference's
  // padding = offset, which is only harmless at offset 0.
  Matrix make_alias() const {
    Matrix m;
    m.alpha_ = alpha_;
    m.y_views_ = {{y_views_.back().offset, 0, y_views_.back().iter_size}};
    m.x_views_ = {{x_views_.back().offset, 0, x_views_.back().iter_size}};
    m.rs_ = rs_;
    m.cs_ = cs_;
    m.buf_ = buf_;
    m.capacity_ = capacity_;
    m.alias_ = true;
    m.where_ = where_;
    return m;
  }
  // Thread 0's buffer becomes every thread's (general_stride.rs:215-220).
  void send_alias(const ThreadInfo& thr) {
    buf_ = static_cast<T*>(thr.broadcast(buf_));
    alias_ = true;
  }

  // ---- RoCM<T> ---------------------------------------------------------------
  size_t get_row_stride() const { return rs_; }
  size_t get_column_stride() const { return cs_; }
  T* get_buffer() const { return buf_ + y_views_.back().offset * rs_ + x_views_.back().offset * cs_; }

  void transpose() {
    if (y_views_.size() != 1 || x_views_.size() != 1)
      throw std::logic_error("can't transpose a submatrix");
    std::swap(y_views_[0], x_views_[0]);
    std::swap(rs_, cs_);
  }

  Residence residence() const { return where_; }
  bool on_device() const { return where_ == Residence::Device; }

  // ---- helpers of src/matrix/matrix.rs:162-226 -----------------------------
  // U[-1, 1) from a seeded engine (the reference's fill_rand is unseeded).
  void fill_rand(std::mt19937_64& rng) {
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (size_t x = 0; x < width(); ++x)
      for (size_t y = 0; y < height(); ++y) set(y, x, (T)u(rng));
  }
  void fill_zero() {
    for (size_t x = 0; x < width(); ++x)
      for (size_t y = 0; y < height(); ++y) set(y, x, T(0));
  }
  T frosqr() const {
    T s = 0;
    for (size_t x = 0; x < width(); ++x)
      for (size_t y = 0; y < height(); ++y) s += get(y, x) * get(y, x);
    return s;
  }
  // this := alpha*o + beta*this (matrix.rs:165-191). With beta == 0 this is
  // not read (BLAS semantics): the reference reads it, so copy_from into a
  // fresh, uninitialized matrix can carry NaN garbage through 0 * NaN.
  void axpby(T alpha, const Matrix& o, T beta) {
    if (width() != o.width() || height() != o.height())
      throw std::invalid_argument("Cannot operate on nonconformal matrices!");
    for (size_t x = 0; x < width(); ++x)
      for (size_t y = 0; y < height(); ++y)
        set(y, x, beta == T(0) ? alpha * o.get(y, x) : alpha * o.get(y, x) + beta * get(y, x));
  }
  void axpy(T alpha, const Matrix& o) { axpby(alpha, o, T(1)); }
  void copy_from(const Matrix& o) { axpby(T(1), o, T(0)); }

  size_t view_depth_y() const { return y_views_.size(); }
  size_t view_depth_x() const { return x_views_.size(); }

 private:
  size_t host_index(size_t y, size_t x) const {
    if (where_ == Residence::Device)
      throw std::logic_error("Matrix::get/set on a device-resident matrix");
    return (y + y_views_.back().offset) * rs_ + (x + x_views_.back().offset) * cs_;
  }
  static size_t push_view(std::vector<MatrixView>& v, size_t blksz) {
    const MatrixView uz = v.back();
    const auto zp = uz.zoomed_size_and_padding(0, blksz);
    v.push_back({uz.offset, zp.second, zp.first});
    return uz.iter_size;
  }
  static void pop(std::vector<MatrixView>& v) {
    if (v.size() < 2) throw std::logic_error("view stack underflow");
    v.pop_back();
  }
  static void slide(std::vector<MatrixView>& v, size_t i, size_t blksz) {
    if (v.size() < 2) throw std::logic_error("slide without a pushed view");
    const MatrixView uz = v[v.size() - 2];
    const auto zp = uz.zoomed_size_and_padding(i, blksz);
    v.back() = {uz.offset + i, zp.second, zp.first};
  }
  static void push_split(std::vector<MatrixView>& v, size_t extent, size_t start, size_t end) {
    const MatrixView uz = v.back();
    const size_t pad = end <= extent ? 0 : end - extent;
    v.push_back({uz.offset + start, pad, end - start});
  }
  void release() {
    if (buf_ && !alias_) {
      if (where_ == Residence::Host) std::free(buf_);
      else (void)hipFree(buf_);
    }
    buf_ = nullptr;
  }

  T alpha_ = T(1);
  std::vector<MatrixView> y_views_;
  std::vector<MatrixView> x_views_;
  size_t rs_ = 1, cs_ = 1;
  T* buf_ = nullptr;
  size_t capacity_ = 0;
  bool alias_ = false;
  Residence where_ = Residence::Host;
};

}  // namespace momms
