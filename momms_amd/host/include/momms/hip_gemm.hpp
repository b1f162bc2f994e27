// GPU leaves of the GemmNode API: a HipGemm node stands in for the whole
// Goto subtree (PartN ... KernelNM -> bli_dgemm_haswell_asm_6x8 of
// src/exper/l3_packing.rs:58-66) and calls the HIP engine through the C ABI
// of include/momms_hip.h.
//
// Semantics follow the reference's nodes: C := c.scalar*C + a.scalar*b.scalar*A*B
// on the current views (alpha/beta as in src/kern/kernel_nm.rs:44-45), and the
// call is synchronous: it returns with C updated. Device-resident operands are
// used in place; host-resident views are staged to HBM and C is copied back.
// Under a team of several threads (below SpawnThreads) only thread 0 launches
// and the team meets at a barrier afterwards (the pattern of
// src/composables/pack.rs:336-349).
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "momms/composables.hpp"
#include "momms/matrix.hpp"
#include "momms_hip.h"

namespace momms {

inline void abi_check(const char* what, int status) {
  if (status != 0)
    throw std::runtime_error(std::string(what) + ": " + momms_last_error_string());
}
inline void hip_check(const char* what, hipError_t e) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

namespace detail {

// Grow-only device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void* get(size_t want) {
    if (want > bytes) {
      if (p) (void)hipFree(p);
      p = nullptr;
      hip_check("hipMalloc", hipMalloc(&p, want));
      bytes = want;
    }
    return p;
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

// A device-side view of one operand: its pointer and strides in HBM.
template <class T>
struct DevView {
  T* ptr;
  int64_t rs, cs;
};

// Makes an h x w view available on the device. Device matrices pass through.
// Host views are copied into `buf` column-major (ld = h) with one 2-D copy
// when a stride is 1, or gathered on the host otherwise.
template <class T, class M>
DevView<T> stage_in(const M& x, size_t h, size_t w, DevBuf& buf, bool copy, hipStream_t st) {
  if (x.on_device())
    return {x.get_buffer(), (int64_t)x.get_row_stride(), (int64_t)x.get_column_stride()};
  T* d = static_cast<T*>(buf.get(std::max<size_t>(h * w * sizeof(T), 256)));
  if (copy && h && w) {
    const T* src = x.get_buffer();
    const size_t rs = x.get_row_stride(), cs = x.get_column_stride();
    if (rs == 1) {
      hip_check("hipMemcpy2DAsync", hipMemcpy2DAsync(d, h * sizeof(T), src, cs * sizeof(T),
                                                      h * sizeof(T), w, hipMemcpyHostToDevice, st));
    } else {
      std::vector<T> tmp(h * w);
      for (size_t j = 0; j < w; ++j)
        for (size_t i = 0; i < h; ++i) tmp[i + j * h] = src[i * rs + j * cs];
      hip_check("hipMemcpyAsync", hipMemcpyAsync(d, tmp.data(), h * w * sizeof(T),
                                                 hipMemcpyHostToDevice, st));
      hip_check("hipStreamSynchronize", hipStreamSynchronize(st));
    }
  }
  return {d, 1, (int64_t)std::max<size_t>(h, 1)};
}

template <class T, class M>
void stage_out(M& x, size_t h, size_t w, const T* d, hipStream_t st) {
  if (x.on_device() || !h || !w) return;
  T* dst = x.get_buffer();
  const size_t rs = x.get_row_stride(), cs = x.get_column_stride();
  if (rs == 1) {
    hip_check("hipMemcpy2DAsync", hipMemcpy2DAsync(dst, cs * sizeof(T), d, h * sizeof(T),
                                                    h * sizeof(T), w, hipMemcpyDeviceToHost, st));
    hip_check("hipStreamSynchronize", hipStreamSynchronize(st));
  } else {
    std::vector<T> tmp(h * w);
    hip_check("hipMemcpyAsync", hipMemcpyAsync(tmp.data(), d, h * w * sizeof(T),
                                               hipMemcpyDeviceToHost, st));
    hip_check("hipStreamSynchronize", hipStreamSynchronize(st));
    for (size_t j = 0; j < w; ++j)
      for (size_t i = 0; i < h; ++i) dst[i * rs + j * cs] = tmp[i + j * h];
  }
}

}  // namespace detail

// Plain GEMM on matrix cores: FP64 (momms_hip_dgemm) or FP32
// (momms_hip_sgemm, the counterpart of the KNM sgemm leaf,
// src/kern/knm_kernel.rs:8-81).
template <class T>
struct HipGemm {
  static_assert(std::is_same<T, double>::value || std::is_same<T, float>::value,
                "HipGemm<T>: T = double or float");
  hipStream_t stream = nullptr;
  detail::DevBuf da, db, dc;

  template <class At, class Bt, class Ct>
  void run(At& a, Bt& b, Ct& c, const ThreadInfo& thr) {
    if (thr.thread_id() == 0) launch(a, b, c);
    if (thr.num_threads() > 1) thr.barrier();
  }
  template <class At, class Bt, class Ct>
  void launch(At& a, Bt& b, Ct& c) {
    const size_t m = c.height(), n = c.width(), k = a.width();
    if (m == 0 || n == 0) return;
    const T alpha = a.get_scalar() * b.get_scalar();
    const T beta = c.get_scalar();
    auto A = detail::stage_in<T>(a, m, k, da, true, stream);
    auto B = detail::stage_in<T>(b, k, n, db, true, stream);
    auto C = detail::stage_in<T>(c, m, n, dc, beta != T(0), stream);
    if constexpr (std::is_same<T, double>::value)
      abi_check("momms_hip_dgemm",
                momms_hip_dgemm((int64_t)m, (int64_t)n, (int64_t)k, &alpha, A.ptr, A.rs, A.cs,
                                B.ptr, B.rs, B.cs, &beta, C.ptr, C.rs, C.cs, stream));
    else
      abi_check("momms_hip_sgemm",
                momms_hip_sgemm((int64_t)m, (int64_t)n, (int64_t)k, &alpha, A.ptr, A.rs, A.cs,
                                B.ptr, B.rs, B.cs, &beta, C.ptr, C.rs, C.cs, stream));
    detail::stage_out<T>(c, m, n, C.ptr, stream);
    hip_check("hipStreamSynchronize", hipStreamSynchronize(stream));
  }
  static std::vector<AlgorithmStep> hierarchy_description() { return {}; }
};

// The libxsmm leaves Xsmm / KernelXsmmA2 (src/kern/kernel_xsmm.rs:8-139) on
// the GPU: a small block in one launch, no packing (momms_hip_dgemm_small).
struct HipXsmm {
  hipStream_t stream = nullptr;
  detail::DevBuf da, db, dc;

  template <class At, class Bt, class Ct>
  void run(At& a, Bt& b, Ct& c, const ThreadInfo& thr) {
    if (thr.thread_id() == 0) launch(a, b, c);
    if (thr.num_threads() > 1) thr.barrier();
  }
  template <class At, class Bt, class Ct>
  void launch(At& a, Bt& b, Ct& c) {
    const size_t m = c.height(), n = c.width(), k = a.width();
    if (m == 0 || n == 0) return;
    const double alpha = a.get_scalar() * b.get_scalar();
    const double beta = c.get_scalar();
    auto A = detail::stage_in<double>(a, m, k, da, true, stream);
    auto B = detail::stage_in<double>(b, k, n, db, true, stream);
    auto C = detail::stage_in<double>(c, m, n, dc, beta != 0.0, stream);
    abi_check("momms_hip_dgemm_small",
              momms_hip_dgemm_small((int64_t)m, (int64_t)n, (int64_t)k, &alpha, A.ptr, A.rs, A.cs,
                                    B.ptr, B.rs, B.cs, &beta, C.ptr, C.rs, C.cs, stream));
    detail::stage_out<double>(c, m, n, C.ptr, stream);
    hip_check("hipStreamSynchronize", hipStreamSynchronize(stream));
  }
  static std::vector<AlgorithmStep> hierarchy_description() { return {}; }
};

// FP64-accurate GEMM from Slices BF16 slices (momms_hip_dgemm_split_bf16).
template <int Slices>
struct HipSplitGemm {
  static_assert(Slices >= 1 && Slices <= 8, "1..8 slices");
  hipStream_t stream = nullptr;
  detail::DevBuf da, db, dc, ws;

  template <class At, class Bt, class Ct>
  void run(At& a, Bt& b, Ct& c, const ThreadInfo& thr) {
    if (thr.thread_id() == 0) launch(a, b, c);
    if (thr.num_threads() > 1) thr.barrier();
  }
  template <class At, class Bt, class Ct>
  void launch(At& a, Bt& b, Ct& c) {
    const size_t m = c.height(), n = c.width(), k = a.width();
    if (m == 0 || n == 0) return;
    const double alpha = a.get_scalar() * b.get_scalar();
    const double beta = c.get_scalar();
    auto A = detail::stage_in<double>(a, m, k, da, true, stream);
    auto B = detail::stage_in<double>(b, k, n, db, true, stream);
    auto C = detail::stage_in<double>(c, m, n, dc, beta != 0.0, stream);
    const size_t wsb = momms_hip_split_bf16_workspace_size((int64_t)m, (int64_t)n, (int64_t)k, Slices);
    void* w = ws.get(wsb);
    abi_check("momms_hip_dgemm_split_bf16",
              momms_hip_dgemm_split_bf16((int64_t)m, (int64_t)n, (int64_t)k, &alpha, A.ptr, A.rs,
                                         A.cs, B.ptr, B.rs, B.cs, &beta, C.ptr, C.rs, C.cs,
                                         Slices, w, wsb, stream));
    detail::stage_out<double>(c, m, n, C.ptr, stream);
    hip_check("hipStreamSynchronize", hipStreamSynchronize(stream));
  }
  static std::vector<AlgorithmStep> hierarchy_description() { return {}; }
};

// FP64-accurate GEMM from BF16 residues modulo Moduli coprime moduli
// (momms_hip_dgemm_crt_bf16; 13 moduli give the FP64-level accuracy at k up
// to 8192).
template <int Moduli>
struct HipCrtGemm {
  static_assert(Moduli >= 1 && Moduli <= 16, "1..16 moduli");
  hipStream_t stream = nullptr;
  detail::DevBuf da, db, dc, ws;

  template <class At, class Bt, class Ct>
  void run(At& a, Bt& b, Ct& c, const ThreadInfo& thr) {
    if (thr.thread_id() == 0) launch(a, b, c);
    if (thr.num_threads() > 1) thr.barrier();
  }
  template <class At, class Bt, class Ct>
  void launch(At& a, Bt& b, Ct& c) {
    const size_t m = c.height(), n = c.width(), k = a.width();
    if (m == 0 || n == 0) return;
    const double alpha = a.get_scalar() * b.get_scalar();
    const double beta = c.get_scalar();
    auto A = detail::stage_in<double>(a, m, k, da, true, stream);
    auto B = detail::stage_in<double>(b, k, n, db, true, stream);
    auto C = detail::stage_in<double>(c, m, n, dc, beta != 0.0, stream);
    const size_t wsb = momms_hip_crt_bf16_workspace_size((int64_t)m, (int64_t)n, (int64_t)k, Moduli);
    if (wsb == 0)
      throw std::invalid_argument("HipCrtGemm: " + std::to_string(Moduli) +
                                  " moduli are too few for k = " + std::to_string(k));
    void* w = ws.get(wsb);
    abi_check("momms_hip_dgemm_crt_bf16",
              momms_hip_dgemm_crt_bf16((int64_t)m, (int64_t)n, (int64_t)k, &alpha, A.ptr, A.rs,
                                       A.cs, B.ptr, B.rs, B.cs, &beta, C.ptr, C.rs, C.cs,
                                       Moduli, w, wsb, stream));
    detail::stage_out<double>(c, m, n, C.ptr, stream);
    hip_check("hipStreamSynchronize", hipStreamSynchronize(stream));
  }
  static std::vector<AlgorithmStep> hierarchy_description() { return {}; }
};

}  // namespace momms
