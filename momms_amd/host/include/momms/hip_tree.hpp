// Control trees on the GPU: the reference's variant trees as tile schedules,
// and its packing nodes over device-resident hierarchical buffers.
//
//   Ukernel<Mr, Nr>          the hierarchy-description shape of KernelNM<Nr, Mr>
//                            (src/kern/kernel_nm.rs:112-117): [M Mr, N Nr]
//   GpuSchedule<Desc>        leaf: the whole product on FP64 matrix cores, its
//                            tiles taken in the order of the tree Desc (the
//                            device form of the loops Desc spells out,
//                            momms_hip_dgemm_tree); Desc is only described,
//                            never run
//   HierMatrix               a device-resident Hierarch<f64> buffer
//                            (src/matrix/hierarch.rs:18-184)
//   PackA/PackB<LH,LW,LRS,LCS, Child>   pack the current A (B) view into a
//                            HierMatrix laid out by Child's hierarchy
//                            description (src/composables/pack.rs:311-419)
//   UnpackC<LH,LW,LRS,LCS, Child>  Child computes alpha*A*B into a zeroed
//                            HierMatrix, which is then added to C
//                            (src/composables/unpack.rs:214-270)
//   DelayedPackA/B<LH,LW,LRS,LCS, Child>, PackPair, UnpairA/B
//                            packing left to the first multiply, for the
//                            overlap trees (src/composables/fused_pack.rs)
//   HipGemmHier<Desc>        leaf on three HierMatrix operands
//                            (momms_hip_dgemm_hier)
//   KnmBlock<Mr, Nr>         the KNM f32 tree's micro-kernel loops over its
//                            packed f32 operands (src/exper/knm.rs:89-97)
// As with HipGemm, under a team of threads only thread 0 launches and the
// team meets at a barrier; host-resident views are staged to HBM.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "momms/composables.hpp"
#include "momms/hip_gemm.hpp"
#include "momms/matrix.hpp"
#include "momms_hip.h"

namespace momms {

// KernelNM<Nr, Mr>'s hierarchy description; never run.
template <size_t Mr, size_t Nr>
struct Ukernel {
  static std::vector<AlgorithmStep> hierarchy_description() {
    return {{AlgorithmStep::M, Mr}, {AlgorithmStep::N, Nr}};
  }
};
// KernelMN<Nr, Mr>'s (src/kern/kernel_mn.rs:91-96): [N Nr, M Mr]; never run.
template <size_t Nr, size_t Mr>
struct UkernelMN {
  static std::vector<AlgorithmStep> hierarchy_description() {
    return {{AlgorithmStep::N, Nr}, {AlgorithmStep::M, Mr}};
  }
};

inline std::vector<momms_algo_step> abi_steps(const std::vector<AlgorithmStep>& d) {
  std::vector<momms_algo_step> s;
  for (const auto& x : d) {
    momms_algo_step t{};
    t.kind = x.kind == AlgorithmStep::M ? MOMMS_STEP_M
             : x.kind == AlgorithmStep::N ? MOMMS_STEP_N
                                          : MOMMS_STEP_K;
    t.bsz = (int64_t)x.bsz;
    s.push_back(t);
  }
  return s;
}

template <class Desc>
struct GpuSchedule {
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
    const auto steps = abi_steps(Desc::hierarchy_description());
    abi_check("momms_hip_dgemm_tree",
              momms_hip_dgemm_tree((int64_t)m, (int64_t)n, (int64_t)k, &alpha, A.ptr, A.rs, A.cs,
                                   B.ptr, B.rs, B.cs, &beta, C.ptr, C.rs, C.cs, steps.data(),
                                   (int64_t)steps.size(), stream));
    detail::stage_out<double>(c, m, n, C.ptr, stream);
    hip_check("hipStreamSynchronize", hipStreamSynchronize(stream));
  }
  static std::vector<AlgorithmStep> hierarchy_description() {
    return Desc::hierarchy_description();
  }
};

// A device-resident hierarchical buffer (Hierarch<T, LH, LW, LRS, LCS>, T =
// double or float): its layout, the device memory (grown, never shrunk: a
// ResizableBuffer, src/matrix/hierarch.rs:318-356) and the scalar.
template <class T>
class HierMatrixT {
 public:
  momms_hier_layout layout{};
  T* ptr() const { return static_cast<T*>(buf_.p); }
  size_t height() const { return (size_t)layout.h; }
  size_t width() const { return (size_t)layout.w; }
  T get_scalar() const { return scalar_; }
  void set_scalar(T s) { scalar_ = s; }
  // Hierarch::resize_to: lays out an h x w matrix along the steps.
  void resize_to(size_t h, size_t w, const std::vector<AlgorithmStep>& desc, int32_t ydim,
                 int32_t xdim, int64_t lh, int64_t lw, int64_t lrs, int64_t lcs) {
    const auto steps = abi_steps(desc);
    abi_check("momms_hier_layout_init",
              momms_hier_layout_init(&layout, (int64_t)h, (int64_t)w, steps.data(),
                                     (int64_t)steps.size(), ydim, xdim, lh, lw, lrs, lcs));
    buf_.get(std::max<size_t>((size_t)layout.capacity * sizeof(T), 256));
  }

 private:
  detail::DevBuf buf_;
  T scalar_ = T(1);
};
using HierMatrix = HierMatrixT<double>;

namespace detail {
// Packs an h x w view (host or device) into P laid out by P.layout.
template <class T, class M>
void pack_view(const M& x, size_t h, size_t w, HierMatrixT<T>& P, DevBuf& stage, hipStream_t st) {
  auto X = stage_in<T>(x, h, w, stage, true, st);
  if constexpr (std::is_same<T, double>::value)
    abi_check("momms_hip_hier_pack",
              momms_hip_hier_pack(&P.layout, X.ptr, X.rs, X.cs, P.ptr(), st));
  else
    abi_check("momms_hip_hier_pack_f32",
              momms_hip_hier_pack_f32(&P.layout, X.ptr, X.rs, X.cs, P.ptr(), st));
}
}  // namespace detail

template <size_t LH, size_t LW, size_t LRS, size_t LCS, class Child, class T = double>
struct PackA {
  Child child;
  HierMatrixT<T> a_pack;
  detail::DevBuf stage;
  hipStream_t stream = nullptr;
  template <class At, class Bt, class Ct>
  void run(At& a, Bt& b, Ct& c, const ThreadInfo& thr) {
    if (thr.thread_id() == 0) {
      a_pack.resize_to(a.height(), a.width(), Child::hierarchy_description(), MOMMS_STEP_M,
                       MOMMS_STEP_K, LH, LW, LRS, LCS);
      detail::pack_view<T>(a, a.height(), a.width(), a_pack, stage, stream);
      a_pack.set_scalar(a.get_scalar());
    }
    if (thr.num_threads() > 1) thr.barrier();
    child.run(a_pack, b, c, thr);
  }
  static std::vector<AlgorithmStep> hierarchy_description() { return Child::hierarchy_description(); }
};

template <size_t LH, size_t LW, size_t LRS, size_t LCS, class Child, class T = double>
struct PackB {
  Child child;
  HierMatrixT<T> b_pack;
  detail::DevBuf stage;
  hipStream_t stream = nullptr;
  template <class At, class Bt, class Ct>
  void run(At& a, Bt& b, Ct& c, const ThreadInfo& thr) {
    if (thr.thread_id() == 0) {
      b_pack.resize_to(b.height(), b.width(), Child::hierarchy_description(), MOMMS_STEP_K,
                       MOMMS_STEP_N, LH, LW, LRS, LCS);
      detail::pack_view<T>(b, b.height(), b.width(), b_pack, stage, stream);
      b_pack.set_scalar(b.get_scalar());
    }
    if (thr.num_threads() > 1) thr.barrier();
    child.run(a, b_pack, c, thr);
  }
  static std::vector<AlgorithmStep> hierarchy_description() { return Child::hierarchy_description(); }
};

// ---------------------------------------------------------------------------
// Packing that waits for the first multiply (src/composables/fused_pack.rs,
// src/matrix/pack_pair.rs): DelayedPackA/B lay the pack buffer out for the
// current view but do not fill it; the operand travels down as a PackPair of
// the view and its buffer. The first leaf that multiplies with the pair fills
// the buffer (the reference packs each leaf panel when its micro-kernel first
// reaches it, pack_pair.rs establish_leaf); UnpairA/B pass the filled buffer
// on, so the nodes below them read the packed operand. The overlap tree
// (src/exper/overlap.rs:71-79) is
//   PartN<Nc, PartK<Kc, DelayedPackB<.., FirstDiffPartM<Mc, Leaf, UnpairB<Leaf>>>>>:
// the first M block multiplies while B is packed, the others reuse it.
template <class Xt, class T = double>
struct PackPair {
  Xt& x;                   // the operand's current view
  HierMatrixT<T>& packed;  // its pack buffer, laid out for that view
  hipStream_t stream;
  detail::DevBuf& stage;
  bool filled = false;
  size_t height() const { return x.height(); }
  size_t width() const { return x.width(); }
  T get_scalar() const { return x.get_scalar(); }
  // The packed operand, filled on first use (thread 0 of a team only, as
  // every device call here).
  HierMatrixT<T>& establish() {
    if (!filled) {
      detail::pack_view<T>(x, x.height(), x.width(), packed, stage, stream);
      packed.set_scalar(x.get_scalar());
      filled = true;
    }
    return packed;
  }
};

namespace detail {
template <class T>
HierMatrixT<T>& established(HierMatrixT<T>& x) { return x; }
template <class Xt, class T>
HierMatrixT<T>& established(PackPair<Xt, T>& p) { return p.establish(); }
}  // namespace detail

template <size_t LH, size_t LW, size_t LRS, size_t LCS, class Child, class T = double>
struct DelayedPackA {
  Child child;
  HierMatrixT<T> a_pack;
  detail::DevBuf stage;
  hipStream_t stream = nullptr;
  template <class At, class Bt, class Ct>
  void run(At& a, Bt& b, Ct& c, const ThreadInfo& thr) {
    if (thr.thread_id() == 0)
      a_pack.resize_to(a.height(), a.width(), Child::hierarchy_description(), MOMMS_STEP_M,
                       MOMMS_STEP_K, LH, LW, LRS, LCS);
    if (thr.num_threads() > 1) thr.barrier();
    PackPair<At, T> pair{a, a_pack, stream, stage};
    child.run(pair, b, c, thr);
  }
  static std::vector<AlgorithmStep> hierarchy_description() { return Child::hierarchy_description(); }
};

template <size_t LH, size_t LW, size_t LRS, size_t LCS, class Child, class T = double>
struct DelayedPackB {
  Child child;
  HierMatrixT<T> b_pack;
  detail::DevBuf stage;
  hipStream_t stream = nullptr;
  template <class At, class Bt, class Ct>
  void run(At& a, Bt& b, Ct& c, const ThreadInfo& thr) {
    if (thr.thread_id() == 0)
      b_pack.resize_to(b.height(), b.width(), Child::hierarchy_description(), MOMMS_STEP_K,
                       MOMMS_STEP_N, LH, LW, LRS, LCS);
    if (thr.num_threads() > 1) thr.barrier();
    PackPair<Bt, T> pair{b, b_pack, stream, stage};
    child.run(a, pair, c, thr);
  }
  static std::vector<AlgorithmStep> hierarchy_description() { return Child::hierarchy_description(); }
};

// The packed side of a pair for the nodes below (fused_pack.rs UnpairA/B):
// the buffer is filled here if no leaf has done it yet, on thread 0, and the
// team meets before the children read it.
template <class Child>
struct UnpairA {
  Child child;
  template <class Xt, class T, class Bt, class Ct>
  void run(PackPair<Xt, T>& a, Bt& b, Ct& c, const ThreadInfo& thr) {
    if (thr.thread_id() == 0) a.establish();
    if (thr.num_threads() > 1) thr.barrier();
    child.run(a.packed, b, c, thr);
  }
  static std::vector<AlgorithmStep> hierarchy_description() { return Child::hierarchy_description(); }
};

template <class Child>
struct UnpairB {
  Child child;
  template <class At, class Xt, class T, class Ct>
  void run(At& a, PackPair<Xt, T>& b, Ct& c, const ThreadInfo& thr) {
    if (thr.thread_id() == 0) b.establish();
    if (thr.num_threads() > 1) thr.barrier();
    child.run(a, b.packed, c, thr);
  }
  static std::vector<AlgorithmStep> hierarchy_description() { return Child::hierarchy_description(); }
};

// C := C + alpha*A*B: Child writes alpha*A*B into a hierarchical buffer with
// scalar 0 (not read), whose elements are then added to C's (Unpacker::add,
// unpack.rs:61-98 and :174-213; C's own scalar is not applied, as there).
template <size_t LH, size_t LW, size_t LRS, size_t LCS, class Child>
struct UnpackC {
  Child child;
  HierMatrix c_pack;
  detail::DevBuf stage;
  hipStream_t stream = nullptr;
  template <class At, class Bt, class Ct>
  void run(At& a, Bt& b, Ct& c, const ThreadInfo& thr) {
    if (thr.thread_id() == 0) {
      c_pack.resize_to(c.height(), c.width(), Child::hierarchy_description(), MOMMS_STEP_M,
                       MOMMS_STEP_N, LH, LW, LRS, LCS);
      c_pack.set_scalar(0.0);
    }
    if (thr.num_threads() > 1) thr.barrier();
    child.run(a, b, c_pack, thr);
    if (thr.thread_id() == 0) {
      const size_t m = c.height(), n = c.width();
      if (m && n) {
        auto C = detail::stage_in<double>(c, m, n, stage, true, stream);
        abi_check("momms_hip_hier_unpack_add",
                  momms_hip_hier_unpack_add(&c_pack.layout, c_pack.ptr(), C.ptr, C.rs, C.cs,
                                            stream));
        detail::stage_out<double>(c, m, n, C.ptr, stream);
      }
      hip_check("hipStreamSynchronize", hipStreamSynchronize(stream));
    }
    if (thr.num_threads() > 1) thr.barrier();
  }
  static std::vector<AlgorithmStep> hierarchy_description() { return Child::hierarchy_description(); }
};

// GemmNode<f64, Hierarch, Hierarch, Hierarch> on the GPU (the l3
// experiment's trees on hierarchical operands, src/exper/l3.rs:46-72).
// An operand may also come as a PackPair of a DelayedPackA/B above: the leaf
// fills its buffer first (the packing of the overlap trees' first block).
template <class Desc>
struct HipGemmHier {
  hipStream_t stream = nullptr;
  template <class At, class Bt>
  void run(At& a_in, Bt& b_in, HierMatrix& c, const ThreadInfo& thr) {
    if (thr.thread_id() == 0) {
      HierMatrix& a = detail::established(a_in);
      HierMatrix& b = detail::established(b_in);
      const double alpha = a.get_scalar() * b.get_scalar();
      const double beta = c.get_scalar();
      abi_check("momms_hip_dgemm_hier",
                momms_hip_dgemm_hier(&a.layout, a.ptr(), &b.layout, b.ptr(), &alpha, &beta,
                                     &c.layout, c.ptr(), stream));
      hip_check("hipStreamSynchronize", hipStreamSynchronize(stream));
    }
    if (thr.num_threads() > 1) thr.barrier();
  }
  static std::vector<AlgorithmStep> hierarchy_description() {
    return Desc::hierarchy_description();
  }
};

// The KNM f32 tree's micro-kernel loops on the GPU: PartN<Nr, PartM<Mr,
// KnmKernel>> of src/exper/knm.rs:89-97 over the packed A (Hierarch<f32,
// Mr, Kr, 1, Mr>) and B (Hierarch<f32, Kr, Nr, 1, Kr>) of its PackA / PackB
// nodes, C a plain matrix. Its hierarchy description is that subtree's,
// [K 4 (knm_kernel.rs:33-37), M Mr, N Nr], so the pack nodes above lay the
// operands out as the reference does. The packed operands are copied out
// column-major (momms_hip_hier_unpack_f32) and multiplied on FP32 matrix
// cores (momms_hip_sgemm).
template <size_t Mr, size_t Nr>
struct KnmBlock {
  hipStream_t stream = nullptr;
  detail::DevBuf da, db, dc;
  template <class Ct>
  void run(HierMatrixT<float>& a, HierMatrixT<float>& b, Ct& c, const ThreadInfo& thr) {
    if (thr.thread_id() == 0) launch(a, b, c);
    if (thr.num_threads() > 1) thr.barrier();
  }
  template <class Ct>
  void launch(HierMatrixT<float>& a, HierMatrixT<float>& b, Ct& c) {
    const size_t m = c.height(), n = c.width(), k = a.width();
    if (m == 0 || n == 0) return;
    const float alpha = a.get_scalar() * b.get_scalar();
    const float beta = c.get_scalar();
    float* A = static_cast<float*>(da.get(std::max<size_t>(m * k * 4, 256)));
    float* B = static_cast<float*>(db.get(std::max<size_t>(k * n * 4, 256)));
    abi_check("momms_hip_hier_unpack_f32",
              momms_hip_hier_unpack_f32(&a.layout, a.ptr(), A, 1, (int64_t)std::max<size_t>(m, 1),
                                        stream));
    abi_check("momms_hip_hier_unpack_f32",
              momms_hip_hier_unpack_f32(&b.layout, b.ptr(), B, 1, (int64_t)std::max<size_t>(k, 1),
                                        stream));
    auto C = detail::stage_in<float>(c, m, n, dc, beta != 0.0f, stream);
    abi_check("momms_hip_sgemm",
              momms_hip_sgemm((int64_t)m, (int64_t)n, (int64_t)k, &alpha, A, 1,
                              (int64_t)std::max<size_t>(m, 1), B, 1,
                              (int64_t)std::max<size_t>(k, 1), &beta, C.ptr, C.rs, C.cs, stream));
    detail::stage_out<float>(c, m, n, C.ptr, stream);
    hip_check("hipStreamSynchronize", hipStreamSynchronize(stream));
  }
  static std::vector<AlgorithmStep> hierarchy_description() {
    return {{AlgorithmStep::K, 4}, {AlgorithmStep::M, Mr}, {AlgorithmStep::N, Nr}};
  }
};

}  // namespace momms
