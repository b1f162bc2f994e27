// GPU nodes dropped into GemmNode trees, checked against TripleLoop.
//
// The harness shape of src/exper/l3_packing.rs:18-47: fresh seeded inputs,
// run the algorithm, then compare. Each tree is the reference's composition
// with the Goto subtree replaced by a HipGemm leaf:
//   HipGemm                                  host Matrix<T> views, staged to HBM
//   PartN<3000, PartK<192, HipGemm>>         beta := 1 after the first k block
//   SpawnThreads<HipGemm> on 4 threads       only thread 0 launches
//   SpawnThreads<ParallelN<.., HipGemm>>     every sub-team launches, at once
//   HipGemm on device-resident matrices      no staging
//   HipSplitGemm<7>                          FP64-accurate from BF16 slices
//   HipCrtGemm<13>                           FP64-accurate from BF16 residues
//   HipGemm<float>                           FP32 (the KNM sgemm counterpart)
//   GpuSchedule<Goto | L3A | L3B | L3C | L4C> the variant trees' tile orders
//   SpawnThreads<PartM<PartN<Barrier<GpuSchedule>>>>  out of core: host blocks
//                                            streamed through the GPU
//   UnpackC<PackB<PackA<HipGemmHier>>>       device Hierarch operands
//   PartN<PartK<DelayedPackB<FirstDiffPartM<Bottom, UnpairB<Bottom>>>>>
//                                            the overlap tree: B packed by the
//                                            first M block's leaf (and A by the
//                                            first N block's, DelayedPackA)
//   PartN<PartK<PackB<PartM<PackA<KnmBlock>>>>> the KNM f32 tree
//   HipXsmm, PartM<PartN<PartK<HipXsmm>>>    the small-GEMM leaf
// Tolerance: |C - C_ref| <= 2*gamma_{k+1}*(|A||B| + |C0|) componentwise, with
// the unit roundoff of T (TripleLoop in T is the reference), plus the
// randomized residual of src/util.rs:55-85.
#include <cmath>
#include <cstdio>

#include "momms/momms.hpp"

using namespace momms;

static int fails = 0;

template <class Algo>
static auto set_threads(Algo& a, size_t n) -> decltype(a.set_n_threads(n), void()) {
  a.set_n_threads(n);
}
static void set_threads(...) {}

// The reference's variant trees (block sizes of src/exper/l3_shapes.rs:50-100,
// src/exper/l4.rs:44-78 and src/exper/out_of_core.rs:52-87, MR = 4, NR = 12),
// as descriptions for GpuSchedule: their M / N loops order the device tiles.
using GotoDesc = PartN<3000, PartK<192, PartM<120, Ukernel<4, 12>>>>;
using L3ADesc = PartM<768, PartK<768, PartN<120, PartK<192, UkernelMN<12, 4>>>>>;
using L3BDesc = PartN<768, PartK<768, PartM<120, PartK<192, Ukernel<4, 12>>>>>;
using L3CDesc = PartN<624, PartM<624, PartK<156, PartM<156, Ukernel<4, 12>>>>>;
using L4CDesc = PartM<3600, PartN<3600, PartK<192, PartM<120, Ukernel<4, 12>>>>>;
// out of core (out_of_core.rs:77-87): blocks of C streamed through the GPU,
// each block's product ordered as the tree's resident-B inner part
using OutOfCore = SpawnThreads<PartM<700, PartN<700, Barrier<GpuSchedule<
    PartK<768, PartN<768, PartM<120, PartK<192, Ukernel<4, 12>>>>>>>>>>;
// the l3 experiment's Goto tree on hierarchical operands (l3.rs:80-108):
// A, B and C packed into device Hierarch buffers laid out by the tree
using GotoHier = UnpackC<4, 12, 12, 1, PackB<192, 12, 12, 1, PackA<4, 192, 1, 4,
                 HipGemmHier<GotoDesc>>>>;
// the KNM sgemm tree (knm.rs:79-97: NC 14400, KC 336, MC 160, MR 16, NR 24,
// KR 4): f32 operands packed into device Hierarch buffers per block, the
// micro-kernel loops of each block on FP32 matrix cores
using KnmTree = PartN<14400, PartK<336, PackB<4, 24, 1, 4, PartM<160, PackA<16, 4, 1, 16,
                KnmBlock<16, 24>, float>>, float>>>;
// the overlap tree (overlap.rs:71-79, MR 4, NR 12, over the l3 experiment's
// hierarchical operands): DelayedPackB lays B's pack out per (N, K) block,
// the first M block's leaf fills it as it multiplies, the later M blocks
// read it through UnpairB
using OverlapBottom = UnpackC<4, 12, 12, 1, PackA<4, 192, 1, 4, HipGemmHier<GotoDesc>>>;
using GotoOverlap = SpawnThreads<PartN<3000, PartK<192, DelayedPackB<192, 12, 12, 1,
                    FirstDiffPartM<120, OverlapBottom, UnpairB<OverlapBottom>>>>>>;
// the mirror image: A's pack delayed to the first N block's leaf
using OverlapBottomA = UnpackC<4, 12, 12, 1, PackB<192, 12, 12, 1, HipGemmHier<GotoDesc>>>;
using GotoOverlapA = PartM<600, PartK<192, DelayedPackA<4, 192, 1, 4,
                     FirstDiffPartN<500, OverlapBottomA, UnpairA<OverlapBottomA>>>>>;
// small blocks straight from the operands (the libxsmm leaves)
using XsmmTree = PartM<96, PartN<64, PartK<128, HipXsmm>>>;

template <class Algo, class T = double>
static void check_tree(const char* name, size_t m, size_t n, size_t k, size_t threads,
                       double tol_scale, bool device) {
  std::mt19937_64 rng(0x6D6F6D6D73 + m + n + k);
  Matrix<T> a(m, k), b(k, n), c0(m, n), cref(m, n), mag(m, n);
  a.fill_rand(rng);
  b.fill_rand(rng);
  c0.fill_rand(rng);
  cref.copy_from(c0);
  auto one = ThreadInfo::single_thread();
  TripleLoop ref;
  ref.run(a, b, cref, one);
  // |A||B| + |C0|
  Matrix<T> aa(m, k), bb(k, n);
  for (size_t x = 0; x < k; ++x)
    for (size_t y = 0; y < m; ++y) aa.set(y, x, std::fabs(a.get(y, x)));
  for (size_t x = 0; x < n; ++x)
    for (size_t y = 0; y < k; ++y) bb.set(y, x, std::fabs(b.get(y, x)));
  for (size_t x = 0; x < n; ++x)
    for (size_t y = 0; y < m; ++y) mag.set(y, x, std::fabs(c0.get(y, x)));
  ref.run(aa, bb, mag, one);

  Matrix<T> c(m, n);
  c.copy_from(c0);
  Algo algo;
  set_threads(algo, threads);
  auto t0 = std::chrono::steady_clock::now();
  if (device) {
    auto da = Matrix<T>::device(m, k), db = Matrix<T>::device(k, n), dc = Matrix<T>::device(m, n);
    const size_t e = sizeof(T);
    hip_check("copy", hipMemcpy(da.get_buffer(), a.get_buffer(), m * k * e, hipMemcpyHostToDevice));
    hip_check("copy", hipMemcpy(db.get_buffer(), b.get_buffer(), k * n * e, hipMemcpyHostToDevice));
    hip_check("copy", hipMemcpy(dc.get_buffer(), c.get_buffer(), m * n * e, hipMemcpyHostToDevice));
    algo.run(da, db, dc, one);
    hip_check("copy", hipMemcpy(c.get_buffer(), dc.get_buffer(), m * n * e, hipMemcpyDeviceToHost));
  } else {
    algo.run(a, b, c, one);
  }
  const double secs = seconds_since(t0);
  const double u = std::ldexp(1.0, std::is_same<T, float>::value ? -24 : -53);
  const double gam = (k + 1) * u / (1 - (k + 1) * u);
  double worst = 0;
  size_t bad = 0;
  for (size_t x = 0; x < n; ++x)
    for (size_t y = 0; y < m; ++y) {
      const double e = std::fabs((double)c.get(y, x) - (double)cref.get(y, x)) / mag.get(y, x);
      worst = std::max(worst, e);
      if (!(e <= tol_scale * gam)) {
        if (!bad)
          std::printf("  first bad entry (%zu, %zu): C %.17g  reference %.17g  magnitude %.17g\n", y,
                      x, (double)c.get(y, x), (double)cref.get(y, x), (double)mag.get(y, x));
        ++bad;
      }
    }
  const double res = std::sqrt((double)test_c_eq_a_b(a, b, c, &c0, T(1), T(1), rng));
  std::printf("%-34s m=%zu n=%zu k=%zu  max|err|/mag=%.3e (bound %.3e)  residual=%.3e  %.2f GFLOP/s  %s\n",
              name, m, n, k, worst, tol_scale * gam, res, gflops(m, n, k, secs),
              bad ? "FAIL" : "PASS");
  if (bad) ++fails;
}

int main() {
  check_tree<HipGemm<double>>("HipGemm (host views)", 300, 211, 517, 1, 2.0, false);
  check_tree<PartN<3000, PartK<192, HipGemm<double>>>>("PartN<3000,PartK<192,HipGemm>>", 257,
                                                        3100, 400, 1, 2.0, false);
  check_tree<SpawnThreads<HipGemm<double>>>("SpawnThreads<HipGemm> x4", 129, 300, 250, 4, 2.0,
                                            false);
  // sub-teams of one thread each launch on the NULL stream at once; their
  // shapes take the stream-K form, whose workspace is per calling thread
  check_tree<SpawnThreads<ParallelN<64, Target<2>, HipGemm<double>>>>(
      "SpawnThreads<ParallelN<64,Target<2>,HipGemm>> x2", 1024, 1024, 2048, 2, 2.0, true);
  check_tree<SpawnThreads<ParallelN<64, TheRest, HipGemm<double>>>>(
      "SpawnThreads<ParallelN<64,TheRest,HipGemm>> x3", 700, 1500, 1800, 3, 2.0, true);
  check_tree<HipGemm<double>>("HipGemm (device matrices)", 512, 384, 640, 1, 2.0, true);
  check_tree<HipSplitGemm<7>>("HipSplitGemm<7>", 300, 200, 700, 1, 2.0, false);
  check_tree<HipCrtGemm<13>>("HipCrtGemm<13>", 300, 200, 700, 1, 2.0, false);
  check_tree<PartN<3000, PartK<192, HipCrtGemm<13>>>>("PartN<3000,PartK<192,HipCrtGemm<13>>>",
                                                      257, 3100, 400, 1, 2.0, false);
  check_tree<HipGemm<float>, float>("HipGemm<float> (host views)", 300, 211, 517, 1, 2.0, false);
  check_tree<GpuSchedule<GotoDesc>>("GpuSchedule<Goto>", 1000, 3100, 600, 1, 2.0, true);
  check_tree<GpuSchedule<L3ADesc>>("GpuSchedule<L3A>", 1000, 900, 600, 1, 2.0, true);
  check_tree<GpuSchedule<L3BDesc>>("GpuSchedule<L3B>", 900, 1000, 600, 1, 2.0, true);
  check_tree<GpuSchedule<L3CDesc>>("GpuSchedule<L3C>", 1300, 1250, 400, 1, 2.0, true);
  check_tree<GpuSchedule<L4CDesc>>("GpuSchedule<L4C> (host views)", 700, 800, 300, 1, 2.0, false);
  check_tree<OutOfCore>("SpawnThreads<PartM<700,PartN<700,Barrier<GpuSchedule>>>> x2", 1500, 1400,
                        500, 2, 2.0, false);
  check_tree<GotoHier>("UnpackC<PackB<PackA<HipGemmHier<Goto>>>>", 700, 500, 450, 1, 2.0, false);
  check_tree<SpawnThreads<GotoHier>>("SpawnThreads<UnpackC<...HipGemmHier>> x3", 300, 260, 200, 3,
                                     2.0, false);
  check_tree<KnmTree, float>("KNM f32 tree (PackB/PackA Hierarch<f32>)", 500, 420, 700, 1, 2.0,
                             false);
  check_tree<GotoOverlap>("SpawnThreads<PartN<PartK<DelayedPackB<FirstDiffPartM<..UnpairB>>>>> x2",
                          700, 3100, 450, 2, 2.0, false);
  check_tree<GotoOverlapA>("PartM<PartK<DelayedPackA<FirstDiffPartN<..UnpairA>>>>", 900, 1300, 400,
                           1, 2.0, false);
  check_tree<HipXsmm>("HipXsmm (device matrices)", 90, 70, 50, 1, 2.0, true);
  check_tree<XsmmTree>("PartM<96,PartN<64,PartK<128,HipXsmm>>>", 300, 200, 260, 1, 2.0, true);
  std::printf("%s (%d failures)\n", fails ? "FAIL" : "PASS", fails);
  return fails ? 1 : 0;
}
