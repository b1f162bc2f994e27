// CPU self-test of the GemmNode host layer (no GPU calls): view arithmetic,
// partitioning nodes, thread team, and parallel composition, each against a
// direct TripleLoop (same summation order, so results must be bit-identical).
#include <cmath>
#include <cstdio>
#include <cstring>

#include "momms/momms.hpp"

using namespace momms;

static int fails = 0;
#define CHECK(cond)                                                   \
  do {                                                                \
    if (!(cond)) {                                                    \
      std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #cond);      \
      ++fails;                                                        \
    }                                                                 \
  } while (0)

template <class Algo>
static bool same_as_triple_loop(size_t m, size_t n, size_t k, size_t threads, bool transpose_c) {
  std::mt19937_64 rng(0x6D6F6D6D73 + m * 31 + n * 7 + k);
  Matrix<double> a(m, k), b(k, n);
  Matrix<double> c1 = transpose_c ? Matrix<double>(n, m) : Matrix<double>(m, n);
  Matrix<double> c2(m, n);
  if (transpose_c) c1.transpose();
  a.fill_rand(rng);
  b.fill_rand(rng);
  c2.fill_rand(rng);
  c1.copy_from(c2);
  SpawnThreads<Algo> algo;
  algo.set_n_threads(threads);
  auto one = ThreadInfo::single_thread();
  algo.run(a, b, c1, one);
  TripleLoop ref;
  ref.run(a, b, c2, one);
  for (size_t x = 0; x < n; ++x)
    for (size_t y = 0; y < m; ++y) {
      const double u = c1.get(y, x), v = c2.get(y, x);
      if (std::memcmp(&u, &v, sizeof(double)) != 0) return false;
    }
  return true;
}

// TripleLoop that counts its calls (which child of a FirstDiff node ran).
template <int Id>
struct CountedLoop {
  static inline int calls = 0;
  TripleLoop loop;
  template <class At, class Bt, class Ct>
  void run(At& a, Bt& b, Ct& c, const ThreadInfo& thr) {
    if (thr.thread_id() == 0) ++calls;
    loop.run(a, b, c, thr);
  }
  static std::vector<AlgorithmStep> hierarchy_description() { return {}; }
};

int main() {
  // --- views (src/matrix/view.rs, general_stride.rs) ---
  {
    MatrixView v{0, 0, 10};
    auto zp = v.zoomed_size_and_padding(8, 4);
    CHECK(zp.first == 2 && zp.second == 0);
    MatrixView p{0, 3, 10};  // 7 physical
    zp = p.zoomed_size_and_padding(4, 4);
    CHECK(zp.first == 4 && zp.second == 1);
    Matrix<double> m(10, 7);
    CHECK(m.get_row_stride() == 1 && m.get_column_stride() == 10);
    for (size_t x = 0; x < 7; ++x)
      for (size_t y = 0; y < 10; ++y) m.set(y, x, (double)(y + 100 * x));
    const size_t h = m.push_y_view(4);
    CHECK(h == 10);
    m.slide_y_view_to(8, 4);
    CHECK(m.height() == 2 && m.get(1, 3) == 9 + 300);
    m.push_x_split(2, 5);
    CHECK(m.width() == 3 && m.get(0, 0) == 8 + 200);
    CHECK(m.get_buffer() == &m.make_alias().get_buffer()[0]);
    m.pop_x_split();
    m.pop_y_view();
    CHECK(m.height() == 10 && m.width() == 7);
    Matrix<double> t(3, 5);
    t.transpose();
    CHECK(t.height() == 5 && t.width() == 3 && t.get_row_stride() == 3 && t.get_column_stride() == 1);
    bool threw = false;
    try {
      Matrix<double> q(2, 2), r(3, 2);
      q.axpby(1.0, r, 1.0);
    } catch (const std::invalid_argument&) {
      threw = true;
    }
    CHECK(threw);
  }
  // --- hierarchy descriptions (part.rs:37-41: innermost first) ---
  {
    using Algo = PartN<3000, PartK<192, PartM<120, TripleLoop>>>;
    auto d = Algo::hierarchy_description();
    CHECK(d.size() == 3 && d[0].kind == AlgorithmStep::M && d[0].bsz == 120 &&
          d[1].kind == AlgorithmStep::K && d[2].kind == AlgorithmStep::N && d[2].bsz == 3000);
  }
  // --- thread team (thread_comm.rs) ---
  {
    auto comm = std::make_shared<ThreadComm>(4);
    std::atomic<int> counter{0};
    std::vector<std::thread> ts;
    std::atomic<int> bad{0};
    int value = 42;
    for (int id = 0; id < 4; ++id)
      ts.emplace_back([&, id] {
        ThreadInfo t(id, comm);
        for (int round = 0; round < 100; ++round) {
          counter.fetch_add(1);
          t.barrier();
          if (counter.load() != 4 * (round + 1)) bad++;
          t.barrier();
        }
        void* p = t.broadcast(id == 0 ? &value : nullptr);
        if (p != &value) bad++;
        ThreadInfo sub = t.split(2);
        if (sub.num_threads() != 2 || sub.thread_id() != (size_t)(id % 2)) bad++;
        sub.barrier();
      });
    for (auto& t : ts) t.join();
    CHECK(bad.load() == 0);
  }
  // --- partitioning and parallel nodes against TripleLoop, bit for bit ---
  CHECK((same_as_triple_loop<PartN<13, PartK<7, PartM<5, TripleLoop>>>>(37, 29, 23, 1, false)));
  CHECK((same_as_triple_loop<PartN<13, PartK<7, PartM<5, TripleLoop>>>>(37, 29, 23, 1, true)));
  CHECK((same_as_triple_loop<PartK<16, ParallelN<8, TheRest, TripleLoop>>>(20, 70, 33, 4, false)));
  CHECK((same_as_triple_loop<ParallelM<6, Target<2>, ParallelN<8, TheRest, Barrier<TripleLoop>>>>(
      50, 41, 9, 4, false)));
  // --- FirstDiffPartM/N/K (part.rs:122-264): the first block takes Child1,
  // the others Child2, the loops and the beta rule as PartM/N/K ---
  {
    using M = FirstDiffPartM<5, CountedLoop<1>, CountedLoop<2>>;
    CHECK((same_as_triple_loop<M>(37, 29, 23, 1, false)));
    CHECK(CountedLoop<1>::calls == 1 && CountedLoop<2>::calls == 7);  // ceil(37/5) blocks
    using N = FirstDiffPartN<13, CountedLoop<3>, CountedLoop<4>>;
    CHECK((same_as_triple_loop<N>(37, 29, 23, 1, true)));
    CHECK(CountedLoop<3>::calls == 1 && CountedLoop<4>::calls == 2);  // ceil(29/13)
    using K = FirstDiffPartK<7, CountedLoop<5>, CountedLoop<6>>;
    CHECK((same_as_triple_loop<K>(37, 29, 23, 1, false)));
    CHECK(CountedLoop<5>::calls == 1 && CountedLoop<6>::calls == 3);  // ceil(23/7)
    CHECK((same_as_triple_loop<PartN<13, FirstDiffPartK<7, PartM<5, TripleLoop>,
                                                        FirstDiffPartM<5, TripleLoop, TripleLoop>>>>(
        40, 31, 30, 1, false)));
    CHECK((same_as_triple_loop<ParallelN<8, TheRest, FirstDiffPartK<7, PartM<5, TripleLoop>,
                                                                    PartM<5, TripleLoop>>>>(
        40, 31, 30, 3, false)));
    auto d = FirstDiffPartM<120, PartK<192, TripleLoop>, TripleLoop>::hierarchy_description();
    CHECK(d.size() == 2 && d[0].kind == AlgorithmStep::K && d[1].kind == AlgorithmStep::M &&
          d[1].bsz == 120);
  }
  std::printf("%s (%d failures)\n", fails ? "FAIL" : "PASS", fails);
  return fails ? 1 : 0;
}
