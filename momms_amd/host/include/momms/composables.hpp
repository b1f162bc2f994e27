// GemmNode algorithm API: nodes nested as template parameters.
//
// Contract of src/composables/mod.rs:24-36: every node N provides
//   void run(At& a, Bt& b, Ct& c, const ThreadInfo& thr);   (C := c.scalar*C + a.scalar*b.scalar*A*B)
//   static std::vector<AlgorithmStep> hierarchy_description(); (innermost step first)
// and is default-constructible. Nodes here:
//   PartM/PartN/PartK<Bsz, Child>   src/composables/part.rs:7-120
//   FirstDiffPartM/N/K<Bsz, Child1, Child2>  src/composables/part.rs:122-264
//   ParallelM/ParallelN<Iota, Nway, Child>, Target<N>, TheRest
//                                   src/composables/parallel_range.rs:8-163
//   SpawnThreads<Child>             src/composables/spawn.rs:37-145
//   Barrier<Child>                  src/composables/barrier.rs:6-27
//   TripleLoop                      src/composables/triple_loop.rs:5-20
// The GPU leaves (HipGemm, HipSplitGemm) are in hip_gemm.hpp.
#pragma once
#include <pthread.h>
#include <sched.h>

#include <functional>
#include <thread>
#include <vector>

#include "momms/matrix.hpp"
#include "momms/thread_comm.hpp"

namespace momms {

struct AlgorithmStep {
  enum Kind { M, N, K } kind;
  size_t bsz;
};

template <size_t Bsz, class Child>
struct PartM {
  Child child;
  template <class At, class Bt, class Ct>
  void run(At& a, Bt& b, Ct& c, const ThreadInfo& thr) {
    const size_t m_save = a.push_y_view(Bsz);
    c.push_y_view(Bsz);
    for (size_t i = 0; i < m_save; i += Bsz) {
      a.slide_y_view_to(i, Bsz);
      c.slide_y_view_to(i, Bsz);
      child.run(a, b, c, thr);
    }
    a.pop_y_view();
    c.pop_y_view();
  }
  static std::vector<AlgorithmStep> hierarchy_description() {
    auto d = Child::hierarchy_description();
    d.push_back({AlgorithmStep::M, Bsz});
    return d;
  }
};

template <size_t Bsz, class Child>
struct PartN {
  Child child;
  template <class At, class Bt, class Ct>
  void run(At& a, Bt& b, Ct& c, const ThreadInfo& thr) {
    const size_t n_save = b.push_x_view(Bsz);
    c.push_x_view(Bsz);
    for (size_t i = 0; i < n_save; i += Bsz) {
      b.slide_x_view_to(i, Bsz);
      c.slide_x_view_to(i, Bsz);
      child.run(a, b, c, thr);
    }
    b.pop_x_view();
    c.pop_x_view();
  }
  static std::vector<AlgorithmStep> hierarchy_description() {
    auto d = Child::hierarchy_description();
    d.push_back({AlgorithmStep::N, Bsz});
    return d;
  }
};

// Loop over k; C's scalar (beta) becomes 1 after the first block and is
// restored at the end (part.rs:96-110).
template <size_t Bsz, class Child>
struct PartK {
  Child child;
  template <class At, class Bt, class Ct>
  void run(At& a, Bt& b, Ct& c, const ThreadInfo& thr) {
    const size_t k_save = a.push_x_view(Bsz);
    b.push_y_view(Bsz);
    const auto beta_save = c.get_scalar();
    for (size_t i = 0; i < k_save; i += Bsz) {
      a.slide_x_view_to(i, Bsz);
      b.slide_y_view_to(i, Bsz);
      child.run(a, b, c, thr);
      c.set_scalar(1);
    }
    a.pop_x_view();
    b.pop_y_view();
    c.set_scalar(beta_save);
  }
  static std::vector<AlgorithmStep> hierarchy_description() {
    auto d = Child::hierarchy_description();
    d.push_back({AlgorithmStep::K, Bsz});
    return d;
  }
};

// The loops of PartM/PartN/PartK with a different child for the first block
// (src/composables/part.rs:122-264): the overlap trees give the first M
// block the leaf that packs B as it multiplies and the later blocks the one
// that reads the packed B. The description is Child1's plus this step.
template <size_t Bsz, class Child1, class Child2>
struct FirstDiffPartM {
  Child1 child1;
  Child2 child2;
  template <class At, class Bt, class Ct>
  void run(At& a, Bt& b, Ct& c, const ThreadInfo& thr) {
    const size_t m_save = a.push_y_view(Bsz);
    c.push_y_view(Bsz);
    for (size_t i = 0; i < m_save; i += Bsz) {
      a.slide_y_view_to(i, Bsz);
      c.slide_y_view_to(i, Bsz);
      if (i == 0) child1.run(a, b, c, thr);
      else child2.run(a, b, c, thr);
    }
    a.pop_y_view();
    c.pop_y_view();
  }
  static std::vector<AlgorithmStep> hierarchy_description() {
    auto d = Child1::hierarchy_description();
    d.push_back({AlgorithmStep::M, Bsz});
    return d;
  }
};

template <size_t Bsz, class Child1, class Child2>
struct FirstDiffPartN {
  Child1 child1;
  Child2 child2;
  template <class At, class Bt, class Ct>
  void run(At& a, Bt& b, Ct& c, const ThreadInfo& thr) {
    const size_t n_save = b.push_x_view(Bsz);
    c.push_x_view(Bsz);
    for (size_t i = 0; i < n_save; i += Bsz) {
      b.slide_x_view_to(i, Bsz);
      c.slide_x_view_to(i, Bsz);
      if (i == 0) child1.run(a, b, c, thr);
      else child2.run(a, b, c, thr);
    }
    b.pop_x_view();
    c.pop_x_view();
  }
  static std::vector<AlgorithmStep> hierarchy_description() {
    auto d = Child1::hierarchy_description();
    d.push_back({AlgorithmStep::N, Bsz});
    return d;
  }
};

// As PartK, C's scalar becomes 1 after the first block and is restored.
template <size_t Bsz, class Child1, class Child2>
struct FirstDiffPartK {
  Child1 child1;
  Child2 child2;
  template <class At, class Bt, class Ct>
  void run(At& a, Bt& b, Ct& c, const ThreadInfo& thr) {
    const size_t k_save = a.push_x_view(Bsz);
    b.push_y_view(Bsz);
    const auto beta_save = c.get_scalar();
    for (size_t i = 0; i < k_save; i += Bsz) {
      a.slide_x_view_to(i, Bsz);
      b.slide_y_view_to(i, Bsz);
      if (i == 0) child1.run(a, b, c, thr);
      else child2.run(a, b, c, thr);
      c.set_scalar(1);
    }
    a.pop_x_view();
    b.pop_y_view();
    c.set_scalar(beta_save);
  }
  static std::vector<AlgorithmStep> hierarchy_description() {
    auto d = Child1::hierarchy_description();
    d.push_back({AlgorithmStep::K, Bsz});
    return d;
  }
};

// How many ways a Parallel node splits its team (parallel_range.rs:8-22).
template <size_t N>
struct Target {
  static size_t n_way(size_t n_threads) { return n_threads % N == 0 ? N : 1; }
};
struct TheRest {
  static size_t n_way(size_t n_threads) { return n_threads; }
};

namespace detail {
struct ParallelInfo {
  bool ready = false;
  std::unique_ptr<ThreadInfo> thr;
  size_t n_way = 1, work_id = 0;
  void init(const ThreadInfo& info, size_t n_way_) {
    n_way = n_way_;
    const size_t sub_n = info.num_threads() / n_way;
    thr = std::make_unique<ThreadInfo>(info.split(n_way));
    work_id = info.thread_id() / sub_n;
    ready = true;
  }
};
}  // namespace detail

template <size_t Iota, class Nway, class Child>
struct ParallelM {
  Child child;
  detail::ParallelInfo par;
  template <class At, class Bt, class Ct>
  void run(At& a, Bt& b, Ct& c, const ThreadInfo& thr) {
    if (!par.ready) par.init(thr, Nway::n_way(thr.num_threads()));
    const size_t n_iotas = a.iter_height() == 0 ? 0 : (a.iter_height() - 1) / Iota + 1;
    const size_t per = n_iotas == 0 ? 0 : (n_iotas - 1) / par.n_way + 1;
    const size_t start = Iota * per * par.work_id;
    const size_t end = start + Iota * per;
    a.push_y_split(start, end);
    c.push_y_split(start, end);
    child.run(a, b, c, *par.thr);
    a.pop_y_split();
    c.pop_y_split();
  }
  static std::vector<AlgorithmStep> hierarchy_description() { return Child::hierarchy_description(); }
};

template <size_t Iota, class Nway, class Child>
struct ParallelN {
  Child child;
  detail::ParallelInfo par;
  template <class At, class Bt, class Ct>
  void run(At& a, Bt& b, Ct& c, const ThreadInfo& thr) {
    if (!par.ready) par.init(thr, Nway::n_way(thr.num_threads()));
    const size_t n_iotas = b.iter_width() == 0 ? 0 : (b.iter_width() - 1) / Iota + 1;
    const size_t per = n_iotas == 0 ? 0 : (n_iotas - 1) / par.n_way + 1;
    const size_t start = Iota * per * par.work_id;
    const size_t end = start + Iota * per;
    b.push_x_split(start, end);
    c.push_x_split(start, end);
    child.run(a, b, c, *par.thr);
    b.pop_x_split();
    c.pop_x_split();
  }
  static std::vector<AlgorithmStep> hierarchy_description() { return Child::hierarchy_description(); }
};

template <class Child>
struct Barrier {
  Child child;
  template <class At, class Bt, class Ct>
  void run(At& a, Bt& b, Ct& c, const ThreadInfo& thr) {
    thr.barrier();
    child.run(a, b, c, thr);
  }
  static std::vector<AlgorithmStep> hierarchy_description() { return Child::hierarchy_description(); }
};

// Fans the tree out to n threads, the caller working as thread 0, each
// thread on aliases of A, B and C and its own cached copy of the child tree
// (spawn.rs:100-135). Threads are pinned to cores 0..n-1 of the affinity mask.
template <class Child>
struct SpawnThreads {
  size_t n_threads = 1;
  std::vector<std::unique_ptr<Child>> trees;

  void set_n_threads(size_t n) {
    n_threads = n < 1 ? 1 : n;
    trees.clear();
  }
  template <class At, class Bt, class Ct>
  void run(At& a, Bt& b, Ct& c, const ThreadInfo& /*thr*/) {
    if (trees.size() != n_threads) {
      trees.clear();
      for (size_t i = 0; i < n_threads; ++i) trees.push_back(std::make_unique<Child>());
    }
    auto comm = std::make_shared<ThreadComm>(n_threads);
    // Aliases are taken here, before thread 0 starts pushing views on a, b, c.
    std::vector<At> as;
    std::vector<Bt> bs;
    std::vector<Ct> cs;
    for (size_t id = 1; id < n_threads; ++id) {
      as.push_back(a.make_alias());
      bs.push_back(b.make_alias());
      cs.push_back(c.make_alias());
    }
    std::vector<std::thread> workers;
    for (size_t id = 1; id < n_threads; ++id) {
      workers.emplace_back([&, id] {
        pin(id);
        ThreadInfo t(id, comm);
        trees[id]->run(as[id - 1], bs[id - 1], cs[id - 1], t);
        t.barrier();
      });
    }
    ThreadInfo t0(0, comm);
    trees[0]->run(a, b, c, t0);
    t0.barrier();
    for (auto& w : workers) w.join();
  }
  static std::vector<AlgorithmStep> hierarchy_description() { return Child::hierarchy_description(); }

 private:
  static void pin(size_t idx) {
    cpu_set_t allowed;
    if (sched_getaffinity(0, sizeof(allowed), &allowed) != 0) return;
    size_t seen = 0;
    for (int cpu = 0; cpu < CPU_SETSIZE; ++cpu) {
      if (!CPU_ISSET(cpu, &allowed)) continue;
      if (seen++ == idx) {
        cpu_set_t one;
        CPU_ZERO(&one);
        CPU_SET(cpu, &one);
        pthread_setaffinity_np(pthread_self(), sizeof(one), &one);
        return;
      }
    }
  }
};

// Naive C += A*B in x, z, y order with c = a*b + c (triple_loop.rs:8-17).
struct TripleLoop {
  template <class At, class Bt, class Ct>
  void run(At& a, Bt& b, Ct& c, const ThreadInfo&) {
    for (size_t x = 0; x < c.width(); ++x)
      for (size_t z = 0; z < a.width(); ++z)
        for (size_t y = 0; y < c.height(); ++y) {
          volatile auto prod = a.get(y, z) * b.get(z, x);
          c.set(y, x, prod + c.get(y, x));
        }
  }
  static std::vector<AlgorithmStep> hierarchy_description() { return {}; }
};

}  // namespace momms
