// Thread team of the host control tree: barrier, pointer broadcast and
// split into sub-teams.
//
// Behaviour of src/thread_comm.rs:7-136 (ThreadComm / ThreadInfo), with
// acquire/release ordering where the reference uses Relaxed atomics (which is
// only correct under x86-TSO). On the GPU this hierarchy is the workgroup /
// wavefront split inside the kernels; on the host it coordinates the threads
// of SpawnThreads, and only thread 0 of a team launches GPU work.
#pragma once
#include <atomic>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <vector>

namespace momms {

class ThreadComm {
 public:
  explicit ThreadComm(size_t n) : n_(n), slot_reads_(n), sub_(n) {}

  size_t size() const { return n_; }

  // Sense-reversing spin barrier (thread_comm.rs:50-66).
  void barrier() {
    if (n_ == 1) return;
    const bool my_sense = sense_.load(std::memory_order_acquire);
    const size_t arrived = arrived_.fetch_add(1, std::memory_order_acq_rel);
    if (arrived == n_ - 1) {
      arrived_.store(0, std::memory_order_relaxed);
      sense_.store(!my_sense, std::memory_order_release);
    } else {
      while (sense_.load(std::memory_order_acquire) == my_sense) {
      }
    }
  }

  // Thread 0's pointer reaches every member (thread_comm.rs:68-89).
  void* broadcast(size_t tid, void* p) {
    if (n_ == 1) return p;
    if (tid == 0) {
      while (slot_reads_.load(std::memory_order_acquire) != n_) {
      }
      slot_.store(p, std::memory_order_relaxed);
      slot_reads_.store(0, std::memory_order_release);
    }
    while (slot_reads_.load(std::memory_order_acquire) == n_) {
    }
    void* r = slot_.load(std::memory_order_relaxed);
    slot_reads_.fetch_add(1, std::memory_order_acq_rel);
    return r;
  }

  // n_way sub-teams of n/n_way threads each (thread_comm.rs:91-107).
  std::shared_ptr<ThreadComm> split(size_t tid, size_t n_way) {
    if (n_way == 0 || n_ % n_way != 0) throw std::invalid_argument("split: n_way must divide n");
    const size_t sub_n = n_ / n_way;
    const size_t which = tid / sub_n;
    barrier();
    if (tid % sub_n == 0) {
      std::lock_guard<std::mutex> g(mu_);
      sub_[which] = std::make_shared<ThreadComm>(sub_n);
    }
    barrier();
    std::lock_guard<std::mutex> g(mu_);
    return sub_[which];
  }

 private:
  size_t n_;
  std::atomic<void*> slot_{nullptr};
  std::atomic<size_t> slot_reads_;
  std::atomic<bool> sense_{false};
  std::atomic<size_t> arrived_{0};
  std::mutex mu_;
  std::vector<std::shared_ptr<ThreadComm>> sub_;
};

class ThreadInfo {
 public:
  ThreadInfo(size_t id, std::shared_ptr<ThreadComm> comm) : id_(id), comm_(std::move(comm)) {}
  static ThreadInfo single_thread() { return ThreadInfo(0, std::make_shared<ThreadComm>(1)); }

  void barrier() const { comm_->barrier(); }
  void* broadcast(void* p) const { return comm_->broadcast(id_, p); }
  size_t num_threads() const { return comm_->size(); }
  size_t thread_id() const { return id_; }
  ThreadInfo split(size_t n_way) const {
    auto sub = comm_->split(id_, n_way);
    return ThreadInfo(id_ % (comm_->size() / n_way), sub);
  }

 private:
  size_t id_;
  std::shared_ptr<ThreadComm> comm_;
};

}  // namespace momms
