// Parallel host memcpy for the download path: the refined mesh leaves the
// device through pinned staging buffers, and the copy from staging into the
// caller's arrays (freshly allocated, so every page faults on first touch)
// runs on several host threads while the copy engine fills the other
// staging buffer.
#pragma once

#include <sched.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace gcdt {

class CopyPool {
 public:
  // workers: helper threads; the calling thread takes one slice as well
  explicit CopyPool(int workers) : n_(workers > 0 ? workers : 0) {
    for (int i = 0; i < n_; i++) th_.emplace_back([this, i] { worker(i); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : th_) t.join();
  }
  CopyPool(const CopyPool &) = delete;
  CopyPool &operator=(const CopyPool &) = delete;

  // Threads the host offers this process (its affinity mask), capped.
  static int default_workers(int cap = 15) {
    cpu_set_t set;
    int n = 1;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) n = CPU_COUNT(&set);
    n -= 1;  // the calling thread copies too
    return n < 0 ? 0 : (n > cap ? cap : n);
  }

  // memcpy(dst, src, bytes) split over the workers and the calling thread.
  void copy(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return;
    if (n_ == 0 || bytes < (size_t(1) << 20)) {
      std::memcpy(dst, src, bytes);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      dst_ = (char *)dst;
      src_ = (const char *)src;
      bytes_ = bytes;
      pending_ = n_;
      gen_++;
    }
    cv_.notify_all();
    slice(n_);
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [this] { return pending_ == 0; });
  }

 private:
  void slice(int id) {
    const size_t parts = (size_t)n_ + 1;
    // 4 KiB-aligned slice boundaries: each page is touched by one thread
    const size_t per = ((bytes_ / parts) + 4095) & ~(size_t)4095;
    const size_t lo = std::min(bytes_, per * (size_t)id);
    const size_t hi = id == n_ ? bytes_ : std::min(bytes_, per * (size_t)(id + 1));
    if (hi > lo) std::memcpy(dst_ + lo, src_ + lo, hi - lo);
  }
  void worker(int id) {
    unsigned long seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
      }
      slice(id);
      std::lock_guard<std::mutex> lk(mu_);
      if (--pending_ == 0) done_.notify_one();
    }
  }

  int n_;
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  char *dst_ = nullptr;
  const char *src_ = nullptr;
  size_t bytes_ = 0;
  int pending_ = 0;
  unsigned long gen_ = 0;
  bool stop_ = false;
};

}  // namespace gcdt
