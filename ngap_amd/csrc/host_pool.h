// Host threads for the copies into pinned staging memory (request bytes in
// the server, the input in ngap_set_input): one thread's copy runs well below
// what the host-to-device link takes, so a large copy is cut into parts that
// helper threads and the calling thread share.
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace ngapamd {

class HostPool {
 public:
  static constexpr size_t kPart = 256u << 10;  // smallest part worth a thread

  // up to `max_helpers` helper threads, no more than the host has CPUs - 1;
  // a thread the system refuses leaves the pool with fewer (the caller's
  // thread always runs parts itself)
  explicit HostPool(unsigned max_helpers = 7) {
    const unsigned hw = std::thread::hardware_concurrency();
    const unsigned n = std::min(max_helpers, hw > 1 ? hw - 1 : 0u);
    th_.reserve(n);
    for (unsigned i = 0; i < n; i++) {
      try {
        th_.emplace_back([this] { loop(); });
      } catch (...) {
        break;
      }
    }
  }
  HostPool(const HostPool &) = delete;
  HostPool &operator=(const HostPool &) = delete;
  ~HostPool() {
    {
      std::lock_guard<std::mutex> g(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : th_) t.join();
  }

  // threads a run can use: the helpers and the caller
  size_t width() const { return th_.size() + 1; }

  // Parts of about n / width() bytes (at least kPart, 4 KiB-aligned) covering
  // [0, n): how run_parts cuts a range.
  size_t part_bytes(size_t n) const {
    const size_t parts = std::max<size_t>(1, std::min(width(), n / kPart));
    return std::max<size_t>(4096, (n / parts + 4095) & ~(size_t)4095);
  }

  // f(i, begin, end) for the parts of [0, n); returns when all have run.
  // Calls from one thread at a time.
  void run_parts(size_t n, const std::function<void(size_t, size_t, size_t)> &f) {
    const size_t step = part_bytes(n);
    const size_t parts = n ? (n + step - 1) / step : 0;
    if (parts <= 1) {
      if (n) f(0, 0, n);
      return;
    }
    auto job = std::make_shared<Job>();
    job->f = &f;
    job->n = n;
    job->step = step;
    job->parts = parts;
    {
      std::lock_guard<std::mutex> g(m_);
      job_ = job;
      gen_++;
    }
    cv_.notify_all();
    work(*job);
    std::unique_lock<std::mutex> l(m_);
    done_cv_.wait(l, [&] { return job->done.load() == job->parts; });
    job_.reset();
  }

  // memcpy(dst, src, n) cut into parts
  void copy(uint8_t *dst, const uint8_t *src, size_t n) {
    run_parts(n, [&](size_t, size_t a, size_t b) { std::copy(src + a, src + b, dst + a); });
  }

 private:
  struct Job {
    const std::function<void(size_t, size_t, size_t)> *f = nullptr;
    size_t n = 0, step = 0, parts = 0;
    std::atomic<size_t> next{0}, done{0};
  };
  void work(Job &j) {
    for (;;) {
      const size_t i = j.next.fetch_add(1);
      if (i >= j.parts) return;
      const size_t a = i * j.step, b = std::min(j.n, a + j.step);
      (*j.f)(i, a, b);
      if (j.done.fetch_add(1) + 1 == j.parts) {
        std::lock_guard<std::mutex> g(m_);
        done_cv_.notify_all();
      }
    }
  }
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      std::shared_ptr<Job> j;
      {
        std::unique_lock<std::mutex> l(m_);
        cv_.wait(l, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        j = job_;
      }
      if (j) work(*j);  // a job already finished has no parts left
    }
  }
  std::vector<std::thread> th_;
  std::mutex m_;
  std::condition_variable cv_, done_cv_;
  std::shared_ptr<Job> job_;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

}  // namespace ngapamd
