// Host worker threads shared by the library's host-side parallel loops (the
// per-reflection setup, geometry.cpp; the risk-tree trainer, risk_fit.cpp).
#pragma once
#include <sys/types.h>
#include <unistd.h>

#include <condition_variable>
#include <cstdint>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace anacor {

// Worker threads kept for the process's lifetime: starting and joining 16
// threads per call cost about 0.6 ms, most of the setup of a 20,000-reflection
// call.  run() splits [0, R) into `parts` contiguous chunks; the caller runs
// chunk 0 and worker w chunk w + 1; it returns when those workers have
// finished their chunks (the other workers wake, read the job's sizes under
// the lock and sleep again, never touching the body), so the job may live on
// the caller's stack.  One job at a time (callers from several threads queue
// on run_mu).  The pool is never destroyed: its threads sleep on the
// condition variable until the process ends.
class WorkerPool {
 public:
  using Body = std::function<void(int64_t, int64_t)>;

  void run(int64_t R, int parts, const Body& body) {
    std::lock_guard<std::mutex> serial(run_mu_);
    int64_t chunk;
    {
      std::unique_lock<std::mutex> lk(mu_);
      try {
        while ((int)threads_.size() < parts - 1) {
          const int w = (int)threads_.size();
          threads_.emplace_back([this, w] { loop(w); });
        }
      } catch (...) {
        // no more threads: the job is split over the ones there are
      }
      const int workers = (int)threads_.size();
      if (parts > workers + 1) parts = workers + 1;
      chunk = (R + parts - 1) / parts;
      body_ = &body;
      R_ = R;
      chunk_ = chunk;
      parts_ = parts;
      acks_ = 0;
      ++gen_;
    }
    cv_.notify_all();
    body((int64_t)0, chunk < R ? chunk : R);
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] { return acks_ == parts - 1; });
    body_ = nullptr;
  }

 private:
  void loop(int w) {
    uint64_t seen = 0;
    {
      std::lock_guard<std::mutex> lk(mu_);
      seen = gen_ - 1;   // created inside run(): take part in the current job
    }
    for (;;) {
      const Body* body;
      int64_t lo, hi;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
        if (w + 1 >= parts_) continue;   // no chunk of this job
        body = body_;
        lo = (int64_t)(w + 1) * chunk_;
        hi = lo + chunk_ < R_ ? lo + chunk_ : R_;
      }
      if (lo < hi) (*body)(lo, hi);
      {
        std::lock_guard<std::mutex> lk(mu_);
        ++acks_;
      }
      done_cv_.notify_one();
    }
  }

  std::mutex run_mu_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::vector<std::thread> threads_;
  uint64_t gen_ = 0;
  const Body* body_ = nullptr;
  int64_t R_ = 0, chunk_ = 0;
  int parts_ = 0;
  int acks_ = 0;
};

// The process's pool (a child after fork() starts its own: the parent's
// threads do not exist there).
inline WorkerPool& host_pool() {
  static std::mutex mu;
  static WorkerPool* p = nullptr;
  static pid_t owner = 0;
  std::lock_guard<std::mutex> lk(mu);
  const pid_t me = getpid();
  if (!p || owner != me) {
    p = new WorkerPool();   // never destroyed (see above)
    owner = me;
  }
  return *p;
}

}  // namespace anacor
