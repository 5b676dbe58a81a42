// Background writers for the VTK snapshots: the time loop hands each output step's files to a
// small pool of threads and goes on stepping the GPU while they are formatted and written. The
// reference writes them inline (src/main.cpp:1090-1103), where text formatting of about 10 M
// numbers per output step costs far more than the GPU steps between two outputs.
#pragma once

#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace lpsio {

class AsyncWriter {
 public:
  // threads == 0: every task runs at once on the caller's thread.
  explicit AsyncWriter(int threads) : cap_(2 * threads) {
    for (int k = 0; k < threads; k++) pool_.emplace_back([this] { work(); });
  }
  ~AsyncWriter() {
    {
      std::lock_guard<std::mutex> g(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (std::thread &t : pool_) t.join();
  }
  AsyncWriter(const AsyncWriter &) = delete;
  AsyncWriter &operator=(const AsyncWriter &) = delete;

  // Queues a task that returns false on failure; blocks while 2 x threads tasks are waiting (the
  // snapshots they hold bound the memory).
  void submit(std::function<bool()> task) {
    if (pool_.empty()) {
      if (!task()) failed_++;
      return;
    }
    std::unique_lock<std::mutex> g(m_);
    room_.wait(g, [this] { return (int)q_.size() < cap_; });
    q_.push_back(std::move(task));
    g.unlock();
    cv_.notify_one();
  }
  // Waits for every queued task; returns the number of failed tasks so far.
  int wait() {
    std::unique_lock<std::mutex> g(m_);
    idle_.wait(g, [this] { return q_.empty() && busy_ == 0; });
    return failed_;
  }

 private:
  void work() {
    for (;;) {
      std::function<bool()> task;
      {
        std::unique_lock<std::mutex> g(m_);
        cv_.wait(g, [this] { return stop_ || !q_.empty(); });
        if (q_.empty()) return;  // stop_ with nothing left
        task = std::move(q_.front());
        q_.pop_front();
        busy_++;
      }
      room_.notify_one();
      const bool ok = task();
      {
        std::lock_guard<std::mutex> g(m_);
        busy_--;
        if (!ok) failed_++;
      }
      idle_.notify_all();
    }
  }

  std::vector<std::thread> pool_;
  std::deque<std::function<bool()>> q_;
  std::mutex m_;
  std::condition_variable cv_, room_, idle_;
  int cap_ = 0, busy_ = 0, failed_ = 0;
  bool stop_ = false;
};

}  // namespace lpsio
