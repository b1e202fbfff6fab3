// A persistent helper thread that runs one task at a time beside its
// owner's thread (host stepper and batch combiner).
#pragma once

#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>

namespace rigs {

class Helper {
  public:
    ~Helper() {
        {
            std::lock_guard<std::mutex> lk(m_);
            quit_ = true;
        }
        cv_.notify_all();
        if (t_.joinable()) t_.join();
    }
    void run(std::function<void()> task) {
        if (!t_.joinable()) t_ = std::thread([this] { loop(); });
        {
            std::lock_guard<std::mutex> lk(m_);
            task_ = std::move(task);
            busy_ = true;
        }
        cv_.notify_all();
    }
    void wait() {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [this] { return !busy_; });
    }

  private:
    void loop() {
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            cv_.wait(lk, [this] { return busy_ || quit_; });
            if (busy_) {
                std::function<void()> task = std::move(task_);
                lk.unlock();
                task();
                lk.lock();
                busy_ = false;
                cv_.notify_all();
            } else if (quit_) {
                return;
            }
        }
    }
    std::thread t_;
    std::mutex m_;
    std::condition_variable cv_;
    std::function<void()> task_;
    bool busy_ = false, quit_ = false;
};

}  // namespace rigs
