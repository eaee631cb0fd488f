// replan_pool.h -- worker threads for the per-step replans (host only).
//
// A step's replans are planned on the host while its ORCA kernels run.
// The threads are made once (and again when the thread count changes)
// instead of once per step: starting a thread costs the host tens of
// microseconds, and the first steps of a process paid for fresh stacks.
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace ecm {

// run() hands the items of one batch out through an atomic index to the
// workers and the calling thread, and returns when every worker has
// finished its share; false if an item threw (the other items still run).
//
// Batches are numbered (gen_).  A worker starts out knowing the number the
// pool had when it was made, so a batch posted before the worker first
// takes the lock is still new to it; every batch is therefore taken by
// every worker, and busy_ (the workers still in the batch) reaches zero.
class ReplanPool {
public:
    ReplanPool() = default;
    ReplanPool(const ReplanPool&) = delete;
    ReplanPool& operator=(const ReplanPool&) = delete;
    ~ReplanPool() { stop(); }

    bool run(int threads, int n, const std::function<void(int)>& fn) {
        failed_.store(false);
        if (threads <= 1 || n <= 1) {
            for (int i = 0; i < n; ++i) call(fn, i);
            return !failed_.load();
        }
        resize((size_t)threads - 1);
        {
            std::lock_guard<std::mutex> lk(m_);
            fn_ = &fn;
            n_ = n;
            next_.store(0);
            busy_ = (int)workers_.size();
            ++gen_;
        }
        wake_.notify_all();
        for (int i; (i = next_.fetch_add(1)) < n;) call(fn, i);
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [&] { return busy_ == 0; });
        fn_ = nullptr;
        return !failed_.load();
    }

    size_t workers() const { return workers_.size(); }

private:
    void call(const std::function<void(int)>& fn, int i) {
        try {
            fn(i);
        } catch (...) {
            failed_.store(true);
        }
    }
    void resize(size_t k) {
        if (workers_.size() == k) return;
        stop();
        uint64_t g0;
        {
            std::lock_guard<std::mutex> lk(m_);
            quit_ = false;
            g0 = gen_;
        }
        for (size_t i = 0; i < k; ++i) workers_.emplace_back([this, g0] { loop(g0); });
    }
    void stop() {
        {
            std::lock_guard<std::mutex> lk(m_);
            quit_ = true;
            ++gen_;
        }
        wake_.notify_all();
        for (auto& t : workers_) t.join();
        workers_.clear();
    }
    void loop(uint64_t seen) {
        std::unique_lock<std::mutex> lk(m_);
        while (true) {
            wake_.wait(lk, [&] { return gen_ != seen; });
            seen = gen_;
            if (quit_) return;
            const std::function<void(int)>* fn = fn_;
            const int n = n_;
            lk.unlock();
            for (int i; (i = next_.fetch_add(1)) < n;) call(*fn, i);
            lk.lock();
            if (--busy_ == 0) done_.notify_one();
        }
    }

    std::vector<std::thread> workers_;
    std::mutex m_;
    std::condition_variable wake_, done_;
    const std::function<void(int)>* fn_ = nullptr;
    int n_ = 0;
    int busy_ = 0;
    uint64_t gen_ = 0;
    bool quit_ = false;
    std::atomic<int> next_{0};
    std::atomic<bool> failed_{false};
};

}  // namespace ecm
