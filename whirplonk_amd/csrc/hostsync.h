// hostsync.h — host-side concurrency helpers of the library (no HIP
// dependency, so the CPU tests build this header alone, under
// ThreadSanitizer):
//   * CopyPool: a few host threads that split one copy job into parts;
//   * WaitHub: waits that sleep instead of spinning, woken by one poller
//     thread that reads the waiters' completion words.
#pragma once

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace whir_gpu {

inline void cpu_relax() { __builtin_ia32_pause(); }

// run(n, fn) calls fn(0 .. n-1) on the pool's threads and the calling thread,
// and returns when every part is done and every worker that joined the job
// has left it (so no worker can claim a part of the next job with this job's
// function).  One job at a time; concurrent callers take turns.
class CopyPool {
public:
    explicit CopyPool(unsigned workers) {
        for (unsigned i = 0; i < workers; ++i) workers_.emplace_back([this] { loop(); });
    }
    // the library's pool: up to 7 workers, half the host's hardware threads
    static CopyPool& get() {
        static CopyPool pool([] {
            const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
            return std::min(7u, hw > 2 ? hw / 2 - 1 : 0u);
        }());
        return pool;
    }
    CopyPool(const CopyPool&) = delete;
    CopyPool& operator=(const CopyPool&) = delete;

    void run(size_t n, const std::function<void(size_t)>& fn) {
        std::lock_guard<std::mutex> one(job_mu_);
        {
            std::lock_guard<std::mutex> lk(mu_);
            fn_ = &fn;
            n_ = n;
            next_.store(0);
            done_ = 0;
            ++gen_;
        }
        cv_.notify_all();
        size_t mine = 0;
        for (size_t i; (i = next_.fetch_add(1)) < n;) {
            fn(i);
            ++mine;
        }
        std::unique_lock<std::mutex> lk(mu_);
        done_ += mine;
        done_cv_.wait(lk, [&] { return done_ == n_ && active_ == 0; });
        fn_ = nullptr;  // a worker that wakes later does not join this job
    }

    ~CopyPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : workers_) t.join();
    }

private:
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(size_t)>* fn;
            size_t n;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || (gen_ != seen && fn_); });
                if (stop_) return;
                seen = gen_;
                fn = fn_;
                n = n_;
                ++active_;
            }
            size_t mine = 0;
            for (size_t i; (i = next_.fetch_add(1)) < n;) {
                (*fn)(i);
                ++mine;
            }
            std::lock_guard<std::mutex> lk(mu_);
            done_ += mine;
            --active_;
            if (done_ == n_ && active_ == 0) done_cv_.notify_all();
        }
    }

    std::vector<std::thread> workers_;
    std::mutex job_mu_, mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(size_t)>* fn_ = nullptr;
    size_t n_ = 0, done_ = 0, active_ = 0;
    std::atomic<size_t> next_{0};
    uint64_t gen_ = 0;
    bool stop_ = false;
};

// Waits that sleep instead of spinning, for hosts with more waiting threads
// than cores (a serving loop's slots): one poller thread reads the completion
// word of every registered wait and wakes each waiter, through its own
// condition variable, once its word has reached the wanted value.  The
// poller sleeps while nothing is registered.  Words are read with acquire
// loads, so what their writer released before the value is visible to the
// woken waiter.
class WaitHub {
public:
    WaitHub() : poller_([this] { poll(); }) {}
    ~WaitHub() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        poller_.join();
    }
    WaitHub(const WaitHub&) = delete;
    WaitHub& operator=(const WaitHub&) = delete;

    // true once *word >= want, false if that was not seen within `timeout`
    bool wait(const uint64_t* word, uint64_t want, std::chrono::microseconds timeout) {
        Waiter w;
        w.word = word;
        w.want = want;
        std::unique_lock<std::mutex> lk(mu_);
        if (__atomic_load_n(word, __ATOMIC_ACQUIRE) >= want) return true;
        waits_.push_back(&w);
        cv_.notify_all();  // the poller, if it sleeps
        if (w.cv.wait_for(lk, timeout, [&] { return w.done; })) return true;
        waits_.erase(std::find(waits_.begin(), waits_.end(), &w));  // still registered: the poller did not see it
        return __atomic_load_n(word, __ATOMIC_ACQUIRE) >= want;
    }

private:
    struct Waiter {
        const uint64_t* word;
        uint64_t want;
        bool done = false;
        std::condition_variable cv;
    };

    void poll() {
        std::unique_lock<std::mutex> lk(mu_);
        while (!stop_) {
            if (waits_.empty()) {
                cv_.wait(lk);
                continue;
            }
            for (size_t i = 0; i < waits_.size();) {
                Waiter* w = waits_[i];
                if (__atomic_load_n(w->word, __ATOMIC_ACQUIRE) >= w->want) {
                    w->done = true;
                    w->cv.notify_one();
                    waits_[i] = waits_.back();
                    waits_.pop_back();
                } else {
                    ++i;
                }
            }
            // let waiters register and the woken ones take the lock
            lk.unlock();
            for (int k = 0; k < 16; ++k) cpu_relax();
            lk.lock();
        }
    }

    std::mutex mu_;
    std::condition_variable cv_;
    std::vector<Waiter*> waits_;
    bool stop_ = false;
    std::thread poller_;  // last member: it starts after the others exist
};

}  // namespace whir_gpu
