// Host build of the library's concurrency helpers (whirplonk_amd/csrc/hostsync.h)
// under stress, for ThreadSanitizer runs (tests/test_host_sync.py):
//   * CopyPool: jobs from several caller threads at once; every part of a job
//     runs exactly once, and none after its run() has returned;
//   * WaitHub: waiters on words that writer threads raise at random times
//     (before and after the wait starts); data written before a word's
//     release is seen by the woken waiter; a wait on a word that never
//     arrives times out, and the hub keeps working afterwards.
// Prints "ok <copy jobs> <waits> <timeouts>" and exits 0 on success.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <memory>
#include <random>
#include <thread>
#include <vector>

#include "hostsync.h"

using whir_gpu::CopyPool;
using whir_gpu::WaitHub;

static std::atomic<long> bad{0};

static void fail(const char* what) {
    if (bad++ < 5) std::printf("FAIL: %s\n", what);
}

int main() {
    // ---------------------------------------------------------- CopyPool
    long jobs = 0;
    {
        // on the heap: the two helpers' mutexes never share a stack address
        auto pool_p = std::make_unique<CopyPool>(3);
        CopyPool& pool = *pool_p;
        std::vector<std::thread> callers;
        std::atomic<long> njobs{0};
        for (int t = 0; t < 4; ++t) {
            callers.emplace_back([&, t] {
                std::mt19937 rng(1234 + t);
                for (int j = 0; j < 300; ++j) {
                    const size_t n = 1 + rng() % 40;
                    std::vector<std::atomic<int>> hits(n);
                    for (auto& h : hits) h.store(0);
                    std::atomic<int> total{0};
                    pool.run(n, [&](size_t i) {
                        hits[i].fetch_add(1);
                        total.fetch_add(1);
                    });
                    for (size_t i = 0; i < n; ++i)
                        if (hits[i].load() != 1) fail("a copy part did not run exactly once");
                    const int after = total.load();
                    std::this_thread::yield();
                    if (total.load() != after || after != (int)n) fail("a copy part ran after its job returned");
                    njobs.fetch_add(1);
                }
            });
        }
        for (auto& c : callers) c.join();
        jobs = njobs.load();
    }

    // ----------------------------------------------------------- WaitHub
    long waits = 0, timeouts = 0;
    {
        auto hub_p = std::make_unique<WaitHub>();
        WaitHub& hub = *hub_p;
        constexpr int W = 6, ROUNDS = 400;
        std::vector<uint64_t> word(W, 0), ack(W, 0), payload(W, 0);
        std::vector<std::thread> threads;
        std::atomic<long> nwaits{0};
        for (int w = 0; w < W; ++w) {
            // writer: payload, then the word with release, at random pauses
            threads.emplace_back([&, w] {
                std::mt19937 rng(99 + w);
                for (uint64_t v = 1; v <= ROUNDS; ++v) {
                    const int pause = rng() % 200;
                    for (int k = 0; k < pause; ++k) whir_gpu::cpu_relax();
                    if (rng() % 16 == 0) std::this_thread::sleep_for(std::chrono::microseconds(rng() % 300));
                    payload[w] = 1000 * v + w;
                    __atomic_store_n(&word[w], v, __ATOMIC_RELEASE);
                    // the next value only after the waiter has seen this one
                    while (__atomic_load_n(&ack[w], __ATOMIC_ACQUIRE) != v) whir_gpu::cpu_relax();
                }
            });
            // waiter: each value in turn, then the acknowledgement
            threads.emplace_back([&, w] {
                for (uint64_t v = 1; v <= ROUNDS; ++v) {
                    if (!hub.wait(&word[w], v, std::chrono::seconds(20))) fail("a wait missed its word");
                    if (payload[w] != 1000 * v + w) fail("data written before the word was not seen");
                    nwaits.fetch_add(1);
                    __atomic_store_n(&ack[w], v, __ATOMIC_RELEASE);
                }
            });
        }
        // waits that time out, interleaved with the others
        uint64_t never = 0;
        for (int k = 0; k < 20; ++k) {
            if (hub.wait(&never, 1, std::chrono::microseconds(200 + 50 * k))) fail("a wait on a silent word returned true");
            ++timeouts;
        }
        for (auto& t : threads) t.join();
        waits = nwaits.load();
        if (waits != (long)W * ROUNDS) fail("not every wait returned");
        // the hub still works after the timeouts
        uint64_t late = 0;
        std::thread raiser([&] {
            std::this_thread::sleep_for(std::chrono::milliseconds(2));
            __atomic_store_n(&late, 7, __ATOMIC_RELEASE);
        });
        if (!hub.wait(&late, 7, std::chrono::seconds(20))) fail("a wait after the timeouts missed its word");
        raiser.join();
    }
    if (bad.load()) return 1;
    std::printf("ok %ld %ld %ld\n", jobs, waits, timeouts);
    return 0;
}
