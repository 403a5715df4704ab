// A small fixed pool of host threads for the factorization: run(f) calls
// f(t) once on every thread t in [0, size()) -- the caller is thread 0 --
// and returns when all calls have returned.  Work split over the pool is
// always by a fixed rule (contiguous ranges, owner = id % size), so results
// never depend on timing.
#pragma once
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace ss {

class HostPool {
public:
    explicit HostPool(int threads) : n_(threads < 1 ? 1 : threads) {
        for (int t = 1; t < n_; t++) th_.emplace_back([this, t] { worker(t); });
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &x : th_) x.join();
    }
    HostPool(const HostPool &) = delete;
    HostPool &operator=(const HostPool &) = delete;

    int size() const { return n_; }

    // f(t) on every thread; `active` < size() runs only threads [0, active).
    void run(const std::function<void(int)> &f, int active = 0) {
        const int a = active <= 0 || active > n_ ? n_ : active;
        if (a == 1) {
            f(0);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &f;
            active_ = a;
            pending_ = a - 1;
            gen_++;
        }
        cv_.notify_all();
        f(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [this] { return pending_ == 0; });
        job_ = nullptr;
    }

    // [lo, hi) of thread t when `count` items are split in `parts` ranges.
    static void range(int64_t count, int parts, int t, int64_t &lo, int64_t &hi) {
        lo = count * t / parts;
        hi = count * (t + 1) / parts;
    }

    // Threads for host factorization: SPEC_SPARS_THREADS if set, else the
    // machine's hardware threads capped at 32 (the OpenMP team the
    // reference's factorization uses, src/example.cc:672).
    static int default_threads() {
        if (const char *e = std::getenv("SPEC_SPARS_THREADS")) {
            const int v = std::atoi(e);
            if (v >= 1) return v > 256 ? 256 : v;
        }
        const unsigned hw = std::thread::hardware_concurrency();
        return hw == 0 ? 1 : (hw > 32 ? 32 : (int)hw);
    }

private:
    void worker(int t) {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(int)> *f = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                if (t >= active_) continue;
                f = job_;
            }
            (*f)(t);
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (--pending_ == 0) done_.notify_one();
            }
        }
    }

    int n_;
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    const std::function<void(int)> *job_ = nullptr;
    uint64_t gen_ = 0;
    int active_ = 0, pending_ = 0;
    bool stop_ = false;
};

// A barrier for the threads of one HostPool::run whose phases are short:
// the last thread to arrive bumps the generation the others spin on (they
// yield after a while, in case the machine is oversubscribed).
class SpinBarrier {
public:
    explicit SpinBarrier(int n) : n_(n) {}
    void wait() {
        if (n_ <= 1) return;
        const unsigned g = gen_.load(std::memory_order_acquire);
        if (count_.fetch_add(1, std::memory_order_acq_rel) == n_ - 1) {
            count_.store(0, std::memory_order_relaxed);
            gen_.fetch_add(1, std::memory_order_acq_rel);
            return;
        }
        for (int spins = 0; gen_.load(std::memory_order_acquire) == g;)
            if (++spins > 4096) std::this_thread::yield();
    }

private:
    const int n_;
    std::atomic<int> count_{0};
    std::atomic<unsigned> gen_{0};
};

}  // namespace ss
