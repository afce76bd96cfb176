// fg_host.h -- a small worker pool for the host-side structure builders
// (CSR and edge patterns, reverse Cuthill-McKee).  These run once per mesh
// on integer data; at 10^7 elements a serial pass takes seconds, so the
// builders split their passes over the host cores.  Every parallel pass is
// written so that its result does not depend on the number of threads or on
// their timing.
#pragma once

#include <stdlib.h>

#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

// worker count: FG_HOST_THREADS if set (1..256), else the hardware threads,
// at most 16
inline int fg_host_threads()
{
    if (const char *s = getenv("FG_HOST_THREADS")) {
        int n = atoi(s);
        if (n >= 1 && n <= 256) return n;
    }
    unsigned h = std::thread::hardware_concurrency();
    if (h == 0) h = 1;
    return h > 16 ? 16 : (int)h;
}

class FgPool {
public:
    FgPool() : n_(fg_host_threads())
    {
        for (int t = 1; t < n_; t++) th_.emplace_back([this, t] { loop(t); });
    }
    ~FgPool()
    {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
            gen_++;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    FgPool(const FgPool &) = delete;
    FgPool &operator=(const FgPool &) = delete;

    int size() const { return n_; }

    // f(tid) on every worker, tid 0 on the calling thread; returns when all
    // are done (the return orders every write of f before what follows)
    void run(const std::function<void(int)> &f)
    {
        if (n_ == 1) {
            f(0);
            return;
        }
        {
            std::lock_guard<std::mutex> g(m_);
            job_ = &f;
            done_ = 0;
            gen_++;
        }
        cv_.notify_all();
        f(0);
        std::unique_lock<std::mutex> l(m_);
        cv_done_.wait(l, [&] { return done_ == n_ - 1; });
        job_ = nullptr;
    }

    // [0, n) in contiguous chunks, one per worker: f(begin, end, tid).
    // Below `serial` items the calling thread does it all.
    template <class F>
    void range(long long n, F &&f, long long serial = 4096)
    {
        if (n <= 0) return;
        if (n < serial || n_ == 1) {
            f(0LL, n, 0);
            return;
        }
        run([&](int t) {
            const long long b = n * t / n_, e = n * (t + 1) / n_;
            if (b < e) f(b, e, t);
        });
    }

private:
    void loop(int t)
    {
        unsigned long long seen = 0;
        for (;;) {
            const std::function<void(int)> *job;
            {
                std::unique_lock<std::mutex> l(m_);
                cv_.wait(l, [&] { return gen_ != seen; });
                seen = gen_;
                if (stop_) return;
                job = job_;
            }
            (*job)(t);
            {
                std::lock_guard<std::mutex> g(m_);
                done_++;
            }
            cv_done_.notify_one();
        }
    }

    int n_;
    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable cv_, cv_done_;
    const std::function<void(int)> *job_ = nullptr;
    unsigned long long gen_ = 0;
    int done_ = 0;
    bool stop_ = false;
};

// exclusive prefix sum in place; returns the total
template <typename T>
inline T fg_exclusive_scan(std::vector<T> &v)
{
    T acc = 0;
    for (auto &x : v) {
        const T c = x;
        x = acc;
        acc += c;
    }
    return acc;
}
