// ts_copy.h -- host-side parallel memcpy between user (pageable) memory and
// the pinned staging buffers.  The host-array entry points move every byte
// twice (user memory <-> pinned staging <-> HBM); one thread's memcpy runs
// well below what PCIe carries, so a copy is split across a few worker
// threads that live as long as the handle.
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <string.h>
#include <condition_variable>
#include <cstdlib>
#include <mutex>
#include <thread>
#include <vector>

namespace ts {

class CopyPool {
  public:
    // nthreads: threads that copy, the caller included (1 = plain memcpy)
    explicit CopyPool(int nthreads)
    {
        nthreads_ = nthreads < 1 ? 1 : (nthreads > 32 ? 32 : nthreads);
        for (int i = 1; i < nthreads_; i++) workers_.emplace_back([this, i] { loop(i); });
    }
    ~CopyPool()
    {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : workers_) t.join();
    }
    CopyPool(const CopyPool &) = delete;
    CopyPool &operator=(const CopyPool &) = delete;

    int threads() const { return nthreads_; }

    // memcpy(dst, src, len) with the work split in nthreads parts of whole
    // 4 KiB pages; small copies stay on the caller.  One copy at a time per
    // pool (the handle is not shared between threads).
    void copy(void *dst, const void *src, size_t len)
    {
        constexpr size_t SMALL = 1u << 20;
        if (nthreads_ == 1 || len < SMALL) {
            memcpy(dst, src, len);
            return;
        }
        {
            std::lock_guard<std::mutex> g(m_);
            dst_ = (char *)dst;
            src_ = (const char *)src;
            len_ = len;
            pending_ = nthreads_ - 1;
            gen_++;
        }
        cv_.notify_all();
        part(0);
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [this] { return pending_ == 0; });
    }

    // threads for a new pool: $TURTLE_COPY_THREADS, else 8, at most the
    // hardware's thread count
    static int default_threads()
    {
        int n = 8;
        if (const char *e = std::getenv("TURTLE_COPY_THREADS")) n = std::atoi(e);
        const int hw = (int)std::thread::hardware_concurrency();
        if (hw > 0 && n > hw) n = hw;
        return n < 1 ? 1 : n;
    }

  private:
    void part(int i)
    {
        const size_t pages = (len_ + 4095) / 4096;
        const size_t per = (pages + nthreads_ - 1) / nthreads_;
        const size_t a = (size_t)i * per * 4096, b = a + per * 4096;
        const size_t e = b < len_ ? b : len_;
        if (a < e) memcpy(dst_ + a, src_ + a, e - a);
    }
    void loop(int i)
    {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            part(i);
            {
                std::lock_guard<std::mutex> g(m_);
                if (--pending_ == 0) done_.notify_one();
            }
        }
    }

    int nthreads_ = 1;
    std::vector<std::thread> workers_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    bool stop_ = false;
    uint64_t gen_ = 0;
    int pending_ = 0;
    char *dst_ = nullptr;
    const char *src_ = nullptr;
    size_t len_ = 0;
};

}  // namespace ts
