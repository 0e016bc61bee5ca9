// dyc_threads.hpp -- the library's host worker threads, free of HIP so that a host-only build
// (and the sanitizer builds of the tests) can run them: the deferred-free helper and the copy
// pool of the host-array calls.
#pragma once

#include <stddef.h>
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

namespace dyc {
namespace host {

// Runs Free(device, p) for each handed-over block on a helper thread. The helper starts on
// first use; drain() waits until every handed-over block is freed and stops it (defer() after
// drain() starts it again).
template <typename Free>
class ReclaimerT {
  public:
    ReclaimerT() = default;
    ReclaimerT(const ReclaimerT &) = delete;
    ReclaimerT &operator=(const ReclaimerT &) = delete;
    ~ReclaimerT() { drain(); }

    void defer(int device, void *p)
    {
        if (!p) return;
        std::lock_guard<std::mutex> lk(m_);
        q_.emplace_back(device, p);
        if (!th_.joinable()) th_ = std::thread([this] { loop(); });
        cv_.notify_one();
    }

    void drain()
    {
        {
            std::lock_guard<std::mutex> lk(m_);
            if (!th_.joinable()) return;
            stop_ = true;
        }
        cv_.notify_one();
        th_.join();
        std::lock_guard<std::mutex> lk(m_);
        th_ = std::thread();
        stop_ = false;
    }

  private:
    void loop()
    {
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            cv_.wait(lk, [this] { return stop_ || !q_.empty(); });
            if (q_.empty()) return;  // stop_ with nothing left
            const std::pair<int, void *> job = q_.front();
            q_.pop_front();
            lk.unlock();
            Free()(job.first, job.second);
            lk.lock();
        }
    }

    std::mutex m_;
    std::condition_variable cv_;
    std::deque<std::pair<int, void *>> q_;
    std::thread th_;
    bool stop_ = false;
};

// Splits a large host memcpy (pageable <-> pinned staging) over worker threads that live as long
// as the staging buffers: starting threads for every copy cost a measurable share of each
// pipelined chunk. copy() returns when every part is copied; one copy() at a time.
class CopyPool {
  public:
    explicit CopyPool(unsigned workers)
    {
        for (unsigned i = 0; i < workers; ++i) th_.emplace_back([this] { loop(); });
    }
    CopyPool(const CopyPool &) = delete;
    CopyPool &operator=(const CopyPool &) = delete;
    ~CopyPool()
    {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }

    unsigned workers() const { return (unsigned)th_.size(); }

    void copy(void *dst, const void *src, size_t bytes)
    {
        constexpr size_t kSplit = 4u << 20;
        const size_t nparts = th_.size() + 1;
        if (bytes < kSplit || nparts == 1) {
            memcpy(dst, src, bytes);
            return;
        }
        const size_t per = ((bytes + nparts - 1) / nparts + 63) & ~(size_t)63;
        {
            std::lock_guard<std::mutex> lk(m_);
            parts_.clear();
            for (size_t lo = per; lo < bytes; lo += per)
                parts_.push_back(Part{static_cast<char *>(dst) + lo, static_cast<const char *>(src) + lo,
                                      std::min(per, bytes - lo)});
            next_ = 0;
            left_ = parts_.size();
        }
        cv_.notify_all();
        memcpy(dst, src, std::min(per, bytes));  // the first part on the calling thread
        std::unique_lock<std::mutex> lk(m_);
        done_cv_.wait(lk, [this] { return left_ == 0; });
        parts_.clear();
    }

  private:
    struct Part {
        char *dst;
        const char *src;
        size_t len;
    };

    void loop()
    {
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            cv_.wait(lk, [this] { return stop_ || next_ < parts_.size(); });
            if (stop_) return;
            const Part p = parts_[next_++];
            lk.unlock();
            memcpy(p.dst, p.src, p.len);
            lk.lock();
            if (--left_ == 0) done_cv_.notify_one();
        }
    }

    std::mutex m_;
    std::condition_variable cv_, done_cv_;
    std::vector<Part> parts_;
    size_t next_ = 0, left_ = 0;
    bool stop_ = false;
    std::vector<std::thread> th_;
};

}  // namespace host
}  // namespace dyc
