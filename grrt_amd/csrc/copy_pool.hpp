// A small pool of host threads for large host-to-host copies (a frame from
// the pinned staging buffer into the caller's pageable buffer).  One copy
// runs at a time; the calling thread takes slices too.
#pragma once
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace grrt_amd {

class CopyPool {
public:
    explicit CopyPool(int workers) {
        for (int i = 0; i < workers; ++i) threads_.emplace_back([this] { loop(); });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : threads_) t.join();
    }
    CopyPool(const CopyPool&) = delete;
    CopyPool& operator=(const CopyPool&) = delete;

    void copy(void* dst, const void* src, size_t bytes) {
        if (threads_.empty() || bytes < kSlice) {
            std::memcpy(dst, src, bytes);
            return;
        }
        {
            std::lock_guard<std::mutex> g(m_);
            dst_ = static_cast<char*>(dst);
            src_ = static_cast<const char*>(src);
            bytes_ = bytes;
            next_.store(0);
            busy_ = static_cast<int>(threads_.size());
            ++gen_;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [this] { return busy_ == 0; });
    }

private:
    static constexpr size_t kSlice = 1 << 20;

    void work() {
        while (true) {
            const size_t o = next_.fetch_add(kSlice);
            if (o >= bytes_) break;
            std::memcpy(dst_ + o, src_ + o, std::min(kSlice, bytes_ - o));
        }
    }
    void loop() {
        unsigned long long seen = 0;
        while (true) {
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            work();
            {
                std::lock_guard<std::mutex> g(m_);
                if (--busy_ == 0) done_.notify_one();
            }
        }
    }

    std::vector<std::thread> threads_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t bytes_ = 0;
    std::atomic<size_t> next_{0};
    int busy_ = 0;
    unsigned long long gen_ = 0;
    bool stop_ = false;
};

}  // namespace grrt_amd
