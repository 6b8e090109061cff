// SPDX-License-Identifier: Apache-2.0
// Host-side runtime pieces: HIP error checks, a parallel host copy pool and
// a pinned double-buffer ring for host <-> device transfers.
//
// Buffers handed to the library by its callers are ordinary pageable memory,
// which a DMA engine cannot address; every transfer therefore goes through
// pinned bounce memory.  A transfer is cut into chunks of the ring's size:
// while the DMA engine moves one chunk, the copy pool fills (or drains) the
// other half of the ring on several host cores, so host copies and DMA
// overlap and the host side is not limited to one core's memcpy rate.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace benes_hip {

struct HipError {
    hipError_t code;
    const char* what;
};

#define BH_CHECK(expr)                                       \
    do {                                                     \
        const hipError_t bh_e_ = (expr);                     \
        if (bh_e_ != hipSuccess) throw HipError{bh_e_, #expr}; \
    } while (0)

// A piece of a gathered or scattered transfer: `len` bytes between host
// address `host` and byte `dev` of a device range.
struct HostPiece {
    uint8_t* host;
    uint64_t dev;
    uint64_t len;
};

// memcpy over a fixed set of worker threads.  One job runs at a time (host
// memory bandwidth is the shared resource); the calling thread works too.
// Each job has its own task list and counters, held by every thread working
// on it, so a worker that comes late to a finished job finds nothing left.
class CopyPool {
public:
    static CopyPool& instance() {
        static CopyPool* pool = new CopyPool();  // never destroyed: workers outlive statics
        return *pool;
    }

    struct Task {
        void* dst;
        const void* src;
        size_t len;
    };

    // Copies every task; returns when all are done.
    void run(const std::vector<Task>& tasks) {
        size_t total = 0;
        for (const Task& t : tasks) total += t.len;
        if (total < kParallelMin || workers_.empty()) {
            for (const Task& t : tasks) std::memcpy(t.dst, t.src, t.len);
            return;
        }
        auto job = std::make_shared<Job>();
        for (const Task& t : tasks) {  // split large tasks so every worker gets a share
            for (size_t off = 0; off < t.len; off += kSlice) {
                const size_t len = std::min(kSlice, t.len - off);
                job->tasks.push_back(Task{static_cast<uint8_t*>(t.dst) + off,
                                          static_cast<const uint8_t*>(t.src) + off, len});
            }
        }
        job->remaining = job->tasks.size();
        std::lock_guard<std::mutex> job_lock(job_mu_);
        {
            std::lock_guard<std::mutex> lock(mu_);
            current_ = job;
            ++generation_;
        }
        cv_.notify_all();
        work(job);
        std::unique_lock<std::mutex> lock(mu_);
        done_cv_.wait(lock, [&] { return job->remaining == 0; });
        current_.reset();
    }

private:
    struct Job {
        std::vector<Task> tasks;
        std::atomic<size_t> next{0};
        size_t remaining = 0;  // guarded by mu_
    };

    static constexpr size_t kSlice = size_t{2} << 20;
    static constexpr size_t kParallelMin = size_t{4} << 20;

    CopyPool() {
        const unsigned hw = std::thread::hardware_concurrency();
        const unsigned n = std::min(8u, hw > 2 ? hw - 2 : 0u);
        for (unsigned i = 0; i < n; ++i) {
            workers_.emplace_back([this] { loop(); });
            workers_.back().detach();
        }
    }

    // Claims and copies tasks of `job` until none is left.
    void work(const std::shared_ptr<Job>& job) {
        size_t done = 0;
        for (size_t i = job->next.fetch_add(1); i < job->tasks.size();
             i = job->next.fetch_add(1)) {
            const Task& t = job->tasks[i];
            std::memcpy(t.dst, t.src, t.len);
            ++done;
        }
        if (done) {
            std::lock_guard<std::mutex> lock(mu_);
            job->remaining -= done;
            if (job->remaining == 0) done_cv_.notify_all();
        }
    }

    void loop() {
        uint64_t seen = 0;
        while (true) {
            std::shared_ptr<Job> job;
            {
                std::unique_lock<std::mutex> lock(mu_);
                cv_.wait(lock, [&] { return generation_ != seen && current_ != nullptr; });
                seen = generation_;
                job = current_;
            }
            work(job);
        }
    }

    std::vector<std::thread> workers_;
    std::mutex job_mu_, mu_;
    std::condition_variable cv_, done_cv_;
    std::shared_ptr<Job> current_;
    uint64_t generation_ = 0;
};

// Two pinned chunks and their events, owned by one stream's user.
class PinnedRing {
public:
    static constexpr uint64_t kChunk = uint64_t{32} << 20;

    ~PinnedRing() {
        for (int b = 0; b < 2; ++b) {
            if (ev_[b]) (void)hipEventDestroy(ev_[b]);
            if (buf_[b]) (void)hipHostFree(buf_[b]);
        }
    }

    // Host pieces -> device range starting at `dev_base`, queued on `s`.
    // Pieces are sorted by device offset and do not overlap; gaps between
    // them are copied as whatever the ring held.  Returns once every host
    // byte has been copied out, so the host buffers may be reused.
    void upload(uint8_t* dev_base, const std::vector<HostPiece>& pieces, hipStream_t s) {
        if (pieces.empty()) return;
        init();
        const uint64_t lo = pieces.front().dev;
        const uint64_t hi = pieces.back().dev + pieces.back().len;
        size_t p = 0;
        for (uint64_t c0 = lo, k = 0; c0 < hi; c0 += kChunk, ++k) {
            const int b = static_cast<int>(k & 1);
            const uint64_t c1 = std::min(hi, c0 + kChunk);
            wait(b);
            std::vector<CopyPool::Task> tasks;
            while (p < pieces.size() && pieces[p].dev + pieces[p].len <= c0) ++p;
            for (size_t q = p; q < pieces.size() && pieces[q].dev < c1; ++q) {
                const uint64_t a = std::max(c0, pieces[q].dev);
                const uint64_t e = std::min(c1, pieces[q].dev + pieces[q].len);
                tasks.push_back({buf_[b] + (a - c0), pieces[q].host + (a - pieces[q].dev), e - a});
            }
            CopyPool::instance().run(tasks);
            BH_CHECK(hipMemcpyAsync(dev_base + c0, buf_[b], c1 - c0, hipMemcpyHostToDevice, s));
            BH_CHECK(hipEventRecord(ev_[b], s));
            pending_[b] = true;
        }
    }

    // Device range -> host pieces, after the work already queued on `s`.
    // Returns when every piece holds its bytes.
    void download(const uint8_t* dev_base, const std::vector<HostPiece>& pieces, hipStream_t s) {
        if (pieces.empty()) return;
        init();
        const uint64_t lo = pieces.front().dev;
        const uint64_t hi = pieces.back().dev + pieces.back().len;
        const uint64_t chunks = (hi - lo + kChunk - 1) / kChunk;
        auto issue = [&](uint64_t k) {
            const int b = static_cast<int>(k & 1);
            const uint64_t c0 = lo + k * kChunk, c1 = std::min(hi, c0 + kChunk);
            wait(b);
            BH_CHECK(hipMemcpyAsync(buf_[b], dev_base + c0, c1 - c0, hipMemcpyDeviceToHost, s));
            BH_CHECK(hipEventRecord(ev_[b], s));
            pending_[b] = true;
        };
        issue(0);
        size_t p = 0;
        for (uint64_t k = 0; k < chunks; ++k) {
            if (k + 1 < chunks) issue(k + 1);
            const int b = static_cast<int>(k & 1);
            const uint64_t c0 = lo + k * kChunk, c1 = std::min(hi, c0 + kChunk);
            wait(b);
            std::vector<CopyPool::Task> tasks;
            while (p < pieces.size() && pieces[p].dev + pieces[p].len <= c0) ++p;
            for (size_t q = p; q < pieces.size() && pieces[q].dev < c1; ++q) {
                const uint64_t a = std::max(c0, pieces[q].dev);
                const uint64_t e = std::min(c1, pieces[q].dev + pieces[q].len);
                tasks.push_back({pieces[q].host + (a - pieces[q].dev), buf_[b] + (a - c0), e - a});
            }
            CopyPool::instance().run(tasks);
        }
    }

private:
    void init() {
        if (buf_[0]) return;
        for (int b = 0; b < 2; ++b) {
            BH_CHECK(hipHostMalloc(reinterpret_cast<void**>(&buf_[b]), kChunk,
                                   hipHostMallocDefault));
            BH_CHECK(hipEventCreateWithFlags(&ev_[b], hipEventDisableTiming));
        }
    }

    // Waits until chunk buffer b is no longer read or written by a DMA.
    void wait(int b) {
        if (!pending_[b]) return;
        BH_CHECK(hipEventSynchronize(ev_[b]));
        pending_[b] = false;
    }

    uint8_t* buf_[2] = {nullptr, nullptr};
    hipEvent_t ev_[2] = {nullptr, nullptr};
    bool pending_[2] = {false, false};
};

}  // namespace benes_hip
