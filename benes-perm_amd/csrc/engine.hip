// SPDX-License-Identifier: Apache-2.0
// Host orchestration of the compact-v1 compression path on one MI355X, and
// the C ABI declared in include/benes_hip.h.
//
// A call compresses a batch (a plain compress is a batch of one): every item
// gets a slice of shared HBM arrays, and each step below is one launch over
// all items that take part in it, driven by a table of per-item segments:
//   validation + inverse  -> per depth above the fusion depth:
//   describe -> union -> route inputs -> route outputs (or one LDS kernel per
//   node up to 2^16 entries) -> pack both sides  -> one subtree launch.
// The host builds the per-item Plans and all segment tables up front, uploads
// them in one copy, queues everything on one stream and reads back one error
// word per item at the end.
#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <map>
#include <thread>
#include <unordered_map>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/benes_hip.h"
#include "host_runtime.hpp"
#include "keyed.hpp"
#include "kernels.hpp"
#include "plan.hpp"
#include "status.hpp"

namespace benes_hip {
namespace {

thread_local double g_device_seconds = 0;

// First-hook binning: -1 by size (the default), 0 never, 1 wherever a node's
// words overflow an XCD's L2 (a test setting).
std::atomic<int> g_hook_binning{-1};

// 64-bit union-find words and level-0 inverse: 0 by size (the default: words
// where a node has 2^31 slots or more, the inverse from 2^32 - 1 entries),
// 1 at every HBM level of every call (a test setting).
std::atomic<int> g_wide_words{0};

// Union pass of the HBM levels whose nodes keep to one XCD: -1 by size (the
// default: plain-store hooks, checked afterwards, where a node holds at most
// kPlainUnionMaxBytes of words), 0 compare-and-swap hooks everywhere, 1
// plain-store hooks at every such depth (a test setting).
std::atomic<int> g_plain_union{-1};
// Measured on uniform permutations of 2^26 uint32 and 2^32 + 12345 uint64
// entries: plain-store hooks pay for their check from nodes of 8 MB of words
// down (2^26: depths 4-9, 27.8 -> 25.7 ms per compress; 16 MB nodes break
// even), with each XCD uniting nodes of about 2 MB of words at a time (whole
// L2-resident nodes of 256 KB with all of an XCD's blocks on one node lose a
// tenth of their hooks to races; interleaving more nodes lets their words
// leave L2).  The bound leaves room for the few extra slots of a node of a
// size that is not a power of two (2^32 + 12345 entries: depth 10's nodes
// hold 2^21 + 6 slots; with them 3.59 -> 3.55 s per compress).
constexpr uint64_t kPlainUnionMaxBytes = uint64_t{9} << 20;

constexpr uint64_t kPlainUnionGroupBytes = uint64_t{2} << 20;

// Colouring of the HBM levels: 0 the union-find (the default), 1 the ruler
// orbit minima (an independent second colouring; the blob is the same).
std::atomic<int> g_colouring{0};

// Optional per-kernel timing: an event pair around every launch, summed per
// kernel name once the call's stream has drained.
struct KernelStat {
    std::string name;
    double ms = 0;
    uint64_t launches = 0;
};
std::mutex g_prof_mu;
bool g_profiling = false;
std::vector<KernelStat> g_prof_stats;

struct Pending {
    const char* name;
    hipEvent_t begin, end;
};

// A fixed range of device memory handed out front to back, reset before each
// call that draws on it (the allocation-free entry point's scratch).
struct Arena {
    uint8_t* base = nullptr;
    size_t cap = 0, used = 0;
    void* take(size_t need) {
        const size_t at = (used + 255) & ~size_t{255};
        if (at > cap || need > cap - at) {
            throw std::length_error("request needs more scratch than the allocation-free entry "
                                    "point holds");
        }
        used = at + need;
        return base + at;
    }
};

// Grow-only device buffer, or (bound to an arena) a slice of the arena taken
// anew by every ensure().
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    Arena* arena = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (ptr && !arena) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    void ensure(size_t need) {
        if (arena) {
            ptr = arena->take(need);
            bytes = need;
            return;
        }
        if (need <= bytes) return;
        if (ptr) BH_CHECK(hipFree(ptr));
        ptr = nullptr;
        bytes = 0;
        BH_CHECK(hipMalloc(&ptr, need));
        bytes = need;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
};

// Device scratch of one stream, reused across calls.
struct Workspace {
    std::mutex mu;
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t ev_begin = nullptr, ev_end = nullptr;
    DevBuf input, level[2], ilevel[2], inv0, uf, srows, trows, desc, plans, items, segs, errs,
        blob, deferred, deferred_counts, bin_fill, rulers, lbits, lost_counts;
    std::vector<int> host_errs;
    PinnedRing ring;  // host <-> device staging of this workspace's stream
    Workspace() = default;
    Workspace(const Workspace&) = delete;
    Workspace& operator=(const Workspace&) = delete;
    ~Workspace() {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        if (ev_begin) (void)hipEventDestroy(ev_begin);
        if (ev_end) (void)hipEventDestroy(ev_end);
    }
    void init(int dev) {
        if (stream) return;
        device = dev;
        BH_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        BH_CHECK(hipEventCreate(&ev_begin));
        BH_CHECK(hipEventCreate(&ev_end));
    }
    // Every device buffer draws on `a` from now on (no stream, events or
    // pinned memory are made for such a workspace).
    void bind(Arena* a) {
        for (DevBuf* b : {&input, &level[0], &level[1], &ilevel[0], &ilevel[1], &inv0, &uf, &srows,
                          &trows, &desc, &plans, &items, &segs, &errs, &blob, &deferred,
                          &deferred_counts, &bin_fill, &rulers, &lbits, &lost_counts}) {
            b->release();
            b->arena = a;
        }
    }
};

Workspace& workspace_for(int dev) {
    static std::mutex mu;
    static std::vector<Workspace*> all;
    std::lock_guard<std::mutex> lock(mu);
    if (static_cast<size_t>(dev) >= all.size()) all.resize(dev + 1, nullptr);
    if (!all[dev]) all[dev] = new Workspace();
    return *all[dev];
}

uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
uint64_t round8(uint64_t a) { return (a + 7) & ~uint64_t{7}; }

// Subtree shapes the subtree kernel is built for: entries log2, threads.
struct SubtreeShape {
    uint32_t log2;
    uint32_t threads;
};
constexpr SubtreeShape kShapes[] = {{14, 1024}, {13, 1024}, {13, 512}, {12, 512}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

// First depth whose largest node fits one LDS subtree of 2^sub_log2 entries.
uint32_t fusion_depth(uint64_t n, uint32_t routing_depth, uint32_t sub_log2) {
    uint32_t d = 0;
    while (ceil_div(n, uint64_t{1} << d) > (uint64_t{1} << sub_log2)) ++d;
    return std::min(d, routing_depth);
}

// Online choice of the subtree shape.  Calls fall into classes by the log2
// of their largest item and of their total entries.  The first
// kTrialRounds * kNumShapes calls of a class are trials that cycle through the
// shapes (any shape gives the same bytes); once they have all reported, the
// class keeps the shape with the least device time over its trials for good,
// so later calls of other inputs in the same class cannot move it.
// BENES_HIP_SUBTREE=<log2>x<threads> pins a shape.
constexpr int kTrialRounds = BENES_HIP_AUTOTUNE_TRIAL_CALLS / kNumShapes;
static_assert(kTrialRounds * kNumShapes == BENES_HIP_AUTOTUNE_TRIAL_CALLS,
              "trial calls are whole rounds over the shapes");

class Tuner {
public:
    static Tuner& instance() {
        static Tuner* t = new Tuner();
        return *t;
    }
    template <typename Jobs>
    static uint32_t class_of(const Jobs& jobs) {
        uint64_t total = 0, most = 0;
        for (const auto& j : jobs) {
            total += j.n;
            most = std::max(most, j.n);
        }
        auto lg = [](uint64_t v) { return v ? 63u - static_cast<uint32_t>(__builtin_clzll(v)) : 0u; };
        return lg(most) * 64 + lg(total);
    }
    // Pins shape `i`, or returns to tuning when i < 0.
    void pin(int i) { forced_.store(i); }
    // The shape for the next call of class `cls`, and whether it is a trial.
    int pick(uint32_t cls, bool& trial) {
        trial = false;
        const int forced = forced_.load();
        if (forced >= 0) return forced;
        std::lock_guard<std::mutex> lock(mu_);
        Entry& e = table_[cls];
        if (e.issued < BENES_HIP_AUTOTUNE_TRIAL_CALLS) {
            trial = true;
            return e.issued++ % kNumShapes;
        }
        return best(e);
    }
    // The shape for a call that runs no trial (it is not timed): the pinned
    // shape, the class's choice once its trials are done, else the first.
    int settled(uint32_t cls) {
        const int forced = forced_.load();
        if (forced >= 0) return forced;
        std::lock_guard<std::mutex> lock(mu_);
        const auto it = table_.find(cls);
        if (it == table_.end() || it->second.reported < BENES_HIP_AUTOTUNE_TRIAL_CALLS) return 0;
        return best(it->second);
    }
    // Device time of a trial call; other calls are not recorded.
    void report(uint32_t cls, int shape, bool trial, double device_ms) {
        if (!trial || device_ms <= 0) return;
        std::lock_guard<std::mutex> lock(mu_);
        Entry& e = table_[cls];
        if (e.ms[shape] == 0 || device_ms < e.ms[shape]) e.ms[shape] = device_ms;
        ++e.reported;
    }
    std::string describe() {
        std::lock_guard<std::mutex> lock(mu_);
        std::string out;
        char line[256];
        const int forced = forced_.load();
        if (forced >= 0) {
            std::snprintf(line, sizeof(line), "pinned %ux%u\n", kShapes[forced].log2,
                          kShapes[forced].threads);
            out += line;
        }
        for (const auto& kv : table_) {
            const Entry& e = kv.second;
            std::snprintf(line, sizeof(line), "largest 2^%u total 2^%u:", kv.first / 64,
                          kv.first % 64);
            out += line;
            for (int i = 0; i < kNumShapes; ++i) {
                std::snprintf(line, sizeof(line), " %ux%u=%.3fms", kShapes[i].log2,
                              kShapes[i].threads, e.ms[i]);
                out += line;
            }
            const int b = best(e);
            if (e.reported < BENES_HIP_AUTOTUNE_TRIAL_CALLS) {
                std::snprintf(line, sizeof(line), " (trials %d of %d)\n", e.reported,
                              BENES_HIP_AUTOTUNE_TRIAL_CALLS);
            } else {
                std::snprintf(line, sizeof(line), " -> %ux%u\n", kShapes[b].log2,
                              kShapes[b].threads);
            }
            out += line;
        }
        return out;
    }

private:
    struct Entry {
        int issued = 0;
        int reported = 0;
        double ms[kNumShapes] = {};
    };
    Tuner() {
        if (const char* v = std::getenv("BENES_HIP_SUBTREE")) {
            unsigned lg = 0, th = 0;
            if (std::sscanf(v, "%ux%u", &lg, &th) == 2) {
                for (int i = 0; i < kNumShapes; ++i) {
                    if (kShapes[i].log2 == lg && kShapes[i].threads == th) forced_.store(i);
                }
            }
        }
    }
    static int best(const Entry& e) {
        int b = 0;
        for (int i = 1; i < kNumShapes; ++i) {
            if (e.ms[i] > 0 && (e.ms[b] == 0 || e.ms[i] < e.ms[b])) b = i;
        }
        return b;
    }
    std::mutex mu_;
    std::map<uint32_t, Entry> table_;
    std::atomic<int> forced_{-1};
};

struct Profiler {
    bool on = false;
    hipStream_t stream = nullptr;
    std::vector<Pending> pending;
    // Device counters read after the call, added to the table as launches.
    struct Counter {
        const char* name;
        const unsigned long long* dev;
        size_t words;
    };
    std::vector<Counter> counters;
    void count(const char* name, const unsigned long long* dev, size_t words) {
        if (on) counters.push_back(Counter{name, dev, words});
    }
    // A host-side count, added to the table as launches.
    std::vector<std::pair<const char*, uint64_t>> host_counts;
    void note(const char* name, uint64_t n) {
        if (on) host_counts.emplace_back(name, n);
    }
    void begin(const char* name) {
        if (!on) return;
        Pending p{name, nullptr, nullptr};
        BH_CHECK(hipEventCreate(&p.begin));
        BH_CHECK(hipEventCreate(&p.end));
        BH_CHECK(hipEventRecord(p.begin, stream));
        pending.push_back(p);
    }
    void end() {
        if (!on) return;
        BH_CHECK(hipEventRecord(pending.back().end, stream));
    }
    // After the stream has been synchronised.
    void collect() {
        if (!on) return;
        std::lock_guard<std::mutex> lock(g_prof_mu);
        for (const Pending& p : pending) {
            float ms = 0;
            BH_CHECK(hipEventElapsedTime(&ms, p.begin, p.end));
            auto it = std::find_if(g_prof_stats.begin(), g_prof_stats.end(),
                                   [&](const KernelStat& k) { return k.name == p.name; });
            if (it == g_prof_stats.end()) {
                g_prof_stats.push_back(KernelStat{p.name, 0, 0});
                it = g_prof_stats.end() - 1;
            }
            it->ms += ms;
            it->launches += 1;
            (void)hipEventDestroy(p.begin);
            (void)hipEventDestroy(p.end);
        }
        pending.clear();
        for (const Counter& c : counters) {
            std::vector<unsigned long long> v(c.words);
            BH_CHECK(hipMemcpy(v.data(), c.dev, sizeof(unsigned long long) * c.words,
                               hipMemcpyDeviceToHost));
            auto it = std::find_if(g_prof_stats.begin(), g_prof_stats.end(),
                                   [&](const KernelStat& k) { return k.name == c.name; });
            if (it == g_prof_stats.end()) {
                g_prof_stats.push_back(KernelStat{c.name, 0, 0});
                it = g_prof_stats.end() - 1;
            }
            for (unsigned long long x : v) it->launches += x;
        }
        counters.clear();
        for (const auto& c : host_counts) {
            auto it = std::find_if(g_prof_stats.begin(), g_prof_stats.end(),
                                   [&](const KernelStat& k) { return k.name == c.first; });
            if (it == g_prof_stats.end()) {
                g_prof_stats.push_back(KernelStat{c.first, 0, 0});
                it = g_prof_stats.end() - 1;
            }
            it->launches += c.second;
        }
        host_counts.clear();
    }
};

// Host-side stages (staging copies) join the same table while profiling,
// named "host:<stage>", with their wall time.
void note_host_time(const char* name, std::chrono::steady_clock::time_point since) {
    const double ms = std::chrono::duration<double, std::milli>(
                          std::chrono::steady_clock::now() - since).count();
    std::lock_guard<std::mutex> lock(g_prof_mu);
    if (!g_profiling) return;
    auto it = std::find_if(g_prof_stats.begin(), g_prof_stats.end(),
                           [&](const KernelStat& k) { return k.name == name; });
    if (it == g_prof_stats.end()) {
        g_prof_stats.push_back(KernelStat{name, 0, 0});
        it = g_prof_stats.end() - 1;
    }
    it->ms += ms;
    it->launches += 1;
}

#define BH_LAUNCH(prof, name, ...)       \
    do {                                 \
        (prof).begin(name);              \
        hipLaunchKernelGGL(__VA_ARGS__); \
        (prof).end();                    \
    } while (0)

// One permutation of a batch as the host sees it.
struct Job {
    const void* dsrc = nullptr;  // device pointer to its entries
    uint64_t n = 0;
    uint32_t width = 4;
    Layout lay;
    uint64_t base = 0;           // entry offset in the shared arrays
    uint64_t blob = 0;           // word offset in the shared blob
    uint32_t df = 0;
};

// A launch: an index range of the uploaded segment table and its grid.
struct Launch {
    uint32_t first = 0, count = 0;
    uint64_t blocks = 0;
};

// Collects the segments of one launch.
struct SegList {
    std::vector<Seg> segs;
    uint64_t blocks = 0;
    void add(uint32_t item, uint64_t blocks_of_item, uint64_t nodes, uint64_t chunks,
             uint64_t desc0, uint64_t row0, uint64_t row_words, bool xcd_align,
             uint32_t depth = 0, uint32_t side = 0, uint32_t narrow_out = 0) {
        if (blocks_of_item == 0) return;
        if (xcd_align) {
            blocks = round8(blocks);
            blocks_of_item = round8(blocks_of_item);
        }
        Seg s{};
        s.block0 = blocks;
        s.blocks = blocks_of_item;
        s.nodes = nodes;
        s.chunks = chunks;
        s.desc0 = desc0;
        s.row0 = row0;
        s.row_words = row_words;
        s.item = item;
        s.depth = depth;
        s.side = side;
        s.narrow_out = narrow_out;
        segs.push_back(s);
        blocks += blocks_of_item;
    }
    // One lane group of a level launch: `lanes` nodes, each its own entry
    // (nodes == 1), sharing 8 * lane_g * max_chunks blocks (level_block).
    struct Lane {
        uint32_t item;
        uint64_t chunks, desc, row0, row_words;
        uint32_t narrow_out;
    };
    void add_lane_group(const std::vector<Lane>& group, uint32_t lane_g) {
        blocks = round8(blocks);
        uint64_t max_chunks = 0;
        for (const Lane& l : group) max_chunks = std::max(max_chunks, l.chunks);
        const uint64_t group_blocks = uint64_t{8} * lane_g * max_chunks;
        for (size_t k = 0; k < group.size(); ++k) {
            Seg s{};
            s.block0 = blocks;
            s.blocks = group_blocks;
            s.nodes = 1;
            s.chunks = group[k].chunks;
            s.desc0 = group[k].desc;
            s.row0 = group[k].row0;
            s.row_words = group[k].row_words;
            s.item = group[k].item;
            s.lane = static_cast<uint32_t>(k);
            s.lanes = static_cast<uint32_t>(group.size());
            s.lane_g = lane_g;
            s.narrow_out = group[k].narrow_out;
            segs.push_back(s);
        }
        blocks += group_blocks;
    }
    // Appends the segments to the batch table.
    Launch commit(std::vector<Seg>& table) const {
        Launch l;
        l.first = static_cast<uint32_t>(table.size());
        l.count = static_cast<uint32_t>(segs.size());
        l.blocks = blocks;
        table.insert(table.end(), segs.begin(), segs.end());
        return l;
    }
};

// Plans every launch of a batch, then queues them on ws.stream.  Results
// (blob words) land in ws.blob at each job's word offset; per-item error flags
// in ws.errs.
// `out` (8-byte aligned, total_words * 8 bytes) receives the blobs, or null
// for the workspace's own blob buffer.
// Work is queued on `st`; the deferred-edge list of each depth holds up to
// `defer_cap` edges (at most kDeferCap).
void enqueue_batch(Workspace& ws, hipStream_t st, std::vector<Job>& jobs, Profiler& prof,
                   uint8_t* out, int shape, uint64_t defer_cap) {
    const SubtreeShape sub = kShapes[shape];
    const int hook_binning = g_hook_binning.load();
    const bool force_wide = g_wide_words.load() == 1;
    const bool rulers = g_colouring.load() == 1;
    const int plain_union = g_plain_union.load();
    const uint32_t count = static_cast<uint32_t>(jobs.size());
    uint64_t total_n = 0, total_words = 0;
    uint32_t max_df = 0;
    bool wide_inverse = false;
    for (Job& j : jobs) {
        j.base = total_n;
        j.blob = total_words;
        j.df = fusion_depth(j.n, j.lay.plan.depth, sub.log2);
        total_n += j.n;
        total_words += ceil_div(j.lay.plan.end_offset, 8);
        max_df = std::max(max_df, j.df);
        wide_inverse = wide_inverse || force_wide || j.n >= 0xFFFFFFFFull;
    }
    // Device copies of plans and item contexts.
    ws.plans.ensure(sizeof(Plan) * count);
    ws.items.ensure(sizeof(ItemCtx) * count);
    ws.errs.ensure(sizeof(int) * (count + 1));  // one flag per item, then the inverse overflow
    if (!out) ws.blob.ensure(total_words * 8);
    uint8_t* blob_ptr = out ? out : ws.blob.as<uint8_t>();
    std::vector<Plan> plans(count);
    std::vector<ItemCtx> items(count);
    for (uint32_t i = 0; i < count; ++i) {
        plans[i] = jobs[i].lay.plan;
        items[i].src = jobs[i].dsrc;
        items[i].plan = ws.plans.as<Plan>() + i;
        items[i].n = jobs[i].n;
        items[i].base = jobs[i].base;
        items[i].blob = jobs[i].blob;
        items[i].width = jobs[i].width;
        items[i].df = jobs[i].df;
        items[i].header_bytes = static_cast<uint32_t>(jobs[i].lay.header_bytes);
        write_header(jobs[i].lay, items[i].header);
    }

    // Segment tables of every launch.  Per-entry kernels cover 2^31 entries
    // per pass at most and loop over the rest.
    const uint64_t per_thread = std::max<uint64_t>(1, ceil_div(total_n, uint64_t{1} << 31));
    std::vector<Seg> table;
    SegList per_entry_list;
    for (uint32_t i = 0; i < count; ++i) {
        per_entry_list.add(i, ceil_div(jobs[i].n, kBlock * per_thread), 0, 0, 0, 0, 0, false);
    }
    const Launch per_entry = per_entry_list.commit(table);
    // k_check_inverse streams kCheckPerThread entries per thread: its blocks
    // do one load each otherwise, and the segment search that starts every
    // block costs more than that in a batch of many items (4096 items: 3.1 ->
    // 0.8 ms).  (k_scatter_inverse keeps one entry per thread: a loop of
    // dependent load-store pairs per thread was slower, 5.4 -> 7.2 ms at 8.)
    constexpr uint64_t kCheckPerThread = 16;
    SegList check_list;
    for (uint32_t i = 0; i < count; ++i) {
        check_list.add(i, ceil_div(jobs[i].n, kBlock * per_thread * kCheckPerThread), 0, 0, 0, 0,
                       0, false);
    }
    const Launch check_entry = check_list.commit(table);
    struct LevelPlan {
        Launch hbm, lds, pack, narrow;  // pack, narrow: both sides' segments
        Launch bin;                     // k_hook_bin: the HBM nodes, 4096 pairs per block
        bool wide_uf = false;
        bool lanes = false;             // small nodes of many items in lane groups, one XCD each
        bool plain = false;             // union hooks by plain stores, then checked
        uint32_t plain_group = 1;       // nodes (~2 MB of words) each XCD unites or routes at a time
        bool binned = false;            // first hooks staged in bins of 2^bin_log2 words
        uint32_t bin_log2 = kHookBinMinLog2;
        uint64_t bins = 0, scatter_blocks = 0;
        uint64_t hbm_entries = 0;  // entries of the items coloured in HBM at this depth
        uint64_t max_slots = 0;    // input slots of the largest HBM node
        uint32_t ruler_rounds = 0; // pointer-jumping rounds covering the longest ruler cycle
    };
    std::vector<LevelPlan> levels(max_df);
    // Descriptors of all depths live side by side (depth d's from desc_base),
    // so one k_describe launch writes them all; rows are per depth.
    SegList describe;
    uint64_t desc_base = 0, max_rows = 1;
    for (uint32_t d = 0; d < max_df; ++d) {
        LevelPlan& lp = levels[d];
        SegList hbm, lds, pack, narrow, bin;
        struct HbmItem {
            uint32_t item;
            uint64_t chunks, desc0, row0, row_words;
            uint32_t narrow_out;  // the children are coloured in LDS: 16-bit entries
        };
        std::vector<HbmItem> hbm_items;
        uint64_t desc_count = desc_base, row_count = 0;
        const uint64_t nodes = uint64_t{1} << d;
        for (uint32_t i = 0; i < count; ++i) {
            if (jobs[i].df <= d) continue;
            const uint64_t cmax = ceil_div(jobs[i].n, nodes);
            const uint64_t maxpairs = cmax / 2;
            const uint64_t row_words = std::max<uint64_t>(1, ceil_div(maxpairs, 64));
            describe.add(i, ceil_div(nodes, 64), nodes, 0, desc_count, row_count, row_words,
                         false, d);
            if (cmax > kLevelLds) {
                const uint64_t chunks = ceil_div(maxpairs + 1, kBlock);
                const uint32_t narrow_out = ceil_div(jobs[i].n, nodes * 2) <= kLevelLds ? 1u : 0u;
                hbm_items.push_back(
                    HbmItem{i, chunks, desc_count, row_count, row_words, narrow_out});
                const uint64_t bin_chunks =
                    ceil_div(maxpairs + 1, uint64_t{kHookBinThreads} * kHookBinPairs);
                bin.add(i, nodes * bin_chunks, nodes, bin_chunks, desc_count, row_count, row_words,
                        true);
                lp.hbm_entries += jobs[i].n;
                lp.max_slots = std::max(lp.max_slots, (cmax + 1) / 2);
                // A node meets at most ceil(cmax / 32) + 1 ruler buckets.
                uint32_t r = 0;
                while ((uint64_t{1} << r) < ceil_div(cmax, uint64_t{1} << kRulerLog2) + 1) ++r;
                lp.ruler_rounds = std::max(lp.ruler_rounds, r);
                lp.wide_uf = lp.wide_uf || force_wide || maxpairs + 1 >= (uint64_t{1} << 31);
            } else {
                lds.add(i, nodes, nodes, 1, desc_count, row_count, row_words, false);
            }
            const Plan& P = jobs[i].lay.plan;
            for (int side = 0; side < 2; ++side) {
                const SideDepth& sd = P.side[side][d];
                const uint32_t span_log2 = sd.group_depth - sd.relative_level - 1;
                const uint64_t packets = ceil_div(maxpairs, uint64_t{1} << span_log2);
                const uint64_t groups = uint64_t{1} << (d - sd.relative_level);
                pack.add(i,
                         ceil_div(groups * packets * pack_words_per_job(sd.group_depth), kBlock),
                         nodes, packets, desc_count, row_count, row_words, false, d, side);
                if (pack_narrow(sd.group_depth, span_log2)) {
                    // One block per (group, kNarrowGroups (node group, row
                    // word) pairs): narrow_shape.
                    const uint64_t node_groups =
                        (uint64_t{1} << sd.relative_level) >> (6 - span_log2);
                    const NarrowShape sh = narrow_shape(node_groups);
                    if (sh.ngb) {
                        narrow.add(i, groups * sh.chunks * ceil_div(row_words, sh.wpb), nodes, 0,
                                   desc_count, row_count, row_words, false, d, side);
                    }
                }
            }
            desc_count += nodes;
            row_count += nodes * row_words;
        }
        // First hooks are binned where the direct stores are bound by address
        // translation: nodes of at least 64 MB of 32-bit words in a level of
        // at least 2^31 entries (measured at 2^26 and 2^32 + 12345 entries:
        // binning gains nothing at 2^26, and 50 ms per depth at 2^32 where
        // nodes hold 64 MB or more).  The test setting bins every node whose
        // words overflow an XCD's L2.
        const uint64_t node_bytes = lp.max_slots * 4;
        lp.binned = !hbm_items.empty() && !lp.wide_uf &&
                    (hook_binning == 1 ? node_bytes > (uint64_t{4} << 20)
                     : hook_binning < 0 ? node_bytes >= (uint64_t{64} << 20) &&
                                              total_n >= (uint64_t{1} << 31)
                                        : false);
        // Many small nodes of several items (a batch): lane groups, so that
        // each node keeps to one XCD at every depth, the shallow ones
        // included, and each XCD interleaves about kPlainUnionGroupBytes of
        // nodes of several items.  One item, or few nodes, keep the node
        // mappings of block_to_node.
        const uint64_t hbm_nodes = hbm_items.size() * nodes;
        lp.lanes = hbm_items.size() >= 2 && hbm_nodes >= 64 && hbm_nodes <= (uint64_t{1} << 16) &&
                   !lp.wide_uf && !lp.binned && !rulers && node_bytes <= kPlainUnionMaxBytes;
        if (lp.lanes) {
            // Largest nodes first, so that a group's lanes have about the
            // same number of chunks.
            std::stable_sort(hbm_items.begin(), hbm_items.end(),
                             [](const HbmItem& a, const HbmItem& b) { return a.chunks > b.chunks; });
            std::vector<SegList::Lane> group;
            uint32_t lane_g = 1;
            for (const HbmItem& h : hbm_items) {
                for (uint64_t j = 0; j < nodes; ++j) {
                    if (group.empty()) {
                        // Nodes of ~kPlainUnionGroupBytes of words per XCD.
                        const uint64_t bytes = h.chunks * kBlock * 4;
                        lane_g = 1;
                        while (lane_g < 64 && bytes * lane_g * 2 <= kPlainUnionGroupBytes) lane_g *= 2;
                    }
                    group.push_back(SegList::Lane{h.item, h.chunks, h.desc0 + j,
                                                  h.row0 + j * h.row_words, h.row_words,
                                                  h.narrow_out});
                    if (group.size() == uint64_t{8} * lane_g) {
                        hbm.add_lane_group(group, lane_g);
                        group.clear();
                    }
                }
            }
            if (!group.empty()) hbm.add_lane_group(group, lane_g);
        } else {
            for (const HbmItem& h : hbm_items) {
                hbm.add(h.item, nodes * h.chunks, nodes, h.chunks, h.desc0, h.row0, h.row_words,
                        true, 0, 0, h.narrow_out);
            }
        }
        lp.hbm = hbm.commit(table);
        // Plain-store hooks where every node keeps to one XCD (8 nodes or
        // more per item, or lane groups) and, by default, its words fit in
        // the XCD's L2.
        lp.plain = lp.hbm.count && (d >= 3 || lp.lanes) && !lp.wide_uf && !lp.binned && !rulers &&
                   (plain_union == 1 ||
                    (plain_union < 0 && node_bytes <= kPlainUnionMaxBytes));
        while (lp.plain_group < 64 && node_bytes * lp.plain_group * 2 <= kPlainUnionGroupBytes) {
            lp.plain_group *= 2;
        }
        if (lp.binned) {
            lp.bin = bin.commit(table);
            // The smallest bins (from 1 MB) a node spans at most kHookBinSpan of.
            while (ceil_div(lp.max_slots, uint64_t{1} << lp.bin_log2) + 1 > kHookBinSpan) {
                ++lp.bin_log2;
            }
            lp.bins = ceil_div(total_n, uint64_t{1} << lp.bin_log2);
            lp.scatter_blocks = ceil_div(lp.bins, kXcds) * kXcds *
                                ((uint64_t{1} << lp.bin_log2) / kHookScatterEntries);
        }
        lp.lds = lds.commit(table);
        lp.pack = pack.commit(table);
        lp.narrow = narrow.commit(table);
        desc_base = desc_count;
        max_rows = std::max(max_rows, row_count);
    }
    const Launch describe_all = describe.commit(table);
    SegList subtree_list;
    for (uint32_t i = 0; i < count; ++i) {
        subtree_list.add(i, uint64_t{1} << jobs[i].df, 0, 0, 0, 0, 0, false);
    }
    const Launch subtree = subtree_list.commit(table);
    // Internal levels inside the largest subtree (rows kept per subtree).
    uint32_t sub_levels = 1;
    for (const Job& j : jobs) sub_levels = std::max(sub_levels, j.lay.plan.depth - j.df);
    // Per-depth bin counters, room for the smallest bins (only when a depth
    // bins its first hooks).
    const uint64_t hook_bins = ceil_div(total_n, uint64_t{1} << kHookBinMinLog2);
    bool any_binned = false;
    for (const LevelPlan& lp : levels) any_binned = any_binned || lp.binned;
    // The level-0 inverse is binned for batches of at least 2^31 entries (the
    // test setting: from 2^20), in the smallest bins (from 2^18 entries) no
    // item spans more than kInvBinSpan of.  Its staging is the union-find
    // array, which then holds 8 bytes per entry.
    const bool inv_binned =
        max_df > 0 && (hook_binning == 1 ? total_n > (uint64_t{1} << 20)
                       : hook_binning < 0 ? total_n >= (uint64_t{1} << 31) : false);
    uint32_t inv_log2 = kHookBinMinLog2;
    Launch inv_bin;
    uint64_t inv_bins = 0, inv_scatter_blocks = 0;
    if (inv_binned) {
        for (const Job& j : jobs) {
            while (((j.base + j.n - 1) >> inv_log2) - (j.base >> inv_log2) + 1 > kInvBinSpan) {
                ++inv_log2;
            }
        }
        SegList inv_list;
        for (uint32_t i = 0; i < count; ++i) {
            inv_list.add(i, ceil_div(jobs[i].n, kInvBinEntries), 0, 0, 0, 0, 0, false);
        }
        inv_bin = inv_list.commit(table);
        inv_bins = ceil_div(total_n, uint64_t{1} << inv_log2);
        inv_scatter_blocks = ceil_div(inv_bins, kXcds) * kXcds *
                             ((uint64_t{1} << inv_log2) / kHookScatterEntries);
    }
    const uint64_t fill_words = (any_binned ? hook_bins * max_df : 0) + inv_bins;
    // One dispatch holds at most 2^32 - 1 work-items.
    uint64_t most = std::max({subtree.blocks * sub.threads, describe_all.blocks * 64,
                              inv_bin.blocks * kInvBinThreads, inv_scatter_blocks * kBlock});
    for (const LevelPlan& lp : levels) {
        most = std::max({most, lp.hbm.blocks * kBlock, lp.lds.blocks * kLevelThreads,
                         lp.pack.blocks * kBlock, lp.narrow.blocks * kNarrowThreads,
                         lp.bin.blocks * kHookBinThreads, lp.scatter_blocks * kBlock});
    }
    if (most >= (uint64_t{1} << 32)) {
        throw std::length_error("batch too large for one call: split it into smaller batches");
    }

    // Ruler buckets of the level arrays (the ruler colouring).
    const uint64_t ruler_buckets = ceil_div(total_n, uint64_t{1} << kRulerLog2) + 1;
    // Scratch.
    ws.segs.ensure(sizeof(Seg) * std::max<size_t>(1, table.size()));
    ws.lbits.ensure(sizeof(uint64_t) * subtree_row_words(sub.log2) * 2 * sub_levels *
                    std::max<uint64_t>(1, subtree.blocks));
    ws.inv0.ensure(total_n * (wide_inverse ? 8 : 4));
    if (max_df > 0) {
        for (int i = 0; i < 2; ++i) {
            ws.level[i].ensure(total_n * 4);
            ws.ilevel[i].ensure(total_n * 4);
        }
        bool any_wide = false;
        for (const LevelPlan& lp : levels) any_wide = any_wide || lp.wide_uf;
        ws.uf.ensure(total_n * (any_wide || inv_binned ? 8 : 4));
        ws.desc.ensure(sizeof(NodeDesc) * std::max<uint64_t>(1, desc_base));
        ws.srows.ensure(8 * max_rows);
        ws.trows.ensure(8 * max_rows);
        ws.deferred.ensure(sizeof(DeferredEdge) * defer_cap);
        // Per depth: 4 counters, then (after all depths) kRejoinCounters
        // counters of the relations k_union_rejoin joined again (counted
        // when profiling).
        ws.deferred_counts.ensure(sizeof(unsigned long long) * (4 + kRejoinCounters) * max_df);
        // One count per check block of the plain-store union pass.
        uint64_t check_blocks = 0;
        for (const LevelPlan& lp : levels) {
            if (lp.plain) check_blocks = std::max<uint64_t>(check_blocks, lp.hbm.blocks);
        }
        if (check_blocks) ws.lost_counts.ensure(sizeof(uint32_t) * check_blocks);
        if (fill_words) ws.bin_fill.ensure(sizeof(uint32_t) * fill_words);
        if (rulers) ws.rulers.ensure((sizeof(uint64_t) + sizeof(uint32_t)) * 2 * ruler_buckets);
    }
    BH_CHECK(hipMemcpyAsync(ws.plans.ptr, plans.data(), sizeof(Plan) * count,
                            hipMemcpyHostToDevice, st));
    BH_CHECK(hipMemcpyAsync(ws.items.ptr, items.data(), sizeof(ItemCtx) * count,
                            hipMemcpyHostToDevice, st));
    BH_CHECK(hipMemcpyAsync(ws.segs.ptr, table.data(), sizeof(Seg) * table.size(),
                            hipMemcpyHostToDevice, st));
    auto fill = [&](void* p, uint64_t words, uint32_t value) {
        if (words == 0) return;
        const uint64_t blocks = std::min<uint64_t>(ceil_div(words, kBlock), uint64_t{1} << 16);
        BH_LAUNCH(prof, "k_fill", k_fill_u32, dim3(static_cast<uint32_t>(blocks)), dim3(kBlock), 0,
                  st, static_cast<uint32_t*>(p), words, value);
    };
    fill(ws.errs.ptr, count + 1, 0);
    fill(blob_ptr, total_words * 2, 0);
    fill(ws.inv0.ptr, total_n * (wide_inverse ? 2 : 1), 0xFFFFFFFFu);
    if (max_df > 0) {
        fill(ws.deferred_counts.ptr, 2 * (4 + kRejoinCounters) * max_df, 0);
        if (fill_words) fill(ws.bin_fill.ptr, fill_words, 0);
    }

    const ItemCtx* ditems = ws.items.as<ItemCtx>();
    const Seg* dsegs = ws.segs.as<Seg>();
    int* errs = ws.errs.as<int>();
    uint64_t* blob = reinterpret_cast<uint64_t*>(blob_ptr);
    auto grid = [](const Launch& l) { return dim3(static_cast<uint32_t>(l.blocks)); };
    auto segp = [&](const Launch& l) { return dsegs + l.first; };

    // Launches of the 64-bit-word instantiations carry their own names, so
    // that a profile shows which one ran.
    auto validate = [&](auto* inv) {
        using VI = std::remove_pointer_t<decltype(inv)>;
        constexpr bool w64 = sizeof(VI) == 8;
        if (inv_binned) {
            uint32_t* const inv_fill = ws.bin_fill.as<uint32_t>() + (fill_words - inv_bins);
            // errs[count] is the overflow flag (a repeated value overfilled a bin).
            int* const overflow = errs + count;
            BH_LAUNCH(prof, w64 ? "k_inverse_bin_w64" : "k_inverse_bin", (k_inverse_bin<VI>),
                      grid(inv_bin), dim3(kInvBinThreads), 0, st, ditems, segp(inv_bin),
                      inv_bin.count, inv_log2, total_n, ws.uf.as<uint64_t>(), inv_fill, overflow,
                      errs);
            BH_LAUNCH(prof, w64 ? "k_inverse_scatter_w64" : "k_inverse_scatter",
                      (k_inverse_scatter<VI>), dim3(static_cast<uint32_t>(inv_scatter_blocks)),
                      dim3(kBlock), 0, st, inv, ws.uf.as<uint64_t>(), inv_fill, inv_bins,
                      inv_log2, total_n);
            // Return at once unless a bin overflowed.
            const dim3 redo_grid(static_cast<uint32_t>(std::min<uint64_t>(per_entry.blocks, 2048)));
            for (int pass = 0; pass < 2; ++pass) {
                BH_LAUNCH(prof, "k_inverse_redo", (k_inverse_redo<VI>), redo_grid, dim3(kBlock), 0,
                          st, ditems, segp(per_entry), per_entry.count, per_entry.blocks, inv,
                          pass, overflow, errs);
            }
        } else {
            BH_LAUNCH(prof, w64 ? "k_scatter_inverse_w64" : "k_scatter_inverse",
                      (k_scatter_inverse<VI>), grid(per_entry), dim3(kBlock), 0, st, ditems,
                      segp(per_entry), per_entry.count, inv, errs);
        }
        BH_LAUNCH(prof, w64 ? "k_check_inverse_w64" : "k_check_inverse", (k_check_inverse<VI>),
                  grid(check_entry), dim3(kBlock), 0, st, ditems, segp(check_entry),
                  check_entry.count, inv, errs);
    };
    hipLaunchKernelGGL(k_headers, dim3(ceil_div(count, 64)), dim3(64), 0, st, ditems, count,
                       blob_ptr);
    if (wide_inverse) validate(ws.inv0.as<uint64_t>());
    else validate(ws.inv0.as<uint32_t>());
    if (describe_all.count) {
        BH_LAUNCH(prof, "k_describe", k_describe, grid(describe_all), dim3(64), 0, st, ditems,
                  segp(describe_all), describe_all.count, ws.desc.as<NodeDesc>(), errs);
    }

    for (const LevelPlan& lp : levels) {
        if (lp.lanes) prof.note("lane_group_depths", 1);
    }
    // Depth d reads level[(d - 1) & 1] and writes level[d & 1].
    for (uint32_t d = 0; d < max_df; ++d) {
        const LevelPlan& lp = levels[d];
        NodeDesc* desc = ws.desc.as<NodeDesc>();
        uint64_t* srows = ws.srows.as<uint64_t>();
        uint64_t* trows = ws.trows.as<uint64_t>();
        const int next = d & 1;
        const uint32_t* plev = ws.level[next ^ 1].as<uint32_t>();
        uint32_t* pdst = ws.level[next].as<uint32_t>();
        uint32_t* idst = ws.ilevel[next].as<uint32_t>();
        const int level0 = d == 0 ? 1 : 0;
        auto run = [&](auto* inv, auto* uf) {
            using VI = std::remove_const_t<std::remove_pointer_t<decltype(inv)>>;
            using W = std::remove_pointer_t<decltype(uf)>;
            constexpr bool w64 = sizeof(W) == 8;
            if (lp.hbm.count && rulers) {
                // Rulers' walks, pointer jumping, then switch values and the
                // routing of P; the owner marks use the union-find words.
                uint32_t* const owner = ws.uf.as<uint32_t>();
                uint64_t* const rmin0 = ws.rulers.as<uint64_t>();
                uint64_t* const rmin1 = rmin0 + ruler_buckets;
                uint32_t* const rnext0 = reinterpret_cast<uint32_t*>(rmin1 + ruler_buckets);
                uint32_t* const rnext1 = rnext0 + ruler_buckets;
                fill(owner, total_n, kUnowned);
                fill(rnext0, ruler_buckets, kRulerUnwalked);
                BH_LAUNCH(prof, "k_ruler_walk", (k_ruler_walk<VI>), grid(lp.hbm), dim3(kBlock), 0,
                          st, ditems, segp(lp.hbm), lp.hbm.count, level0, plev, inv, desc, owner,
                          rnext0, rmin0, errs);
                const dim3 jump_grid(static_cast<uint32_t>(
                    std::min<uint64_t>(ceil_div(ruler_buckets, kBlock), uint64_t{1} << 16)));
                for (uint32_t r = 0; r < lp.ruler_rounds; ++r) {
                    BH_LAUNCH(prof, "k_ruler_jump", k_ruler_jump, jump_grid, dim3(kBlock), 0, st,
                              ruler_buckets, r & 1 ? rnext1 : rnext0, r & 1 ? rmin1 : rmin0,
                              r & 1 ? rnext0 : rnext1, r & 1 ? rmin0 : rmin1);
                }
                const bool odd_rounds = (lp.ruler_rounds & 1) != 0;
                BH_LAUNCH(prof, "k_ruler_route", (k_ruler_route<VI>), grid(lp.hbm), dim3(kBlock),
                          0, st, ditems, segp(lp.hbm), lp.hbm.count, level0, plev, inv, pdst, desc,
                          owner, odd_rounds ? rnext1 : rnext0, odd_rounds ? rmin1 : rmin0, srows,
                          errs);
                BH_LAUNCH(prof, sizeof(VI) == 8 ? "k_route_outputs_w64" : "k_route_outputs",
                          (k_route_outputs<VI>), grid(lp.hbm), dim3(kBlock), 0, st, segp(lp.hbm),
                          lp.hbm.count, inv, idst, desc, srows, trows, nullptr, 0u, errs);
            } else if (lp.hbm.count) {
                // Per depth: edges deferred, flatten blocks finished, hooks
                // under the slot just below.
                unsigned long long* const dcount =
                    ws.deferred_counts.as<unsigned long long>() + 4 * d;
                // Fewer than 8 nodes: each node's blocks spread over several
                // XCDs.  Coherent walks pay only there, and only when chains
                // of neighbouring slots make up more than 1/64 of the slots
                // (counted in every eighth block).
                const bool spread = d < 3 && !lp.lanes;
                const unsigned long long chain_threshold =
                    spread ? lp.hbm_entries / 2 / 64 / 8 : ~0ull;
                // The first hooks (direct or binned) store only the words that
                // get a parent: the others start as root words, set at the
                // root here and below it by the previous depth's
                // k_route_outputs.
                if (d == 0) fill(uf, total_n * sizeof(W) / 4, 0xFFFFFFFFu);
                if constexpr (std::is_same_v<W, uint32_t>) {
                    if (lp.binned) {
                        // Staged in this depth's output level arrays, which
                        // the routing kernels write only after the union.
                        uint32_t* const fill = ws.bin_fill.as<uint32_t>() + hook_bins * d;
                        if (spread) {
                            BH_LAUNCH(prof, "k_hook_bin", (k_hook_bin<VI, true>), grid(lp.bin),
                                      dim3(kHookBinThreads), 0, st, segp(lp.bin), lp.bin.count,
                                      inv, desc, lp.bin_log2, pdst, idst, fill, dcount + 2,
                                      chain_threshold, errs);
                        } else {
                            BH_LAUNCH(prof, "k_hook_bin", (k_hook_bin<VI, false>), grid(lp.bin),
                                      dim3(kHookBinThreads), 0, st, segp(lp.bin), lp.bin.count,
                                      inv, desc, lp.bin_log2, pdst, idst, fill, dcount + 2,
                                      chain_threshold, errs);
                        }
                        BH_LAUNCH(prof, "k_hook_scatter", k_hook_scatter,
                                  dim3(static_cast<uint32_t>(lp.scatter_blocks)), dim3(kBlock), 0,
                                  st, uf, pdst, idst, fill, lp.bins, lp.bin_log2);
                    }
                }
                if (lp.binned) {
                    // Hooked above.
                } else if (spread) {
                    BH_LAUNCH(prof, w64 ? "k_hook_first_w64" : "k_hook_first",
                              (k_hook_first<VI, W, true>), grid(lp.hbm),
                              dim3(kBlock), 0, st, segp(lp.hbm), lp.hbm.count, inv, uf, desc,
                              dcount + 2, chain_threshold, 63u, 0u, errs);
                } else {
                    // Each XCD hooks its nodes one after another; a node whose
                    // words exceed half an XCD's L2 (4 MB) but fit in it is
                    // hooked in two passes of half the slots each, so that a
                    // pass's random stores stay L2-resident.
                    const uint64_t node_bytes = lp.max_slots * sizeof(W);
                    const bool halves = node_bytes > (uint64_t{2} << 20) &&
                                        node_bytes <= (uint64_t{4} << 20);
                    uint32_t shift = 63;
                    if (halves) {
                        shift = 0;
                        while ((uint64_t{2} << shift) < lp.max_slots) ++shift;
                    }
                    for (uint32_t part = 0; part < (halves ? 2u : 1u); ++part) {
                        BH_LAUNCH(prof, w64 ? "k_hook_first_w64" : "k_hook_first",
                                  (k_hook_first<VI, W, false>), grid(lp.hbm),
                                  dim3(kBlock), 0, st, segp(lp.hbm), lp.hbm.count, inv, uf, desc,
                                  dcount + 2, chain_threshold, shift, part, errs);
                    }
                }
                DeferredEdge* const deferred = ws.deferred.as<DeferredEdge>();
                bool plain_done = false;
                if constexpr (std::is_same_v<W, uint32_t>) {
                    if (lp.plain) {
                        // Hook records in this depth's inverse output, which
                        // only k_route_outputs and k_level_lds write, later.
                        uint64_t* const rec = reinterpret_cast<uint64_t*>(idst);
                        unsigned long long* const rejoined =
                            ws.deferred_counts.as<unsigned long long>() + 4 * max_df +
                            kRejoinCounters * d;
                        BH_LAUNCH(prof, "k_union_plain", (k_union_plain<VI>), grid(lp.hbm),
                                  dim3(kBlock), 0, st, segp(lp.hbm), lp.hbm.count, inv, uf, desc,
                                  rec, deferred, dcount, defer_cap, lp.plain_group, errs);
                        uint32_t* const lost = ws.lost_counts.as<uint32_t>();
                        BH_LAUNCH(prof, "k_union_verify", (k_union_verify<VI>), grid(lp.hbm),
                                  dim3(kVerifyThreads), 0, st, segp(lp.hbm), lp.hbm.count, uf,
                                  desc, rec, lost, lp.plain_group, errs);
                        // kRejoinSpan check blocks of one XCD per block.
                        const uint64_t per_xcd = ceil_div(lp.hbm.blocks, uint64_t{kXcds});
                        const uint64_t rejoin_blocks = kXcds * ceil_div(per_xcd, uint64_t{kRejoinSpan});
                        BH_LAUNCH(prof, "k_union_rejoin", k_union_rejoin,
                                  dim3(static_cast<uint32_t>(rejoin_blocks)), dim3(kBlock), 0, st,
                                  segp(lp.hbm), lp.hbm.count, uf, desc, rec, lost, lp.hbm.blocks,
                                  lp.plain_group, deferred, dcount, defer_cap,
                                  prof.on ? rejoined : nullptr);
                        prof.count("union_rejoined", rejoined, kRejoinCounters);
                        plain_done = true;
                    }
                }
                if (!plain_done) {
                    BH_LAUNCH(prof, w64 ? "k_union_w64" : "k_union", (k_union<VI, W>),
                              grid(lp.hbm), dim3(kBlock), 0, st, segp(lp.hbm), lp.hbm.count, inv,
                              uf, desc, deferred, dcount, defer_cap, errs);
                }
                // Returns at once unless k_union deferred an edge.
                const dim3 flat_grid(static_cast<uint32_t>(std::min<uint64_t>(lp.hbm.blocks, 2048)));
                if (spread) {
                    BH_LAUNCH(prof, w64 ? "k_flatten_w64" : "k_flatten",
                              (k_flatten<W, CoherentForest>), flat_grid, dim3(kBlock), 0, st,
                              segp(lp.hbm), lp.hbm.count, lp.hbm.blocks, uf, desc, deferred,
                              dcount, defer_cap, errs);
                } else {
                    BH_LAUNCH(prof, w64 ? "k_flatten_w64" : "k_flatten",
                              (k_flatten<W, DeviceForest>), flat_grid, dim3(kBlock), 0, st,
                              segp(lp.hbm), lp.hbm.count, lp.hbm.blocks, uf, desc, deferred,
                              dcount, defer_cap, errs);
                }
                BH_LAUNCH(prof, w64 ? "k_route_inputs_w64" : "k_route_inputs", (k_route_inputs<W>),
                          grid(lp.hbm), dim3(kBlock),
                          0, st, ditems, segp(lp.hbm), lp.hbm.count, level0, plev, pdst, uf, desc,
                          srows, dcount + 2, chain_threshold, spread ? 0u : lp.plain_group,
                          errs);
                // Root words for the next depth's first hooks.
                const bool next_hooks = d + 1 < max_df && levels[d + 1].hbm.count;
                BH_LAUNCH(prof, sizeof(VI) == 8 ? "k_route_outputs_w64" : "k_route_outputs",
                          (k_route_outputs<VI>), grid(lp.hbm),
                          dim3(kBlock), 0, st, segp(lp.hbm), lp.hbm.count, inv, idst, desc, srows,
                          trows, next_hooks ? ws.uf.ptr : nullptr,
                          next_hooks && levels[d + 1].wide_uf ? 1u : 0u, errs);
            }
            if (lp.lds.count) {
                BH_LAUNCH(prof, sizeof(VI) == 8 ? "k_level_lds_w64" : "k_level_lds",
                          (k_level_lds<VI>), grid(lp.lds),
                          dim3(kLevelThreads), 0, st, ditems, segp(lp.lds), lp.lds.count, level0,
                          plev, inv, pdst, idst, desc, srows, trows, errs);
            }
        };
        if (d == 0) {
            if (wide_inverse) {
                if (lp.wide_uf) run(ws.inv0.as<uint64_t>(), ws.uf.as<uint64_t>());
                else run(ws.inv0.as<uint64_t>(), ws.uf.as<uint32_t>());
            } else {
                if (lp.wide_uf) run(ws.inv0.as<uint32_t>(), ws.uf.as<uint64_t>());
                else run(ws.inv0.as<uint32_t>(), ws.uf.as<uint32_t>());
            }
        } else {
            const uint32_t* isrc = ws.ilevel[next ^ 1].as<uint32_t>();
            if (lp.wide_uf) run(isrc, ws.uf.as<uint64_t>());
            else run(isrc, ws.uf.as<uint32_t>());
        }
        // Both sides in one launch each (segments carry their side).
        if (lp.pack.count) {
            BH_LAUNCH(prof, "k_pack", k_pack, grid(lp.pack), dim3(kBlock), 0, st, ditems,
                      segp(lp.pack), lp.pack.count, d, desc, srows, trows, blob, errs);
        }
        if (lp.narrow.count) {
            BH_LAUNCH(prof, "k_pack_narrow", k_pack_narrow, grid(lp.narrow), dim3(kNarrowThreads),
                      0, st, ditems, segp(lp.narrow), lp.narrow.count, d, desc, srows, trows,
                      blob, errs);
        }
    }
    // An item with fusion depth df was last written at depth df - 1, into
    // level[(df - 1) & 1]; items with df == 0 read their own source.
    // Then the subtrees' switch rows (left in ws.lbits) are packed.
    auto subtree_kernel = [&](auto kernel, auto emit) {
        BH_LAUNCH(prof, "k_subtree", kernel, grid(subtree), dim3(sub.threads), 0, st, ditems,
                  segp(subtree), subtree.count, ws.level[0].as<uint32_t>(),
                  ws.level[1].as<uint32_t>(), ws.lbits.as<uint64_t>(), sub_levels, blob, errs);
        BH_LAUNCH(prof, "k_subtree_emit", emit, grid(subtree), dim3(emit_threads(sub.log2)), 0, st, ditems,
                  segp(subtree), subtree.count, ws.lbits.as<uint64_t>(), sub_levels, blob, errs);
    };
    switch (sub.log2 * 10000 + sub.threads) {
        case 141024: subtree_kernel(k_subtree<14, 1024>, k_subtree_emit<14>); break;
        case 131024: subtree_kernel(k_subtree<13, 1024>, k_subtree_emit<13>); break;
        case 130512: subtree_kernel(k_subtree<13, 512>, k_subtree_emit<13>); break;
        default: subtree_kernel(k_subtree<12, 512>, k_subtree_emit<12>); break;
    }
    BH_CHECK(hipGetLastError());
}

uint64_t batch_words(const std::vector<Job>& jobs) {
    uint64_t w = 0;
    for (const Job& j : jobs) w += ceil_div(j.lay.plan.end_offset, 8);
    return w;
}

int finish(const Status& st, char* err, size_t err_len) {
    if (!st.ok() && err && err_len) {
        std::strncpy(err, st.message.c_str(), err_len - 1);
        err[err_len - 1] = 0;
    }
    return st.code;
}

bool check_input(uint64_t n, uint32_t width, Status& st) {
    if (width != 4 && width != 8) {
        st.set(kInvalidArgument, "input width must be 4 or 8 bytes");
        return false;
    }
    if (n == 0) {
        st.set(kInvalidArgument, "permutation size is out of range");
        return false;
    }
    if (width == 4 && n > 0xFFFFFFFFull) {
        st.set(kInvalidArgument, "permutation size requires 64-bit source values");
        return false;
    }
    if (n > (uint64_t{1} << 33)) {  // beyond one device's memory and dispatch limits
        st.set(kLength, "permutation too large for one device (at most 2^33 entries)");
        return false;
    }
    return true;
}

bool format_args(const benes_hip_format* f, FormatArgs& out, Status& st) {
    if (!f) return true;  // defaults
    out.middle_group_log2 = f->middle_group_log2;
    out.placement = f->middle_placement;
    if ((f->n_input_clusters && !f->input_clusters) ||
        (f->n_output_clusters && !f->output_clusters)) {
        st.set(kInvalidArgument, "cluster list pointer is null");
        return false;
    }
    out.input_clusters.assign(f->input_clusters, f->input_clusters + f->n_input_clusters);
    out.output_clusters.assign(f->output_clusters, f->output_clusters + f->n_output_clusters);
    return true;
}

// Host-side checks of one request and its layout.
bool prepare_job(const void* perm, uint64_t n, uint32_t width, const benes_hip_format* format,
                 Job& job, Status& st) {
    if (!check_input(n, width, st)) return false;
    FormatArgs fa;
    if (!format_args(format, fa, st)) return false;
    if (!build_layout(n, fa, job.lay, st)) return false;
    job.dsrc = perm;
    job.n = n;
    job.width = width;
    return true;
}

int current_device(Status& st) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) {
        st.set(kRuntime, "hip backend is unavailable");
        return -1;
    }
    return dev;
}

// Device generation of the permutations computable entry by entry.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_generate(uint32_t kind, uint64_t n, uint64_t seed,
                                                     T* __restrict__ out) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += stride) {
        uint64_t v = i;
        if (kind == BENES_HIP_PERM_REVERSE) v = n - 1 - i;
        else if (kind == BENES_HIP_PERM_KEYED) v = keyed_perm(i, n, seed);
        out[i] = static_cast<T>(v);
    }
}

const char* kNotPermutation = "input must contain every value from zero through size minus one";

// Orders `after` behind the work already queued on `before`.
void stream_after(hipStream_t after, hipStream_t before) {
    hipEvent_t e;
    BH_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    BH_CHECK(hipEventRecord(e, before));
    BH_CHECK(hipStreamWaitEvent(after, e, 0));
    BH_CHECK(hipEventDestroy(e));
}

// Runs a batch whose inputs are on the device.  `out` receives the blobs at
// their word offsets (null: the workspace blob).  Per-item error flags are
// returned in `flags`.
void run_batch(Workspace& ws, std::vector<Job>& jobs, uint8_t* out, hipStream_t user_stream,
               std::vector<int>& flags, double* device_seconds) {
    const hipStream_t s = ws.stream;
    if (user_stream) stream_after(s, user_stream);
    Profiler prof;
    {
        std::lock_guard<std::mutex> plock(g_prof_mu);
        prof.on = g_profiling;
    }
    prof.stream = s;
    const uint32_t cls = Tuner::class_of(jobs);
    bool trial = false;
    const int shape = Tuner::instance().pick(cls, trial);
    BH_CHECK(hipEventRecord(ws.ev_begin, s));
    enqueue_batch(ws, s, jobs, prof, out, shape, kDeferCap);
    BH_CHECK(hipEventRecord(ws.ev_end, s));
    flags.assign(jobs.size(), 0);
    BH_CHECK(hipMemcpyAsync(flags.data(), ws.errs.ptr, sizeof(int) * jobs.size(),
                            hipMemcpyDeviceToHost, s));
    BH_CHECK(hipStreamSynchronize(s));
    prof.collect();
    float ms = 0;
    BH_CHECK(hipEventElapsedTime(&ms, ws.ev_begin, ws.ev_end));
    g_device_seconds = ms / 1e3;
    Tuner::instance().report(cls, shape, trial, ms);
    if (device_seconds) *device_seconds = g_device_seconds;
    if (user_stream) stream_after(user_stream, s);
}

template <typename Body>
int with_workspace(Status& st, Body&& body) {
    if (!benes_hip_available()) {
        st.set(kRuntime, "hip backend is unavailable");
        return st.code;
    }
    const int dev = current_device(st);
    if (dev < 0) return st.code;
    Workspace& ws = workspace_for(dev);
    std::lock_guard<std::mutex> lock(ws.mu);
    try {
        ws.init(dev);
        body(ws);
    } catch (const HipError& e) {
        st.set(kRuntime, std::string("hip error: ") + hipGetErrorString(e.code) + " in " + e.what);
    } catch (const std::length_error& e) {
        st.set(kLength, e.what());
    } catch (const std::bad_alloc&) {
        st.set(kLength, "allocation failed");
    }
    return st.code;
}

// Lookups of `count` values (d_values null: 0..count-1) in a device blob, on
// the workspace's stream ordered after `user_stream`.
void run_lookup(Workspace& ws, const uint8_t* d_blob, uint64_t blob_size,
                const uint64_t* d_values, uint64_t count, int inverse, void* d_out,
                uint32_t width, hipStream_t user_stream, Status& st) {
    const hipStream_t s = ws.stream;
    if (user_stream) stream_after(s, user_stream);
    uint8_t hdr[20 + 64];
    const uint64_t avail = std::min<uint64_t>(blob_size, sizeof(hdr));
    BH_CHECK(hipMemcpyAsync(hdr, d_blob, avail, hipMemcpyDeviceToHost, s));
    BH_CHECK(hipStreamSynchronize(s));
    Layout lay;
    if (!parse_header(hdr, avail, blob_size, lay, st)) return;
    if (width == 4 && lay.plan.n > 0x100000000ull) {
        st.set(kInvalidArgument, "lookup results need 64-bit entries for this size");
        return;
    }
    if (count == 0) return;
    ws.plans.ensure(sizeof(Plan));
    ws.errs.ensure(sizeof(int));
    BH_CHECK(hipMemcpyAsync(ws.plans.ptr, &lay.plan, sizeof(Plan), hipMemcpyHostToDevice, s));
    BH_CHECK(hipMemsetAsync(ws.errs.ptr, 0, sizeof(int), s));
    const uint64_t blocks = std::min<uint64_t>(ceil_div(count, kBlock), uint64_t{1} << 20);
    if (width == 4) {
        hipLaunchKernelGGL(k_lookup<uint32_t>, dim3(static_cast<uint32_t>(blocks)), dim3(kBlock),
                           0, s, ws.plans.as<Plan>(), d_blob, d_values, count, inverse,
                           static_cast<uint32_t*>(d_out), ws.errs.as<int>());
    } else {
        hipLaunchKernelGGL(k_lookup<uint64_t>, dim3(static_cast<uint32_t>(blocks)), dim3(kBlock),
                           0, s, ws.plans.as<Plan>(), d_blob, d_values, count, inverse,
                           static_cast<uint64_t*>(d_out), ws.errs.as<int>());
    }
    BH_CHECK(hipGetLastError());
    int flag = 0;
    BH_CHECK(hipMemcpyAsync(&flag, ws.errs.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
    BH_CHECK(hipStreamSynchronize(s));
    if (flag) st.set(kOutOfRange, "permutation lookup is out of range");
    if (user_stream) stream_after(user_stream, s);
}

// Whether the n entries at device address `dsrc` are a permutation of
// 0..n-1: the validation kernels of a compress, alone, on stream `s`.
bool device_is_permutation(Workspace& ws, hipStream_t s, const void* dsrc, uint64_t n,
                           uint32_t width) {
    const bool wide = n >= 0xFFFFFFFFull;
    ItemCtx item{};
    item.src = dsrc;
    item.n = n;
    item.width = width;
    const uint64_t per_thread = std::max<uint64_t>(1, ceil_div(n, uint64_t{1} << 31));
    Seg seg{};
    seg.blocks = ceil_div(n, kBlock * per_thread);
    ws.items.ensure(sizeof(ItemCtx));
    ws.segs.ensure(sizeof(Seg));
    ws.errs.ensure(sizeof(int));
    ws.inv0.ensure(n * (wide ? 8 : 4));
    BH_CHECK(hipMemcpyAsync(ws.items.ptr, &item, sizeof(item), hipMemcpyHostToDevice, s));
    BH_CHECK(hipMemcpyAsync(ws.segs.ptr, &seg, sizeof(seg), hipMemcpyHostToDevice, s));
    const auto fill = [&](void* p, uint64_t words, uint32_t value) {
        const uint64_t blocks = std::min<uint64_t>(ceil_div(words, kBlock), uint64_t{1} << 16);
        hipLaunchKernelGGL(k_fill_u32, dim3(static_cast<uint32_t>(blocks)), dim3(kBlock), 0, s,
                           static_cast<uint32_t*>(p), words, value);
    };
    fill(ws.errs.ptr, 1, 0);
    fill(ws.inv0.ptr, n * (wide ? 2 : 1), 0xFFFFFFFFu);
    const dim3 grid(static_cast<uint32_t>(seg.blocks));
    const ItemCtx* it = ws.items.as<ItemCtx>();
    const Seg* sg = ws.segs.as<Seg>();
    if (wide) {
        hipLaunchKernelGGL(k_scatter_inverse<uint64_t>, grid, dim3(kBlock), 0, s, it, sg, 1u,
                           ws.inv0.as<uint64_t>(), ws.errs.as<int>());
        hipLaunchKernelGGL(k_check_inverse<uint64_t>, grid, dim3(kBlock), 0, s, it, sg, 1u,
                           ws.inv0.as<uint64_t>(), ws.errs.as<int>());
    } else {
        hipLaunchKernelGGL(k_scatter_inverse<uint32_t>, grid, dim3(kBlock), 0, s, it, sg, 1u,
                           ws.inv0.as<uint32_t>(), ws.errs.as<int>());
        hipLaunchKernelGGL(k_check_inverse<uint32_t>, grid, dim3(kBlock), 0, s, it, sg, 1u,
                           ws.inv0.as<uint32_t>(), ws.errs.as<int>());
    }
    BH_CHECK(hipGetLastError());
    int flag = 0;
    BH_CHECK(hipMemcpyAsync(&flag, ws.errs.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
    BH_CHECK(hipStreamSynchronize(s));
    return flag == 0;
}

// The reference validates the permutation before it looks at the format
// (src/benes.cpp:361-363, src/cuda/benes_cuda.cu:1158): when the format has
// been rejected (`st` holds that error), the input is checked on the device
// and a non-permutation reports that instead.
void permutation_first(const void* perm, bool on_device, uint64_t n, uint32_t width,
                       hipStream_t user_stream, Status& st) {
    if (!perm) return;
    Status vs;
    bool bad = false;
    with_workspace(vs, [&](Workspace& ws) {
        const void* dsrc = perm;
        if (!on_device) {
            ws.input.ensure(n * width);
            ws.ring.upload(ws.input.as<uint8_t>(),
                           {HostPiece{static_cast<uint8_t*>(const_cast<void*>(perm)), 0, n * width}},
                           ws.stream);
            dsrc = ws.input.ptr;
        }
        if (user_stream) stream_after(ws.stream, user_stream);
        bad = !device_is_permutation(ws, ws.stream, dsrc, n, width);
        if (user_stream) stream_after(user_stream, ws.stream);
    });
    if (bad) {
        st = Status();
        st.set(kInvalidArgument, kNotPermutation);
    }
}

int compress_one(const void* perm, bool on_device, uint64_t n, uint32_t width,
                 const benes_hip_format* format, uint8_t* dst_host, uint8_t* dst_dev,
                 uint64_t capacity, uint64_t* out_size, hipStream_t user_stream,
                 double* device_seconds, char* err, size_t err_len) {
    Status st;
    std::vector<Job> jobs(1);
    if (!check_input(n, width, st)) return finish(st, err, err_len);
    if (!prepare_job(perm, n, width, format, jobs[0], st)) {
        permutation_first(perm, on_device, n, width, user_stream, st);
        return finish(st, err, err_len);
    }
    const uint64_t end = jobs[0].lay.plan.end_offset;
    if (out_size) *out_size = end;
    if ((!dst_host && !dst_dev) || capacity < end) {
        st.set(kBufferTooSmall, "output buffer is too small");
        return finish(st, err, err_len);
    }
    with_workspace(st, [&](Workspace& ws) {
        if (!on_device) {
            const auto t0 = std::chrono::steady_clock::now();
            ws.input.ensure(n * width);
            ws.ring.upload(ws.input.as<uint8_t>(),
                           {HostPiece{static_cast<uint8_t*>(const_cast<void*>(perm)), 0, n * width}},
                           ws.stream);
            jobs[0].dsrc = ws.input.ptr;
            note_host_time("host:upload", t0);
        }
        // Write straight into the caller's buffer when it can take whole words.
        const bool direct = dst_dev && (reinterpret_cast<uintptr_t>(dst_dev) & 7) == 0 &&
                            capacity >= round8(end);
        std::vector<int> flags;
        const auto t_run = std::chrono::steady_clock::now();
        run_batch(ws, jobs, direct ? dst_dev : nullptr, user_stream, flags, device_seconds);
        note_host_time("host:run", t_run);
        if (flags[0]) {
            st.set(kInvalidArgument, kNotPermutation);
            return;
        }
        if (dst_dev && !direct) {
            BH_CHECK(hipMemcpyAsync(dst_dev, ws.blob.ptr, end, hipMemcpyDeviceToDevice,
                                    ws.stream));
            BH_CHECK(hipStreamSynchronize(ws.stream));
        }
        if (dst_host) {
            const auto t0 = std::chrono::steady_clock::now();
            ws.ring.download(ws.blob.as<uint8_t>(), {HostPiece{dst_host, 0, end}}, ws.stream);
            note_host_time("host:download", t0);
        }
    });
    return finish(st, err, err_len);
}

// ---------------------------------------------------------------------------
// Allocation-free compress (benes_hip_compress_device_noalloc): every device
// buffer a call needs beside its input and output is a slice of `scratch`, a
// fixed array that each device holds from the time the library's code object
// is loaded on it until the process ends.  The host-side name of the array
// is not a device address; the device's copy is located once per device with
// hipGetSymbolAddress and only that address is used afterwards.  Calls on one
// device share the array: a call clears what must start at known values
// (blob, inverse, flags, counters) before its kernels run, and calls that
// overlap on one device race.  Work goes on the caller's stream; no stream,
// event, pinned or device memory is created.

constexpr uint64_t kNoAllocMaxEntries = BENES_HIP_NOALLOC_MAX_ENTRIES;
constexpr uint64_t kNoAllocScratchBytes = BENES_HIP_NOALLOC_SCRATCH_BYTES;
// Deferred long walks per depth (an edge past it walks its own path).
constexpr uint64_t kNoAllocDeferCap = uint64_t{1} << 14;

// Worst case at 2^20 entries (the smallest subtree shape, 64-bit words
// forced): 8 MiB inverse, four 4 MiB level arrays, 8 MiB of union-find
// words, 384 KiB of deferred edges, under 300 KiB of rows, descriptors,
// segments, plans and the union check's counts.
static_assert(kNoAllocScratchBytes >= 33 * (uint64_t{1} << 20) &&
                  kNoAllocMaxEntries == (uint64_t{1} << 20),
              "scratch sized for 2^20 entries");

static __device__ __align__(256) uint64_t scratch[kNoAllocScratchBytes / 8];

struct NoAllocState {
    std::mutex mu;
    bool ready = false;
    uint8_t* ptr = nullptr;  // this device's copy of `scratch`
    Arena arena;
    Workspace ws;            // every buffer bound to the arena
};

NoAllocState& noalloc_for(int dev) {
    static std::mutex mu;
    static std::vector<NoAllocState*> all;
    std::lock_guard<std::mutex> lock(mu);
    if (static_cast<size_t>(dev) >= all.size()) all.resize(dev + 1, nullptr);
    if (!all[dev]) all[dev] = new NoAllocState();
    return *all[dev];
}

int compress_noalloc(const void* d_perm, uint64_t n, uint32_t width,
                     const benes_hip_format* format, uint8_t* d_out, uint64_t capacity,
                     uint64_t* out_size, hipStream_t stream, char* err, size_t err_len) {
    Status st;
    std::vector<Job> jobs(1);
    if (!check_input(n, width, st)) return finish(st, err, err_len);
    const bool format_ok = prepare_job(d_perm, n, width, format, jobs[0], st);
    const uint64_t end = format_ok ? jobs[0].lay.plan.end_offset : 0;
    if (out_size) *out_size = end;
    if (n > kNoAllocMaxEntries) {
        st = Status();
        st.set(kLength, "permutation too large for the allocation-free entry point (at most 2^20 "
                        "entries)");
        return finish(st, err, err_len);
    }
    if (!d_perm) {
        st.set(kInvalidArgument, "permutation pointer is null");
        return finish(st, err, err_len);
    }
    if (format_ok &&
        (!d_out || (reinterpret_cast<uintptr_t>(d_out) & 7) != 0 || capacity < round8(end))) {
        st.set(kBufferTooSmall, "output buffer must be 8-byte aligned and hold the blob rounded up "
                                "to whole 8-byte words");
        return finish(st, err, err_len);
    }
    if (!benes_hip_available()) {
        st.set(kRuntime, "hip backend is unavailable");
        return finish(st, err, err_len);
    }
    const int dev = current_device(st);
    if (dev < 0) return finish(st, err, err_len);
    NoAllocState& na = noalloc_for(dev);
    std::lock_guard<std::mutex> lock(na.mu);
    try {
        if (!na.ready) {
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                st.set(kRuntime, std::string("cannot locate the allocation-free scratch: ") +
                                     hipGetErrorString(e));
                return finish(st, err, err_len);
            }
            na.ptr = static_cast<uint8_t*>(p);
            na.arena.base = na.ptr;
            na.arena.cap = kNoAllocScratchBytes;
            na.ws.bind(&na.arena);
            na.ready = true;
        }
        na.arena.used = 0;
        if (!format_ok) {
            // The format was rejected: the permutation is checked first, as
            // the reference does (src/benes.cpp:361-363).
            if (!device_is_permutation(na.ws, stream, d_perm, n, width)) {
                st = Status();
                st.set(kInvalidArgument, kNotPermutation);
            }
            return finish(st, err, err_len);
        }
        Profiler prof;  // never on here: its events would be allocations
        const int shape = Tuner::instance().settled(Tuner::class_of(jobs));
        enqueue_batch(na.ws, stream, jobs, prof, d_out, shape, kNoAllocDeferCap);
        int flag = 0;
        BH_CHECK(hipMemcpyAsync(&flag, na.ws.errs.ptr, sizeof(int), hipMemcpyDeviceToHost, stream));
        BH_CHECK(hipStreamSynchronize(stream));
        if (flag) st.set(kInvalidArgument, kNotPermutation);
    } catch (const HipError& e) {
        st.set(kRuntime, std::string("hip error: ") + hipGetErrorString(e.code) + " in " + e.what);
    } catch (const std::length_error& e) {
        st.set(kLength, e.what());
    }
    return finish(st, err, err_len);
}

// Whole-range positional reads and writes of a file descriptor.
bool pread_all(int fd, uint8_t* dst, uint64_t len, uint64_t off) {
    while (len) {
        const ssize_t r = ::pread(fd, dst, std::min<uint64_t>(len, uint64_t{1} << 30),
                                  static_cast<off_t>(off));
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return false;
        dst += r;
        off += static_cast<uint64_t>(r);
        len -= static_cast<uint64_t>(r);
    }
    return true;
}

bool pwrite_all(int fd, const uint8_t* src, uint64_t len, uint64_t off) {
    while (len) {
        const ssize_t r = ::pwrite(fd, src, std::min<uint64_t>(len, uint64_t{1} << 30),
                                   static_cast<off_t>(off));
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return false;
        src += r;
        off += static_cast<uint64_t>(r);
        len -= static_cast<uint64_t>(r);
    }
    return true;
}

struct FdCloser {
    int fd;
    ~FdCloser() {
        if (fd >= 0) ::close(fd);
    }
};

// compress_file (src/benes.cpp:498-553) over the host path of
// benes_hip_compress: the file is read whole into host memory, compressed as
// a host array of little-endian values (the staging ring takes it to HBM,
// where it is validated), and the blob is written out.  Checks and messages
// follow the reference in its order: the input file, its size, the read,
// then compression (with compress's own error order), then the output,
// created or truncated only once compression has succeeded.
int compress_file(const char* input_path, const char* output_path, uint32_t width,
                  const benes_hip_format* format, uint64_t* out_size, double* device_seconds,
                  char* err, size_t err_len) {
    Status st;
    if (out_size) *out_size = 0;
    FdCloser in{input_path ? ::open(input_path, O_RDONLY | O_CLOEXEC) : -1};
    struct stat sb {};
    if (in.fd < 0 || ::fstat(in.fd, &sb) != 0) {
        st.set(kRuntime, "cannot open permutation input file");
        return finish(st, err, err_len);
    }
    const uint64_t length = sb.st_size > 0 ? static_cast<uint64_t>(sb.st_size) : 0;
    if ((width != 4 && width != 8) || length == 0 || length % width != 0) {
        st.set(kInvalidArgument, "permutation input has an invalid element width or size");
        return finish(st, err, err_len);
    }
    std::vector<uint8_t> raw;
    std::vector<uint8_t> blob;
    try {
        raw.resize(length);
    } catch (const std::bad_alloc&) {
        st.set(kLength, "allocation failed");
        return finish(st, err, err_len);
    }
    if (!pread_all(in.fd, raw.data(), length, 0)) {
        st.set(kRuntime, "failed to read permutation input file");
        return finish(st, err, err_len);
    }
    ::close(in.fd);
    in.fd = -1;
    const uint64_t n = length / width;
    // A size query, then the compress into a host buffer of that size.
    uint64_t size = 0;
    int rc = compress_one(raw.data(), false, n, width, format, nullptr, nullptr, 0, &size,
                          nullptr, nullptr, err, err_len);
    if (rc != BENES_HIP_BUFFER_TOO_SMALL) {
        if (rc != BENES_HIP_OK) return rc;
        st.set(kLogic, "size query failed");
        return finish(st, err, err_len);
    }
    try {
        blob.resize(size);
    } catch (const std::bad_alloc&) {
        st.set(kLength, "allocation failed");
        return finish(st, err, err_len);
    }
    rc = compress_one(raw.data(), false, n, width, format, blob.data(), nullptr, size, &size,
                      nullptr, device_seconds, err, err_len);
    if (rc != BENES_HIP_OK) return rc;
    raw = std::vector<uint8_t>();
    const auto t0 = std::chrono::steady_clock::now();
    FdCloser out{output_path ? ::open(output_path, O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0666)
                             : -1};
    bool ok = out.fd >= 0 && pwrite_all(out.fd, blob.data(), size, 0);
    if (out.fd >= 0) {
        ok = ::close(out.fd) == 0 && ok;
        out.fd = -1;
    }
    note_host_time("host:file_write", t0);
    if (!ok) {
        st.set(kRuntime, "failed to write compressed permutation file");
        return finish(st, err, err_len);
    }
    if (out_size) *out_size = size;
    return finish(st, err, err_len);
}

// Prepares the jobs of a batch; items failing host checks get their status
// and are left out (`index` maps jobs back to items).
void prepare_batch(benes_hip_item* items, uint64_t count, std::vector<Job>& jobs,
                   std::vector<uint64_t>& index) {
    // A job holds its 8 KB plan: each is built in place, and the vector is
    // sized up front so that growing it copies none.
    jobs.reserve(jobs.size() + count);
    index.reserve(index.size() + count);
    for (uint64_t i = 0; i < count; ++i) {
        Status st;
        Job& job = jobs.emplace_back();
        items[i].out_size = 0;
        items[i].status = BENES_HIP_OK;
        items[i].message[0] = 0;
        if (!prepare_job(items[i].perm, items[i].n, items[i].width_bytes, &items[i].format, job,
                         st)) {
            jobs.pop_back();
            items[i].status = st.code;
            std::strncpy(items[i].message, st.message.c_str(), sizeof(items[i].message) - 1);
            items[i].message[sizeof(items[i].message) - 1] = 0;
            continue;
        }
        items[i].out_size = job.lay.plan.end_offset;
        index.push_back(i);
    }
}

void mark_invalid(benes_hip_item& item) {
    item.status = BENES_HIP_INVALID_ARGUMENT;
    std::strncpy(item.message, kNotPermutation, sizeof(item.message) - 1);
    item.message[sizeof(item.message) - 1] = 0;
}


// ---------------------------------------------------------------------------
// Request loop (continuous batching)
//
// Requests queue in arrival order.  The server owns `lanes` lanes, each a
// stream with its own workspace (the routing state of its batch stays in
// HBM), its own pinned staging ring and its own host thread.  A scheduler
// thread hands the longest prefix of the queue that fits a lane's share of
// the HBM budget and the batch cap to the first idle lane; the lane thread
// stages the inputs, queues the whole batch, drains the blobs into the
// requests' buffers and wakes their waiters.  Requests that arrive meanwhile
// go to the next idle lane, so the staging of one batch overlaps the device
// work of another, and batches on different lanes overlap on the device.

struct Request {
    uint64_t id = 0;
    Job job;
    const void* host_perm = nullptr;
    uint8_t* out = nullptr;
    uint64_t out_capacity = 0;
    uint64_t footprint = 0;
    std::chrono::steady_clock::time_point submitted, done;
    int status = -1;  // -1 pending
    std::string message;
};

struct Lane {
    Workspace ws;
    std::vector<Request*> batch;  // set by the scheduler, emptied by the lane thread
    std::vector<Job> jobs;
    std::vector<int> outcome;     // per request of the batch, published under the lock
    std::thread thread;
    std::condition_variable cv;  // wakes the lane thread (waits on Server::mu)
    bool busy = false;           // a batch is assigned and not yet finished
    uint64_t held = 0;           // budget bytes of the batch in flight
};

uint64_t request_footprint(uint64_t n, uint32_t width, uint64_t blob_bytes) {
    return n * (width + 28) + n / 4 + round8(blob_bytes) + (64u << 10);
}

}  // namespace

struct Server {
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::unordered_map<uint64_t, Request*> requests;
    std::vector<Lane*> lanes;
    std::thread thread;
    bool stop = false;        // no more requests: drain and end the scheduler
    bool stop_lanes = false;  // end the lane threads (after the scheduler)
    int device = 0;
    uint64_t budget = 0, lane_cap = 0, max_batch = 0, next_id = 1;
    uint64_t in_flight_bytes = 0, peak_bytes = 0, batches = 0, completed = 0, entries = 0;
    uint64_t max_concurrent = 0;
    std::string fatal;

    void run();
    bool assign(Lane& lane);
    void serve(Lane& lane);
    void process(Lane& lane);
};

// A batch closes once it holds this many entries: small requests are
// grouped until their launches are worth it, while large ones go alone, so
// that the staging of one lane's batch overlaps the device work of another's
// (measured over the served workload: 1894 -> 2258 Mentries/s at 3 lanes when
// batches were capped at 16 requests instead of 64).
constexpr uint64_t kBatchEntryTarget = uint64_t{1} << 22;

bool Server::assign(Lane& lane) {
    // Caller holds mu.  Take the queue prefix that fits this lane.
    uint64_t bytes = 0, entries = 0;
    lane.batch.clear();
    while (!queue.empty() && lane.batch.size() < max_batch && entries < kBatchEntryTarget) {
        Request* r = queue.front();
        if (bytes + r->footprint > lane_cap || in_flight_bytes + bytes + r->footprint > budget) {
            break;
        }
        bytes += r->footprint;
        entries += r->job.n;
        lane.batch.push_back(r);
        queue.pop_front();
    }
    if (lane.batch.empty()) return false;
    lane.held = bytes;
    in_flight_bytes += bytes;
    peak_bytes = std::max(peak_bytes, in_flight_bytes);
    ++batches;
    lane.busy = true;
    lane.cv.notify_one();
    return true;
}

// Runs one batch on the lane (without holding mu): stage in, queue, drain.
void Server::process(Lane& lane) {
    std::vector<Request*>& batch = lane.batch;
    BH_CHECK(hipSetDevice(device));
    lane.ws.init(device);
    uint64_t in_bytes = 0;
    for (Request* r : batch) in_bytes += round8(r->job.n * r->job.width);
    lane.ws.input.ensure(in_bytes);
    lane.ws.errs.ensure(sizeof(int) * batch.size());
    lane.jobs.clear();
    std::vector<HostPiece> in;
    uint64_t at = 0;
    for (Request* r : batch) {
        const uint64_t len = r->job.n * r->job.width;
        in.push_back(HostPiece{static_cast<uint8_t*>(const_cast<void*>(r->host_perm)), at, len});
        Job j = r->job;
        j.dsrc = lane.ws.input.as<uint8_t>() + at;
        lane.jobs.push_back(std::move(j));
        at += round8(len);
    }
    const hipStream_t s = lane.ws.stream;
    auto t0 = std::chrono::steady_clock::now();
    lane.ws.ring.upload(lane.ws.input.as<uint8_t>(), in, s);
    note_host_time("host:serve_upload", t0);
    Profiler prof;
    {
        std::lock_guard<std::mutex> plock(g_prof_mu);
        prof.on = g_profiling;
    }
    prof.stream = s;
    t0 = std::chrono::steady_clock::now();
    const uint32_t cls = Tuner::class_of(lane.jobs);
    bool trial = false;
    const int shape = Tuner::instance().pick(cls, trial);
    BH_CHECK(hipEventRecord(lane.ws.ev_begin, s));
    enqueue_batch(lane.ws, s, lane.jobs, prof, nullptr, shape, kDeferCap);
    BH_CHECK(hipEventRecord(lane.ws.ev_end, s));
    note_host_time("host:serve_enqueue", t0);
    std::vector<int> flags(batch.size());
    BH_CHECK(hipMemcpyAsync(flags.data(), lane.ws.errs.ptr, sizeof(int) * batch.size(),
                            hipMemcpyDeviceToHost, s));
    BH_CHECK(hipStreamSynchronize(s));
    float batch_ms = 0;
    BH_CHECK(hipEventElapsedTime(&batch_ms, lane.ws.ev_begin, lane.ws.ev_end));
    Tuner::instance().report(cls, shape, trial, batch_ms);
    note_host_time("host:serve_device", t0);
    prof.collect();
    // Outcomes stay with the lane until the blobs are in the callers' buffers;
    // a request's status is only set under the server lock afterwards, so a
    // waiter never sees it finished while its buffer is still being written.
    std::vector<HostPiece> out;
    lane.outcome.assign(batch.size(), kOk);
    for (size_t k = 0; k < batch.size(); ++k) {
        const Request* r = batch[k];
        const uint64_t end = r->job.lay.plan.end_offset;
        if (flags[k]) {
            lane.outcome[k] = kInvalidArgument;
        } else if (!r->out || r->out_capacity < end) {
            lane.outcome[k] = kBufferTooSmall;
        } else {
            out.push_back(HostPiece{r->out, lane.jobs[k].blob * 8, end});
        }
    }
    t0 = std::chrono::steady_clock::now();
    lane.ws.ring.download(lane.ws.blob.as<uint8_t>(), out, s);
    note_host_time("host:serve_download", t0);
}

void Server::serve(Lane& lane) {
    std::unique_lock<std::mutex> lock(mu);
    while (true) {
        lane.cv.wait(lock, [&] { return stop_lanes || !lane.batch.empty(); });
        if (lane.batch.empty()) break;  // stop_lanes with nothing assigned
        std::string error;
        lock.unlock();
        try {
            process(lane);
        } catch (const HipError& e) {
            error = std::string("hip error: ") + hipGetErrorString(e.code) + " in " + e.what;
        } catch (const std::length_error& e) {
            error = e.what();
        } catch (const std::bad_alloc&) {
            error = "allocation failed";
        }
        lock.lock();
        const auto now = std::chrono::steady_clock::now();
        for (size_t k = 0; k < lane.batch.size(); ++k) {
            Request* r = lane.batch[k];
            if (!error.empty()) {
                r->status = kRuntime;
                r->message = error;
            } else {
                r->status = lane.outcome[k];
                if (r->status == kOk) entries += r->job.n;
                if (r->status == kInvalidArgument) r->message = kNotPermutation;
                if (r->status == kBufferTooSmall) r->message = "output buffer is too small";
            }
            r->done = now;
            ++completed;
        }
        if (!error.empty() && fatal.empty()) fatal = error;
        lane.batch.clear();
        lane.busy = false;
        in_flight_bytes -= lane.held;
        lane.held = 0;
        cv_done.notify_all();
        cv_work.notify_one();
    }
}

void Server::run() {
    std::unique_lock<std::mutex> lock(mu);
    while (true) {
        const bool any_busy =
            std::any_of(lanes.begin(), lanes.end(), [](Lane* l) { return l->busy; });
        if (stop && queue.empty() && !any_busy) break;
        if (!fatal.empty()) {
            // Fail everything still waiting; lanes in flight finish normally.
            for (Request* r : queue) {
                r->status = kRuntime;
                r->message = fatal;
                r->done = std::chrono::steady_clock::now();
                ++completed;
            }
            queue.clear();
            cv_done.notify_all();
        }
        for (Lane* lane : lanes) {
            if (!lane->busy && !queue.empty()) assign(*lane);
        }
        uint64_t busy = 0;
        for (Lane* lane : lanes) busy += lane->busy ? 1 : 0;
        max_concurrent = std::max(max_concurrent, busy);
        cv_work.wait(lock);
    }
}

}  // namespace benes_hip

using namespace benes_hip;

extern "C" {

int benes_hip_available(void) {
    static int cached = -1;
    static std::once_flag once;
    std::call_once(once, [] {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
            cached = 0;
            return;
        }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
            cached = 0;
            return;
        }
        cached = std::strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
    });
    return cached;
}

int benes_hip_compress(const void* perm, uint64_t n, uint32_t width_bytes,
                       const benes_hip_format* format, uint8_t* out, uint64_t out_capacity,
                       uint64_t* out_size, double* device_seconds, char* err, size_t err_len) {
    return compress_one(perm, false, n, width_bytes, format, out, nullptr, out_capacity,
                        out_size, nullptr, device_seconds, err, err_len);
}

int benes_hip_compress_file(const char* input_path, const char* output_path,
                            uint32_t width_bytes, const benes_hip_format* format,
                            uint64_t* out_size, double* device_seconds, char* err,
                            size_t err_len) {
    return compress_file(input_path, output_path, width_bytes, format, out_size, device_seconds,
                         err, err_len);
}

int benes_hip_compress_device(const void* d_perm, uint64_t n, uint32_t width_bytes,
                              const benes_hip_format* format, uint8_t* d_out,
                              uint64_t out_capacity, uint64_t* out_size, void* stream, char* err,
                              size_t err_len) {
    return compress_one(d_perm, true, n, width_bytes, format, nullptr, d_out, out_capacity,
                        out_size, static_cast<hipStream_t>(stream), nullptr, err, err_len);
}

int benes_hip_compress_device_noalloc(const void* d_perm, uint64_t n, uint32_t width_bytes,
                                      const benes_hip_format* format, uint8_t* d_out,
                                      uint64_t out_capacity, uint64_t* out_size, void* stream,
                                      char* err, size_t err_len) {
    return compress_noalloc(d_perm, n, width_bytes, format, d_out, out_capacity, out_size,
                            static_cast<hipStream_t>(stream), err, err_len);
}

int benes_hip_batch_layout(benes_hip_item* items, uint64_t count, uint64_t* total_bytes) {
    std::vector<Job> jobs;
    std::vector<uint64_t> index;
    prepare_batch(items, count, jobs, index);
    uint64_t words = 0;
    for (size_t k = 0; k < jobs.size(); ++k) {
        items[index[k]].out_offset = words * 8;
        words += ceil_div(jobs[k].lay.plan.end_offset, 8);
    }
    if (total_bytes) *total_bytes = words * 8;
    return BENES_HIP_OK;
}

int benes_hip_compress_batch(benes_hip_item* items, uint64_t count, double* device_seconds,
                             char* err, size_t err_len) {
    Status st;
    std::vector<Job> jobs;
    std::vector<uint64_t> index;
    prepare_batch(items, count, jobs, index);
    if (jobs.empty()) return finish(st, err, err_len);
    with_workspace(st, [&](Workspace& ws) {
        uint64_t bytes = 0;
        std::vector<uint64_t> at(jobs.size());
        for (size_t k = 0; k < jobs.size(); ++k) {
            at[k] = bytes;
            bytes += round8(jobs[k].n * jobs[k].width);
        }
        ws.input.ensure(bytes);
        std::vector<HostPiece> in;
        for (size_t k = 0; k < jobs.size(); ++k) {
            in.push_back(HostPiece{static_cast<uint8_t*>(const_cast<void*>(jobs[k].dsrc)), at[k],
                                   jobs[k].n * jobs[k].width});
            jobs[k].dsrc = ws.input.as<uint8_t>() + at[k];
        }
        ws.ring.upload(ws.input.as<uint8_t>(), in, ws.stream);
        std::vector<int> flags;
        run_batch(ws, jobs, nullptr, nullptr, flags, device_seconds);
        std::vector<HostPiece> out;
        for (size_t k = 0; k < jobs.size(); ++k) {
            benes_hip_item& item = items[index[k]];
            if (flags[k]) {
                mark_invalid(item);
                continue;
            }
            if (!item.out || item.out_capacity < item.out_size) {
                item.status = BENES_HIP_BUFFER_TOO_SMALL;
                continue;
            }
            out.push_back(HostPiece{item.out, jobs[k].blob * 8, item.out_size});
        }
        ws.ring.download(ws.blob.as<uint8_t>(), out, ws.stream);
    });
    return finish(st, err, err_len);
}

int benes_hip_compress_batch_device(benes_hip_item* items, uint64_t count, uint8_t* d_out,
                                    uint64_t out_capacity, void* stream, double* device_seconds,
                                    char* err, size_t err_len) {
    Status st;
    std::vector<Job> jobs;
    std::vector<uint64_t> index;
    prepare_batch(items, count, jobs, index);
    if (jobs.empty()) return finish(st, err, err_len);
    const uint64_t need = batch_words(jobs) * 8;
    if (!d_out || out_capacity < need || (reinterpret_cast<uintptr_t>(d_out) & 7) != 0) {
        st.set(kBufferTooSmall, "batch output buffer is too small or not 8-byte aligned");
        return finish(st, err, err_len);
    }
    with_workspace(st, [&](Workspace& ws) {
        std::vector<int> flags;
        run_batch(ws, jobs, d_out, static_cast<hipStream_t>(stream), flags, device_seconds);
        for (size_t k = 0; k < jobs.size(); ++k) {
            benes_hip_item& item = items[index[k]];
            item.out_offset = jobs[k].blob * 8;
            if (flags[k]) mark_invalid(item);
        }
    });
    return finish(st, err, err_len);
}

int benes_hip_generate_device(uint32_t kind, uint64_t n, uint64_t seed, void* d_out,
                              uint32_t width_bytes, void* stream) {
    if (kind != BENES_HIP_PERM_IDENTITY && kind != BENES_HIP_PERM_REVERSE &&
        kind != BENES_HIP_PERM_KEYED) {
        return BENES_HIP_INVALID_ARGUMENT;
    }
    if ((width_bytes != 4 && width_bytes != 8) || (width_bytes == 4 && n > 0x100000000ull) ||
        (n && !d_out)) {
        return BENES_HIP_INVALID_ARGUMENT;
    }
    if (n == 0) return BENES_HIP_OK;
    const auto s = static_cast<hipStream_t>(stream);
    const uint64_t blocks = std::min<uint64_t>(ceil_div(n, kBlock), uint64_t{1} << 20);
    if (width_bytes == 4) {
        hipLaunchKernelGGL(k_generate<uint32_t>, dim3(static_cast<uint32_t>(blocks)), dim3(kBlock),
                           0, s, kind, n, seed, static_cast<uint32_t*>(d_out));
    } else {
        hipLaunchKernelGGL(k_generate<uint64_t>, dim3(static_cast<uint32_t>(blocks)), dim3(kBlock),
                           0, s, kind, n, seed, static_cast<uint64_t*>(d_out));
    }
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
        return BENES_HIP_RUNTIME;
    }
    return BENES_HIP_OK;
}

int benes_hip_lookup_device(const uint8_t* d_blob, uint64_t blob_size, const uint64_t* d_values,
                            uint64_t count, int inverse, void* d_out, uint32_t width_bytes,
                            void* stream, char* err, size_t err_len) {
    Status st;
    if ((width_bytes != 4 && width_bytes != 8) || !d_blob || (count && !d_out)) {
        st.set(kInvalidArgument, "invalid lookup arguments");
        return finish(st, err, err_len);
    }
    with_workspace(st, [&](Workspace& ws) {
        run_lookup(ws, d_blob, blob_size, d_values, count, inverse, d_out, width_bytes,
                   static_cast<hipStream_t>(stream), st);
    });
    return finish(st, err, err_len);
}

int benes_hip_lookup(const uint8_t* blob, uint64_t blob_size, const uint64_t* values,
                     uint64_t count, int inverse, void* out, uint32_t width_bytes, char* err,
                     size_t err_len) {
    Status st;
    if ((width_bytes != 4 && width_bytes != 8) || !blob || (count && (!out || !values))) {
        st.set(kInvalidArgument, "invalid lookup arguments");
        return finish(st, err, err_len);
    }
    with_workspace(st, [&](Workspace& ws) {
        ws.blob.ensure(std::max<uint64_t>(blob_size, 1));
        ws.input.ensure(std::max<uint64_t>(count * 8, 8));
        ws.inv0.ensure(std::max<uint64_t>(count * width_bytes, 8));
        ws.ring.upload(ws.blob.as<uint8_t>(),
                       {HostPiece{const_cast<uint8_t*>(blob), 0, blob_size}}, ws.stream);
        if (count) {
            ws.ring.upload(ws.input.as<uint8_t>(),
                           {HostPiece{reinterpret_cast<uint8_t*>(const_cast<uint64_t*>(values)),
                                      0, count * 8}},
                           ws.stream);
        }
        run_lookup(ws, ws.blob.as<uint8_t>(), blob_size, ws.input.as<uint64_t>(), count,
                   inverse, ws.inv0.ptr, width_bytes, nullptr, st);
        if (st.ok() && count) {
            ws.ring.download(ws.inv0.as<uint8_t>(),
                             {HostPiece{static_cast<uint8_t*>(out), 0, count * width_bytes}},
                             ws.stream);
        }
    });
    return finish(st, err, err_len);
}

int benes_hip_set_subtree_shape(uint32_t log2_entries, uint32_t threads) {
    if (log2_entries == 0 && threads == 0) {
        Tuner::instance().pin(-1);
        return BENES_HIP_OK;
    }
    for (int i = 0; i < kNumShapes; ++i) {
        if (kShapes[i].log2 == log2_entries && kShapes[i].threads == threads) {
            Tuner::instance().pin(i);
            return BENES_HIP_OK;
        }
    }
    return BENES_HIP_INVALID_ARGUMENT;
}

int benes_hip_set_hook_binning(int mode) {
    if (mode < -1 || mode > 1) return BENES_HIP_INVALID_ARGUMENT;
    g_hook_binning.store(mode);
    return BENES_HIP_OK;
}

int benes_hip_set_plain_union(int mode) {
    if (mode < -1 || mode > 1) return BENES_HIP_INVALID_ARGUMENT;
    g_plain_union.store(mode);
    return BENES_HIP_OK;
}

int benes_hip_set_colouring(int mode) {
    if (mode < 0 || mode > 1) return BENES_HIP_INVALID_ARGUMENT;
    g_colouring.store(mode);
    return BENES_HIP_OK;
}

int benes_hip_set_wide_words(int mode) {
    if (mode < 0 || mode > 1) return BENES_HIP_INVALID_ARGUMENT;
    g_wide_words.store(mode);
    return BENES_HIP_OK;
}

int benes_hip_autotune_report(char* out, size_t len) {
    const std::string text = Tuner::instance().describe();
    if (out && len) {
        std::strncpy(out, text.c_str(), len - 1);
        out[len - 1] = 0;
    }
    return static_cast<int>(text.size());
}

void benes_hip_release_workspace(void) {
    int dev = -1;
    if (!benes_hip_available() || hipGetDevice(&dev) != hipSuccess) return;
    Workspace& ws = workspace_for(dev);
    std::lock_guard<std::mutex> lock(ws.mu);
    if (ws.stream) (void)hipStreamSynchronize(ws.stream);
    for (DevBuf* b : {&ws.input, &ws.level[0], &ws.level[1], &ws.ilevel[0], &ws.ilevel[1],
                      &ws.inv0, &ws.uf, &ws.srows, &ws.trows, &ws.desc, &ws.plans, &ws.items,
                      &ws.segs, &ws.errs, &ws.blob, &ws.deferred, &ws.deferred_counts,
                      &ws.bin_fill, &ws.rulers, &ws.lbits}) {
        b->release();
    }
}

int benes_hip_encoded_size(uint64_t n, const benes_hip_format* format, uint64_t* out_size,
                           char* err, size_t err_len) {
    Status st;
    FormatArgs fa;
    Layout lay;
    if (format_args(format, fa, st) && build_layout(n, fa, lay, st)) {
        *out_size = lay.plan.end_offset;
    }
    return finish(st, err, err_len);
}

double benes_hip_last_device_seconds(void) { return g_device_seconds; }

void benes_hip_set_profiling(int enable) {
    std::lock_guard<std::mutex> lock(g_prof_mu);
    g_profiling = enable != 0;
    g_prof_stats.clear();
}

int benes_hip_kernel_stats(int index, char* name, size_t name_len, double* total_ms,
                           uint64_t* launches) {
    std::lock_guard<std::mutex> lock(g_prof_mu);
    if (index < 0 || static_cast<size_t>(index) >= g_prof_stats.size()) return 0;
    const KernelStat& k = g_prof_stats[index];
    if (name && name_len) {
        std::strncpy(name, k.name.c_str(), name_len - 1);
        name[name_len - 1] = 0;
    }
    if (total_ms) *total_ms = k.ms;
    if (launches) *launches = k.launches;
    return 1;
}

int benes_hip_layout_of(uint64_t n, const benes_hip_format* format, benes_hip_layout* out,
                        char* err, size_t err_len) {
    Status st;
    FormatArgs fa;
    Layout lay;
    if (format_args(format, fa, st) && build_layout(n, fa, lay, st)) {
        std::memset(out, 0, sizeof(*out));
        out->size = n;
        out->network_size = lay.network_size;
        out->input_offset = lay.plan.input_offset;
        out->output_offset = lay.plan.output_offset;
        out->middle_offset = lay.plan.middle_offset;
        out->end_offset = lay.plan.end_offset;
        out->switch_bits = lay.plan.switch_bits_total;
        out->middle_bits = lay.plan.middle_bits_total;
        out->routing_depth = lay.plan.depth;
        out->middle_group_log2 = static_cast<uint8_t>(lay.plan.leaf_log2);
        out->middle_placement = static_cast<uint8_t>(lay.plan.placement);
        out->n_input_clusters = static_cast<uint8_t>(lay.input_clusters.size());
        out->n_output_clusters = static_cast<uint8_t>(lay.output_clusters.size());
        std::memcpy(out->input_clusters, lay.input_clusters.data(), lay.input_clusters.size());
        std::memcpy(out->output_clusters, lay.output_clusters.data(), lay.output_clusters.size());
    }
    return finish(st, err, err_len);
}

int benes_hip_node_address(uint64_t n, const benes_hip_format* format, int side, uint32_t depth,
                           uint64_t path, benes_hip_address* out, char* err, size_t err_len) {
    Status st;
    FormatArgs fa;
    Layout lay;
    if (!format_args(format, fa, st) || !build_layout(n, fa, lay, st)) {
        return finish(st, err, err_len);
    }
    const Plan& p = lay.plan;
    if (depth > p.depth || (depth > 0 && node_count(n, depth - 1, path >> 1) <= p.leaf) ||
        (depth < 64 && path >= (uint64_t{1} << depth))) {
        st.set(kInvalidArgument, "no such node");
        return finish(st, err, err_len);
    }
    std::memset(out, 0, sizeof(*out));
    out->count = node_count(n, depth, path);
    out->offset = node_offset(n, depth, path);
    if (out->count > p.leaf) {
        const SideAddr a = node_address(p, side, depth, path);
        out->packet_base = a.packet_base;
        out->full_packet_bits = a.full_packet_bits;
        out->full_packet_prefix = a.full_prefix;
        out->tail_packet_prefix = a.tail_prefix;
        out->pair_span = uint64_t{1} << a.span_log2;
        out->tail_packet = a.tail_packet;
    } else {
        out->rank_bit = leaf_rank_bit(p, depth, path);
    }
    return finish(st, err, err_len);
}

uint64_t benes_hip_address_bit(const benes_hip_address* a, uint64_t pair) {
    const uint64_t packet = pair / a->pair_span;
    return a->packet_base + packet * a->full_packet_bits +
           (packet == a->tail_packet ? a->tail_packet_prefix : a->full_packet_prefix) +
           pair % a->pair_span;
}

struct benes_hip_server : benes_hip::Server {};

benes_hip_server* benes_hip_server_create(uint64_t hbm_budget_bytes, uint32_t max_batch,
                                          uint32_t lanes, char* err, size_t err_len) {
    Status st;
    if (!benes_hip_available()) {
        st.set(kRuntime, "hip backend is unavailable");
        finish(st, err, err_len);
        return nullptr;
    }
    const int dev = current_device(st);
    if (dev < 0) {
        finish(st, err, err_len);
        return nullptr;
    }
    auto* srv = new benes_hip_server();
    srv->device = dev;
    srv->budget = hbm_budget_bytes;
    lanes = std::max<uint32_t>(1, lanes);
    srv->lane_cap = hbm_budget_bytes / lanes;
    srv->max_batch = std::max<uint32_t>(1, max_batch);
    for (uint32_t i = 0; i < lanes; ++i) srv->lanes.push_back(new Lane());
    for (Lane* lane : srv->lanes) lane->thread = std::thread([srv, lane] { srv->serve(*lane); });
    srv->thread = std::thread([srv] { srv->run(); });
    return srv;
}

uint64_t benes_hip_server_submit(benes_hip_server* srv, const void* perm, uint64_t n,
                                 uint32_t width_bytes, const benes_hip_format* format,
                                 uint8_t* out, uint64_t out_capacity, int* status, char* err,
                                 size_t err_len) {
    Status st;
    auto* r = new Request();
    if (!prepare_job(perm, n, width_bytes, format, r->job, st)) {
        delete r;
        if (status) *status = finish(st, err, err_len);
        return 0;
    }
    r->host_perm = perm;
    r->out = out;
    r->out_capacity = out_capacity;
    r->footprint = request_footprint(n, width_bytes, r->job.lay.plan.end_offset);
    r->submitted = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> lock(srv->mu);
    if (r->footprint > srv->lane_cap) {
        delete r;
        st.set(kLength, "request exceeds the server's HBM budget");
        if (status) *status = finish(st, err, err_len);
        return 0;
    }
    r->id = srv->next_id++;
    srv->requests[r->id] = r;
    srv->queue.push_back(r);
    srv->cv_work.notify_one();
    if (status) *status = BENES_HIP_OK;
    return r->id;
}

int benes_hip_server_wait(benes_hip_server* srv, uint64_t id, int64_t timeout_ms,
                          uint64_t* out_size, double* latency_seconds, char* err,
                          size_t err_len) {
    Status st;
    std::unique_lock<std::mutex> lock(srv->mu);
    auto it = srv->requests.find(id);
    if (it == srv->requests.end()) {
        st.set(kInvalidArgument, "unknown request id");
        return finish(st, err, err_len);
    }
    Request* r = it->second;
    auto ready = [&] { return r->status != -1; };
    if (timeout_ms < 0) {
        srv->cv_done.wait(lock, ready);
    } else if (!srv->cv_done.wait_for(lock, std::chrono::milliseconds(timeout_ms), ready)) {
        st.set(kPending, "request still running");
        return finish(st, err, err_len);
    }
    if (out_size) *out_size = r->job.lay.plan.end_offset;
    if (latency_seconds) {
        *latency_seconds = std::chrono::duration<double>(r->done - r->submitted).count();
    }
    st.set(r->status, r->message);
    srv->requests.erase(it);
    delete r;
    return finish(st, err, err_len);
}

int benes_hip_server_stats(benes_hip_server* srv, benes_hip_server_counters* out) {
    std::lock_guard<std::mutex> lock(srv->mu);
    out->completed = srv->completed;
    out->entries = srv->entries;
    out->batches = srv->batches;
    out->peak_bytes = srv->peak_bytes;
    out->max_concurrent_batches = srv->max_concurrent;
    out->queued = srv->queue.size();
    out->budget_bytes = srv->budget;
    return BENES_HIP_OK;
}

void benes_hip_server_destroy(benes_hip_server* srv) {
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lock(srv->mu);
        srv->stop = true;
        srv->cv_work.notify_all();
    }
    srv->thread.join();  // returns once the queue is empty and no lane is busy
    {
        std::lock_guard<std::mutex> lock(srv->mu);
        srv->stop_lanes = true;
        for (Lane* lane : srv->lanes) lane->cv.notify_all();
    }
    for (Lane* lane : srv->lanes) {
        lane->thread.join();
        delete lane;
    }
    for (auto& kv : srv->requests) delete kv.second;
    delete srv;
}

}  // extern "C"

