// SPDX-License-Identifier: Apache-2.0
// Host orchestration of the compact-v1 compression path on one MI355X, and
// the C ABI declared in include/benes_hip.h.
//
// One call = validate on device, then for each global level
// describe -> union -> resolve -> pack(in) -> pack(out), then one
// k_subtree launch for everything below the fusion depth, all queued on one
// stream; the host only builds the per-depth Plan (a few KB) and reads back a
// single error word at the end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/benes_hip.h"
#include "kernels.hpp"
#include "plan.hpp"
#include "status.hpp"

namespace benes_hip {
namespace {

thread_local double g_device_seconds = 0;

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

struct HipError {
    hipError_t code;
    const char* what;
};

#define BH_CHECK(expr)                                       \
    do {                                                     \
        const hipError_t e_ = (expr);                        \
        if (e_ != hipSuccess) throw HipError{e_, #expr};     \
    } while (0)

// Grow-only device buffer.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    void ensure(size_t need) {
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

// Per-device scratch reused across calls (calls on one device serialise on
// its mutex; separate devices run independently).
struct Workspace {
    std::mutex mu;
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t ev_begin = nullptr, ev_end = nullptr;
    DevBuf input, level[2], ilevel[2], inv0, uf, srows, trows, desc, plan, err, blob;
    void init(int dev) {
        if (stream) return;
        device = dev;
        BH_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        BH_CHECK(hipEventCreate(&ev_begin));
        BH_CHECK(hipEventCreate(&ev_end));
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

uint32_t fusion_depth(uint64_t n) {
    uint32_t d = 0;
    while (ceil_div(n, uint64_t{1} << d) > kSub) ++d;
    return d;
}

struct Profiler {
    bool on = false;
    hipStream_t stream = nullptr;
    std::vector<Pending> pending;
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
    }
};

#define BH_LAUNCH(prof, name, ...)      \
    do {                                \
        (prof).begin(name);             \
        hipLaunchKernelGGL(__VA_ARGS__); \
        (prof).end();                   \
    } while (0)

// Queues the whole pipeline on ws.stream.  `src` holds n entries of V on the
// device; the blob (rounded up to whole words) lands in ws.blob.
template <typename V>
void enqueue_pipeline(Workspace& ws, const V* src, const Layout& lay, Profiler& prof) {
    const Plan& plan = lay.plan;
    const uint64_t n = plan.n;
    const hipStream_t st = ws.stream;
    const uint64_t blob_words = ceil_div(plan.end_offset, 8);
    ws.blob.ensure(blob_words * 8);
    ws.plan.ensure(sizeof(Plan));
    ws.err.ensure(sizeof(int));
    BH_CHECK(hipMemcpyAsync(ws.plan.ptr, &plan, sizeof(Plan), hipMemcpyHostToDevice, st));
    BH_CHECK(hipMemsetAsync(ws.err.ptr, 0, sizeof(int), st));
    BH_CHECK(hipMemsetAsync(ws.blob.ptr, 0, blob_words * 8, st));
    int* err = ws.err.as<int>();
    const Plan* dplan = ws.plan.as<Plan>();
    uint64_t* blob = ws.blob.as<uint64_t>();
    const uint32_t df = std::min<uint32_t>(fusion_depth(n), plan.depth);
    // Validation and the inverse: 32-bit while every index fits below the
    // all-ones sentinel.
    const bool wide_inverse = n >= 0xFFFFFFFFull;
    const uint32_t blocks_n = static_cast<uint32_t>(ceil_div(n, kBlock));
    ws.inv0.ensure(n * (wide_inverse ? 8 : 4));
    BH_CHECK(hipMemsetAsync(ws.inv0.ptr, 0xFF, n * (wide_inverse ? 8 : 4), st));
    auto validate = [&](auto* inv) {
        using VI = std::remove_pointer_t<decltype(inv)>;
        BH_LAUNCH(prof, "k_scatter_inverse", (k_scatter_inverse<V, VI>), dim3(blocks_n),
                  dim3(kBlock), 0, st, src, n, inv, err);
        BH_LAUNCH(prof, "k_check_inverse", (k_check_inverse<VI>), dim3(blocks_n), dim3(kBlock), 0,
                  st, inv, n, err);
    };
    if (wide_inverse) validate(ws.inv0.as<uint64_t>());
    else validate(ws.inv0.as<uint32_t>());
    if (df > 0) {
        for (int i = 0; i < 2; ++i) {
            ws.level[i].ensure(n * 4);
            ws.ilevel[i].ensure(n * 4);
        }
    }
    int next = 0;
    for (uint32_t d = 0; d < df; ++d) {
        const uint64_t nodes = uint64_t{1} << d;
        const uint64_t cmax = ceil_div(n, nodes);
        const uint64_t maxpairs = cmax / 2;
        const uint64_t row_words = std::max<uint64_t>(1, ceil_div(maxpairs, 64));
        ws.desc.ensure(nodes * sizeof(NodeDesc));
        ws.srows.ensure(nodes * row_words * 8);
        ws.trows.ensure(nodes * row_words * 8);
        NodeDesc* desc = ws.desc.as<NodeDesc>();
        BH_LAUNCH(prof, "k_describe", k_describe, dim3(static_cast<uint32_t>(ceil_div(nodes, 64))),
                  dim3(64), 0, st, dplan, d, row_words, desc, err);
        // Input slots of a node number ceil(c/2) < 2^31 for 32-bit words.
        const bool wide = cmax / 2 + 1 >= (uint64_t{1} << 31);
        ws.uf.ensure(n * (wide ? 8 : 4));
        if (cmax > kLevelLds) BH_CHECK(hipMemsetAsync(ws.uf.ptr, 0xFF, n * (wide ? 8 : 4), st));
        const uint64_t chunks = ceil_div(maxpairs + 1, kBlock);
        const dim3 grid(static_cast<uint32_t>(chunks * nodes));
        uint32_t* pdst = ws.level[next].as<uint32_t>();
        uint32_t* idst = ws.ilevel[next].as<uint32_t>();
        uint64_t* srows = ws.srows.as<uint64_t>();
        uint64_t* trows = ws.trows.as<uint64_t>();
        const bool in_lds = cmax <= kLevelLds;
        auto run = [&](auto psrc, auto isrc, auto uf) {
            using VS = std::remove_const_t<std::remove_pointer_t<decltype(psrc)>>;
            using VI = std::remove_const_t<std::remove_pointer_t<decltype(isrc)>>;
            using W = std::remove_pointer_t<decltype(uf)>;
            if (in_lds) {
                BH_LAUNCH(prof, "k_level_lds", (k_level_lds<VS, VI>),
                          dim3(static_cast<uint32_t>(nodes)), dim3(kLevelThreads), 0, st, psrc,
                          isrc, pdst, idst, desc, srows, trows, err);
                return;
            }
            BH_LAUNCH(prof, "k_union", (k_union<VI, W>), grid, dim3(kBlock), 0, st, isrc, uf, desc,
                      nodes, chunks, err);
            BH_LAUNCH(prof, "k_route_inputs", (k_route_inputs<VS, W>), grid, dim3(kBlock), 0, st,
                      psrc, pdst, uf, desc, nodes, chunks, srows, err);
            BH_LAUNCH(prof, "k_route_outputs", (k_route_outputs<VI>), grid, dim3(kBlock), 0, st,
                      isrc, idst, desc, nodes, chunks, srows, trows, err);
        };
        if (d == 0) {
            if (wide_inverse) {
                if (wide) run(src, ws.inv0.as<uint64_t>(), ws.uf.as<uint64_t>());
                else run(src, ws.inv0.as<uint64_t>(), ws.uf.as<uint32_t>());
            } else {
                if (wide) run(src, ws.inv0.as<uint32_t>(), ws.uf.as<uint64_t>());
                else run(src, ws.inv0.as<uint32_t>(), ws.uf.as<uint32_t>());
            }
        } else {
            const uint32_t* ps = ws.level[next ^ 1].as<uint32_t>();
            const uint32_t* is = ws.ilevel[next ^ 1].as<uint32_t>();
            if (wide) run(ps, is, ws.uf.as<uint64_t>());
            else run(ps, is, ws.uf.as<uint32_t>());
        }
        for (int side = 0; side < 2; ++side) {
            const SideDepth& sd = plan.side[side][d];
            const uint32_t span_log2 = sd.group_depth - sd.relative_level - 1;
            const uint64_t packets = ceil_div(maxpairs, uint64_t{1} << span_log2);
            const uint64_t groups = uint64_t{1} << (d - sd.relative_level);
            const uint64_t waves = groups * packets;
            BH_LAUNCH(prof, "k_pack", k_pack,
                      dim3(static_cast<uint32_t>(ceil_div(waves, kBlock / 64))), dim3(kBlock), 0,
                      st, dplan, desc, side, d, packets, side == 0 ? srows : trows, blob, err);
        }
        next ^= 1;
    }
    const uint64_t subtrees = uint64_t{1} << df;
    if (df == 0) {
        BH_LAUNCH(prof, "k_subtree", k_subtree<V>, dim3(static_cast<uint32_t>(subtrees)),
                  dim3(kSubThreads), 0, st, src, dplan, df, blob, err);
    } else {
        BH_LAUNCH(prof, "k_subtree", k_subtree<uint32_t>, dim3(static_cast<uint32_t>(subtrees)),
                  dim3(kSubThreads), 0, st, ws.level[next ^ 1].as<uint32_t>(), dplan, df, blob,
                  err);
    }
    BH_CHECK(hipGetLastError());
}

void copy_err(Status& st, const char* msg, size_t len, char* err, size_t err_len) {
    (void)len;
    if (err && err_len) {
        std::strncpy(err, msg, err_len - 1);
        err[err_len - 1] = 0;
    }
    (void)st;
}

int finish(const Status& st, char* err, size_t err_len) {
    if (!st.ok()) copy_err(const_cast<Status&>(st), st.message.c_str(), st.message.size(), err,
                           err_len);
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

int current_device(Status& st) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) {
        st.set(kRuntime, "hip backend is unavailable");
        return -1;
    }
    return dev;
}

const char* kNotPermutation = "input must contain every value from zero through size minus one";

// Common body of both entry points.  `dst_host` (host API) or `dst_dev`
// (device API) receives the blob.
int compress_common(const void* perm, bool perm_on_device, uint64_t n, uint32_t width,
                    const benes_hip_format* format, uint8_t* dst_host, uint8_t* dst_dev,
                    uint64_t capacity, uint64_t* out_size, hipStream_t user_stream,
                    double* device_seconds, char* err, size_t err_len) {
    Status st;
    if (!check_input(n, width, st)) return finish(st, err, err_len);
    FormatArgs fa;
    if (!format_args(format, fa, st)) return finish(st, err, err_len);
    Layout lay;
    if (!build_layout(n, fa, lay, st)) return finish(st, err, err_len);
    if (out_size) *out_size = lay.plan.end_offset;
    if ((!dst_host && !dst_dev) || capacity < lay.plan.end_offset) {
        st.set(kBufferTooSmall, "output buffer is too small");
        return finish(st, err, err_len);
    }
    if (!benes_hip_available()) {
        st.set(kRuntime, "hip backend is unavailable");
        return finish(st, err, err_len);
    }
    const int dev = current_device(st);
    if (dev < 0) return finish(st, err, err_len);
    Workspace& ws = workspace_for(dev);
    std::lock_guard<std::mutex> lock(ws.mu);
    try {
        ws.init(dev);
        const hipStream_t s = ws.stream;
        if (user_stream) {
            // Order after the caller's pending work on its stream.
            hipEvent_t e;
            BH_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            BH_CHECK(hipEventRecord(e, user_stream));
            BH_CHECK(hipStreamWaitEvent(s, e, 0));
            BH_CHECK(hipEventDestroy(e));
        }
        const void* dsrc = perm;
        if (!perm_on_device) {
            ws.input.ensure(n * width);
            BH_CHECK(hipMemcpyAsync(ws.input.ptr, perm, n * width, hipMemcpyHostToDevice, s));
            dsrc = ws.input.ptr;
        }
        Profiler prof;
        {
            std::lock_guard<std::mutex> plock(g_prof_mu);
            prof.on = g_profiling;
        }
        prof.stream = s;
        BH_CHECK(hipEventRecord(ws.ev_begin, s));
        if (width == 4) {
            enqueue_pipeline(ws, static_cast<const uint32_t*>(dsrc), lay, prof);
        } else {
            enqueue_pipeline(ws, static_cast<const uint64_t*>(dsrc), lay, prof);
        }
        std::vector<uint8_t> header(lay.header_bytes);
        write_header(lay, header.data());
        if (dst_dev) {
            BH_CHECK(hipMemcpyAsync(dst_dev, ws.blob.ptr, lay.plan.end_offset,
                                    hipMemcpyDeviceToDevice, s));
            BH_CHECK(hipMemcpyAsync(dst_dev, header.data(), header.size(), hipMemcpyHostToDevice,
                                    s));
        }
        BH_CHECK(hipEventRecord(ws.ev_end, s));
        int flag = 0;
        BH_CHECK(hipMemcpyAsync(&flag, ws.err.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
        if (dst_host) {
            BH_CHECK(hipMemcpyAsync(dst_host, ws.blob.ptr, lay.plan.end_offset,
                                    hipMemcpyDeviceToHost, s));
        }
        BH_CHECK(hipStreamSynchronize(s));
        prof.collect();
        float ms = 0;
        BH_CHECK(hipEventElapsedTime(&ms, ws.ev_begin, ws.ev_end));
        g_device_seconds = ms / 1e3;
        if (device_seconds) *device_seconds = g_device_seconds;
        if (user_stream) {
            hipEvent_t e;
            BH_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            BH_CHECK(hipEventRecord(e, s));
            BH_CHECK(hipStreamWaitEvent(user_stream, e, 0));
            BH_CHECK(hipEventDestroy(e));
        }
        if (flag) {
            st.set(kInvalidArgument, kNotPermutation);
            return finish(st, err, err_len);
        }
        if (dst_host) std::memcpy(dst_host, header.data(), header.size());
    } catch (const HipError& e) {
        st.set(kRuntime, std::string("hip error: ") + hipGetErrorString(e.code) + " in " + e.what);
    } catch (const std::bad_alloc&) {
        st.set(kLength, "allocation failed");
    }
    return finish(st, err, err_len);
}

}  // namespace
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
    return compress_common(perm, false, n, width_bytes, format, out, nullptr, out_capacity,
                           out_size, nullptr, device_seconds, err, err_len);
}

int benes_hip_compress_device(const void* d_perm, uint64_t n, uint32_t width_bytes,
                              const benes_hip_format* format, uint8_t* d_out,
                              uint64_t out_capacity, uint64_t* out_size, void* stream, char* err,
                              size_t err_len) {
    return compress_common(d_perm, true, n, width_bytes, format, nullptr, d_out, out_capacity,
                           out_size, static_cast<hipStream_t>(stream), nullptr, err, err_len);
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

}  // extern "C"
