// Library-wide entry points: error reporting, version, device synchronisation and the
// sampled kernel timer used for live roofline measurements.
#include "pp_common.hpp"

#include <algorithm>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

namespace pp {
static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
const char *get_error() { return g_last_error.c_str(); }

struct TimerSlot {
    hipEvent_t start = nullptr, stop = nullptr;
    int kernel = 0, rows = 0;
};

struct TimerStats {
    double total_ms = 0.0;
    int64_t sampled = 0, launches = 0;
};

static struct {
    std::mutex mu;
    bool enabled = false;
    int every = 1;
    std::map<std::pair<int, int>, TimerStats> stats;
    std::vector<TimerSlot *> pending, pool;
} g_timer;

bool timer_enabled() { return g_timer.enabled; }

TimerSlot *timer_begin(int kernel_id, int rows, hipStream_t s) {
    if (!g_timer.enabled) return nullptr;
    std::lock_guard<std::mutex> lock(g_timer.mu);
    TimerStats &st = g_timer.stats[{kernel_id, rows}];
    // sample a pseudo-random 1/every of the launches (a multiplicative hash of the launch count):
    // a fixed stride would alias with the launch pattern of a solve (e.g. 4 phases per solve)
    const uint32_t h = (uint32_t)(st.launches++) * 2654435761u;
    const bool sample = g_timer.every == 1 || (h >> 16) % (uint32_t)g_timer.every == 0;
    if (!sample) return nullptr;
    TimerSlot *slot;
    if (!g_timer.pool.empty()) {
        slot = g_timer.pool.back();
        g_timer.pool.pop_back();
    } else {
        slot = new TimerSlot;
        if (hipEventCreate(&slot->start) != hipSuccess || hipEventCreate(&slot->stop) != hipSuccess) {
            delete slot;
            return nullptr;
        }
    }
    slot->kernel = kernel_id;
    slot->rows = rows;
    if (hipEventRecord(slot->start, s) != hipSuccess) {
        g_timer.pool.push_back(slot);
        return nullptr;
    }
    return slot;
}

void timer_end(TimerSlot *slot, hipStream_t s) {
    std::lock_guard<std::mutex> lock(g_timer.mu);
    if (hipEventRecord(slot->stop, s) != hipSuccess) {
        g_timer.pool.push_back(slot);
        return;
    }
    g_timer.pending.push_back(slot);
}

// Fold completed samples into the statistics (blocks until every recorded stop event is done).
static int timer_drain() {
    for (TimerSlot *slot : g_timer.pending) {
        PP_HIP(hipEventSynchronize(slot->stop));
        float ms = 0.0f;
        PP_HIP(hipEventElapsedTime(&ms, slot->start, slot->stop));
        TimerStats &st = g_timer.stats[{slot->kernel, slot->rows}];
        st.total_ms += ms;
        st.sampled += 1;
        g_timer.pool.push_back(slot);
    }
    g_timer.pending.clear();
    return PP_OK;
}

// grid-stride copy of 16-byte words; every byte is read once and written once
__global__ void __launch_bounds__(256) k_copy_bytes(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t words) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < words; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}
} // namespace pp

extern "C" {
const char *pp_last_error(void) { return pp::get_error(); }
int pp_version(void) { return 1; }
int pp_device_sync(void) {
    PP_HIP(hipDeviceSynchronize());
    return PP_OK;
}

int pp_copy_bytes(const void *src, void *dst, int64_t bytes, void *stream) {
    PP_CHECK_ARG(src && dst && bytes >= 0 && bytes % 16 == 0, "pp_copy_bytes: null pointer or size not a multiple of 16");
    if (bytes == 0) return PP_OK;
    const size_t words = (size_t)bytes / 16;
    const unsigned blocks = (unsigned)std::min<size_t>((words + 255) / 256, 8192);
    hipLaunchKernelGGL(pp::k_copy_bytes, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), words);
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_kernel_timer_start(int32_t sample_every) {
    PP_CHECK_ARG(sample_every >= 1, "sample_every must be >= 1");
    std::lock_guard<std::mutex> lock(pp::g_timer.mu);
    if (int e = pp::timer_drain()) return e;
    pp::g_timer.stats.clear();
    pp::g_timer.every = sample_every;
    pp::g_timer.enabled = true;
    return PP_OK;
}

int pp_kernel_timer_stop(void) {
    std::lock_guard<std::mutex> lock(pp::g_timer.mu);
    pp::g_timer.enabled = false;
    return pp::timer_drain();
}

int pp_kernel_timer_read(int32_t capacity, int32_t *kernel_ids, int32_t *rows, double *total_ms, int64_t *sampled,
                         int64_t *launches, int32_t *count) {
    PP_CHECK_ARG(count != nullptr && capacity >= 0, "bad arguments");
    std::lock_guard<std::mutex> lock(pp::g_timer.mu);
    if (int e = pp::timer_drain()) return e;
    int i = 0;
    for (const auto &kv : pp::g_timer.stats) {
        if (i < capacity) {
            if (kernel_ids) kernel_ids[i] = kv.first.first;
            if (rows) rows[i] = kv.first.second;
            if (total_ms) total_ms[i] = kv.second.total_ms;
            if (sampled) sampled[i] = kv.second.sampled;
            if (launches) launches[i] = kv.second.launches;
        }
        ++i;
    }
    *count = i;
    return PP_OK;
}
}
