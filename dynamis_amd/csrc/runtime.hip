// Plumbing entry points: errors, device selection, memory, streams and events. These stand in
// for the reference's wgpu wrapper (crates/dynamis-gpu/src/context.rs:185, src/buffer.rs:6,
// src/timing.rs:22) on the HIP runtime.
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "common.h"

namespace dyn {

static thread_local std::string g_error;

void set_error(const std::string& message) { g_error = message; }
void clear_error() { g_error.clear(); }

namespace {

struct Timed {
    const char* name;
    hipEvent_t start;
    hipEvent_t stop;
};

struct Profile {
    std::mutex lock;
    bool enabled = false;
    std::string watch;  // when set, only launches of this kernel are bracketed
    std::vector<Timed> pending;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<std::string, std::pair<uint32_t, double>>> collected;

    hipEvent_t take() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
};

Profile& profile() {
    static Profile p;
    return p;
}

}  // namespace

bool profiling() {
    Profile& p = profile();
    std::lock_guard<std::mutex> guard(p.lock);
    return p.enabled;
}

ProfileScope::ProfileScope(const char* name, hipStream_t stream) : name_(name), stream_(stream) {
    Profile& p = profile();
    std::lock_guard<std::mutex> guard(p.lock);
    if (!p.enabled) return;
    if (!p.watch.empty() && p.watch != name_) return;
    hipEvent_t e = p.take();
    if (e && hipEventRecord(e, stream_) == hipSuccess) start_ = e;
}

ProfileScope::~ProfileScope() {
    if (!start_) return;
    Profile& p = profile();
    std::lock_guard<std::mutex> guard(p.lock);
    hipEvent_t e = p.take();
    if (e && hipEventRecord(e, stream_) == hipSuccess) {
        p.pending.push_back(Timed{name_, static_cast<hipEvent_t>(start_), e});
    }
}

}  // namespace dyn

extern "C" {

int dyn_version(void) { return 1; }

const char* dyn_last_error(void) { return dyn::g_error.c_str(); }

int dyn_device_count(int* count) {
    if (!count) return dyn::fail(DYN_E_INVALID, "dyn_device_count: null output");
    int n = 0;
    hipError_t err = hipGetDeviceCount(&n);
    if (err != hipSuccess) {
        // A host without a GPU reports an error here; that is a count of zero, not a failure.
        (void)hipGetLastError();
        n = 0;
    }
    *count = n;
    return DYN_OK;
}

int dyn_set_device(int ordinal) {
    DYN_HIP(hipSetDevice(ordinal));
    return DYN_OK;
}

int dyn_device_name(char* out, uint32_t capacity) {
    if (!out || capacity == 0) return dyn::fail(DYN_E_INVALID, "dyn_device_name: null output");
    int device = 0;
    DYN_HIP(hipGetDevice(&device));
    hipDeviceProp_t prop;
    DYN_HIP(hipGetDeviceProperties(&prop, device));
    std::string name = std::string(prop.name) + "|" + prop.gcnArchName;
    const size_t n = name.size() < capacity - 1 ? name.size() : capacity - 1;
    std::memcpy(out, name.data(), n);
    out[n] = '\0';
    return DYN_OK;
}

int dyn_malloc(void** d_ptr, size_t bytes) {
    if (!d_ptr) return dyn::fail(DYN_E_INVALID, "dyn_malloc: null output");
    *d_ptr = nullptr;
    if (bytes == 0) bytes = 256;
    DYN_HIP(hipMalloc(d_ptr, bytes));
    return DYN_OK;
}

int dyn_free(void* d_ptr) {
    if (d_ptr) DYN_HIP(hipFree(d_ptr));
    return DYN_OK;
}

int dyn_memcpy_htod(void* d_dst, const void* src, size_t bytes) {
    if (bytes) DYN_HIP(hipMemcpy(d_dst, src, bytes, hipMemcpyHostToDevice));
    return DYN_OK;
}

int dyn_memcpy_dtoh(void* dst, const void* d_src, size_t bytes) {
    if (bytes) DYN_HIP(hipMemcpy(dst, d_src, bytes, hipMemcpyDeviceToHost));
    return DYN_OK;
}

int dyn_memcpy_dtod(void* d_dst, const void* d_src, size_t bytes) {
    if (bytes) DYN_HIP(hipMemcpy(d_dst, d_src, bytes, hipMemcpyDeviceToDevice));
    return DYN_OK;
}

int dyn_memset(void* d_dst, int value, size_t bytes) {
    if (bytes) DYN_HIP(hipMemset(d_dst, value, bytes));
    return DYN_OK;
}

int dyn_stream_create(void** stream) {
    if (!stream) return dyn::fail(DYN_E_INVALID, "dyn_stream_create: null output");
    hipStream_t s = nullptr;
    DYN_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return DYN_OK;
}

int dyn_stream_destroy(void* stream) {
    if (stream) DYN_HIP(hipStreamDestroy(static_cast<hipStream_t>(stream)));
    return DYN_OK;
}

int dyn_stream_synchronize(void* stream) {
    DYN_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    return DYN_OK;
}

int dyn_event_create(void** event) {
    if (!event) return dyn::fail(DYN_E_INVALID, "dyn_event_create: null output");
    hipEvent_t e = nullptr;
    DYN_HIP(hipEventCreate(&e));
    *event = e;
    return DYN_OK;
}

int dyn_event_destroy(void* event) {
    if (event) DYN_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
    return DYN_OK;
}

int dyn_event_record(void* event, void* stream) {
    DYN_HIP(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)));
    return DYN_OK;
}

int dyn_profile_enable(int on) {
    dyn::Profile& p = dyn::profile();
    std::lock_guard<std::mutex> guard(p.lock);
    p.enabled = on != 0;
    return DYN_OK;
}

int dyn_profile_watch(const char* kernel) {
    dyn::Profile& p = dyn::profile();
    std::lock_guard<std::mutex> guard(p.lock);
    p.watch = kernel ? kernel : "";
    return DYN_OK;
}

int dyn_profile_collect(int* kernels) {
    if (!kernels) return dyn::fail(DYN_E_INVALID, "dyn_profile_collect: null output");
    dyn::Profile& p = dyn::profile();
    std::lock_guard<std::mutex> guard(p.lock);
    std::map<std::string, std::pair<uint32_t, double>> sums;
    for (const dyn::Timed& t : p.pending) {
        float ms = 0.0f;
        DYN_HIP(hipEventSynchronize(t.stop));
        DYN_HIP(hipEventElapsedTime(&ms, t.start, t.stop));
        auto& slot = sums[t.name];
        slot.first += 1;
        slot.second += ms;
        p.pool.push_back(t.start);
        p.pool.push_back(t.stop);
    }
    p.pending.clear();
    p.collected.assign(sums.begin(), sums.end());
    *kernels = static_cast<int>(p.collected.size());
    return DYN_OK;
}

int dyn_profile_entry(int i, char* name, size_t name_capacity, uint32_t* launches,
                      double* total_ms) {
    dyn::Profile& p = dyn::profile();
    std::lock_guard<std::mutex> guard(p.lock);
    if (i < 0 || static_cast<size_t>(i) >= p.collected.size()) {
        return dyn::fail(DYN_E_INVALID, "dyn_profile_entry: index out of range");
    }
    const auto& entry = p.collected[static_cast<size_t>(i)];
    if (name && name_capacity) {
        size_t n = entry.first.size() < name_capacity - 1 ? entry.first.size() : name_capacity - 1;
        std::memcpy(name, entry.first.data(), n);
        name[n] = '\0';
    }
    if (launches) *launches = entry.second.first;
    if (total_ms) *total_ms = entry.second.second;
    return DYN_OK;
}

int dyn_event_elapsed_ms(float* ms, void* start, void* stop) {
    if (!ms) return dyn::fail(DYN_E_INVALID, "dyn_event_elapsed_ms: null output");
    DYN_HIP(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
    DYN_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
    return DYN_OK;
}

}  // extern "C"
