// Shared plumbing of the HIP library: error reporting across the C ABI, launch helpers and the
// small wave64 / workgroup primitives the kernels build on. gfx950 only: wave64, 64 KiB LDS per
// workgroup budget, 256 CUs in 8 XCDs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "dynamis.h"

namespace dyn {

// Thread-local text of the last failure, returned by dyn_last_error().
void set_error(const std::string& message);
void clear_error();

inline int fail(int code, const std::string& message) {
    set_error(message);
    return code;
}

inline int hip_fail(hipError_t err, const char* what) {
    return fail(DYN_E_HIP, std::string(what) + ": " + hipGetErrorString(err));
}

#define DYN_HIP(call)                                   \
    do {                                                \
        hipError_t dyn_err_ = (call);                   \
        if (dyn_err_ != hipSuccess) {                   \
            return ::dyn::hip_fail(dyn_err_, #call);    \
        }                                               \
    } while (0)

// Checks the launch that was just enqueued.
#define DYN_LAUNCHED(name)                                          \
    do {                                                            \
        hipError_t dyn_err_ = hipGetLastError();                    \
        if (dyn_err_ != hipSuccess) {                               \
            return ::dyn::hip_fail(dyn_err_, "launch of " name);    \
        }                                                           \
    } while (0)

#define DYN_TRY(expr)                  \
    do {                               \
        int dyn_rc_ = (expr);          \
        if (dyn_rc_ != DYN_OK) {       \
            return dyn_rc_;            \
        }                              \
    } while (0)

constexpr int kWave = 64;

__host__ __device__ inline uint32_t ceil_div(uint64_t a, uint64_t b) {
    return static_cast<uint32_t>((a + b - 1) / b);
}
inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Carves consecutive 256-byte aligned regions out of one caller-provided scratch block.
struct Carve {
    char* base;
    size_t used = 0;
    explicit Carve(void* p) : base(static_cast<char*>(p)) {}
    template <typename T>
    T* take(size_t count) {
        T* out = reinterpret_cast<T*>(base ? base + used : nullptr);
        used = align_up(used + count * sizeof(T), 256);
        return out;
    }
};

// ---- device helpers ----------------------------------------------------------------------

__device__ inline uint32_t counter_word(uint32_t slot) { return slot * DYN_COUNTER_STRIDE_WORDS; }

__device__ inline uint32_t atomic_load_agent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void atomic_store_agent(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline uint32_t lane_id() { return __lane_id(); }

// Inclusive prefix sum over the 64 lanes of a wave.
__device__ inline uint32_t wave_inclusive_sum(uint32_t v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int offset = 1; offset < kWave; offset <<= 1) {
        uint32_t up = __shfl_up(v, offset, kWave);
        if (lane >= static_cast<uint32_t>(offset)) v += up;
    }
    return v;
}

__device__ inline uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int offset = kWave / 2; offset > 0; offset >>= 1) v += __shfl_xor(v, offset, kWave);
    return v;
}

__device__ inline uint32_t wave_max(uint32_t v) {
#pragma unroll
    for (int offset = kWave / 2; offset > 0; offset >>= 1) v = max(v, __shfl_xor(v, offset, kWave));
    return v;
}

__device__ inline uint32_t wave_or(uint32_t v) {
#pragma unroll
    for (int offset = kWave / 2; offset > 0; offset >>= 1) v |= __shfl_xor(v, offset, kWave);
    return v;
}

// Exclusive prefix sum over a workgroup of THREADS lanes (one value per lane). `scratch` must
// hold THREADS/64 words of LDS. Returns the exclusive prefix; *total receives the sum.
template <int THREADS>
__device__ inline uint32_t block_exclusive_sum(uint32_t v, uint32_t* scratch, uint32_t* total) {
    constexpr int kWaves = THREADS / kWave;
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / kWave;
    uint32_t inclusive = wave_inclusive_sum(v);
    if (lane == kWave - 1) scratch[wave] = inclusive;
    __syncthreads();
    uint32_t before = 0;
    uint32_t sum = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        uint32_t s = scratch[w];
        if (w < static_cast<int>(wave)) before += s;
        sum += s;
    }
    __syncthreads();
    if (total) *total = sum;
    return before + inclusive - v;
}

}  // namespace dyn

namespace dyn {

// Whether per-kernel profiling (dyn_profile_enable) is on.
bool profiling();

// Brackets one kernel launch with events on its stream while profiling is enabled
// (dyn_profile_enable); a no-op otherwise.
class ProfileScope {
public:
    ProfileScope(const char* name, hipStream_t stream);
    ~ProfileScope();
    ProfileScope(const ProfileScope&) = delete;
    ProfileScope& operator=(const ProfileScope&) = delete;

private:
    const char* name_;
    hipStream_t stream_;
    void* start_ = nullptr;
};

}  // namespace dyn
