// Stable LSD radix sort for gfx950: 8-bit digits, one histogram read for all digits, then one
// "onesweep" pass per digit. Each pass ranks a 4096-key tile with wave64 ballots (a key's rank
// among equal digits is the popcount of the lanes below it that hold the same digit, plus the
// counts of earlier items and earlier waves), publishes the tile's digit counts for a decoupled
// look-back across tiles, stages the ranked tile in LDS and writes it out digit-run by digit-run
// so that the global stores coalesce.
//
// Behaviour follows `RadixSort::sort` (crates/dynamis-sort/src/radix.rs:397-476): records
// [base, base+len) of (major, minor, payload), len read on the device, digits of the minor word
// first then the major word (radix.rs:312-317), ties kept in input order, result in the primaries.
// The reference ranks with per-unit serial prefix loops (sort_scatter.wgsl:45-104); the order
// produced is the same stable order.
#include "sort.h"

#include <map>
#include <mutex>
#include <utility>
#include <vector>

namespace dyn {

namespace {

constexpr uint32_t kFlagAggregate = 1u << 30;
constexpr uint32_t kFlagPrefix = 2u << 30;
constexpr uint32_t kFlagValue = (1u << 30) - 1;
constexpr uint32_t kSpinLimit = 1u << 22;

}  // namespace

// Look-back waits that gave up since the count was last cleared, over every sort this process ran
// on the device (never expected: a tile's predecessors always publish). A timed-out tile scatters
// with a partial prefix, so a non-zero count means some sort's order is wrong; dyn_sort_faults
// reads it.
__device__ uint32_t g_sort_faults = 0u;

namespace {

struct DigitList {
    uint32_t major[kSortMaxPasses];
    uint32_t shift[kSortMaxPasses];
    uint32_t count;
};

__global__ void __launch_bounds__(256) sort_prepare_kernel(uint32_t* hist, uint32_t* ctl,
                                                           const uint32_t* d_count,
                                                           const uint32_t* d_base,
                                                           uint32_t capacity) {
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < kSortMaxPasses * kSortBins; i += blockDim.x) hist[i] = 0;
    if (t >= 2 && t < kCtlWords) ctl[t] = 0;
    if (t == 0) {
        uint32_t base = d_base ? d_base[0] : 0u;
        if (base > capacity) base = capacity;
        uint32_t count = d_count ? d_count[0] : capacity;
        uint32_t room = capacity - base;
        ctl[kCtlLen] = count < room ? count : room;
        ctl[kCtlBase] = base;
    }
}

__global__ void __launch_bounds__(256) sort_histogram_kernel(const uint32_t* __restrict__ major,
                                                             const uint32_t* __restrict__ minor,
                                                             uint32_t* hist, const uint32_t* ctl,
                                                             uint32_t* flags0, DigitList digits) {
    __shared__ uint32_t s_hist[kSortMaxPasses][kSortBins];
    const uint32_t t = threadIdx.x;
    for (uint32_t p = 0; p < digits.count; ++p) s_hist[p][t] = 0;
    __syncthreads();
    const uint32_t len = ctl[kCtlLen];
    const uint32_t base = ctl[kCtlBase];
    bool uses_major = false, uses_minor = false;
    for (uint32_t p = 0; p < digits.count; ++p) {
        uses_major |= digits.major[p] != 0;
        uses_minor |= digits.major[p] == 0;
    }
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + t; i < len; i += stride) {
        const uint32_t hi = uses_major ? major[base + i] : 0u;
        const uint32_t lo = uses_minor ? minor[base + i] : 0u;
        for (uint32_t p = 0; p < digits.count; ++p) {
            const uint32_t word = digits.major[p] ? hi : lo;
            atomicAdd(&s_hist[p][(word >> digits.shift[p]) & 0xFFu], 1u);
        }
    }
    __syncthreads();
    for (uint32_t p = 0; p < digits.count; ++p) {
        const uint32_t c = s_hist[p][t];
        if (c) atomicAdd(&hist[p * kSortBins + t], c);
    }
    const uint32_t words = ceil_div(len, kSortMinTile) * kSortBins;
    for (uint32_t i = blockIdx.x * blockDim.x + t; i < words; i += stride) flags0[i] = 0;
}

template <int ITEMS, bool MOVE_OTHER>
__global__ void __launch_bounds__(256) sort_pass_kernel(
    const uint32_t* __restrict__ in_sort, const uint32_t* __restrict__ in_other,
    const uint32_t* __restrict__ in_pay, uint32_t* __restrict__ out_sort,
    uint32_t* __restrict__ out_other, uint32_t* __restrict__ out_pay,
    const uint32_t* __restrict__ hist_pass, uint32_t* flags_pass, uint32_t* flags_next,
    uint32_t* ctl, uint32_t pass, uint32_t shift) {
    constexpr int TILE = kSortThreads * ITEMS;
    constexpr int WAVES = kSortThreads / kWave;
    __shared__ uint32_t s_key[TILE];
    __shared__ uint32_t s_other[MOVE_OTHER ? TILE : 1];
    __shared__ uint32_t s_pay[TILE];
    __shared__ uint32_t s_wave[WAVES][kSortBins];
    __shared__ uint32_t s_start[kSortBins];
    __shared__ uint32_t s_global[kSortBins];
    __shared__ uint32_t s_digit_base[kSortBins];
    __shared__ uint32_t s_scan[WAVES];
    __shared__ uint32_t s_tile;

    const uint32_t t = threadIdx.x;
    const uint32_t lane = lane_id();
    const uint32_t wave = t / kWave;
    const uint32_t len = ctl[kCtlLen];
    const uint32_t base = ctl[kCtlBase];
    const uint32_t ntiles = ceil_div(len, TILE);
    const uint64_t below = (1ull << lane) - 1ull;

    bool based = false;
    for (;;) {
        if (t == 0) s_tile = atomicAdd(&ctl[kCtlTiles + pass], 1u);
#pragma unroll
        for (int w = 0; w < WAVES; ++w) s_wave[w][t] = 0;
        __syncthreads();
        const uint32_t tile = s_tile;
        if (tile >= ntiles) break;
        if (!based) {
            // Global start of every digit: exclusive scan of this pass's histogram (only
            // workgroups that hold a tile need it).
            s_digit_base[t] = block_exclusive_sum<kSortThreads>(hist_pass[t], s_scan, nullptr);
            based = true;
        }
        const uint32_t first = tile * TILE;
        const uint32_t tile_n = min(static_cast<uint32_t>(TILE), len - first);
        if (flags_next) flags_next[tile * kSortBins + t] = 0;

        uint32_t key[ITEMS], other[ITEMS], pay[ITEMS], rank[ITEMS];
        const uint32_t wave_first = wave * (ITEMS * kWave);
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const uint32_t local = wave_first + i * kWave + lane;
            const bool valid = local < tile_n;
            const uint32_t at = base + first + local;
            key[i] = valid ? in_sort[at] : 0u;
            pay[i] = valid ? in_pay[at] : 0u;
            if (MOVE_OTHER) other[i] = valid ? in_other[at] : 0u;
        }
        // Rank within the wave: items in order, lanes in order, which is input order.
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const uint32_t local = wave_first + i * kWave + lane;
            const bool valid = local < tile_n;
            const uint32_t d = (key[i] >> shift) & 0xFFu;
            uint64_t peers = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const bool bit = (d >> b) & 1u;
                const uint64_t m = __ballot(bit);
                peers &= bit ? m : ~m;
            }
            const uint32_t before = __popcll(peers & below);
            const uint32_t total = __popcll(peers);
            uint32_t prior = 0;
            if (valid) prior = s_wave[wave][d];
            if (valid && before == 0) s_wave[wave][d] = prior + total;
            rank[i] = prior + before;
        }
        __syncthreads();

        // Digit t: per-wave offsets, tile count, tile-local start.
        uint32_t count = 0;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) {
            const uint32_t c = s_wave[w][t];
            s_wave[w][t] = count;
            count += c;
        }
        const uint32_t start = block_exclusive_sum<kSortThreads>(count, s_scan, nullptr);
        s_start[t] = start;

        // Decoupled look-back over earlier tiles for digit t.
        uint32_t* mine = &flags_pass[tile * kSortBins + t];
        uint32_t excl = 0;
        if (tile == 0) {
            atomic_store_agent(mine, kFlagPrefix | count);
        } else {
            atomic_store_agent(mine, kFlagAggregate | count);
            int64_t j = static_cast<int64_t>(tile) - 1;
            uint32_t spins = 0;
            while (j >= 0) {
                const uint32_t v = atomic_load_agent(&flags_pass[j * kSortBins + t]);
                const uint32_t status = v >> 30;
                if (status == 0) {
                    if (++spins > kSpinLimit) {
                        atomicOr(&ctl[kCtlError], 1u);
                        atomicAdd(&g_sort_faults, 1u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                excl += v & kFlagValue;
                if (status == 2) break;
                --j;
            }
            atomic_store_agent(mine, kFlagPrefix | ((excl + count) & kFlagValue));
        }
        s_global[t] = s_digit_base[t] + excl - start;
        __syncthreads();

#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const uint32_t local = wave_first + i * kWave + lane;
            if (local < tile_n) {
                const uint32_t d = (key[i] >> shift) & 0xFFu;
                const uint32_t lp = s_start[d] + s_wave[wave][d] + rank[i];
                s_key[lp] = key[i];
                s_pay[lp] = pay[i];
                if (MOVE_OTHER) s_other[lp] = other[i];
            }
        }
        __syncthreads();
        for (uint32_t j = t; j < tile_n; j += kSortThreads) {
            const uint32_t k = s_key[j];
            const uint32_t place = s_global[(k >> shift) & 0xFFu] + j;
            // Always below len when the look-back completed; the guard keeps a timed-out
            // look-back (ctl error word set) from writing outside the stream.
            if (place >= len) continue;
            const uint32_t at = base + place;
            out_sort[at] = k;
            out_pay[at] = s_pay[j];
            if (MOVE_OTHER) out_other[at] = s_other[j];
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) sort_copy_kernel(const uint32_t* __restrict__ from_major,
                                                        const uint32_t* __restrict__ from_minor,
                                                        const uint32_t* __restrict__ from_pay,
                                                        uint32_t* to_major, uint32_t* to_minor,
                                                        uint32_t* to_pay, const uint32_t* ctl) {
    const uint32_t len = ctl[kCtlLen];
    const uint32_t base = ctl[kCtlBase];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
        to_major[base + i] = from_major[base + i];
        to_pay[base + i] = from_pay[base + i];
        if (to_minor) to_minor[base + i] = from_minor[base + i];
    }
}

DigitList make_list(const SortDigit* digits, uint32_t passes) {
    DigitList list{};
    list.count = passes;
    for (uint32_t p = 0; p < passes; ++p) {
        list.major[p] = digits[p].major;
        list.shift[p] = digits[p].shift;
    }
    return list;
}

uint32_t stream_groups(uint32_t capacity) {
    const uint32_t want = ceil_div(capacity ? capacity : 1u, 4096u);
    return want < 1024u ? want : 1024u;
}

}  // namespace

const SortNames kGenericSortNames = {"sort_prepare_kernel", "sort_histogram_kernel",
                                     "sort_pass_kernel", "sort_copy_kernel"};
const SortNames kEntrySortNames = {"entry_sort_prepare", "entry_sort_histogram",
                                   "entry_sort_pass", "entry_sort_copy"};
const SortNames kPairSortNames = {"pair_sort_prepare", "pair_sort_histogram", "pair_sort_pass",
                                  "pair_sort_copy"};

size_t sort_scratch_layout(uint32_t capacity, uint32_t passes, bool minor_lane, void* base,
                           SortScratch* out) {
    Carve carve(base);
    SortScratch s;
    const size_t words = capacity ? capacity : 1u;
    s.major = carve.take<uint32_t>(words);
    s.minor = minor_lane ? carve.take<uint32_t>(words) : nullptr;
    s.payload = carve.take<uint32_t>(words);
    s.hist = carve.take<uint32_t>(kSortMaxPasses * kSortBins);
    s.ctl = carve.take<uint32_t>(kCtlWords);
    s.max_tiles = ceil_div(words, kSortMinTile);
    s.flags = carve.take<uint32_t>(static_cast<size_t>(passes ? passes : 1u) * s.max_tiles *
                                   kSortBins);
    if (out) *out = s;
    return carve.used;
}

int sort_digits(uint32_t major_words, uint32_t minor_words, SortDigit* digits) {
    uint32_t n = 0;
    for (uint32_t w = 0; w < minor_words; ++w) digits[n++] = SortDigit{0u, 8u * w};
    for (uint32_t w = 0; w < major_words; ++w) digits[n++] = SortDigit{1u, 8u * w};
    return static_cast<int>(n);
}

int sort_prepare(hipStream_t stream, const SortScratch& s, const uint32_t* d_count,
                 const uint32_t* d_base, uint32_t capacity, const SortNames& names) {
    {
        ProfileScope scope(names.prepare, stream);
        sort_prepare_kernel<<<1, 256, 0, stream>>>(s.hist, s.ctl, d_count, d_base, capacity);
    }
    DYN_LAUNCHED("sort_prepare_kernel");
    return DYN_OK;
}

int sort_histogram(hipStream_t stream, const uint32_t* major, const uint32_t* minor,
                   const SortScratch& s, const SortDigit* digits, uint32_t passes,
                   uint32_t capacity, const SortNames& names) {
    {
        ProfileScope scope(names.histogram, stream);
        sort_histogram_kernel<<<stream_groups(capacity), 256, 0, stream>>>(
            major, minor, s.hist, s.ctl, s.flags, make_list(digits, passes));
    }
    DYN_LAUNCHED("sort_histogram_kernel");
    return DYN_OK;
}

template <int ITEMS>
static int launch_passes(hipStream_t stream, uint32_t* major, uint32_t* minor, uint32_t* payload,
                         const SortScratch& s, const SortDigit* digits, uint32_t passes,
                         uint32_t capacity, const SortNames& names, uint32_t group_cap) {
    constexpr int TILE = kSortThreads * ITEMS;
    const uint32_t tiles = ceil_div(capacity ? capacity : 1u, TILE);
    // Persistent workgroups take tiles in order (any count makes progress: a tile's look-back
    // only waits on earlier tiles, which running groups hold).
    const uint32_t groups = tiles < group_cap ? tiles : group_cap;
    // The look-back words are laid out for the narrowest tiling, so every tiling fits.
    const size_t flag_stride = static_cast<size_t>(s.max_tiles) * kSortBins;
    if (static_cast<size_t>(tiles) * kSortBins > flag_stride) {
        return fail(DYN_E_INVALID, "sort: tiling exceeds the look-back layout");
    }
    uint32_t* src[3] = {major, minor, payload};
    uint32_t* dst[3] = {s.major, s.minor, s.payload};
    const bool carry_minor = minor != nullptr;
    for (uint32_t p = 0; p < passes; ++p) {
        const bool on_major = digits[p].major != 0;
        if (!on_major && !carry_minor) return fail(DYN_E_INVALID, "sort: minor digit without minor lane");
        uint32_t* in_sort = on_major ? src[0] : src[1];
        uint32_t* in_other = on_major ? src[1] : src[0];
        uint32_t* out_sort = on_major ? dst[0] : dst[1];
        uint32_t* out_other = on_major ? dst[1] : dst[0];
        uint32_t* flags_pass = s.flags + p * flag_stride;
        uint32_t* flags_next = (p + 1 < passes) ? s.flags + (p + 1) * flag_stride : nullptr;
        if (carry_minor) {
            {
                ProfileScope scope(names.pass, stream);
                sort_pass_kernel<ITEMS, true><<<groups, kSortThreads, 0, stream>>>(
                    in_sort, in_other, src[2], out_sort, out_other, dst[2], s.hist + p * kSortBins,
                    flags_pass, flags_next, s.ctl, p, digits[p].shift);
            }
        } else {
            {
                ProfileScope scope(names.pass, stream);
                sort_pass_kernel<ITEMS, false><<<groups, kSortThreads, 0, stream>>>(
                    in_sort, nullptr, src[2], out_sort, nullptr, dst[2], s.hist + p * kSortBins,
                    flags_pass, flags_next, s.ctl, p, digits[p].shift);
            }
        }
        DYN_LAUNCHED("sort_pass_kernel");
        for (int l = 0; l < 3; ++l) {
            uint32_t* tmp = src[l];
            src[l] = dst[l];
            dst[l] = tmp;
        }
    }
    if (passes % 2 == 1) {
        {
            ProfileScope scope(names.copy, stream);
            sort_copy_kernel<<<stream_groups(capacity), 256, 0, stream>>>(
                s.major, s.minor, s.payload, major, minor, payload, s.ctl);
        }
        DYN_LAUNCHED("sort_copy_kernel");
    }
    return DYN_OK;
}

// ---- tiling: the autotuner's table ---------------------------------------------------------------
//
// A pass's tiling is keys per lane (8: 2048-key tiles, 16: 4096-key tiles) and the cap on its
// persistent workgroups. The tuner times every candidate on the device for a capacity bucket
// (ceil(log2(capacity))) and keeps the fastest per (device, bucket); untuned buckets use the
// default: narrow tiles up to 2^21 keys (more of the chip busy on shorter streams), 1024 groups
// (about four per CU).

struct SortTiling {
    uint32_t items, groups;
};

constexpr SortTiling kTilings[] = {{8u, 512u}, {8u, 1024u}, {8u, 2048u}, {16u, 512u}, {16u, 1024u}, {16u, 2048u}};
constexpr uint32_t kTilingCount = sizeof(kTilings) / sizeof(kTilings[0]);

std::mutex& tune_mutex() {
    static std::mutex m;
    return m;
}
std::map<std::pair<int, uint32_t>, SortTiling>& tune_table() {
    static std::map<std::pair<int, uint32_t>, SortTiling> t;
    return t;
}

uint32_t tune_bucket(uint32_t capacity) {
    return capacity <= 1u ? 0u : 32u - static_cast<uint32_t>(__builtin_clz(capacity - 1u));
}

SortTiling default_tiling(uint32_t capacity) {
    return SortTiling{capacity <= (1u << 21) ? 8u : 16u, 1024u};
}

SortTiling tiling_for(uint32_t capacity) {
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return default_tiling(capacity);
    std::lock_guard<std::mutex> lock(tune_mutex());
    const auto it = tune_table().find({device, tune_bucket(capacity)});
    return it == tune_table().end() ? default_tiling(capacity) : it->second;
}

int passes_with(hipStream_t stream, uint32_t* major, uint32_t* minor, uint32_t* payload, const SortScratch& s,
                const SortDigit* digits, uint32_t passes, uint32_t capacity, const SortNames& names,
                SortTiling t) {
    if (t.items == 8u) return launch_passes<8>(stream, major, minor, payload, s, digits, passes, capacity, names, t.groups);
    return launch_passes<16>(stream, major, minor, payload, s, digits, passes, capacity, names, t.groups);
}

int sort_passes(hipStream_t stream, uint32_t* major, uint32_t* minor, uint32_t* payload,
                const SortScratch& s, const SortDigit* digits, uint32_t passes,
                uint32_t capacity, const SortNames& names) {
    return passes_with(stream, major, minor, payload, s, digits, passes, capacity, names, tiling_for(capacity));
}

namespace {

__global__ void __launch_bounds__(256) sort_tune_fill(uint32_t* major, uint32_t* minor, uint32_t* payload,
                                                      uint32_t count, uint32_t seed) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t h = (i + seed) * 0x9E3779B1u;
    h ^= h >> 15;
    h *= 0x85EBCA77u;
    h ^= h >> 13;
    major[i] = h;
    minor[i] = h * 0xC2B2AE3Du;
    payload[i] = i;
}

}  // namespace

}  // namespace dyn

extern "C" {

uint32_t dyn_sort_units_for(uint32_t live) {
    uint64_t units = (static_cast<uint64_t>(live) + 255u) / 256u + 64u;
    if (units < 1u) units = 1u;
    if (units > 512u) units = 512u;
    return static_cast<uint32_t>(units);
}

uint32_t dyn_sort_key_words(uint32_t elements) {
    const uint32_t last = elements ? elements - 1u : 0u;
    const uint32_t bits = last ? 32u - static_cast<uint32_t>(__builtin_clz(last)) : 0u;
    uint32_t words = (bits + 7u) / 8u;
    if (words < 1u) words = 1u;
    if (words > 4u) words = 4u;
    return words;
}

size_t dyn_sort_scratch_bytes(uint32_t capacity, uint32_t digit_passes) {
    return dyn::sort_scratch_layout(capacity, digit_passes, true, nullptr, nullptr);
}

int dyn_sort_autotune(void* stream, uint32_t capacity, uint32_t digit_passes, uint32_t* items_out,
                      uint32_t* groups_out) {
    using namespace dyn;
    if (capacity == 0u || capacity > kSortMaxCapacity) return fail(DYN_E_INVALID, "dyn_sort_autotune: capacity out of range");
    if (digit_passes == 0u || digit_passes > kSortMaxPasses) {
        return fail(DYN_E_INVALID, "dyn_sort_autotune: digit_passes must be 1..8");
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    const size_t scratch_bytes = sort_scratch_layout(capacity, digit_passes, true, nullptr, nullptr);
    std::vector<void*> owned;
    auto cleanup = [&](int rc) {
        (void)hipStreamSynchronize(st);
        for (void* p : owned) (void)hipFree(p);
        return rc;
    };
    uint32_t* lanes[3] = {nullptr, nullptr, nullptr};
    void* scratch = nullptr;
    for (auto*& lane : lanes) {
        if (hipMalloc(&lane, 4u * static_cast<size_t>(capacity)) != hipSuccess) return cleanup(fail(DYN_E_HIP, "dyn_sort_autotune: lanes"));
        owned.push_back(lane);
    }
    if (hipMalloc(&scratch, scratch_bytes) != hipSuccess) return cleanup(fail(DYN_E_HIP, "dyn_sort_autotune: scratch"));
    owned.push_back(scratch);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        return cleanup(fail(DYN_E_HIP, "dyn_sort_autotune: events"));
    }
    SortScratch s;
    sort_scratch_layout(capacity, digit_passes, true, scratch, &s);
    SortDigit digits[kSortMaxPasses];
    sort_digits(digit_passes > 4u ? 4u : digit_passes, digit_passes > 4u ? digit_passes - 4u : 0u, digits);
    const SortNames& names = kGenericSortNames;
    float best_ms = 0.0f;
    uint32_t best = kTilingCount;
    int rc = DYN_OK;
    for (uint32_t round = 0; round < 4 && rc == DYN_OK; ++round) {  // round 0 warms every variant
        for (uint32_t c = 0; c < kTilingCount && rc == DYN_OK; ++c) {
            sort_tune_fill<<<ceil_div(capacity, 256u), 256, 0, st>>>(lanes[0], lanes[1], lanes[2], capacity, round);
            rc = sort_prepare(st, s, nullptr, nullptr, capacity, names);
            if (rc == DYN_OK) rc = sort_histogram(st, lanes[0], lanes[1], s, digits, digit_passes, capacity, names);
            if (rc != DYN_OK || hipEventRecord(e0, st) != hipSuccess) break;
            rc = passes_with(st, lanes[0], lanes[1], lanes[2], s, digits, digit_passes, capacity, names, kTilings[c]);
            if (rc != DYN_OK || hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) break;
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (round > 0 && (best == kTilingCount || ms < best_ms)) {
                best_ms = ms;
                best = c;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != DYN_OK) return cleanup(rc);
    if (best == kTilingCount) return cleanup(fail(DYN_E_HIP, "dyn_sort_autotune: timing failed"));
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return cleanup(fail(DYN_E_HIP, "dyn_sort_autotune: device"));
    {
        std::lock_guard<std::mutex> lock(tune_mutex());
        tune_table()[{device, tune_bucket(capacity)}] = kTilings[best];
    }
    if (items_out) *items_out = kTilings[best].items;
    if (groups_out) *groups_out = kTilings[best].groups;
    return cleanup(DYN_OK);
}

int dyn_sort_tuning_export(uint32_t* out, uint32_t capacity, uint32_t* count) {
    using namespace dyn;
    if (!count) return fail(DYN_E_INVALID, "dyn_sort_tuning_export: null count");
    std::lock_guard<std::mutex> lock(tune_mutex());
    *count = static_cast<uint32_t>(tune_table().size());
    uint32_t i = 0;
    for (const auto& kv : tune_table()) {
        if (i >= capacity || !out) break;
        out[4 * i + 0] = static_cast<uint32_t>(kv.first.first);
        out[4 * i + 1] = kv.first.second;
        out[4 * i + 2] = kv.second.items;
        out[4 * i + 3] = kv.second.groups;
        ++i;
    }
    return DYN_OK;
}

int dyn_sort_tuning_import(const uint32_t* entries, uint32_t count) {
    using namespace dyn;
    if (count && !entries) return fail(DYN_E_INVALID, "dyn_sort_tuning_import: null entries");
    for (uint32_t i = 0; i < count; ++i) {
        const SortTiling t{entries[4 * i + 2], entries[4 * i + 3]};
        bool known = false;
        for (const SortTiling& k : kTilings) known |= k.items == t.items && k.groups == t.groups;
        if (!known || entries[4 * i + 1] > 31u) return fail(DYN_E_INVALID, "dyn_sort_tuning_import: not a tiling of this build");
    }
    std::lock_guard<std::mutex> lock(tune_mutex());
    tune_table().clear();
    for (uint32_t i = 0; i < count; ++i) {
        tune_table()[{static_cast<int>(entries[4 * i]), entries[4 * i + 1]}] = SortTiling{entries[4 * i + 2], entries[4 * i + 3]};
    }
    return DYN_OK;
}

int dyn_sort_faults(uint32_t* out, int clear) {
    if (!out) return dyn::fail(DYN_E_INVALID, "dyn_sort_faults: null output");
    DYN_HIP(hipDeviceSynchronize());
    DYN_HIP(hipMemcpyFromSymbol(out, HIP_SYMBOL(dyn::g_sort_faults), sizeof(uint32_t), 0, hipMemcpyDeviceToHost));
    if (clear) {
        const uint32_t zero = 0u;
        DYN_HIP(hipMemcpyToSymbol(HIP_SYMBOL(dyn::g_sort_faults), &zero, sizeof(uint32_t), 0, hipMemcpyHostToDevice));
    }
    return DYN_OK;
}

int dyn_sort_pairs(void* stream, uint32_t* d_major, uint32_t* d_minor, uint32_t* d_payload,
                   uint32_t capacity, const uint32_t* d_count, const uint32_t* d_base,
                   uint32_t plan, uint32_t major_words, uint32_t minor_words, void* d_scratch,
                   size_t scratch_bytes) {
    using namespace dyn;
    (void)plan;
    if (major_words > 4u || minor_words > 4u) {
        return fail(DYN_E_INVALID, "dyn_sort_pairs: a key has at most 4 byte digits per word");
    }
    if (major_words + minor_words == 0u) {
        return fail(DYN_E_INVALID,
                    "dyn_sort_pairs: a sort needs at least one key digit to permute the payload");
    }
    if (!d_major || !d_minor || !d_payload || !d_scratch) {
        return fail(DYN_E_INVALID, "dyn_sort_pairs: null lane or scratch");
    }
    if (capacity > kSortMaxCapacity) {
        return fail(DYN_E_INVALID, "dyn_sort_pairs: capacity above 2^30-1 records");
    }
    const uint32_t passes = major_words + minor_words;
    SortScratch s;
    const size_t need = sort_scratch_layout(capacity, passes, true, d_scratch, &s);
    if (scratch_bytes < need) {
        return fail(DYN_E_CAPACITY, "dyn_sort_pairs: scratch smaller than dyn_sort_scratch_bytes");
    }
    SortDigit digits[kSortMaxPasses];
    sort_digits(major_words, minor_words, digits);
    hipStream_t st = static_cast<hipStream_t>(stream);
    DYN_TRY(sort_prepare(st, s, d_count, d_base, capacity));
    DYN_TRY(sort_histogram(st, d_major, d_minor, s, digits, passes, capacity));
    DYN_TRY(sort_passes(st, d_major, d_minor, d_payload, s, digits, passes, capacity));
    return DYN_OK;
}

}  // extern "C"

// ---- the sort with no allocation: its scratch is a file-scope device array ----------------------

namespace {

// The scratch of dyn_sort_pairs_static: one array per device (each GPU the library runs on holds
// its own copy until the process ends), large enough for 2^20 records at any digit count.
constexpr size_t kStaticScratchBytes = size_t{17} << 20;
__device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kStaticDevices = 64;
std::mutex static_mutex;
void* static_scratch[kStaticDevices] = {};

// The current device's address of `scratch`, looked up on the device's first call (the array's
// host-side name is no device address).
int static_scratch_of_device(void** out) {
    int device = 0;
    DYN_HIP(hipGetDevice(&device));
    if (device < 0 || device >= kStaticDevices) return dyn::fail(DYN_E_INVALID, "dyn_sort_pairs_static: device ordinal out of range");
    std::lock_guard<std::mutex> guard(static_mutex);
    if (!static_scratch[device]) {
        void* ptr = nullptr;
        const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (rc != hipSuccess || !ptr) return dyn::hip_fail(rc, "dyn_sort_pairs_static: address of the static scratch");
        static_scratch[device] = ptr;
    }
    *out = static_scratch[device];
    return DYN_OK;
}

}  // namespace

extern "C" {

size_t dyn_sort_static_scratch_bytes(void) { return kStaticScratchBytes; }

int dyn_sort_pairs_static(void* stream, uint32_t* d_major, uint32_t* d_minor, uint32_t* d_payload,
                          uint32_t capacity, const uint32_t* d_count, const uint32_t* d_base, uint32_t plan,
                          uint32_t major_words, uint32_t minor_words) {
    using namespace dyn;
    if (major_words > 4u || minor_words > 4u || major_words + minor_words == 0u) {
        return dyn_sort_pairs(stream, d_major, d_minor, d_payload, capacity, d_count, d_base, plan, major_words,
                              minor_words, nullptr, 0u);  // the same argument errors
    }
    if (capacity > kSortMaxCapacity ||
        sort_scratch_layout(capacity, major_words + minor_words, true, nullptr, nullptr) > kStaticScratchBytes) {
        return fail(DYN_E_CAPACITY, "dyn_sort_pairs_static: capacity beyond the static scratch (2^20 records)");
    }
    void* base = nullptr;
    DYN_TRY(static_scratch_of_device(&base));
    return dyn_sort_pairs(stream, d_major, d_minor, d_payload, capacity, d_count, d_base, plan, major_words,
                          minor_words, base, kStaticScratchBytes);
}

}  // extern "C"
