// Broadphase stage on gfx950: base-cell resolution, grid-entry emission, the region sort, and
// candidate-pair emission from same-level cell mates and coarser-level links.
//
// Semantics follow the reference (crates/dynamis-rigid/shaders/broadphase_aabb.wgsl:21-24,
// collider_entries.wgsl:9-35, collider_emit.wgsl:6-63, crates/dynamis-broadphase/src/
// passes.rs:76-160, shaders/cell_pairs.wgsl, level_links.wgsl, pair_emit.wgsl). The layout is
// chosen for the device:
//   - entry slots come from a prefix sum over per-collider entry counts instead of one atomic
//     counter, so entries are laid out in collider order and equal keys leave the stable sort in
//     collider order (the reference's order within a key is whatever its atomics produced);
//   - the sort's digit histogram is counted while the entries are written, saving a read;
//   - the sorted keys get a per-level directory of 2^16 hash buckets, so a coarser-level probe
//     finds its cell's run with two directory reads and a short search inside one bucket instead
//     of a binary search over the whole region (grid_index.wgsl:231-253);
//   - a same-level entry finds its cell's run by walking its sorted neighbours: its own key is
//     the cell key it would search for.
#include "grid.h"
#include "sort.h"

namespace dyn {

namespace {

constexpr int kBpThreads = 256;
constexpr uint32_t kEntryPasses = 4;  // 4 key bytes, passes.rs:88
constexpr uint32_t kResolutionBlocks = 1024;
constexpr uint32_t kSparseLevel = 32;  // coarser levels of at most this many entries are read whole
constexpr uint32_t kLinkStage = 64;    // link pairs a collider's counting pass keeps for the emitting pass
constexpr uint32_t kGatherRun = 8;     // mean run length of a coarser level from which records are gathered
constexpr uint32_t kLinkLanesBelow = 32768;  // up to this many colliders the link passes lack threads: 8 lanes
constexpr uint32_t kLinkModeLanes = 1u;      // link mode word: eight lanes per collider
constexpr uint32_t kLinkModeGathered = 2u;   // link mode word: records gathered into key order

__device__ __forceinline__ uint32_t* ctr(uint32_t* counters, uint32_t slot) {
    return counters + slot * DYN_COUNTER_STRIDE_WORDS;
}



__global__ void bp_reset_kernel(uint32_t* counters, uint32_t* level_keys) {
    const uint32_t slots[] = {DYN_COUNTER_ENTRIES,     DYN_COUNTER_PAIRS,
                              DYN_COUNTER_REFUSED_PAIRS, DYN_COUNTER_ENTRY_FAULTS,
                              DYN_COUNTER_GRID_LEVELS, DYN_COUNTER_COARSE_ACTIVE,
                              DYN_COUNTER_GRID_SCALE,  DYN_COUNTER_GRID_EXTENT};
    const uint32_t t = threadIdx.x;
    if (t < sizeof(slots) / sizeof(slots[0])) *ctr(counters, slots[t]) = 0;
    if (t < kGridLevels) level_keys[t] = 0u;
}

// broadphase_aabb.wgsl:21-24: the base cell is the smallest resolution box's largest side.
// Each block reduces a grid-strided share to one (scale, extent) partial; a single workgroup then
// folds the partials. No two blocks ever update the same word.
__global__ void __launch_bounds__(kBpThreads) bp_resolution_kernel(const dyn_aabb* __restrict__ boxes,
                                                                   uint32_t n, uint32_t* partials) {
    __shared__ uint32_t s_red[2][kBpThreads / kWave];
    uint32_t scale = 0, extent = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 lo = reinterpret_cast<const float4*>(boxes)[2 * i];
        const float4 hi = reinterpret_cast<const float4*>(boxes)[2 * i + 1];
        const float3 mn = make_float3(lo.x, lo.y, lo.z);
        const float3 mx = make_float3(hi.x, hi.y, hi.z);
        scale = max(scale, __float_as_uint(grid_scale_of(mn, mx)));
        extent = max(extent, __float_as_uint(fmaxf(grid_extent_of(mn, mx), 0.0f)));
    }
    scale = wave_max(scale);
    extent = wave_max(extent);
    const uint32_t wave = threadIdx.x / kWave;
    if (lane_id() == 0) {
        s_red[0][wave] = scale;
        s_red[1][wave] = extent;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBpThreads / kWave; ++w) {
            scale = max(scale, s_red[0][w]);
            extent = max(extent, s_red[1][w]);
        }
        partials[2 * blockIdx.x] = scale;
        partials[2 * blockIdx.x + 1] = extent;
    }
}

// Folds the resolution partials into GRID_SCALE / GRID_EXTENT (both are float bits of
// non-negative values, so the integer maximum is the float maximum).
__global__ void __launch_bounds__(1024) bp_resolution_fold_kernel(const uint32_t* __restrict__ partials,
                                                                  uint32_t blocks, uint32_t* counters) {
    __shared__ uint32_t s_red[2][1024 / kWave];
    uint32_t scale = 0, extent = 0;
    for (uint32_t b = threadIdx.x; b < blocks; b += blockDim.x) {
        scale = max(scale, partials[2 * b]);
        extent = max(extent, partials[2 * b + 1]);
    }
    scale = wave_max(scale);
    extent = wave_max(extent);
    const uint32_t wave = threadIdx.x / kWave;
    if (lane_id() == 0) {
        s_red[0][wave] = scale;
        s_red[1][wave] = extent;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 1024 / kWave; ++w) {
            scale = max(scale, s_red[0][w]);
            extent = max(extent, s_red[1][w]);
        }
        *ctr(counters, DYN_COUNTER_GRID_SCALE) = scale;
        *ctr(counters, DYN_COUNTER_GRID_EXTENT) = extent;
    }
}

// Per collider: level, cell count, emitted entries (<= 8). Per block: the entry total and the
// block's share of the level mask, cell faults and coarse activity (folded by bp_scan_kernel).
__global__ void __launch_bounds__(kBpThreads) bp_count_kernel(const dyn_aabb* __restrict__ aabbs,
                                                              const uint32_t* __restrict__ flags,
                                                              uint32_t n, const uint32_t* counters,
                                                              uint32_t* cnt_level,
                                                              uint32_t* block_totals,
                                                              uint32_t* block_stats) {
    __shared__ uint32_t s_scan[kBpThreads / kWave];
    __shared__ uint32_t s_stats[3][kBpThreads / kWave];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const float base = grid_cell_size(counters[DYN_COUNTER_GRID_SCALE * DYN_COUNTER_STRIDE_WORDS],
                                      counters[DYN_COUNTER_GRID_EXTENT * DYN_COUNTER_STRIDE_WORDS]);
    uint32_t emitted = 0, levels = 0, faults = 0, coarse = 0;
    if (i < n) {
        const float4 lo = reinterpret_cast<const float4*>(aabbs)[2 * i];
        const float4 hi = reinterpret_cast<const float4*>(aabbs)[2 * i + 1];
        const float3 mn = make_float3(lo.x, lo.y, lo.z);
        const float3 mx = make_float3(hi.x, hi.y, hi.z);
        const uint32_t f = flags ? flags[i] : (DYN_BP_MOBILE | DYN_BP_AWAKE);
        const uint32_t level = shape_levels(mn, mx, base);
        const Cells cells = grid_cells(mn, mx, level_cell_size(level, base));
        const uint32_t count = cell_count(cells);
        emitted = min(count, DYN_MAX_CELLS_PER_COLLIDER);
        faults = count - emitted;
        levels = 1u << level;
        coarse = (level > 0u && (f & DYN_BP_AWAKE)) ? 1u : 0u;
        cnt_level[i] = emitted | (level << 8);
    }
    levels = wave_or(levels);
    faults = wave_sum(faults);
    coarse = wave_sum(coarse);
    const uint32_t wave = threadIdx.x / kWave;
    if (lane_id() == 0) {
        s_stats[0][wave] = levels;
        s_stats[1][wave] = faults;
        s_stats[2][wave] = coarse;
    }
    uint32_t total = 0;
    block_exclusive_sum<kBpThreads>(emitted, s_scan, &total);
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBpThreads / kWave; ++w) {
            levels |= s_stats[0][w];
            faults += s_stats[1][w];
            coarse += s_stats[2][w];
        }
        block_totals[blockIdx.x] = total;
        block_stats[3 * blockIdx.x] = levels;
        block_stats[3 * blockIdx.x + 1] = faults;
        block_stats[3 * blockIdx.x + 2] = coarse;
    }
}

// One workgroup: exclusive scan of the block totals; entry count; sort control words.
__global__ void __launch_bounds__(1024) bp_scan_kernel(const uint32_t* __restrict__ block_totals,
                                                       const uint32_t* __restrict__ block_stats,
                                                       uint32_t blocks, uint32_t* block_offsets,
                                                       uint32_t* counters, uint32_t* sort_hist,
                                                       uint32_t* sort_ctl,
                                                       uint32_t entry_capacity) {
    __shared__ uint32_t s_scan[1024 / kWave];
    __shared__ uint32_t s_stats[3][1024 / kWave];
    __shared__ uint32_t s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    uint32_t levels = 0, faults = 0, coarse = 0;
    for (uint32_t b = threadIdx.x; b < blocks; b += blockDim.x) {
        levels |= block_stats[3 * b];
        faults += block_stats[3 * b + 1];
        coarse += block_stats[3 * b + 2];
    }
    levels = wave_or(levels);
    faults = wave_sum(faults);
    coarse = wave_sum(coarse);
    if (lane_id() == 0) {
        s_stats[0][threadIdx.x / kWave] = levels;
        s_stats[1][threadIdx.x / kWave] = faults;
        s_stats[2][threadIdx.x / kWave] = coarse;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 1024 / kWave; ++w) {
            levels |= s_stats[0][w];
            faults += s_stats[1][w];
            coarse += s_stats[2][w];
        }
        *ctr(counters, DYN_COUNTER_GRID_LEVELS) = levels;
        *ctr(counters, DYN_COUNTER_ENTRY_FAULTS) = faults;
        *ctr(counters, DYN_COUNTER_COARSE_ACTIVE) = coarse;
    }
    for (uint32_t first = 0; first < blocks; first += 1024) {
        const uint32_t b = first + threadIdx.x;
        const uint32_t v = b < blocks ? block_totals[b] : 0u;
        uint32_t total = 0;
        const uint32_t excl = block_exclusive_sum<1024>(v, s_scan, &total);
        if (b < blocks) block_offsets[b] = s_carry + excl;
        __syncthreads();
        if (threadIdx.x == 0) s_carry += total;
        __syncthreads();
    }
    for (uint32_t k = threadIdx.x; k < kSortMaxPasses * kSortBins; k += blockDim.x) sort_hist[k] = 0;
    if (threadIdx.x == 0) {
        const uint32_t total = s_carry;
        *ctr(counters, DYN_COUNTER_ENTRIES) = total;
        sort_ctl[kCtlLen] = total < entry_capacity ? total : entry_capacity;
        sort_ctl[kCtlBase] = 0;
    }
    if (threadIdx.x >= 2 && threadIdx.x < kCtlWords) sort_ctl[threadIdx.x] = 0;
}

// collider_emit.wgsl:6-63 with slots from the scan; also counts the sort's key digits.
__global__ void __launch_bounds__(kBpThreads) bp_emit_kernel(
    const dyn_aabb* __restrict__ aabbs, const uint32_t* __restrict__ groups,
    const uint32_t* __restrict__ flags, uint32_t n, uint32_t* counters,
    const uint32_t* __restrict__ cnt_level, const uint32_t* __restrict__ block_offsets,
    uint32_t* __restrict__ entry_keys, uint32_t* __restrict__ entry_order,
    dyn_grid_entry* __restrict__ entries, uint32_t capacity, uint32_t* sort_hist,
    const uint32_t* sort_ctl, uint32_t* flags0, const uint32_t* __restrict__ worlds,
    uint32_t* __restrict__ first_slot) {
    __shared__ uint32_t s_hist[kEntryPasses][kSortBins];
    __shared__ uint32_t s_scan[kBpThreads / kWave];
    const uint32_t t = threadIdx.x;
#pragma unroll
    for (uint32_t p = 0; p < kEntryPasses; ++p) s_hist[p][t] = 0;
    const uint32_t i = blockIdx.x * blockDim.x + t;
    const uint32_t packed = i < n ? cnt_level[i] : 0u;
    const uint32_t emitted = packed & 0xFFu;
    const uint32_t level = packed >> 8;
    const uint32_t first = block_offsets[blockIdx.x] +
                           block_exclusive_sum<kBpThreads>(emitted, s_scan, nullptr);
    if (i < n) first_slot[i] = first;  // the primary entry's slot, for the level links
    uint32_t faults = 0;
    if (emitted) {
        const float base = grid_cell_size(*ctr(counters, DYN_COUNTER_GRID_SCALE),
                                          *ctr(counters, DYN_COUNTER_GRID_EXTENT));
        const float4 lo = reinterpret_cast<const float4*>(aabbs)[2 * i];
        const float4 hi = reinterpret_cast<const float4*>(aabbs)[2 * i + 1];
        const float3 mn = make_float3(lo.x, lo.y, lo.z);
        const float3 mx = make_float3(hi.x, hi.y, hi.z);
        const uint32_t f = flags ? flags[i] : (DYN_BP_MOBILE | DYN_BP_AWAKE);
        const uint32_t group = groups ? groups[i] : i;
        const Cells cells = grid_cells(mn, mx, level_cell_size(level, base));
        const uint32_t salt = worlds ? world_salt(worlds[i]) : 0u;
        uint32_t info_base = i | (DYN_ENTRY_KIND_COLLIDER << DYN_ENTRY_KIND_SHIFT);
        if (f & DYN_BP_AWAKE) info_base |= DYN_ENTRY_AWAKE;
        if (f & DYN_BP_MOBILE) info_base |= DYN_ENTRY_MOBILE;
        for (uint32_t ordinal = 0; ordinal < emitted; ++ordinal) {
            const uint32_t slot = first + ordinal;
            if (slot >= capacity) {
                ++faults;
                continue;
            }
            int32_t coord[3];
            cell_at(cells, ordinal, coord);
            const uint32_t offset = cell_offset(cells, coord);
            uint32_t info = info_base | (offset << DYN_ENTRY_CELL_SHIFT);
            if (offset == 0u) info |= DYN_ENTRY_PRIMARY;
            const uint32_t key = cell_key(level, coord[0], coord[1], coord[2]) ^ salt;
            entry_keys[slot] = key;
            entry_order[slot] = slot;
            float4* rec = reinterpret_cast<float4*>(entries + slot);
            rec[0] = make_float4(lo.x, lo.y, lo.z, __uint_as_float(group));
            rec[1] = make_float4(hi.x, hi.y, hi.z, __uint_as_float(info));
#pragma unroll
            for (uint32_t p = 0; p < kEntryPasses; ++p) atomicAdd(&s_hist[p][(key >> (8 * p)) & 0xFFu], 1u);
        }
    }
    faults = wave_sum(faults);
    if (lane_id() == 0 && faults) atomicAdd(ctr(counters, DYN_COUNTER_ENTRY_FAULTS), faults);
    __syncthreads();
#pragma unroll
    for (uint32_t p = 0; p < kEntryPasses; ++p) {
        const uint32_t c = s_hist[p][t];
        if (c) atomicAdd(&sort_hist[p * kSortBins + t], c);
    }
    // Clear the first sort pass's look-back words.
    const uint32_t words = ceil_div(sort_ctl[kCtlLen], kSortMinTile) * kSortBins;
    for (uint32_t k = blockIdx.x * blockDim.x + t; k < words; k += gridDim.x * blockDim.x) flags0[k] = 0;
}

// Whether to read entry records in sorted-key order (gathered once after the sort) or through the
// sorted slots. Gathering pays when some coarser level holds long runs of equal keys, which every
// finer collider probing that level scans; with short runs the gather's random reads cost more
// than they save. Either way the same records are read.
__device__ __forceinline__ bool gather_pays(const uint32_t* __restrict__ dir, uint32_t occupied,
                                            const uint32_t* __restrict__ level_keys) {
    const uint32_t coarser = occupied & (occupied - 1u);  // every occupied level but the finest
    for (uint32_t m = coarser; m != 0u; m &= m - 1u) {
        const uint32_t level = static_cast<uint32_t>(__ffs(m)) - 1u;
        const uint32_t entries = dir[level * kDirStride + kDirBuckets] - dir[level * kDirStride];
        if (entries > kSparseLevel && entries > kGatherRun * level_keys[level]) return true;
    }
    return false;
}

// Distinct keys of every level, one count per run start.
__global__ void __launch_bounds__(kBpThreads) bp_level_keys_kernel(const uint32_t* __restrict__ keys,
                                                                   const uint32_t* sort_ctl,
                                                                   uint32_t* level_keys) {
    __shared__ uint32_t s_levels[kGridLevels];
    if (threadIdx.x < kGridLevels) s_levels[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t len = sort_ctl[kCtlLen];
    // A few blocks stride over the keys, so few adds reach each level's global word.
    for (uint32_t first = blockIdx.x * kBpThreads; first < len; first += gridDim.x * kBpThreads) {
        const uint32_t e = first + threadIdx.x;
        const bool start = e < len && (e == 0u || keys[e - 1] != keys[e]);
        const uint32_t level = start ? keys[e] >> DYN_LEVEL_KEY_SHIFT : 0u;
        // one shared add per level present in the wave
        uint64_t pending = __ballot(start);
        while (pending != 0u) {
            const uint32_t leader = static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(pending))) - 1u;
            const uint32_t target = __shfl(level, leader, kWave);
            const uint64_t same = __ballot(start && level == target);
            if (lane_id() == leader) atomicAdd(&s_levels[target], static_cast<uint32_t>(__popcll(same)));
            pending &= ~same;
        }
    }
    __syncthreads();
    if (threadIdx.x < kGridLevels && s_levels[threadIdx.x]) atomicAdd(&level_keys[threadIdx.x], s_levels[threadIdx.x]);
}

// The entry records in sorted-key order, so the pair walks read each cell's run contiguously.
__global__ void __launch_bounds__(kBpThreads) bp_gather_kernel(const uint32_t* __restrict__ order,
                                                               const dyn_grid_entry* __restrict__ entries,
                                                               const uint32_t* sort_ctl,
                                                               const uint32_t* __restrict__ dir,
                                                               const uint32_t* counters,
                                                               const uint32_t* __restrict__ level_keys,
                                                               dyn_grid_entry* __restrict__ sorted,
                                                               uint32_t colliders, uint32_t* link_mode) {
    const bool pays = gather_pays(dir, counters[DYN_COUNTER_GRID_LEVELS * DYN_COUNTER_STRIDE_WORDS], level_keys);
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // the level links' lanes, and whether records are gathered
        *link_mode = (pays || colliders <= kLinkLanesBelow ? kLinkModeLanes : 0u) | (pays ? kLinkModeGathered : 0u);
    }
    if (!pays) return;
    const uint32_t len = sort_ctl[kCtlLen];
    for (uint32_t e = blockIdx.x * kBpThreads + threadIdx.x; e < len; e += gridDim.x * kBpThreads) {
        const float4* src = reinterpret_cast<const float4*>(entries + order[e]);
        float4* dst = reinterpret_cast<float4*>(sorted + e);
        dst[0] = src[0];
        dst[1] = src[1];
    }
}

// First sorted index of each 2^16-bucket of every occupied level (lower bound of the bucket's
// smallest key); bucket kDirBuckets of a level holds the level's end.
__global__ void __launch_bounds__(256) bp_directory_kernel(const uint32_t* __restrict__ keys,
                                                          const uint32_t* counters,
                                                          const uint32_t* sort_ctl, uint32_t* dir) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= kGridLevels * kDirStride) return;
    const uint32_t level = g / kDirStride;
    const uint32_t bucket = g % kDirStride;
    const uint32_t occupied = counters[DYN_COUNTER_GRID_LEVELS * DYN_COUNTER_STRIDE_WORDS];
    if (!((occupied >> level) & 1u)) return;
    const uint64_t probe64 = (static_cast<uint64_t>(level) << DYN_LEVEL_KEY_SHIFT) +
                             (static_cast<uint64_t>(bucket) << (DYN_LEVEL_KEY_SHIFT - kDirBits));
    const uint32_t len = sort_ctl[kCtlLen];
    uint32_t lo = 0, hi = len;
    if (probe64 > 0xFFFFFFFFull) {
        lo = len;
    } else {
        const uint32_t probe = static_cast<uint32_t>(probe64);
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (keys[mid] < probe) lo = mid + 1; else hi = mid;
        }
    }
    dir[g] = lo;
}

// Candidate pairs are written in two passes (pair_emit.wgsl:10-31 appends each pair with a device
// atomic on one counter, which serializes): a counting pass records how many pairs every
// collider's level links and every sorted entry's cell mates produce, one scan turns the counts
// into offsets, and an emitting pass repeats the same walk and writes each pair at its offset.
// Link pairs come first, then cell-mate pairs; the canonical sort orders them afterwards.
struct PairCounter {
    uint32_t n = 0;
    __device__ __forceinline__ void operator()(uint32_t, uint32_t) { ++n; }
};

// The level-links passes give each collider an aligned group of kLanes lanes (8 where the
// coarser levels hold long runs, else 1), which test kLanes consecutive run entries at a time; a
// group's matches are numbered in entry order (ballot ranks), so its pairs come out in the order
// of a one-thread walk. Group sinks are called by every lane of the group together, each with
// its own entry's outcome.
constexpr uint32_t kLinkLanes = 8;
constexpr uint32_t kLinkColliders = kBpThreads / kLinkLanes;  // colliders per block, 8 lanes each

template <uint32_t kLanes>
__device__ __forceinline__ uint32_t link_group_bits(bool p) {
    const uint32_t lane = lane_id();
    return static_cast<uint32_t>(__ballot(p) >> (lane & ~(kLanes - 1u))) & ((1u << kLanes) - 1u);
}

// Lower bound of `key` in the sorted keys[lo, hi) by the lane group: rounds of kLanes probes at
// evenly spaced positions narrow the range about (kLanes + 1)-fold (a bisection for one lane),
// then the group reads the last few keys at once.
template <uint32_t kLanes>
__device__ __forceinline__ uint32_t group_lower_bound(const uint32_t* __restrict__ keys, uint32_t lo,
                                                      uint32_t hi, uint32_t key) {
    const uint32_t q = lane_id() % kLanes;
    const auto at = [&](uint32_t span, uint32_t lane) {
        return lo + static_cast<uint32_t>(static_cast<uint64_t>(span) * (lane + 1u) / (kLanes + 1u));
    };
    while (hi - lo > kLanes) {  // probe positions strictly increase across the lanes
        const uint32_t span = hi - lo;
        const uint32_t below = __popc(link_group_bits<kLanes>(keys[at(span, q)] < key));
        const uint32_t next_lo = below == 0u ? lo : at(span, below - 1u) + 1u;
        hi = below == kLanes ? hi : at(span, below);
        lo = next_lo;
    }
    const uint32_t e = lo + q;
    return lo + __popc(link_group_bits<kLanes>(e < hi && keys[e] < key));
}

// Counts the group's pairs and keeps the first kLinkStage, which the emitting pass then copies
// instead of walking again.
template <uint32_t kLanes>
struct GroupStage {
    uint32_t n = 0;
    uint2* stage;
    __device__ __forceinline__ void put(bool ok, uint32_t a, uint32_t b) {
        const uint32_t bits = link_group_bits<kLanes>(ok);
        const uint32_t at = n + __popc(bits & ((1u << (lane_id() % kLanes)) - 1u));
        if (ok && at < kLinkStage) stage[at] = make_uint2(a, b);
        n += __popc(bits);
    }
};

// Writes the group's pairs from `at` on, in order, up to the capacity.
template <uint32_t kLanes>
struct GroupWriter {
    uint32_t at;
    uint32_t capacity;
    uint32_t* major;
    uint32_t* minor;
    __device__ __forceinline__ void put(bool ok, uint32_t a, uint32_t b) {
        const uint32_t bits = link_group_bits<kLanes>(ok);
        const uint32_t to = at + __popc(bits & ((1u << (lane_id() % kLanes)) - 1u));
        if (ok && to < capacity) {
            major[to] = a;
            minor[to] = b;
        }
        at += __popc(bits);
    }
};

struct PairWriter {
    uint32_t at;
    uint32_t capacity;
    uint32_t* major;
    uint32_t* minor;
    __device__ __forceinline__ void operator()(uint32_t a, uint32_t b) {
        if (at < capacity) {
            major[at] = a;
            minor[at] = b;
        }
        ++at;
    }
};

// pair_emit.wgsl:10-31 (node identity is excluded by the callers): true with the canonical
// pair in *lo, *hi when the two entries make a candidate pair.
__device__ __forceinline__ bool pair_filter(uint32_t group_a, uint32_t info_a, uint32_t group_b, uint32_t info_b,
                                            const uint32_t* __restrict__ worlds, uint32_t* lo, uint32_t* hi) {
    if (((info_a & DYN_ENTRY_KIND_MASK) >> DYN_ENTRY_KIND_SHIFT) != DYN_ENTRY_KIND_COLLIDER) return false;
    if (((info_b & DYN_ENTRY_KIND_MASK) >> DYN_ENTRY_KIND_SHIFT) != DYN_ENTRY_KIND_COLLIDER) return false;
    if (group_a == group_b) return false;
    if (!(info_a & DYN_ENTRY_MOBILE) && !(info_b & DYN_ENTRY_MOBILE)) return false;
    const uint32_t a = info_a & DYN_ENTRY_INDEX_MASK;
    const uint32_t b = info_b & DYN_ENTRY_INDEX_MASK;
    if (worlds && worlds[a] != worlds[b]) return false;  // a batch's worlds never meet
    *lo = min(a, b);
    *hi = max(a, b);
    return true;
}

template <typename Sink>
__device__ __forceinline__ void emit_pair(Sink& sink, uint32_t group_a, uint32_t info_a, uint32_t group_b,
                                          uint32_t info_b, const uint32_t* __restrict__ worlds) {
    uint32_t a, b;
    if (pair_filter(group_a, info_a, group_b, info_b, worlds, &a, &b)) sink(a, b);
}

struct EntryView {
    float3 mn, mx;
    uint32_t group, info;
};

__device__ __forceinline__ EntryView load_entry(const dyn_grid_entry* __restrict__ entries,
                                                uint32_t node) {
    const float4* rec = reinterpret_cast<const float4*>(entries + node);
    const float4 a = rec[0];
    const float4 b = rec[1];
    EntryView e;
    e.mn = make_float3(a.x, a.y, a.z);
    e.mx = make_float3(b.x, b.y, b.z);
    e.group = __float_as_uint(a.w);
    e.info = __float_as_uint(b.w);
    return e;
}

struct EntryRecords {
    const dyn_grid_entry* sorted;   // records in key order, when gathered
    const dyn_grid_entry* entries;  // records by slot
    const uint32_t* order;          // slot of each sorted key
    const uint32_t* level_keys;     // distinct keys of every level
    bool gathered;
    __device__ __forceinline__ EntryView at(uint32_t e) const {
        return gathered ? load_entry(sorted, e) : load_entry(entries, order[e]);
    }
};

__device__ __forceinline__ bool same3(const int32_t a[3], const int32_t b[3]) {
    return a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
}

// The pair tests of cell_pairs.wgsl:6-18 / level_links.wgsl:27-37: boxes overlap, the other
// entry sits in `cell`, and the overlap's min corner lies in `cell`.
__device__ __forceinline__ bool mate_in_cell(const EntryView& a, const EntryView& b, float cs,
                                             const int32_t cell[3]) {
    if (!aabb_overlaps(a.mn, a.mx, b.mn, b.mx)) return false;
    int32_t oc[3];
    entry_cell(b.mn, b.info, cs, oc);
    if (!same3(oc, cell)) return false;
    int32_t mc[3] = {to_i32(floorf(fmaxf(a.mn.x, b.mn.x) / cs)),
                     to_i32(floorf(fmaxf(a.mn.y, b.mn.y) / cs)),
                     to_i32(floorf(fmaxf(a.mn.z, b.mn.z) / cs))};
    return same3(mc, cell);
}

// Same-level cell mates (cell_pairs.wgsl:20-50) of one sorted entry. An entry's own key is the
// key of the cell it searches, so its run is the block of equal keys around it; most entries have
// no equal neighbour and finish after reading three coalesced keys.
template <typename Sink>
__device__ __forceinline__ void cell_mates(uint32_t index, uint32_t len,
                                           const uint32_t* __restrict__ keys,
                                           const EntryRecords& rec, float grid,
                                           const uint32_t* __restrict__ worlds, Sink& sink) {
    const uint32_t key = keys[index];
    const bool before = index > 0 && keys[index - 1] == key;
    const bool after = index + 1 < len && keys[index + 1] == key;
    if (!before && !after) return;
    const EntryView self = rec.at(index);
    if (((self.info & DYN_ENTRY_KIND_MASK) >> DYN_ENTRY_KIND_SHIFT) != DYN_ENTRY_KIND_COLLIDER) return;
    if (!(self.info & DYN_ENTRY_AWAKE)) return;
    // The entry's level is the level bits of its own key (emitted from shape_levels of this box).
    const float cs = level_cell_size(key >> DYN_LEVEL_KEY_SHIFT, grid);
    int32_t cell[3];
    entry_cell(self.mn, self.info, cs, cell);
    // Earlier mates in the run: only entries that are not awake pair from this side.
    for (uint32_t e = index; e-- > 0 && keys[e] == key;) {
        const EntryView o = rec.at(e);
        if (o.info & DYN_ENTRY_AWAKE) continue;
        if (mate_in_cell(self, o, cs, cell)) emit_pair(sink, self.group, self.info, o.group, o.info, worlds);
    }
    for (uint32_t e = index + 1; e < len && keys[e] == key; ++e) {
        const EntryView o = rec.at(e);
        if (mate_in_cell(self, o, cs, cell)) emit_pair(sink, self.group, self.info, o.group, o.info, worlds);
    }
}

// Cell mates, counting pass: one thread per sorted entry.
__global__ void __launch_bounds__(kBpThreads) bp_mates_count_kernel(
    const uint32_t* __restrict__ keys, EntryRecords rec, const uint32_t* __restrict__ dir, const uint32_t* counters,
    const uint32_t* sort_ctl, const uint32_t* __restrict__ worlds, uint32_t* mate_counts,
    uint32_t* mate_blocks) {
    __shared__ uint32_t s_scan[kBpThreads / kWave];
    const uint32_t len = sort_ctl[kCtlLen];
    const float grid = grid_cell_size(counters[DYN_COUNTER_GRID_SCALE * DYN_COUNTER_STRIDE_WORDS],
                                      counters[DYN_COUNTER_GRID_EXTENT * DYN_COUNTER_STRIDE_WORDS]);
    const uint32_t index = blockIdx.x * kBpThreads + threadIdx.x;
    PairCounter counter;
    rec.gathered = gather_pays(dir, counters[DYN_COUNTER_GRID_LEVELS * DYN_COUNTER_STRIDE_WORDS], rec.level_keys);
    if (index < len) cell_mates(index, len, keys, rec, grid, worlds, counter);
    if (index < len) mate_counts[index] = counter.n;
    uint32_t total = 0;
    block_exclusive_sum<kBpThreads>(counter.n, s_scan, &total);
    if (threadIdx.x == 0) mate_blocks[blockIdx.x] = total;
}

// Cell mates, emitting pass: the same walk, pairs written at their scanned offsets.
__global__ void __launch_bounds__(kBpThreads) bp_mates_emit_kernel(
    const uint32_t* __restrict__ keys, EntryRecords rec, const uint32_t* __restrict__ dir, const uint32_t* counters,
    const uint32_t* sort_ctl, const uint32_t* __restrict__ worlds,
    const uint32_t* __restrict__ mate_counts, const uint32_t* __restrict__ mate_offsets,
    uint32_t* pair_major, uint32_t* pair_minor, uint32_t pair_capacity) {
    __shared__ uint32_t s_scan[kBpThreads / kWave];
    const uint32_t len = sort_ctl[kCtlLen];
    const float grid = grid_cell_size(counters[DYN_COUNTER_GRID_SCALE * DYN_COUNTER_STRIDE_WORDS],
                                      counters[DYN_COUNTER_GRID_EXTENT * DYN_COUNTER_STRIDE_WORDS]);
    const uint32_t index = blockIdx.x * kBpThreads + threadIdx.x;
    const uint32_t mine = index < len ? mate_counts[index] : 0u;
    const uint32_t at = mate_offsets[blockIdx.x] + block_exclusive_sum<kBpThreads>(mine, s_scan, nullptr);
    if (mine == 0u || at >= pair_capacity) return;
    PairWriter writer{at, pair_capacity, pair_major, pair_minor};
    rec.gathered = gather_pays(dir, counters[DYN_COUNTER_GRID_LEVELS * DYN_COUNTER_STRIDE_WORDS], rec.level_keys);
    cell_mates(index, len, keys, rec, grid, worlds, writer);
}

// Coarser-level links (level_links.wgsl:43-65) of one collider, walked by its lane group. Every
// collider's primary entry (cell offset 0) is its first slot, so the probe starts from the
// collider's own box, flags and level instead of from the sorted array; a primary beyond capacity
// probes nothing, as in the reference. Only strictly coarser levels are probed, which hold none
// of the collider's own entries, so the reference's skip of the probing entry itself never
// applies. Entry records are read in sorted order (bp_gather_kernel), so a run is read
// contiguously, the group's lanes reading neighbouring records.
template <uint32_t kLanes, typename Sink>
__device__ __forceinline__ void level_links(uint32_t i, uint32_t packed, uint32_t slot,
                                            const dyn_aabb* __restrict__ aabbs,
                                            const uint32_t* __restrict__ groups,
                                            const uint32_t* __restrict__ flags,
                                            const uint32_t* __restrict__ keys,
                                            const EntryRecords& rec,
                                            const uint32_t* __restrict__ dir, bool coarse_active,
                                            uint32_t occupied, float grid, uint32_t capacity,
                                            const uint32_t* __restrict__ worlds, Sink& sink) {
    const uint32_t q = lane_id() % kLanes;
    const uint32_t emitted = packed & 0xFFu;
    if (emitted == 0u || slot >= capacity) return;
    const uint32_t f = flags ? flags[i] : (DYN_BP_MOBILE | DYN_BP_AWAKE);
    const bool awake = (f & DYN_BP_AWAKE) != 0u;
    if (!awake && !coarse_active) return;
    const uint32_t level = packed >> 8;
    uint32_t coarser = occupied & ~((1u << min(level + 1u, 31u)) - 1u);
    if (coarser == 0u) return;
    EntryView self;
    const float4 lo = reinterpret_cast<const float4*>(aabbs)[2 * i];
    const float4 hi = reinterpret_cast<const float4*>(aabbs)[2 * i + 1];
    self.mn = make_float3(lo.x, lo.y, lo.z);
    self.mx = make_float3(hi.x, hi.y, hi.z);
    self.group = groups ? groups[i] : i;
    self.info = i | (DYN_ENTRY_KIND_COLLIDER << DYN_ENTRY_KIND_SHIFT) | DYN_ENTRY_PRIMARY;
    if (awake) self.info |= DYN_ENTRY_AWAKE;
    if (f & DYN_BP_MOBILE) self.info |= DYN_ENTRY_MOBILE;
    const uint32_t salt = worlds ? world_salt(worlds[i]) : 0u;
    while (coarser != 0u) {
        const uint32_t link = 31u - static_cast<uint32_t>(__clz(coarser));
        coarser &= ~(1u << link);
        const float cs = level_cell_size(link, grid);
        const Cells cells = grid_cells(self.mn, self.mx, cs);
        const uint32_t level_first = dir[link * kDirStride], level_end = dir[link * kDirStride + kDirBuckets];
        if (level_end - level_first <= kSparseLevel) {
            // A sparse level is read whole: an entry is in the run of one of this box's cells c
            // exactly when its own cell is c and its key is c's key in this world, which
            // mate_in_cell and the key test below check, so the same pairs come out.
            for (uint32_t base = level_first; base < level_end; base += kLanes) {
                const uint32_t e = base + q;
                bool ok = false;
                uint32_t a = 0u, b = 0u;
                if (e < level_end) {
                    const EntryView o = rec.at(e);
                    if (awake || (o.info & DYN_ENTRY_AWAKE)) {
                        int32_t oc[3];
                        entry_cell(o.mn, o.info, cs, oc);
                        if (oc[0] >= cells.mn[0] && oc[0] <= cells.mx[0] && oc[1] >= cells.mn[1] &&
                            oc[1] <= cells.mx[1] && oc[2] >= cells.mn[2] && oc[2] <= cells.mx[2] &&
                            keys[e] == (cell_key(link, oc[0], oc[1], oc[2]) ^ salt) &&
                            mate_in_cell(self, o, cs, oc)) {
                            ok = pair_filter(self.group, self.info, o.group, o.info, worlds, &a, &b);
                        }
                    }
                }
                sink.put(ok, a, b);
            }
            continue;
        }
        const uint32_t count = cell_count(cells);
        for (uint32_t ordinal = 0; ordinal < count; ++ordinal) {
            int32_t cell[3];
            cell_at(cells, ordinal, cell);
            // the cell's run: from the lower bound of its key in the key's directory bucket to the
            // first other key, found chunk by chunk as the run is read
            const uint32_t key = cell_key(link, cell[0], cell[1], cell[2]) ^ salt;
            const uint32_t level_of_key = key >> DYN_LEVEL_KEY_SHIFT;
            const uint32_t bucket = (key & DYN_CELL_HASH_MASK) >> (DYN_LEVEL_KEY_SHIFT - kDirBits);
            const uint32_t* d = dir + level_of_key * kDirStride + bucket;
            const uint32_t bucket_end = d[1];
            for (uint32_t base = group_lower_bound<kLanes>(keys, d[0], bucket_end, key); base < bucket_end;
                 base += kLanes) {
                const uint32_t e = base + q;
                const bool in_run = e < bucket_end && keys[e] == key;
                bool ok = false;
                uint32_t a = 0u, b = 0u;
                if (in_run) {
                    const EntryView o = rec.at(e);
                    if ((awake || (o.info & DYN_ENTRY_AWAKE)) && mate_in_cell(self, o, cs, cell)) {
                        ok = pair_filter(self.group, self.info, o.group, o.info, worlds, &a, &b);
                    }
                }
                sink.put(ok, a, b);
                if (link_group_bits<kLanes>(in_run) != (1u << kLanes) - 1u) break;  // the run ended here
            }
        }
    }
}

// Whether the level links give a collider kLinkLanes lanes: where some coarser level holds long
// runs of equal keys (the records are then gathered into key order too, gather_pays), which every
// finer collider probing that level reads, or where the colliders are too few to fill the GPU one
// lane each; otherwise the lanes would mostly wait, and one lane walks a collider's probes.
// bp_gather_kernel writes the choice and the gather decision to one word (kLinkModeLanes,
// kLinkModeGathered), which both link passes and the offsets read. Each pass is launched in both
// shapes, a workgroup per chunk; the workgroups of the shape not chosen read the word and return.

template <uint32_t kLanes>
__device__ __forceinline__ uint32_t link_chunk_colliders() { return kBpThreads / kLanes; }

__host__ __device__ __forceinline__ uint32_t link_chunks(uint32_t n, uint32_t lanes) {
    return ceil_div(n ? n : 1u, kBpThreads / lanes);
}

struct LinkArgs {
    const dyn_aabb* aabbs;
    const uint32_t* groups;
    const uint32_t* flags;
    uint32_t n;
    const uint32_t* cnt_level;
    const uint32_t* first_slot;
    const uint32_t* keys;
    EntryRecords rec;
    const uint32_t* dir;
    bool coarse_active;
    uint32_t occupied;
    float grid;
    uint32_t capacity;
    const uint32_t* worlds;
};

__device__ __forceinline__ LinkArgs link_args(const dyn_aabb* aabbs, const uint32_t* groups, const uint32_t* flags,
                                              uint32_t n, const uint32_t* cnt_level, const uint32_t* first_slot,
                                              const uint32_t* keys, EntryRecords rec, const uint32_t* dir,
                                              const uint32_t* counters, uint32_t capacity, const uint32_t* worlds,
                                              uint32_t mode) {
    rec.gathered = (mode & kLinkModeGathered) != 0u;
    return LinkArgs{aabbs, groups, flags, n, cnt_level, first_slot, keys, rec, dir,
                    counters[DYN_COUNTER_COARSE_ACTIVE * DYN_COUNTER_STRIDE_WORDS] != 0u,
                    counters[DYN_COUNTER_GRID_LEVELS * DYN_COUNTER_STRIDE_WORDS],
                    grid_cell_size(counters[DYN_COUNTER_GRID_SCALE * DYN_COUNTER_STRIDE_WORDS],
                                   counters[DYN_COUNTER_GRID_EXTENT * DYN_COUNTER_STRIDE_WORDS]),
                    capacity, worlds};
}

// Level links, counting pass: chunks of kBpThreads / kLanes consecutive colliders, a chunk per
// workgroup, kLanes lanes per collider; link_blocks holds each chunk's pair total.
template <uint32_t kLanes>
// Eight waves a SIMD (64 registers, a few spilled): the walks are latency-bound (the batch's
// count 1.63 -> 1.42 ms; the emit pass, which writes, was slower so).
__global__ void __launch_bounds__(kBpThreads, 8) bp_links_count_kernel(
    const dyn_aabb* __restrict__ aabbs, const uint32_t* __restrict__ groups,
    const uint32_t* __restrict__ flags, uint32_t n, const uint32_t* __restrict__ cnt_level,
    const uint32_t* __restrict__ first_slot, const uint32_t* __restrict__ keys,
    EntryRecords rec,
    const uint32_t* __restrict__ dir, const uint32_t* counters, uint32_t capacity,
    const uint32_t* __restrict__ worlds, const uint32_t* link_mode, uint32_t* link_counts,
    uint32_t* link_blocks, uint2* link_stage) {
    __shared__ uint32_t s_scan[kBpThreads / kWave];
    const uint32_t mode = *link_mode;
    if (((mode & kLinkModeLanes) != 0u) != (kLanes > 1u)) return;
    const LinkArgs a = link_args(aabbs, groups, flags, n, cnt_level, first_slot, keys, rec, dir, counters,
                                 capacity, worlds, mode);
    const uint32_t chunks = link_chunks(n, kLanes);
    for (uint32_t chunk = blockIdx.x; chunk < chunks; chunk += gridDim.x) {
        const uint32_t i = chunk * link_chunk_colliders<kLanes>() + threadIdx.x / kLanes;
        const bool leader = threadIdx.x % kLanes == 0u && i < n;
        GroupStage<kLanes> counter{0u, link_stage + static_cast<size_t>(i < n ? i : 0u) * kLinkStage};
        if (i < n) {
            level_links<kLanes>(i, cnt_level[i], first_slot[i], a.aabbs, a.groups, a.flags, a.keys, a.rec, a.dir,
                                a.coarse_active, a.occupied, a.grid, a.capacity, a.worlds, counter);
            if (leader) link_counts[i] = counter.n;
        }
        uint32_t total = 0;
        block_exclusive_sum<kBpThreads>(leader ? counter.n : 0u, s_scan, &total);
        if (threadIdx.x == 0) link_blocks[chunk] = total;
    }
}

// Level links, emitting pass: the counting pass's chunks and lanes.
template <uint32_t kLanes>
__global__ void __launch_bounds__(kBpThreads) bp_links_emit_kernel(
    const dyn_aabb* __restrict__ aabbs, const uint32_t* __restrict__ groups,
    const uint32_t* __restrict__ flags, uint32_t n, const uint32_t* __restrict__ cnt_level,
    const uint32_t* __restrict__ first_slot, const uint32_t* __restrict__ keys,
    EntryRecords rec,
    const uint32_t* __restrict__ dir, const uint32_t* counters, uint32_t capacity,
    const uint32_t* __restrict__ worlds, const uint32_t* link_mode, const uint32_t* __restrict__ link_counts,
    const uint32_t* __restrict__ link_offsets, const uint2* __restrict__ link_stage,
    uint32_t* pair_major, uint32_t* pair_minor, uint32_t pair_capacity) {
    __shared__ uint32_t s_scan[kBpThreads / kWave];
    __shared__ uint32_t s_at[kBpThreads / kLanes];
    const uint32_t mode = *link_mode;
    if (((mode & kLinkModeLanes) != 0u) != (kLanes > 1u)) return;
    const LinkArgs a = link_args(aabbs, groups, flags, n, cnt_level, first_slot, keys, rec, dir, counters,
                                 capacity, worlds, mode);
    const uint32_t chunks = link_chunks(n, kLanes);
    const uint32_t g = threadIdx.x / kLanes, q = threadIdx.x % kLanes;
    for (uint32_t chunk = blockIdx.x; chunk < chunks; chunk += gridDim.x) {
        const uint32_t i = chunk * link_chunk_colliders<kLanes>() + g;
        const uint32_t mine = i < n ? link_counts[i] : 0u;
        const uint32_t excl = block_exclusive_sum<kBpThreads>(q == 0u ? mine : 0u, s_scan, nullptr);
        if (q == 0u) s_at[g] = link_offsets[chunk] + excl;
        __syncthreads();
        const uint32_t at = s_at[g];
        if (mine != 0u && at < pair_capacity) {
            if (mine <= kLinkStage) {  // the counting pass kept them all
                const uint2* kept = link_stage + static_cast<size_t>(i) * kLinkStage;
                for (uint32_t j = q; j < mine && at + j < pair_capacity; j += kLanes) {
                    const uint2 p = kept[j];
                    pair_major[at + j] = p.x;
                    pair_minor[at + j] = p.y;
                }
            } else {
                GroupWriter<kLanes> writer{at, pair_capacity, pair_major, pair_minor};
                level_links<kLanes>(i, cnt_level[i], first_slot[i], a.aabbs, a.groups, a.flags, a.keys, a.rec,
                                    a.dir, a.coarse_active, a.occupied, a.grid, a.capacity, a.worlds, writer);
            }
        }
        __syncthreads();  // s_at is reused by the next chunk
    }
}

// One workgroup: exclusive offsets of the link blocks, then of the mate blocks after them, the
// pair count and the refused count.
__global__ void __launch_bounds__(1024) bp_pair_offsets_kernel(uint32_t* link_blocks, uint32_t colliders,
                                                               const uint32_t* link_mode,
                                                               uint32_t* mate_blocks, uint32_t mate_n,
                                                               uint32_t* counters, uint32_t pair_capacity) {
    const uint32_t link_n = link_chunks(colliders, (*link_mode & kLinkModeLanes) ? kLinkLanes : 1u);
    constexpr uint32_t kItems = 8;  // block totals per thread and round
    __shared__ uint32_t s_scan[1024 / kWave];
    __shared__ uint32_t s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    for (int part = 0; part < 2; ++part) {
        uint32_t* blocks = part == 0 ? link_blocks : mate_blocks;
        const uint32_t count = part == 0 ? link_n : mate_n;
        for (uint32_t first = 0; first < count; first += 1024u * kItems) {
            __syncthreads();
            const uint32_t base = first + threadIdx.x * kItems;
            uint32_t v[kItems];
            uint32_t sum = 0;
#pragma unroll
            for (uint32_t j = 0; j < kItems; ++j) {
                v[j] = base + j < count ? blocks[base + j] : 0u;
                sum += v[j];
            }
            uint32_t total = 0;
            uint32_t run = s_carry + block_exclusive_sum<1024>(sum, s_scan, &total);
#pragma unroll
            for (uint32_t j = 0; j < kItems; ++j) {
                if (base + j < count) blocks[base + j] = run;
                run += v[j];
            }
            __syncthreads();
            if (threadIdx.x == 0) s_carry += total;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t total = s_carry;
        counters[DYN_COUNTER_PAIRS * DYN_COUNTER_STRIDE_WORDS] = total;
        counters[DYN_COUNTER_REFUSED_PAIRS * DYN_COUNTER_STRIDE_WORDS] =
            total > pair_capacity ? total - pair_capacity : 0u;
    }
}

struct BpLayout {
    uint32_t* cnt_level;
    uint32_t* block_totals;
    uint32_t* block_stats;
    uint32_t* block_offsets;
    uint32_t* res_partials;
    uint32_t* link_counts;
    uint32_t* first_slot;
    uint32_t* link_mode;
    uint32_t* link_blocks;
    uint32_t* mate_counts;
    uint32_t* mate_blocks;
    uint32_t* dir;
    uint32_t* pair_payload;
    dyn_grid_entry* sorted;
    uint2* link_stage;
    uint32_t* level_keys;
    SortScratch entry_sort;
    SortScratch pair_sort;
    size_t bytes;
};

BpLayout bp_layout(uint32_t colliders, uint32_t entry_capacity, uint32_t pair_capacity,
                   void* base) {
    BpLayout l{};
    Carve carve(base);
    const uint32_t blocks = ceil_div(colliders ? colliders : 1u, kBpThreads);
    l.cnt_level = carve.take<uint32_t>(colliders ? colliders : 1u);
    l.block_totals = carve.take<uint32_t>(blocks);
    l.block_stats = carve.take<uint32_t>(3 * static_cast<size_t>(blocks));
    l.block_offsets = carve.take<uint32_t>(blocks);
    l.res_partials = carve.take<uint32_t>(2 * kResolutionBlocks);
    const uint32_t entry_blocks = ceil_div(entry_capacity ? entry_capacity : 1u, kBpThreads);
    l.link_counts = carve.take<uint32_t>(colliders ? colliders : 1u);
    l.first_slot = carve.take<uint32_t>(colliders ? colliders : 1u);
    l.link_mode = carve.take<uint32_t>(1u);
    l.link_blocks = carve.take<uint32_t>(ceil_div(colliders ? colliders : 1u, kLinkColliders));
    l.mate_counts = carve.take<uint32_t>(entry_capacity ? entry_capacity : 1u);
    l.mate_blocks = carve.take<uint32_t>(entry_blocks);
    l.dir = carve.take<uint32_t>(static_cast<size_t>(kGridLevels) * kDirStride);
    l.pair_payload = carve.take<uint32_t>(pair_capacity ? pair_capacity : 1u);
    l.sorted = carve.take<dyn_grid_entry>(entry_capacity ? entry_capacity : 1u);
    l.link_stage = carve.take<uint2>(static_cast<size_t>(colliders ? colliders : 1u) * kLinkStage);
    l.level_keys = carve.take<uint32_t>(kGridLevels);
    const uint32_t pair_passes = 2u * dyn_sort_key_words(colliders ? colliders : 1u);
    char* sort_base = base ? static_cast<char*>(base) + carve.used : nullptr;
    const size_t entry_bytes = sort_scratch_layout(entry_capacity, kEntryPasses, false, sort_base,
                                                   &l.entry_sort);
    char* pair_base = base ? sort_base + entry_bytes : nullptr;
    const size_t pair_bytes = sort_scratch_layout(pair_capacity, pair_passes, true, pair_base,
                                                  &l.pair_sort);
    l.bytes = carve.used + entry_bytes + pair_bytes;
    return l;
}

}  // namespace

}  // namespace dyn

extern "C" {

size_t dyn_broadphase_scratch_bytes(uint32_t colliders, uint32_t entry_capacity,
                                    uint32_t pair_capacity) {
    return dyn::bp_layout(colliders, entry_capacity, pair_capacity, nullptr).bytes;
}

int dyn_broadphase(void* stream, const dyn_broadphase_io* io) {
    using namespace dyn;
    if (!io) return fail(DYN_E_INVALID, "dyn_broadphase: null io");
    const uint32_t n = io->colliders;
    if (n > DYN_ENTRY_INDEX_MASK + 1u) {
        return fail(DYN_E_INVALID, "dyn_broadphase: collider indices must fit 24 bits");
    }
    if (io->entry_capacity > kSortMaxCapacity || io->pair_capacity > kSortMaxCapacity) {
        return fail(DYN_E_INVALID, "dyn_broadphase: capacity above 2^30-1");
    }
    if (!io->d_counters || !io->d_scratch ||
        (n && (!io->d_aabbs || !io->d_entry_keys || !io->d_entry_order || !io->d_entries ||
               !io->d_pair_major || !io->d_pair_minor))) {
        return fail(DYN_E_INVALID, "dyn_broadphase: null buffer");
    }
    BpLayout l = bp_layout(n, io->entry_capacity, io->pair_capacity, io->d_scratch);
    if (io->scratch_bytes < l.bytes) {
        return fail(DYN_E_CAPACITY, "dyn_broadphase: scratch smaller than dyn_broadphase_scratch_bytes");
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    uint32_t* counters = io->d_counters;
    {
        ProfileScope scope("bp_reset_kernel", st);
        bp_reset_kernel<<<1, 64, 0, st>>>(counters, l.level_keys);
    }
    DYN_LAUNCHED("bp_reset_kernel");
    const uint32_t blocks = ceil_div(n ? n : 1u, kBpThreads);
    if (n) {
        const dyn_aabb* res = io->d_resolution ? io->d_resolution : io->d_aabbs;
        {
            ProfileScope scope("bp_resolution_kernel", st);
            bp_resolution_kernel<<<min(blocks, kResolutionBlocks), kBpThreads, 0, st>>>(
                res, n, l.res_partials);
        }
        DYN_LAUNCHED("bp_resolution_kernel");
        {
            ProfileScope scope("bp_resolution_fold_kernel", st);
            bp_resolution_fold_kernel<<<1, 1024, 0, st>>>(l.res_partials,
                                                          min(blocks, kResolutionBlocks), counters);
        }
        DYN_LAUNCHED("bp_resolution_fold_kernel");
        {
            ProfileScope scope("bp_count_kernel", st);
            bp_count_kernel<<<blocks, kBpThreads, 0, st>>>(io->d_aabbs, io->d_flags, n, counters,
                                                           l.cnt_level, l.block_totals,
                                                           l.block_stats);
        }
        DYN_LAUNCHED("bp_count_kernel");
    }
    {
        ProfileScope scope("bp_scan_kernel", st);
        bp_scan_kernel<<<1, 1024, 0, st>>>(l.block_totals, l.block_stats, n ? blocks : 0u,
                                           l.block_offsets,
                                           counters, l.entry_sort.hist, l.entry_sort.ctl,
                                           io->entry_capacity);
    }
    DYN_LAUNCHED("bp_scan_kernel");
    if (n == 0 || io->entry_capacity == 0) return DYN_OK;
    {
        ProfileScope scope("bp_emit_kernel", st);
        bp_emit_kernel<<<blocks, kBpThreads, 0, st>>>(
            io->d_aabbs, io->d_groups, io->d_flags, n, counters, l.cnt_level, l.block_offsets,
            io->d_entry_keys, io->d_entry_order, io->d_entries, io->entry_capacity, l.entry_sort.hist,
            l.entry_sort.ctl, l.entry_sort.flags, io->d_worlds, l.first_slot);
    }
    DYN_LAUNCHED("bp_emit_kernel");
    SortDigit digits[kSortMaxPasses];
    sort_digits(kEntryPasses, 0u, digits);
    DYN_TRY(sort_passes(st, io->d_entry_keys, nullptr, io->d_entry_order, l.entry_sort, digits,
                        kEntryPasses, io->entry_capacity, kEntrySortNames));
    const uint32_t entry_blocks = ceil_div(io->entry_capacity, kBpThreads);
    const uint32_t dir_threads = kGridLevels * kDirStride;
    {
        ProfileScope scope("bp_directory_kernel", st);
        bp_directory_kernel<<<ceil_div(dir_threads, 256), 256, 0, st>>>(io->d_entry_keys, counters,
                                                                        l.entry_sort.ctl, l.dir);
    }
    DYN_LAUNCHED("bp_directory_kernel");
    {
        ProfileScope scope("bp_level_keys_kernel", st);
        bp_level_keys_kernel<<<min(entry_blocks, 1024u), kBpThreads, 0, st>>>(io->d_entry_keys, l.entry_sort.ctl,
                                                                  l.level_keys);
    }
    DYN_LAUNCHED("bp_level_keys_kernel");
    {
        ProfileScope scope("bp_gather_kernel", st);
        bp_gather_kernel<<<min(entry_blocks, 2048u), kBpThreads, 0, st>>>(io->d_entry_order, io->d_entries,
                                                              l.entry_sort.ctl, l.dir, counters,
                                                              l.level_keys, l.sorted, n, l.link_mode);
    }
    DYN_LAUNCHED("bp_gather_kernel");
    const EntryRecords rec{l.sorted, io->d_entries, io->d_entry_order, l.level_keys, false};
    const uint32_t thread_groups = link_chunks(n, 1u);
    const uint32_t lane_groups = link_chunks(n, kLinkLanes);
    {
        ProfileScope scope("bp_links_count_kernel", st);
        bp_links_count_kernel<1><<<thread_groups, kBpThreads, 0, st>>>(
            io->d_aabbs, io->d_groups, io->d_flags, n, l.cnt_level, l.first_slot,
            io->d_entry_keys, rec, l.dir, counters,
            io->entry_capacity, io->d_worlds, l.link_mode, l.link_counts, l.link_blocks, l.link_stage);
        bp_links_count_kernel<kLinkLanes><<<lane_groups, kBpThreads, 0, st>>>(
            io->d_aabbs, io->d_groups, io->d_flags, n, l.cnt_level, l.first_slot,
            io->d_entry_keys, rec, l.dir, counters,
            io->entry_capacity, io->d_worlds, l.link_mode, l.link_counts, l.link_blocks, l.link_stage);
    }
    DYN_LAUNCHED("bp_links_count_kernel");
    {
        ProfileScope scope("bp_mates_count_kernel", st);
        bp_mates_count_kernel<<<entry_blocks, kBpThreads, 0, st>>>(
            io->d_entry_keys, rec, l.dir, counters, l.entry_sort.ctl,
            io->d_worlds, l.mate_counts, l.mate_blocks);
    }
    DYN_LAUNCHED("bp_mates_count_kernel");
    {
        ProfileScope scope("bp_pair_offsets_kernel", st);
        bp_pair_offsets_kernel<<<1, 1024, 0, st>>>(l.link_blocks, n, l.link_mode, l.mate_blocks, entry_blocks,
                                                   counters, io->pair_capacity);
    }
    DYN_LAUNCHED("bp_pair_offsets_kernel");
    if (io->pair_capacity) {
        {
            ProfileScope scope("bp_links_emit_kernel", st);
            bp_links_emit_kernel<1><<<thread_groups, kBpThreads, 0, st>>>(
                io->d_aabbs, io->d_groups, io->d_flags, n, l.cnt_level, l.first_slot,
                io->d_entry_keys, rec, l.dir, counters,
                io->entry_capacity, io->d_worlds, l.link_mode, l.link_counts, l.link_blocks, l.link_stage,
                io->d_pair_major, io->d_pair_minor, io->pair_capacity);
            bp_links_emit_kernel<kLinkLanes><<<lane_groups, kBpThreads, 0, st>>>(
                io->d_aabbs, io->d_groups, io->d_flags, n, l.cnt_level, l.first_slot,
                io->d_entry_keys, rec, l.dir, counters,
                io->entry_capacity, io->d_worlds, l.link_mode, l.link_counts, l.link_blocks, l.link_stage,
                io->d_pair_major, io->d_pair_minor, io->pair_capacity);
        }
        DYN_LAUNCHED("bp_links_emit_kernel");
        {
            ProfileScope scope("bp_mates_emit_kernel", st);
            bp_mates_emit_kernel<<<entry_blocks, kBpThreads, 0, st>>>(
                io->d_entry_keys, rec, l.dir, counters, l.entry_sort.ctl,
                io->d_worlds, l.mate_counts, l.mate_blocks, io->d_pair_major, io->d_pair_minor,
                io->pair_capacity);
        }
        DYN_LAUNCHED("bp_mates_emit_kernel");
    }
    if (io->sort_pairs && io->pair_capacity) {
        const uint32_t words = dyn_sort_key_words(n);
        sort_digits(words, words, digits);
        const uint32_t* pairs = counters + DYN_COUNTER_PAIRS * DYN_COUNTER_STRIDE_WORDS;
        DYN_TRY(sort_prepare(st, l.pair_sort, pairs, nullptr, io->pair_capacity, kPairSortNames));
        DYN_TRY(sort_histogram(st, io->d_pair_major, io->d_pair_minor, l.pair_sort, digits,
                               2u * words, io->pair_capacity, kPairSortNames));
        DYN_TRY(sort_passes(st, io->d_pair_major, io->d_pair_minor, l.pair_payload, l.pair_sort,
                            digits, 2u * words, io->pair_capacity, kPairSortNames));
    }
    return DYN_OK;
}

}  // extern "C"
