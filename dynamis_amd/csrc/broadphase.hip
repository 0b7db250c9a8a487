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

__device__ __forceinline__ uint32_t* ctr(uint32_t* counters, uint32_t slot) {
    return counters + slot * DYN_COUNTER_STRIDE_WORDS;
}



__global__ void bp_reset_kernel(uint32_t* counters) {
    const uint32_t slots[] = {DYN_COUNTER_ENTRIES,     DYN_COUNTER_PAIRS,
                              DYN_COUNTER_REFUSED_PAIRS, DYN_COUNTER_ENTRY_FAULTS,
                              DYN_COUNTER_GRID_LEVELS, DYN_COUNTER_COARSE_ACTIVE,
                              DYN_COUNTER_GRID_SCALE,  DYN_COUNTER_GRID_EXTENT};
    const uint32_t t = threadIdx.x;
    if (t < sizeof(slots) / sizeof(slots[0])) *ctr(counters, slots[t]) = 0;
}

// broadphase_aabb.wgsl:21-24: the base cell is the smallest resolution box's largest side.
__global__ void __launch_bounds__(kBpThreads) bp_resolution_kernel(const dyn_aabb* __restrict__ boxes,
                                                                   uint32_t n, uint32_t* counters) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t scale = 0, extent = 0;
    if (i < n) {
        const float4 lo = reinterpret_cast<const float4*>(boxes)[2 * i];
        const float4 hi = reinterpret_cast<const float4*>(boxes)[2 * i + 1];
        const float3 mn = make_float3(lo.x, lo.y, lo.z);
        const float3 mx = make_float3(hi.x, hi.y, hi.z);
        scale = __float_as_uint(grid_scale_of(mn, mx));
        extent = __float_as_uint(fmaxf(grid_extent_of(mn, mx), 0.0f));
    }
    scale = wave_max(scale);
    extent = wave_max(extent);
    if (lane_id() == 0) {
        atomicMax(ctr(counters, DYN_COUNTER_GRID_SCALE), scale);
        atomicMax(ctr(counters, DYN_COUNTER_GRID_EXTENT), extent);
    }
}

// Per collider: level, cell count, emitted entries (<= 8), faults, coarse activity, levels.
__global__ void __launch_bounds__(kBpThreads) bp_count_kernel(const dyn_aabb* __restrict__ aabbs,
                                                              const uint32_t* __restrict__ flags,
                                                              uint32_t n, uint32_t* counters,
                                                              uint32_t* cnt_level,
                                                              uint32_t* block_totals) {
    __shared__ uint32_t s_scan[kBpThreads / kWave];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const float base = grid_cell_size(*ctr(counters, DYN_COUNTER_GRID_SCALE),
                                      *ctr(counters, DYN_COUNTER_GRID_EXTENT));
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
    if (lane_id() == 0) {
        if (levels) atomicOr(ctr(counters, DYN_COUNTER_GRID_LEVELS), levels);
        if (faults) atomicAdd(ctr(counters, DYN_COUNTER_ENTRY_FAULTS), faults);
        if (coarse) atomicAdd(ctr(counters, DYN_COUNTER_COARSE_ACTIVE), coarse);
    }
    uint32_t total = 0;
    block_exclusive_sum<kBpThreads>(emitted, s_scan, &total);
    if (threadIdx.x == 0) block_totals[blockIdx.x] = total;
}

// One workgroup: exclusive scan of the block totals; entry count; sort control words.
__global__ void __launch_bounds__(1024) bp_scan_kernel(const uint32_t* __restrict__ block_totals,
                                                       uint32_t blocks, uint32_t* block_offsets,
                                                       uint32_t* counters, uint32_t* sort_hist,
                                                       uint32_t* sort_ctl,
                                                       uint32_t entry_capacity) {
    __shared__ uint32_t s_scan[1024 / kWave];
    __shared__ uint32_t s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
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
    const uint32_t* sort_ctl, uint32_t* flags0) {
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
            const uint32_t key = cell_key(level, coord[0], coord[1], coord[2]);
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

// Half-open run [first, end) of `key` in the sorted keys, via the level directory.
__device__ __forceinline__ void key_run(const uint32_t* __restrict__ keys,
                                        const uint32_t* __restrict__ dir, uint32_t key,
                                        uint32_t* first, uint32_t* end) {
    const uint32_t level = key >> DYN_LEVEL_KEY_SHIFT;
    const uint32_t bucket = (key & DYN_CELL_HASH_MASK) >> (DYN_LEVEL_KEY_SHIFT - kDirBits);
    const uint32_t* d = dir + level * kDirStride + bucket;
    uint32_t lo = d[0];
    const uint32_t bucket_end = d[1];
    uint32_t hi = bucket_end;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (keys[mid] < key) lo = mid + 1; else hi = mid;
    }
    *first = lo;
    uint32_t e = lo;
    while (e < bucket_end && keys[e] == key) ++e;
    *end = e;
}

__device__ __forceinline__ void emit_pair(uint32_t* counters, uint32_t* pair_major,
                                          uint32_t* pair_minor, uint32_t pair_capacity,
                                          uint32_t group_a, uint32_t info_a, uint32_t group_b,
                                          uint32_t info_b) {
    // pair_emit.wgsl:10-31 (node identity is excluded by the callers).
    if (((info_a & DYN_ENTRY_KIND_MASK) >> DYN_ENTRY_KIND_SHIFT) != DYN_ENTRY_KIND_COLLIDER) return;
    if (((info_b & DYN_ENTRY_KIND_MASK) >> DYN_ENTRY_KIND_SHIFT) != DYN_ENTRY_KIND_COLLIDER) return;
    if (group_a == group_b) return;
    if (!(info_a & DYN_ENTRY_MOBILE) && !(info_b & DYN_ENTRY_MOBILE)) return;
    const uint32_t slot = atomicAdd(ctr(counters, DYN_COUNTER_PAIRS), 1u);
    if (slot < pair_capacity) {
        const uint32_t a = info_a & DYN_ENTRY_INDEX_MASK;
        const uint32_t b = info_b & DYN_ENTRY_INDEX_MASK;
        pair_minor[slot] = max(a, b);
        pair_major[slot] = min(a, b);
    } else {
        atomicAdd(ctr(counters, DYN_COUNTER_REFUSED_PAIRS), 1u);
    }
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

// One thread per sorted entry: cell mates (cell_pairs.wgsl:20-50) when awake, coarser-level
// links (level_links.wgsl:43-65) when primary.
__global__ void __launch_bounds__(256) bp_pairs_kernel(
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ order,
    const dyn_grid_entry* __restrict__ entries, const uint32_t* __restrict__ dir,
    uint32_t* counters, const uint32_t* sort_ctl, uint32_t* pair_major, uint32_t* pair_minor,
    uint32_t pair_capacity) {
    const uint32_t len = sort_ctl[kCtlLen];
    const uint32_t index = blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= len) return;
    const uint32_t key = keys[index];
    const uint32_t node = order[index];
    const EntryView self = load_entry(entries, node);
    if (((self.info & DYN_ENTRY_KIND_MASK) >> DYN_ENTRY_KIND_SHIFT) != DYN_ENTRY_KIND_COLLIDER) return;
    const float grid = grid_cell_size(*ctr(counters, DYN_COUNTER_GRID_SCALE),
                                      *ctr(counters, DYN_COUNTER_GRID_EXTENT));
    const bool awake = (self.info & DYN_ENTRY_AWAKE) != 0u;
    // The entry's level is the level bits of its own key (emitted from shape_levels of this box).
    const uint32_t level = key >> DYN_LEVEL_KEY_SHIFT;

    if (awake) {
        const float cs = level_cell_size(level, grid);
        int32_t cell[3];
        entry_cell(self.mn, self.info, cs, cell);
        // Earlier mates in the run: only entries that are not awake pair from this side.
        for (uint32_t e = index; e-- > 0 && keys[e] == key;) {
            const uint32_t other = order[e];
            const EntryView o = load_entry(entries, other);
            if (o.info & DYN_ENTRY_AWAKE) continue;
            if (mate_in_cell(self, o, cs, cell))
                emit_pair(counters, pair_major, pair_minor, pair_capacity, self.group, self.info,
                          o.group, o.info);
        }
        for (uint32_t e = index + 1; e < len && keys[e] == key; ++e) {
            const uint32_t other = order[e];
            const EntryView o = load_entry(entries, other);
            if (mate_in_cell(self, o, cs, cell))
                emit_pair(counters, pair_major, pair_minor, pair_capacity, self.group, self.info,
                          o.group, o.info);
        }
    }

    if (self.info & DYN_ENTRY_PRIMARY) {
        if (!awake && *ctr(counters, DYN_COUNTER_COARSE_ACTIVE) == 0u) return;
        const uint32_t occupied = *ctr(counters, DYN_COUNTER_GRID_LEVELS);
        const uint32_t shift = min(level + 1u, 31u);
        uint32_t coarser = occupied & ~((1u << shift) - 1u);
        while (coarser != 0u) {
            const uint32_t link = 31u - static_cast<uint32_t>(__clz(coarser));
            coarser &= ~(1u << link);
            const float cs = level_cell_size(link, grid);
            const Cells cells = grid_cells(self.mn, self.mx, cs);
            const uint32_t count = cell_count(cells);
            for (uint32_t ordinal = 0; ordinal < count; ++ordinal) {
                int32_t cell[3];
                cell_at(cells, ordinal, cell);
                uint32_t first, end;
                key_run(keys, dir, cell_key(link, cell[0], cell[1], cell[2]), &first, &end);
                for (uint32_t e = first; e < end; ++e) {
                    const uint32_t other = order[e];
                    if (other == node) continue;
                    const EntryView o = load_entry(entries, other);
                    if (!awake && !(o.info & DYN_ENTRY_AWAKE)) continue;
                    if (mate_in_cell(self, o, cs, cell))
                        emit_pair(counters, pair_major, pair_minor, pair_capacity, self.group,
                                  self.info, o.group, o.info);
                }
            }
        }
    }
}

struct BpLayout {
    uint32_t* cnt_level;
    uint32_t* block_totals;
    uint32_t* block_offsets;
    uint32_t* dir;
    uint32_t* pair_payload;
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
    l.block_offsets = carve.take<uint32_t>(blocks);
    l.dir = carve.take<uint32_t>(static_cast<size_t>(kGridLevels) * kDirStride);
    l.pair_payload = carve.take<uint32_t>(pair_capacity ? pair_capacity : 1u);
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
        bp_reset_kernel<<<1, 64, 0, st>>>(counters);
    }
    DYN_LAUNCHED("bp_reset_kernel");
    const uint32_t blocks = ceil_div(n ? n : 1u, kBpThreads);
    if (n) {
        const dyn_aabb* res = io->d_resolution ? io->d_resolution : io->d_aabbs;
        {
            ProfileScope scope("bp_resolution_kernel", st);
            bp_resolution_kernel<<<blocks, kBpThreads, 0, st>>>(res, n, counters);
        }
        DYN_LAUNCHED("bp_resolution_kernel");
        {
            ProfileScope scope("bp_count_kernel", st);
            bp_count_kernel<<<blocks, kBpThreads, 0, st>>>(io->d_aabbs, io->d_flags, n, counters,
                                                           l.cnt_level, l.block_totals);
        }
        DYN_LAUNCHED("bp_count_kernel");
    }
    {
        ProfileScope scope("bp_scan_kernel", st);
        bp_scan_kernel<<<1, 1024, 0, st>>>(l.block_totals, n ? blocks : 0u, l.block_offsets,
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
            l.entry_sort.ctl, l.entry_sort.flags);
    }
    DYN_LAUNCHED("bp_emit_kernel");
    SortDigit digits[kSortMaxPasses];
    sort_digits(kEntryPasses, 0u, digits);
    DYN_TRY(sort_passes(st, io->d_entry_keys, nullptr, io->d_entry_order, l.entry_sort, digits,
                        kEntryPasses, io->entry_capacity, kEntrySortNames));
    const uint32_t dir_threads = kGridLevels * kDirStride;
    {
        ProfileScope scope("bp_directory_kernel", st);
        bp_directory_kernel<<<ceil_div(dir_threads, 256), 256, 0, st>>>(io->d_entry_keys, counters,
                                                                        l.entry_sort.ctl, l.dir);
    }
    DYN_LAUNCHED("bp_directory_kernel");
    {
        ProfileScope scope("bp_pairs_kernel", st);
        bp_pairs_kernel<<<ceil_div(io->entry_capacity, 256), 256, 0, st>>>(
            io->d_entry_keys, io->d_entry_order, io->d_entries, l.dir, counters, l.entry_sort.ctl,
            io->d_pair_major, io->d_pair_minor, io->pair_capacity);
    }
    DYN_LAUNCHED("bp_pairs_kernel");
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
