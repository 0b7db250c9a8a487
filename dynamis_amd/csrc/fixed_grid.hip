// The literal fixed-cell broadphase: one uniform grid of a given cell size (4 m in the isolated
// sort + broadphase workload), each box entered in every cell its AABB touches with the key
// (cell hash, box index), an LSD radix sort of the keys, the cell runs of the sorted keys, and the
// overlapping pairs emitted once each and canonicalised (first < second, sorted).
//
// It is the alternative to the reference's hierarchical grid (crates/dynamis-broadphase: levels
// of cells sized from the boxes, `Broadphase::record` passes.rs:76-160) that a fixed 4 m grid with
// "cell hash, box index" keys describes. Both report every pair of overlapping boxes exactly once,
// so the canonical pair set is the same as the hierarchical grid's on the same boxes; the cell
// hash is the grid's (grid_index.wgsl:28-33), cells are floor(coordinate / cell) (grid_index.wgsl
// :47-52), the overlap test and the once-per-pair rule (the overlap's min corner lies in the cell)
// are cell_pairs.wgsl's.
//
// Kernels: per box its cell count; per 1024 boxes a block sum, one scan of the block sums; per box
// its entries written at its scanned offset (cell order x-major, as grid_index.wgsl:63-72); the
// sort (dyn_sort_pairs: cell hash major, box index minor, cell ordinal payload); per sorted entry a
// scan of the later entries of its hash run for the pair test (wave-aggregated appends); the pair
// sort.
#include "common.h"
#include "grid.h"
#include "sort.h"

namespace dyn {
namespace {

constexpr uint32_t kFixedThreads = 256;
constexpr uint32_t kFixedItems = 4;
constexpr uint32_t kFixedBlock = kFixedThreads * kFixedItems;  // boxes per scan block
constexpr uint32_t kMaxCellsPerBox = 64;                       // a box touching more cells is refused

struct FixedScratch {
    uint32_t* counts;        // per box: its cells (0 when refused)
    uint32_t* block_sums;    // per scan block: sum, then exclusive offset
    uint32_t* pair_payload;  // the pair sort's payload lane
    void* entry_sort;
    size_t entry_sort_bytes;
    void* pair_sort;
    size_t pair_sort_bytes;
};

size_t fixed_layout(uint32_t boxes, uint32_t entry_capacity, uint32_t pair_capacity, void* base,
                    FixedScratch* out) {
    const uint32_t blocks = ceil_div(boxes ? boxes : 1u, kFixedBlock);
    const uint32_t box_words = dyn_sort_key_words(boxes ? boxes : 1u);
    const size_t entry_sort = dyn_sort_scratch_bytes(entry_capacity ? entry_capacity : 1u, 4u + box_words);
    const size_t pair_sort = dyn_sort_scratch_bytes(pair_capacity ? pair_capacity : 1u, 2u * box_words);
    auto align = [](size_t b) { return (b + 255u) & ~static_cast<size_t>(255u); };
    size_t at = 0;
    const size_t counts = at;
    at += align(4u * static_cast<size_t>(boxes ? boxes : 1u));
    const size_t sums = at;
    at += align(4u * static_cast<size_t>(blocks) + 4u);
    const size_t payload = at;
    at += align(4u * static_cast<size_t>(pair_capacity ? pair_capacity : 1u));
    const size_t es = at;
    at += align(entry_sort);
    const size_t ps = at;
    at += align(pair_sort);
    if (base && out) {
        char* b = static_cast<char*>(base);
        out->counts = reinterpret_cast<uint32_t*>(b + counts);
        out->block_sums = reinterpret_cast<uint32_t*>(b + sums);
        out->pair_payload = reinterpret_cast<uint32_t*>(b + payload);
        out->entry_sort = b + es;
        out->entry_sort_bytes = entry_sort;
        out->pair_sort = b + ps;
        out->pair_sort_bytes = pair_sort;
    }
    return at;
}

__device__ __forceinline__ Cells box_cells(const dyn_aabb& b, float cell) {
    return grid_cells(make_float3(b.min[0], b.min[1], b.min[2]), make_float3(b.max[0], b.max[1], b.max[2]), cell);
}

// A box's cell count, or 0 (refused) past kMaxCellsPerBox; spans are computed in 64 bits.
__device__ __forceinline__ uint32_t box_cell_count(const Cells& c) {
    uint64_t n = 1;
    for (int a = 0; a < 3; ++a) {
        const int64_t s = static_cast<int64_t>(c.mx[a]) - static_cast<int64_t>(c.mn[a]) + 1;
        if (s <= 0 || s > kMaxCellsPerBox) return 0u;
        n *= static_cast<uint64_t>(s);
    }
    return n <= kMaxCellsPerBox ? static_cast<uint32_t>(n) : 0u;
}

// counts: [entries, pairs, refused boxes, refused pairs]
__global__ void __launch_bounds__(kFixedThreads) k_fixed_reset(uint32_t* counts) {
    if (threadIdx.x < 4) counts[threadIdx.x] = 0u;
}

__global__ void __launch_bounds__(kFixedThreads) k_fixed_count(const dyn_aabb* __restrict__ boxes, uint32_t n,
                                                               float cell, uint32_t* __restrict__ per_box,
                                                               uint32_t* __restrict__ block_sums,
                                                               uint32_t* __restrict__ counts) {
    __shared__ uint32_t s_sum;
    if (threadIdx.x == 0) s_sum = 0u;
    __syncthreads();
    uint32_t local = 0, refused = 0;
#pragma unroll
    for (uint32_t k = 0; k < kFixedItems; ++k) {
        const uint32_t i = blockIdx.x * kFixedBlock + k * kFixedThreads + threadIdx.x;
        if (i >= n) continue;
        const uint32_t c = box_cell_count(box_cells(boxes[i], cell));
        per_box[i] = c;
        local += c;
        refused += c == 0u ? 1u : 0u;
    }
    local = wave_sum(local);
    refused = wave_sum(refused);
    if (lane_id() == 0) {
        atomicAdd(&s_sum, local);
        if (refused) atomicAdd(&counts[2], refused);
    }
    __syncthreads();
    if (threadIdx.x == 0) block_sums[blockIdx.x] = s_sum;
}

// One workgroup: exclusive scan of the block sums in place; the total into counts[0] (clamped to
// the entry capacity later by the emit kernel's bounds).
__global__ void __launch_bounds__(kFixedThreads) k_fixed_scan_sums(uint32_t* block_sums, uint32_t blocks,
                                                                   uint32_t* counts) {
    __shared__ uint32_t s_wave[kFixedThreads / kWave];
    __shared__ uint32_t s_carry;
    if (threadIdx.x == 0) s_carry = 0u;
    __syncthreads();
    for (uint32_t base = 0; base < blocks; base += kFixedThreads) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < blocks ? block_sums[i] : 0u;
        const uint32_t incl = wave_inclusive_sum(v);
        if (lane_id() == kWave - 1) s_wave[threadIdx.x / kWave] = incl;
        __syncthreads();
        uint32_t before = s_carry;
        for (uint32_t w = 0; w < threadIdx.x / kWave; ++w) before += s_wave[w];
        if (i < blocks) block_sums[i] = before + incl - v;
        __syncthreads();
        if (threadIdx.x == kFixedThreads - 1) {
            uint32_t total = s_carry;
            for (uint32_t w = 0; w < kFixedThreads / kWave; ++w) total += s_wave[w];
            s_carry = total;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[0] = s_carry;
}

__global__ void __launch_bounds__(kFixedThreads) k_fixed_emit(const dyn_aabb* __restrict__ boxes, uint32_t n,
                                                              float cell, const uint32_t* __restrict__ per_box,
                                                              const uint32_t* __restrict__ block_sums,
                                                              uint32_t capacity, uint32_t* __restrict__ hash,
                                                              uint32_t* __restrict__ box, uint32_t* __restrict__ ordinal) {
    __shared__ uint32_t s_wave[kFixedThreads / kWave];
    // each thread owns kFixedItems consecutive boxes of the block, so the offsets follow box order
    const uint32_t first = blockIdx.x * kFixedBlock + threadIdx.x * kFixedItems;
    uint32_t c[kFixedItems];
    uint32_t mine = 0;
#pragma unroll
    for (uint32_t k = 0; k < kFixedItems; ++k) {
        c[k] = first + k < n ? per_box[first + k] : 0u;
        mine += c[k];
    }
    const uint32_t incl = wave_inclusive_sum(mine);
    if (lane_id() == kWave - 1) s_wave[threadIdx.x / kWave] = incl;
    __syncthreads();
    uint32_t at = block_sums[blockIdx.x] + incl - mine;
    for (uint32_t w = 0; w < threadIdx.x / kWave; ++w) at += s_wave[w];
    for (uint32_t k = 0; k < kFixedItems; ++k) {
        if (c[k] == 0u) continue;
        const uint32_t i = first + k;
        const Cells cells = box_cells(boxes[i], cell);
        for (uint32_t o = 0; o < c[k]; ++o, ++at) {
            if (at >= capacity) continue;
            int32_t xyz[3];
            cell_at(cells, o, xyz);
            hash[at] = cell_hash(xyz[0], xyz[1], xyz[2]);
            box[at] = i;
            ordinal[at] = o;
        }
    }
}

// Per sorted entry: the later entries of its hash run that sit in the same cell and overlap it,
// kept where the overlap's min corner lies in the cell (cell_pairs.wgsl:6-18).
__global__ void __launch_bounds__(kFixedThreads) k_fixed_pairs(const dyn_aabb* __restrict__ boxes, float cell,
                                                               const uint32_t* __restrict__ hash,
                                                               const uint32_t* __restrict__ box,
                                                               const uint32_t* __restrict__ ordinal,
                                                               uint32_t* __restrict__ counts, uint32_t entry_capacity,
                                                               uint32_t* __restrict__ first_out,
                                                               uint32_t* __restrict__ second_out, uint32_t pair_capacity) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t len = min(counts[0], entry_capacity);
    uint32_t found[8];
    uint32_t nfound = 0;
    uint32_t a = 0;
    bool more = false;  // a run longer than the local list: appended one at a time below
    if (e + 1 < len && hash[e + 1] == hash[e]) {
        const uint32_t key = hash[e];
        a = box[e];
        const dyn_aabb ba = boxes[a];
        const Cells ca = box_cells(ba, cell);
        int32_t here[3];
        cell_at(ca, ordinal[e], here);
        const float3 amn = make_float3(ba.min[0], ba.min[1], ba.min[2]);
        const float3 amx = make_float3(ba.max[0], ba.max[1], ba.max[2]);
        for (uint32_t f = e + 1; f < len && hash[f] == key; ++f) {
            const uint32_t b = box[f];
            const dyn_aabb bb = boxes[b];
            const float3 bmn = make_float3(bb.min[0], bb.min[1], bb.min[2]);
            const float3 bmx = make_float3(bb.max[0], bb.max[1], bb.max[2]);
            if (b == a || !aabb_overlaps(amn, amx, bmn, bmx)) continue;
            int32_t there[3];
            cell_at(box_cells(bb, cell), ordinal[f], there);
            if (there[0] != here[0] || there[1] != here[1] || there[2] != here[2]) continue;  // a hash collision
            const int32_t mc[3] = {to_i32(floorf(fmaxf(amn.x, bmn.x) / cell)), to_i32(floorf(fmaxf(amn.y, bmn.y) / cell)),
                                   to_i32(floorf(fmaxf(amn.z, bmn.z) / cell))};
            if (mc[0] != here[0] || mc[1] != here[1] || mc[2] != here[2]) continue;
            if (nfound < 8u) {
                found[nfound++] = b;
            } else {
                more = true;
                const uint32_t slot = atomicAdd(&counts[1], 1u);
                if (slot < pair_capacity) {
                    first_out[slot] = min(a, b);
                    second_out[slot] = max(a, b);
                } else {
                    atomicAdd(&counts[3], 1u);
                }
            }
        }
    }
    (void)more;
    // wave-aggregated append of the listed pairs
    const uint32_t before = wave_inclusive_sum(nfound) - nfound;
    const uint32_t total = wave_sum(nfound);
    uint32_t base = 0;
    if (lane_id() == 0 && total) base = atomicAdd(&counts[1], total);
    base = __shfl(base, 0);
    for (uint32_t k = 0; k < nfound; ++k) {
        const uint32_t slot = base + before + k;
        if (slot < pair_capacity) {
            first_out[slot] = min(a, found[k]);
            second_out[slot] = max(a, found[k]);
        } else {
            atomicAdd(&counts[3], 1u);
        }
    }
}

__global__ void __launch_bounds__(kFixedThreads) k_fixed_clamp(uint32_t* counts, uint32_t entry_capacity,
                                                               uint32_t pair_capacity, uint32_t* sort_counts) {
    if (threadIdx.x != 0) return;
    sort_counts[0] = min(counts[0], entry_capacity);
    sort_counts[1] = min(counts[1], pair_capacity);
}

}  // namespace
}  // namespace dyn

extern "C" {

size_t dyn_fixed_grid_scratch_bytes(uint32_t boxes, uint32_t entry_capacity, uint32_t pair_capacity) {
    return dyn::fixed_layout(boxes, entry_capacity, pair_capacity, nullptr, nullptr) + 256u;
}

int dyn_fixed_grid(void* stream, const dyn_aabb* d_boxes, uint32_t boxes, float cell_size, uint32_t* d_entry_hash,
                   uint32_t* d_entry_box, uint32_t* d_entry_ordinal, uint32_t entry_capacity, uint32_t* d_pair_first,
                   uint32_t* d_pair_second, uint32_t pair_capacity, uint32_t* d_counts, void* d_scratch,
                   size_t scratch_bytes) {
    using namespace dyn;
    if ((boxes && !d_boxes) || !d_entry_hash || !d_entry_box || !d_entry_ordinal || !d_pair_first ||
        !d_pair_second || !d_counts || !d_scratch) {
        return fail(DYN_E_INVALID, "dyn_fixed_grid: null argument");
    }
    if (!(cell_size > 0.0f)) return fail(DYN_E_INVALID, "dyn_fixed_grid: cell_size must be positive");
    if (entry_capacity > kSortMaxCapacity || pair_capacity > kSortMaxCapacity) {
        return fail(DYN_E_INVALID, "dyn_fixed_grid: capacities above 2^30-1");
    }
    if (scratch_bytes < dyn_fixed_grid_scratch_bytes(boxes, entry_capacity, pair_capacity)) {
        return fail(DYN_E_CAPACITY, "dyn_fixed_grid: scratch smaller than dyn_fixed_grid_scratch_bytes");
    }
    FixedScratch s;
    fixed_layout(boxes, entry_capacity, pair_capacity, d_scratch, &s);
    // the two clamped lengths the sorts read live past the layout (the +256 above)
    uint32_t* sort_counts = reinterpret_cast<uint32_t*>(static_cast<char*>(d_scratch) +
                                                        fixed_layout(boxes, entry_capacity, pair_capacity, nullptr, nullptr));
    hipStream_t st = static_cast<hipStream_t>(stream);
    const uint32_t blocks = ceil_div(boxes ? boxes : 1u, kFixedBlock);
    {
        ProfileScope scope("k_fixed_reset", st);
        k_fixed_reset<<<1, kFixedThreads, 0, st>>>(d_counts);
    }
    DYN_LAUNCHED("k_fixed_reset");
    if (boxes) {
        {
            ProfileScope scope("k_fixed_count", st);
            k_fixed_count<<<blocks, kFixedThreads, 0, st>>>(d_boxes, boxes, cell_size, s.counts, s.block_sums, d_counts);
        }
        DYN_LAUNCHED("k_fixed_count");
        {
            ProfileScope scope("k_fixed_scan_sums", st);
            k_fixed_scan_sums<<<1, kFixedThreads, 0, st>>>(s.block_sums, blocks, d_counts);
        }
        DYN_LAUNCHED("k_fixed_scan_sums");
        {
            ProfileScope scope("k_fixed_emit", st);
            k_fixed_emit<<<blocks, kFixedThreads, 0, st>>>(d_boxes, boxes, cell_size, s.counts, s.block_sums,
                                                            entry_capacity, d_entry_hash, d_entry_box, d_entry_ordinal);
        }
        DYN_LAUNCHED("k_fixed_emit");
    }
    {
        ProfileScope scope("k_fixed_clamp", st);
        k_fixed_clamp<<<1, 64, 0, st>>>(d_counts, entry_capacity, pair_capacity, sort_counts);
    }
    DYN_LAUNCHED("k_fixed_clamp");
    const uint32_t box_words = dyn_sort_key_words(boxes ? boxes : 1u);
    // entries by (cell hash, box index); the cell ordinal rides as the payload
    DYN_TRY(dyn_sort_pairs(st, d_entry_hash, d_entry_box, d_entry_ordinal, entry_capacity, &sort_counts[0], nullptr, 1u,
                           4u, box_words, s.entry_sort, s.entry_sort_bytes));
    {
        ProfileScope scope("k_fixed_pairs", st);
        k_fixed_pairs<<<ceil_div(entry_capacity ? entry_capacity : 1u, kFixedThreads), kFixedThreads, 0, st>>>(
            d_boxes, cell_size, d_entry_hash, d_entry_box, d_entry_ordinal, d_counts, entry_capacity, d_pair_first,
            d_pair_second, pair_capacity);
    }
    DYN_LAUNCHED("k_fixed_pairs");
    {
        ProfileScope scope("k_fixed_clamp", st);
        k_fixed_clamp<<<1, 64, 0, st>>>(d_counts, entry_capacity, pair_capacity, sort_counts);
    }
    DYN_LAUNCHED("k_fixed_clamp");
    // canonical pair order: (first, second) lexicographic
    DYN_TRY(dyn_sort_pairs(st, d_pair_first, d_pair_second, s.pair_payload, pair_capacity, &sort_counts[1], nullptr, 1u,
                           box_words, box_words, s.pair_sort, s.pair_sort_bytes));
    return DYN_OK;
}

}  // extern "C"
