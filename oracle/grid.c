/*
 * grid.c - CPU restatement of the reference's hierarchical hashed-grid broadphase.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows, function by function:
 *   cell_hash / cell_key         crates/dynamis-shader/shaders/grid_index.wgsl:28-37
 *   level_cell_size              grid_index.wgsl:39-41
 *   grid_cells / span / count    grid_index.wgsl:47-61
 *   grid_cell_at / offset        grid_index.wgsl:63-77
 *   shape_levels                 grid_index.wgsl:79-89
 *   grid_coarser_levels          grid_index.wgsl:109-111
 *   grid_overlap_cell            grid_index.wgsl:113-115
 *   entry_cell                   grid_index.wgsl:225-229
 *   entry_bounds_in              grid_index.wgsl:231-253 (binary search of a key's run)
 *   grid_scale_of / extent       crates/dynamis-shader/shaders/core.wgsl:109-119
 *   grid_cell_size               core.wgsl:121-126
 *   aabb_overlaps                core.wgsl:128-130
 *   record_grid_resolution       crates/dynamis-rigid/shaders/broadphase_aabb.wgsl:21-24
 *   entry emission               crates/dynamis-rigid/shaders/collider_entries.wgsl:9-35,
 *                                collider_emit.wgsl:6-63
 *   region sort                  crates/dynamis-broadphase/src/passes.rs:76-89 (4 key bytes,
 *                                payload = entry_order)
 *   level links                  crates/dynamis-broadphase/shaders/level_links.wgsl:6-65
 *   cell mates                   crates/dynamis-broadphase/shaders/cell_pairs.wgsl:6-50
 *   pair emission + filter       crates/dynamis-broadphase/shaders/pair_emit.wgsl:3-31
 *   pair sort                    crates/dynamis-rigid/src/narrowphase.rs:116-125 (key_words of
 *                                the collider count on both words, src/lib.rs:75)
 * The stage models one awake region holding every collider (no resting or immovable entries),
 * with entries appended in collider order.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static uint32_t f2u(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

static float u2f(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

/* WGSL f32 -> i32 conversion: round toward zero, saturate at the range ends. */
static int32_t to_i32(float f) {
    if (f != f) return 0;
    if (f >= 2147483648.0f) return 2147483647;
    if (f < -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)f;
}

uint32_t oracle_cell_hash(int32_t cx, int32_t cy, int32_t cz) {
    uint32_t x = (uint32_t)cx * 0x9E3779B9u;
    uint32_t y = (uint32_t)cy * 0x85EBCA77u;
    uint32_t z = (uint32_t)cz * 0xC2B2AE3Du;
    return (x ^ y ^ z ^ (x << 7) ^ (y >> 3) ^ (z << 11)) & DYN_CELL_HASH_MASK;
}

uint32_t oracle_cell_key(uint32_t level, int32_t x, int32_t y, int32_t z) {
    return (level << DYN_LEVEL_KEY_SHIFT) | oracle_cell_hash(x, y, z);
}

static float level_cell_size(uint32_t level, float cell_size) {
    return cell_size * (float)(1u << level);
}

float oracle_grid_cell_size(uint32_t scale_bits, uint32_t extent_bits) {
    float scale = u2f(scale_bits);
    float extent = u2f(extent_bits);
    float finest = scale > 0.0f ? 1.0f / scale : DYN_GRID_SCALE_LIMIT;
    float coarse = extent * ldexpf(1.0f, -(int)DYN_GRID_LEVEL_LIMIT);
    return finest > coarse ? finest : coarse;
}

static float grid_extent_of(const dyn_aabb* b) {
    float dx = b->max[0] - b->min[0];
    float dy = b->max[1] - b->min[1];
    float dz = b->max[2] - b->min[2];
    float m = dx > dy ? dx : dy;
    return m > dz ? m : dz;
}

uint32_t oracle_grid_scale_bits(const dyn_aabb* box) {
    float extent = grid_extent_of(box);
    if (extent <= 0.0f) return f2u(0.0f);
    float s = 1.0f / extent;
    if (s < 1.0f / DYN_GRID_SCALE_LIMIT) s = 1.0f / DYN_GRID_SCALE_LIMIT;
    if (s > DYN_GRID_SCALE_LIMIT) s = DYN_GRID_SCALE_LIMIT;
    return f2u(s);
}

uint32_t oracle_grid_extent_bits(const dyn_aabb* box) {
    float extent = grid_extent_of(box);
    return f2u(extent > 0.0f ? extent : 0.0f);
}

typedef struct {
    int32_t min[3];
    int32_t max[3];
} cells_t;

static cells_t grid_cells(const float* mn, const float* mx, float cell) {
    cells_t c;
    for (int a = 0; a < 3; a++) {
        c.min[a] = to_i32(floorf(mn[a] / cell));
        c.max[a] = to_i32(floorf(mx[a] / cell));
    }
    return c;
}

static void cell_span(const cells_t* c, uint32_t span[3]) {
    for (int a = 0; a < 3; a++) span[a] = (uint32_t)(c->max[a] - c->min[a] + 1);
}

static uint32_t cell_count(const cells_t* c) {
    uint32_t s[3];
    cell_span(c, s);
    return s[0] * s[1] * s[2];
}

static void cell_at(const cells_t* c, uint32_t ordinal, int32_t out[3]) {
    uint32_t s[3];
    cell_span(c, s);
    uint32_t rows = s[1] * s[2];
    out[0] = c->min[0] + (int32_t)(ordinal / rows);
    out[1] = c->min[1] + (int32_t)((ordinal / s[2]) % s[1]);
    out[2] = c->min[2] + (int32_t)(ordinal % s[2]);
}

static uint32_t cell_offset(const cells_t* c, const int32_t coord[3]) {
    return (uint32_t)(coord[0] - c->min[0]) | ((uint32_t)(coord[1] - c->min[1]) << 1) |
           ((uint32_t)(coord[2] - c->min[2]) << 2);
}

static uint32_t levels_of(const float* mn, const float* mx, float base) {
    uint32_t level = 0;
    for (;;) {
        cells_t c = grid_cells(mn, mx, level_cell_size(level, base));
        uint32_t s[3];
        cell_span(&c, s);
        uint32_t m = s[0] > s[1] ? s[0] : s[1];
        m = m > s[2] ? m : s[2];
        if (m <= DYN_MAX_CELLS_PER_AXIS || level >= DYN_GRID_LEVEL_LIMIT) break;
        level++;
    }
    return level;
}

uint32_t oracle_shape_levels(const dyn_aabb* box, float base) {
    return levels_of(box->min, box->max, base);
}

static int overlaps(const dyn_grid_entry* a, const dyn_grid_entry* b) {
    for (int k = 0; k < 3; k++) {
        if (!(a->min[k] <= b->max[k])) return 0;
        if (!(b->min[k] <= a->max[k])) return 0;
    }
    return 1;
}

static void entry_cell(const dyn_grid_entry* e, float cell, int32_t out[3]) {
    uint32_t bits = (e->info & DYN_ENTRY_CELL_MASK) >> DYN_ENTRY_CELL_SHIFT;
    out[0] = to_i32(floorf(e->min[0] / cell)) + (int32_t)(bits & 1u);
    out[1] = to_i32(floorf(e->min[1] / cell)) + (int32_t)((bits >> 1) & 1u);
    out[2] = to_i32(floorf(e->min[2] / cell)) + (int32_t)((bits >> 2) & 1u);
}

static int same_cell(const int32_t a[3], const int32_t b[3]) {
    return a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
}

static void overlap_cell(const dyn_grid_entry* a, const dyn_grid_entry* b, float cell,
                         int32_t out[3]) {
    for (int k = 0; k < 3; k++) {
        float m = a->min[k] > b->min[k] ? a->min[k] : b->min[k];
        out[k] = to_i32(floorf(m / cell));
    }
}

typedef struct {
    const uint32_t* keys;
    const uint32_t* order;
    const dyn_grid_entry* entries;
    uint32_t live;
    uint32_t* counters;
    uint32_t* pair_major;
    uint32_t* pair_minor;
    uint32_t pair_capacity;
} pass_t;

static uint32_t* ctr(uint32_t* counters, uint32_t slot) {
    return &counters[slot * DYN_COUNTER_STRIDE_WORDS];
}

static void key_run(const pass_t* p, uint32_t key, uint32_t* first, uint32_t* end) {
    uint32_t lo = 0, hi = p->live;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2u;
        if (p->keys[mid] < key) lo = mid + 1u; else hi = mid;
    }
    *first = lo;
    hi = p->live;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2u;
        if (p->keys[mid] <= key) lo = mid + 1u; else hi = mid;
    }
    *end = lo;
}

static void emit_pair(pass_t* p, uint32_t first, uint32_t second) {
    const dyn_grid_entry* a = &p->entries[first];
    const dyn_grid_entry* b = &p->entries[second];
    uint32_t ka = (a->info & DYN_ENTRY_KIND_MASK) >> DYN_ENTRY_KIND_SHIFT;
    uint32_t kb = (b->info & DYN_ENTRY_KIND_MASK) >> DYN_ENTRY_KIND_SHIFT;
    if (ka != DYN_ENTRY_KIND_COLLIDER || kb != DYN_ENTRY_KIND_COLLIDER) return;
    if (first == second || a->group == b->group) return;
    if (!(a->info & DYN_ENTRY_MOBILE) && !(b->info & DYN_ENTRY_MOBILE)) return;
    uint32_t slot = (*ctr(p->counters, DYN_COUNTER_PAIRS))++;
    if (slot < p->pair_capacity) {
        uint32_t ia = a->info & DYN_ENTRY_INDEX_MASK;
        uint32_t ib = b->info & DYN_ENTRY_INDEX_MASK;
        p->pair_minor[slot] = ia > ib ? ia : ib;
        p->pair_major[slot] = ia < ib ? ia : ib;
    } else {
        (*ctr(p->counters, DYN_COUNTER_REFUSED_PAIRS))++;
    }
}

static int entry_awake(const pass_t* p, uint32_t node) {
    return (p->entries[node].info & DYN_ENTRY_AWAKE) != 0u;
}

static void link_level(pass_t* p, uint32_t node, uint32_t level, int awake, float grid) {
    const dyn_grid_entry* box = &p->entries[node];
    float cs = level_cell_size(level, grid);
    cells_t cells = grid_cells(box->min, box->max, cs);
    uint32_t count = cell_count(&cells);
    for (uint32_t ordinal = 0; ordinal < count; ordinal++) {
        int32_t cell[3];
        cell_at(&cells, ordinal, cell);
        uint32_t first, end;
        key_run(p, oracle_cell_key(level, cell[0], cell[1], cell[2]), &first, &end);
        for (uint32_t e = first; e < end; e++) {
            uint32_t other = p->order[e];
            if (other == node) continue;
            if (!awake && !entry_awake(p, other)) continue;
            const dyn_grid_entry* ob = &p->entries[other];
            if (!overlaps(box, ob)) continue;
            int32_t oc[3], mc[3];
            entry_cell(ob, cs, oc);
            if (!same_cell(oc, cell)) continue;
            overlap_cell(box, ob, cs, mc);
            if (!same_cell(mc, cell)) continue;
            emit_pair(p, node, other);
        }
    }
}

static void level_links(pass_t* p, uint32_t index, float grid, uint32_t occupied) {
    uint32_t node = p->order[index];
    uint32_t info = p->entries[node].info;
    if (((info & DYN_ENTRY_KIND_MASK) >> DYN_ENTRY_KIND_SHIFT) != DYN_ENTRY_KIND_COLLIDER) return;
    if (!(info & DYN_ENTRY_PRIMARY)) return;
    int awake = entry_awake(p, node);
    if (!awake && *ctr(p->counters, DYN_COUNTER_COARSE_ACTIVE) == 0u) return;
    uint32_t level = levels_of(p->entries[node].min, p->entries[node].max, grid);
    uint32_t shift = level + 1u < 31u ? level + 1u : 31u;
    uint32_t coarser = occupied & ~((1u << shift) - 1u);
    while (coarser != 0u) {
        uint32_t link = 31u - (uint32_t)__builtin_clz(coarser);
        coarser &= ~(1u << link);
        link_level(p, node, link, awake, grid);
    }
}

static void cell_pairs(pass_t* p, uint32_t index, float grid) {
    uint32_t node = p->order[index];
    uint32_t info = p->entries[node].info;
    if (((info & DYN_ENTRY_KIND_MASK) >> DYN_ENTRY_KIND_SHIFT) != DYN_ENTRY_KIND_COLLIDER) return;
    if (!entry_awake(p, node)) return;
    const dyn_grid_entry* box = &p->entries[node];
    uint32_t level = levels_of(box->min, box->max, grid);
    float cs = level_cell_size(level, grid);
    int32_t cell[3];
    entry_cell(box, cs, cell);
    uint32_t first, end;
    key_run(p, oracle_cell_key(level, cell[0], cell[1], cell[2]), &first, &end);
    for (uint32_t e = first; e < end; e++) {
        if (e == index) continue;
        uint32_t other = p->order[e];
        if (entry_awake(p, other) && e < index) continue;
        const dyn_grid_entry* ob = &p->entries[other];
        if (!overlaps(box, ob)) continue;
        int32_t oc[3], mc[3];
        entry_cell(ob, cs, oc);
        if (!same_cell(oc, cell)) continue;
        overlap_cell(box, ob, cs, mc);
        if (!same_cell(mc, cell)) continue;
        emit_pair(p, node, other);
    }
}

uint32_t oracle_broadphase(const dyn_aabb* aabbs, const uint32_t* groups, const uint32_t* flags,
                           const dyn_aabb* resolution, uint32_t colliders,
                           uint32_t* entry_keys, uint32_t* entry_order, dyn_grid_entry* entries,
                           uint32_t entry_capacity, uint32_t* pair_major, uint32_t* pair_minor,
                           uint32_t pair_capacity, uint32_t* counters, uint32_t sort_pairs) {
    memset(counters, 0, sizeof(uint32_t) * DYN_COUNTER_BLOCK_WORDS);
    /* broadphase_aabb.wgsl:21-24 */
    uint32_t scale = 0, extent = 0;
    for (uint32_t i = 0; i < colliders; i++) {
        const dyn_aabb* r = resolution ? &resolution[i] : &aabbs[i];
        uint32_t s = oracle_grid_scale_bits(r);
        uint32_t e = oracle_grid_extent_bits(r);
        if (s > scale) scale = s;
        if (e > extent) extent = e;
    }
    *ctr(counters, DYN_COUNTER_GRID_SCALE) = scale;
    *ctr(counters, DYN_COUNTER_GRID_EXTENT) = extent;
    float base = oracle_grid_cell_size(scale, extent);

    /* collider_entries.wgsl:9-35 + collider_emit.wgsl:6-63, appended in collider order */
    for (uint32_t i = 0; i < colliders; i++) {
        uint32_t f = flags ? flags[i] : (DYN_BP_MOBILE | DYN_BP_AWAKE);
        int awake = (f & DYN_BP_AWAKE) != 0u;
        const dyn_aabb* box = &aabbs[i];
        uint32_t level = levels_of(box->min, box->max, base);
        *ctr(counters, DYN_COUNTER_GRID_LEVELS) |= 1u << level;
        cells_t cells = grid_cells(box->min, box->max, level_cell_size(level, base));
        uint32_t count = cell_count(&cells);
        uint32_t emitted = count < DYN_MAX_CELLS_PER_COLLIDER ? count : DYN_MAX_CELLS_PER_COLLIDER;
        if (count > emitted) *ctr(counters, DYN_COUNTER_ENTRY_FAULTS) += count - emitted;
        if (level > 0u && awake) (*ctr(counters, DYN_COUNTER_COARSE_ACTIVE))++;
        for (uint32_t ordinal = 0; ordinal < emitted; ordinal++) {
            uint32_t slot = (*ctr(counters, DYN_COUNTER_ENTRIES))++;
            if (slot >= entry_capacity) {
                (*ctr(counters, DYN_COUNTER_ENTRY_FAULTS))++;
                continue;
            }
            int32_t coord[3];
            cell_at(&cells, ordinal, coord);
            uint32_t offset = cell_offset(&cells, coord);
            uint32_t info = i | (offset << DYN_ENTRY_CELL_SHIFT) |
                            (DYN_ENTRY_KIND_COLLIDER << DYN_ENTRY_KIND_SHIFT);
            if (awake) info |= DYN_ENTRY_AWAKE;
            if (f & DYN_BP_MOBILE) info |= DYN_ENTRY_MOBILE;
            if (offset == 0u) info |= DYN_ENTRY_PRIMARY;
            dyn_grid_entry* rec = &entries[slot];
            memcpy(rec->min, box->min, 12);
            memcpy(rec->max, box->max, 12);
            rec->group = groups ? groups[i] : i;
            rec->info = info;
            entry_keys[slot] = oracle_cell_key(level, coord[0], coord[1], coord[2]);
            entry_order[slot] = slot;
        }
    }
    uint32_t total = *ctr(counters, DYN_COUNTER_ENTRIES);
    uint32_t live = total < entry_capacity ? total : entry_capacity;

    /* passes.rs:76-89: sort the region's keys (4 bytes), payload = entry_order */
    uint32_t* dummy = (uint32_t*)calloc(live ? live : 1u, 4u);
    oracle_sort_pairs(entry_keys, dummy, entry_order, live, live, 0u, 4u, 0u);
    free(dummy);

    pass_t p = {entry_keys, entry_order, entries, live, counters, pair_major, pair_minor,
                pair_capacity};
    uint32_t occupied = *ctr(counters, DYN_COUNTER_GRID_LEVELS);
    for (uint32_t index = 0; index < live; index++) level_links(&p, index, base, occupied);
    for (uint32_t index = 0; index < live; index++) cell_pairs(&p, index, base);

    if (sort_pairs) {
        uint32_t pairs = *ctr(counters, DYN_COUNTER_PAIRS);
        uint32_t n = pairs < pair_capacity ? pairs : pair_capacity;
        uint32_t words = oracle_key_words(colliders > 0u ? colliders : 1u);
        uint32_t* payload = (uint32_t*)calloc(n ? n : 1u, 4u);
        oracle_sort_pairs(pair_major, pair_minor, payload, n, n, 0u, words, words);
        free(payload);
    }
    return live;
}
