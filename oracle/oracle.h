/*
 * oracle.h - CPU restatement of the reference's rigid-body step algorithms.
 *
 * TEST INFRASTRUCTURE ONLY. The product (libdynamis_amd.so) never links, loads or calls this
 * code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and only as
 * the checker or as the timed CPU baseline. Every function names the reference file:line it
 * follows. Plain C11, serial, compiled with -ffp-contract=off so float expressions round in the
 * order they are written.
 *
 * Pinning (what each function is checked against besides the HIP path):
 *   oracle_sort_pairs   - numpy stable lexsort on the masked key bytes, and the literal cases
 *                         of crates/dynamis-sort/tests/gpu.rs (golden fixture)
 *   oracle_cell_hash    - known-answer values worked from grid_index.wgsl:28-33 by hand-checked
 *                         big-integer arithmetic in the golden fixture
 *   oracle_broadphase   - brute-force O(K^2) overlap pairs (the canonical pair set), and the
 *                         pair counts of crates/dynamis-world/tests/suite/broadphase.rs
 */
#ifndef DYN_ORACLE_H
#define DYN_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/dyn_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- sort (crates/dynamis-sort/src/radix.rs) ---- */
uint32_t oracle_units_for(uint32_t live);
uint32_t oracle_key_words(uint32_t elements);
/* Returns 0, or -1 when no digit is requested (the reference asserts). */
int oracle_sort_pairs(uint32_t* major, uint32_t* minor, uint32_t* payload, uint32_t capacity,
                      uint32_t count, uint32_t base, uint32_t major_words, uint32_t minor_words);

/* ---- grid (crates/dynamis-shader/shaders/grid_index.wgsl, core.wgsl) ---- */
uint32_t oracle_cell_hash(int32_t x, int32_t y, int32_t z);
uint32_t oracle_cell_key(uint32_t level, int32_t x, int32_t y, int32_t z);
float oracle_grid_cell_size(uint32_t scale_bits, uint32_t extent_bits);
uint32_t oracle_grid_scale_bits(const dyn_aabb* box);
uint32_t oracle_grid_extent_bits(const dyn_aabb* box);
uint32_t oracle_shape_levels(const dyn_aabb* box, float base);

/* One broadphase pass of the awake region over K boxes (see dynamis.h dyn_broadphase).
 * Outputs: entries in emission order (keys, order = slot, records), then sorted; pairs in
 * emission order, then canonicalised when sort_pairs != 0. counters has
 * DYN_COUNTER_BLOCK_WORDS words. Returns the number of entries emitted (<= entry_capacity). */
uint32_t oracle_broadphase(const dyn_aabb* aabbs, const uint32_t* groups, const uint32_t* flags,
                           const dyn_aabb* resolution, uint32_t colliders,
                           uint32_t* entry_keys, uint32_t* entry_order, dyn_grid_entry* entries,
                           uint32_t entry_capacity, uint32_t* pair_major, uint32_t* pair_minor,
                           uint32_t pair_capacity, uint32_t* counters, uint32_t sort_pairs);

#ifdef __cplusplus
}
#endif

#endif
