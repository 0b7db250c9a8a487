// Device-side grid arithmetic of the hierarchical hashed grid. The semantics are the
// reference's (crates/dynamis-shader/shaders/grid_index.wgsl:28-115, core.wgsl:109-130); every
// float operation is IEEE-rounded (division and floor), so keys and cells are bit-exact.
#pragma once

#include "common.h"

namespace dyn {

constexpr uint32_t kGridLevels = DYN_GRID_LEVEL_LIMIT + 1;  // levels 0..30
constexpr uint32_t kDirBits = 16;                          // directory buckets per level: 2^16
constexpr uint32_t kDirBuckets = 1u << kDirBits;
constexpr uint32_t kDirStride = kDirBuckets + 1;           // +1 end sentinel per level

// WGSL f32 -> i32: round toward zero, saturate, NaN -> 0.
__device__ __forceinline__ int32_t to_i32(float f) {
    if (!(f == f)) return 0;
    if (f >= 2147483648.0f) return 2147483647;
    if (f < -2147483648.0f) return INT32_MIN;
    return static_cast<int32_t>(f);
}

// grid_index.wgsl:28-33
__device__ __forceinline__ uint32_t cell_hash(int32_t cx, int32_t cy, int32_t cz) {
    const uint32_t x = static_cast<uint32_t>(cx) * 0x9E3779B9u;
    const uint32_t y = static_cast<uint32_t>(cy) * 0x85EBCA77u;
    const uint32_t z = static_cast<uint32_t>(cz) * 0xC2B2AE3Du;
    return (x ^ y ^ z ^ (x << 7) ^ (y >> 3) ^ (z << 11)) & DYN_CELL_HASH_MASK;
}

// grid_index.wgsl:35-37
__device__ __forceinline__ uint32_t cell_key(uint32_t level, int32_t x, int32_t y, int32_t z) {
    return (level << DYN_LEVEL_KEY_SHIFT) | cell_hash(x, y, z);
}

// The hash salt of a world of a batch: its cells' keys are the single-world keys with the hash
// bits XORed by it, so worlds sharing space fall in different hash buckets. World 0 (and a
// single world) has salt 0, the reference's keys.
__device__ __forceinline__ uint32_t world_salt(uint32_t world) {
    const uint32_t h = world * 0x27D4EB2Fu;
    return (h ^ (h >> 15)) & DYN_CELL_HASH_MASK;
}

// grid_index.wgsl:39-41
__device__ __forceinline__ float level_cell_size(uint32_t level, float base) {
    return base * static_cast<float>(1u << level);
}

// core.wgsl:121-126
__device__ __forceinline__ float grid_cell_size(uint32_t scale_bits, uint32_t extent_bits) {
    const float scale = __uint_as_float(scale_bits);
    const float extent = __uint_as_float(extent_bits);
    const float finest = scale > 0.0f ? 1.0f / scale : DYN_GRID_SCALE_LIMIT;
    const float coarse = extent * 9.313225746154785e-10f;  // exp2(-30), exact
    return fmaxf(finest, coarse);
}

// core.wgsl:117-119
__device__ __forceinline__ float grid_extent_of(float3 mn, float3 mx) {
    return fmaxf(fmaxf(mx.x - mn.x, mx.y - mn.y), mx.z - mn.z);
}

// core.wgsl:109-115
__device__ __forceinline__ float grid_scale_of(float3 mn, float3 mx) {
    const float extent = grid_extent_of(mn, mx);
    if (extent <= 0.0f) return 0.0f;
    return fminf(fmaxf(1.0f / extent, 1.0f / DYN_GRID_SCALE_LIMIT), DYN_GRID_SCALE_LIMIT);
}

struct Cells {
    int32_t mn[3];
    int32_t mx[3];
};

// grid_index.wgsl:47-52
__device__ __forceinline__ Cells grid_cells(float3 mn, float3 mx, float cell) {
    Cells c;
    c.mn[0] = to_i32(floorf(mn.x / cell));
    c.mn[1] = to_i32(floorf(mn.y / cell));
    c.mn[2] = to_i32(floorf(mn.z / cell));
    c.mx[0] = to_i32(floorf(mx.x / cell));
    c.mx[1] = to_i32(floorf(mx.y / cell));
    c.mx[2] = to_i32(floorf(mx.z / cell));
    return c;
}

__device__ __forceinline__ uint32_t span_of(const Cells& c, int a) {
    return static_cast<uint32_t>(c.mx[a] - c.mn[a] + 1);
}

// grid_index.wgsl:58-61
__device__ __forceinline__ uint32_t cell_count(const Cells& c) {
    return span_of(c, 0) * span_of(c, 1) * span_of(c, 2);
}

// grid_index.wgsl:63-72 (x-major, then y, then z)
__device__ __forceinline__ void cell_at(const Cells& c, uint32_t ordinal, int32_t out[3]) {
    const uint32_t sy = span_of(c, 1), sz = span_of(c, 2);
    const uint32_t rows = sy * sz;
    out[0] = c.mn[0] + static_cast<int32_t>(ordinal / rows);
    out[1] = c.mn[1] + static_cast<int32_t>((ordinal / sz) % sy);
    out[2] = c.mn[2] + static_cast<int32_t>(ordinal % sz);
}

// grid_index.wgsl:74-77
__device__ __forceinline__ uint32_t cell_offset(const Cells& c, const int32_t coord[3]) {
    return static_cast<uint32_t>(coord[0] - c.mn[0]) |
           (static_cast<uint32_t>(coord[1] - c.mn[1]) << 1) |
           (static_cast<uint32_t>(coord[2] - c.mn[2]) << 2);
}

// grid_index.wgsl:79-89: the finest level at which the box spans at most 2 cells per axis.
__device__ __forceinline__ uint32_t shape_levels(float3 mn, float3 mx, float base) {
    uint32_t level = 0;
    for (;;) {
        const Cells c = grid_cells(mn, mx, level_cell_size(level, base));
        const uint32_t m = max(max(span_of(c, 0), span_of(c, 1)), span_of(c, 2));
        if (m <= DYN_MAX_CELLS_PER_AXIS || level >= DYN_GRID_LEVEL_LIMIT) break;
        ++level;
    }
    return level;
}

// grid_index.wgsl:225-229
__device__ __forceinline__ void entry_cell(float3 mn, uint32_t info, float cell, int32_t out[3]) {
    const uint32_t bits = (info & DYN_ENTRY_CELL_MASK) >> DYN_ENTRY_CELL_SHIFT;
    out[0] = to_i32(floorf(mn.x / cell)) + static_cast<int32_t>(bits & 1u);
    out[1] = to_i32(floorf(mn.y / cell)) + static_cast<int32_t>((bits >> 1) & 1u);
    out[2] = to_i32(floorf(mn.z / cell)) + static_cast<int32_t>((bits >> 2) & 1u);
}

// core.wgsl:128-130
__device__ __forceinline__ bool aabb_overlaps(float3 amn, float3 amx, float3 bmn, float3 bmx) {
    return amn.x <= bmx.x && amn.y <= bmx.y && amn.z <= bmx.z && bmn.x <= amx.x &&
           bmn.y <= amx.y && bmn.z <= amx.z;
}

}  // namespace dyn
