// Support cells of hull shape sources: for each cell of a cube map over local directions, the
// hull vertices that can be the support vertex of some direction in the cell. The device's hull
// support scans a direction's cell's candidates (in ascending vertex order) instead of every
// vertex, and returns the same vertex as the linear scan of convex.wgsl:1-85 (the first vertex of
// largest dot(v, d), the dot evaluated the same way).
//
// Why the result is the same: a vertex is left out of a cell only when another vertex exceeds it
// by more than kSupportMargin x (sum of |coordinate| over the hull's largest vertex) at every
// corner of the cell, widened by kSupportWiden (so everywhere in it, the difference of two dots
// being linear in the direction). The single-precision dot of a vertex is within 2u (1 + 2u) of
// that sum of the exact dot (u = 2^-24), relative to the direction's largest component, so
// every vertex that attains the largest single-precision dot stays a candidate, and the first of
// them in ascending order is the linear scan's answer. Directions whose largest component is
// outside [1e-30, 1e30] (zero, tiny, huge, or not finite) take the linear scan.
#pragma once

#include <stdint.h>

namespace dyn {

constexpr uint32_t kSupportFaceCells = 8;  // cells per face edge
constexpr uint32_t kSupportCells = 6u * kSupportFaceCells * kSupportFaceCells;
constexpr uint32_t kSupportMinVertices = 12;  // smaller hulls keep the linear scan
constexpr uint32_t kSupportMaxVertices = 4096;  // larger hulls keep it too
// A cell's record, two words: word 0 holds the candidate count in bits 0-7 (kSupportNoCell: the
// cell, or its whole source, keeps the linear scan) and the first two candidates' vertex indices in
// bits 8-19 and 20-31; word 1 the third and fourth in bits 0-11 and 12-23 when there are at most
// kSupportInline candidates, else the offset of the cell's ascending index list. Candidates are
// in ascending vertex order; vertex indices fit 12 bits (sources of at most 4096 vertices).
constexpr uint32_t kSupportNoCell = 0xFFu;
constexpr uint32_t kSupportInline = 4;
constexpr uint32_t kSupportMaxList = 254;  // longer lists keep the linear scan for their cell
constexpr double kSupportMargin = 1e-5;
constexpr double kSupportWiden = 1e-3;  // in face coordinates (a face spans [-1, 1])
constexpr float kSupportMinComponent = 1e-30f;
constexpr float kSupportMaxComponent = 1e30f;

// The cube-map cell of a direction: the face of its largest component (x before y before z on
// ties; +axis when that component is not negative), face coordinates (u, v) = the other two
// components in (x, y, z) order over the largest (with an approximate reciprocal: the cells are
// widened far beyond its error). Returns false when the direction takes the linear scan.
__device__ inline bool support_cell_of(float x, float y, float z, uint32_t* cell) {
    const float ax = fabsf(x), ay = fabsf(y), az = fabsf(z);
    uint32_t axis;
    float m, a, b, c;
    if (ax >= ay && ax >= az) {
        axis = 0u, m = ax, c = x, a = y, b = z;
    } else if (ay >= az) {
        axis = 1u, m = ay, c = y, a = x, b = z;
    } else {
        axis = 2u, m = az, c = z, a = x, b = y;
    }
    if (!(m >= kSupportMinComponent) || !(m <= kSupportMaxComponent)) return false;
    const uint32_t face = 2u * axis + (c < 0.0f ? 1u : 0u);
    const float half = static_cast<float>(kSupportFaceCells) * 0.5f;
    const float inv = __builtin_amdgcn_rcpf(m);
    int iu = static_cast<int>((a * inv + 1.0f) * half), iv = static_cast<int>((b * inv + 1.0f) * half);
    const int top = static_cast<int>(kSupportFaceCells) - 1;
    iu = iu < 0 ? 0 : (iu > top ? top : iu);
    iv = iv < 0 ? 0 : (iv > top ? top : iv);
    *cell = (face * kSupportFaceCells + static_cast<uint32_t>(iu)) * kSupportFaceCells + static_cast<uint32_t>(iv);
    return true;
}

}  // namespace dyn
