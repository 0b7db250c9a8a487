// Host construction of the support cells of hull shape sources (support_cells.h): per source and
// cube-map cell, the ascending list of vertices that can be the support vertex of a direction in
// the cell. Built once per shape upload, in FP64, over the sources in parallel.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "common.h"
#include "support_cells.h"

namespace {

using dyn::kSupportCells;
using dyn::kSupportInline;
using dyn::kSupportMaxList;
using dyn::kSupportNoCell;
using dyn::kSupportFaceCells;

struct Dir {
    double d[3];
};

// The direction of face coordinates (u, v) on a face: the face's axis component +-1, the other
// two components u and v in (x, y, z) order (the mapping of support_cell_of).
Dir face_dir(uint32_t face, double u, double v) {
    const uint32_t axis = face / 2u;
    const double sign = (face & 1u) ? -1.0 : 1.0;
    Dir r{};
    r.d[axis] = sign;
    const uint32_t a = axis == 0u ? 1u : 0u;
    const uint32_t b = axis == 2u ? 1u : 2u;
    r.d[a] = u;
    r.d[b] = v;
    return r;
}

double dotd(const float* v, const Dir& d) {
    return static_cast<double>(v[0]) * d.d[0] + static_cast<double>(v[1]) * d.d[1] + static_cast<double>(v[2]) * d.d[2];
}

uint32_t argmax(const float* verts, uint32_t count, const Dir& d) {
    uint32_t best = 0;
    double value = dotd(verts, d);
    for (uint32_t i = 1; i < count; ++i) {
        const double s = dotd(verts + 4u * i, d);
        if (s > value) {
            value = s;
            best = i;
        }
    }
    return best;
}

// The candidate lists of one source: for each cell, the vertices no sampled support vertex
// exceeds by more than the margin at all four corners of the widened cell.
void build_source(const float* verts, uint32_t count, std::vector<std::vector<uint16_t>>* lists) {
    double reach = 0.0;  // the largest sum of |coordinates| of a vertex
    for (uint32_t i = 0; i < count; ++i) {
        const float* v = verts + 4u * i;
        reach = std::max(reach, std::fabs(static_cast<double>(v[0])) + std::fabs(static_cast<double>(v[1])) +
                                    std::fabs(static_cast<double>(v[2])));
    }
    const double margin = dyn::kSupportMargin * reach;
    const double width = 2.0 / kSupportFaceCells;
    lists->assign(kSupportCells, {});
    std::vector<uint32_t> sampled;
    for (uint32_t face = 0; face < 6u; ++face) {
        for (uint32_t iu = 0; iu < kSupportFaceCells; ++iu) {
            for (uint32_t iv = 0; iv < kSupportFaceCells; ++iv) {
                const double u0 = -1.0 + iu * width - dyn::kSupportWiden, u1 = -1.0 + (iu + 1u) * width + dyn::kSupportWiden;
                const double v0 = -1.0 + iv * width - dyn::kSupportWiden, v1 = -1.0 + (iv + 1u) * width + dyn::kSupportWiden;
                const Dir corners[4] = {face_dir(face, u0, v0), face_dir(face, u0, v1), face_dir(face, u1, v0),
                                        face_dir(face, u1, v1)};
                sampled.clear();
                for (int su = 0; su <= 2; ++su) {
                    for (int sv = 0; sv <= 2; ++sv) {
                        const Dir d = face_dir(face, u0 + 0.5 * su * (u1 - u0), v0 + 0.5 * sv * (v1 - v0));
                        const uint32_t w = argmax(verts, count, d);
                        if (std::find(sampled.begin(), sampled.end(), w) == sampled.end()) sampled.push_back(w);
                    }
                }
                std::vector<uint16_t>& out = (*lists)[(face * kSupportFaceCells + iu) * kSupportFaceCells + iv];
                for (uint32_t i = 0; i < count; ++i) {
                    bool dominated = false;
                    for (uint32_t w : sampled) {
                        if (w == i) continue;
                        bool all = true;
                        for (const Dir& c : corners) {
                            if (!(dotd(verts + 4u * w, c) - dotd(verts + 4u * i, c) > margin)) {
                                all = false;
                                break;
                            }
                        }
                        if (all) {
                            dominated = true;
                            break;
                        }
                    }
                    if (!dominated) out.push_back(static_cast<uint16_t>(i));
                }
            }
        }
    }
}

}  // namespace

extern "C" {

int dyn_hull_support_cells(const dyn_shape_source* sources, uint32_t source_count, const float* vertices,
                           uint32_t vertex_count, uint32_t* cells, uint16_t* candidates, uint32_t capacity,
                           uint32_t* candidate_count) {
    using namespace dyn;
    if ((source_count && (!sources || !cells)) || (vertex_count && !vertices) || !candidate_count) {
        return fail(DYN_E_INVALID, "dyn_hull_support_cells: null argument");
    }
    std::vector<std::vector<std::vector<uint16_t>>> all(source_count);
    std::vector<char> tabled(source_count, 0);
    for (uint32_t s = 0; s < source_count; ++s) {
        const dyn_shape_source& src = sources[s];
        if (static_cast<uint64_t>(src.vertex_offset) + src.vertex_count > vertex_count) {
            return fail(DYN_E_INVALID, "dyn_hull_support_cells: a source reaches past the vertex array");
        }
        bool bounded = true;
        for (uint32_t i = 0; i < src.vertex_count && bounded; ++i) {
            for (int k = 0; k < 3; ++k) {
                const float x = vertices[4u * (static_cast<size_t>(src.vertex_offset) + i) + k];
                bounded = bounded && std::isfinite(x) && std::fabs(x) < 1e6f;
            }
        }
        tabled[s] = src.vertex_count >= kSupportMinVertices && src.vertex_count <= dyn::kSupportMaxVertices && bounded;
    }
    const uint32_t threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < threads; ++t) {
        pool.emplace_back([&, t]() {
            for (uint32_t s = t; s < source_count; s += threads) {
                if (tabled[s]) build_source(vertices + 4u * static_cast<size_t>(sources[s].vertex_offset),
                                            sources[s].vertex_count, &all[s]);
            }
        });
    }
    for (std::thread& th : pool) th.join();
    uint64_t total = 0;
    for (uint32_t s = 0; s < source_count; ++s) {
        for (uint32_t c = 0; c < kSupportCells; ++c) {
            uint32_t* cell = cells + 2u * (static_cast<size_t>(s) * kSupportCells + c);
            cell[0] = kSupportNoCell;
            cell[1] = 0u;
            if (!tabled[s]) continue;
            const std::vector<uint16_t>& list = all[s][c];
            if (list.empty() || list.size() > kSupportMaxList) continue;  // the linear scan
            if (total > 0xFFFFFFFFull - list.size()) return fail(DYN_E_CAPACITY, "dyn_hull_support_cells: above 2^32 candidates");
            uint32_t idx[4] = {0u, 0u, 0u, 0u};
            for (size_t k = 0; k < list.size() && k < kSupportInline; ++k) idx[k] = list[k];
            cell[0] = static_cast<uint32_t>(list.size()) | (idx[0] << 8) | (idx[1] << 20);
            cell[1] = list.size() <= kSupportInline ? (idx[2] | (idx[3] << 12)) : static_cast<uint32_t>(total);
            if (candidates && total + list.size() <= capacity) {
                std::copy(list.begin(), list.end(), candidates + total);
            }
            total += list.size();
        }
    }
    *candidate_count = static_cast<uint32_t>(total);
    if (candidates && total > capacity) {
        return fail(DYN_E_CAPACITY, "dyn_hull_support_cells: candidate capacity smaller than candidate_count");
    }
    return DYN_OK;
}

}  // extern "C"
