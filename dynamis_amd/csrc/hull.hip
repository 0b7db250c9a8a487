// Host construction of convex hull shape sources: quickhull over a point set
// (crates/dynamis-hull/src/hull.rs:51-288, as used by World::add_hull_from_mesh,
// crates/dynamis-world/src/shape.rs:50-57) and the hull's solid geometry in FP64
// (crates/dynamis-hull/src/solid.rs:1-105). Host code only; FP32 and FP64 expressions keep the
// reference's operation order.
#include <array>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include "common.h"

namespace {

using P3 = std::array<float, 3>;

float dot3f(const P3& a, const P3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
P3 sub3f(const P3& a, const P3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
float len3f(const P3& a) { return std::sqrt(dot3f(a, a)); }

// hull.rs:267-288
P3 face_normal(const std::vector<P3>& pts, size_t a, size_t b, size_t c) {
    const P3 ab = sub3f(pts[b], pts[a]);
    const P3 ac = sub3f(pts[c], pts[a]);
    const P3 cr = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2],
                   ab[0] * ac[1] - ab[1] * ac[0]};
    const float len = len3f(cr);
    if (len < 1e-12f) return {0.0f, 0.0f, 0.0f};
    return {cr[0] / len, cr[1] / len, cr[2] / len};
}

struct Face {
    size_t a, b, c;
    P3 normal;
    float offset;
    std::vector<size_t> outside;
    bool dead = false;
    float distance(const P3& p) const { return dot3f(normal, p) - offset; }
};

Face make_face(size_t a, size_t b, size_t c, const std::vector<P3>& pts) {
    Face f;
    f.a = a;
    f.b = b;
    f.c = c;
    f.normal = face_normal(pts, a, b, c);
    f.offset = dot3f(f.normal, pts[a]);
    return f;
}

// hull.rs:195-201: a point joins the first live face that sees it
void assign_outside(const std::vector<P3>& pts, std::vector<Face>& faces, const std::vector<size_t>* restrict_to) {
    std::vector<size_t> candidates;
    if (restrict_to) {
        candidates = *restrict_to;
    } else {
        for (size_t i = 0; i < pts.size(); ++i) candidates.push_back(i);
    }
    for (size_t p : candidates) {
        for (Face& f : faces) {
            if (f.dead) continue;
            if (f.distance(pts[p]) > 1e-6f) {
                f.outside.push_back(p);
                break;
            }
        }
    }
}

// Rust's Iterator::max_by keeps the last of equal maxima; these helpers do the same.
template <typename Key>
size_t last_max(const std::vector<size_t>& items, Key key) {
    size_t best = items[0];
    float best_key = key(items[0]);
    for (size_t i = 1; i < items.size(); ++i) {
        const float k = key(items[i]);
        if (!(k < best_key)) {  // k >= best_key: a later equal maximum wins
            best = items[i];
            best_key = k;
        }
    }
    return best;
}

// hull.rs:88-168. Returns false when the points are degenerate (the reference asserts).
bool quickhull(const std::vector<P3>& pts, std::vector<size_t>* out_vertices,
               std::vector<std::array<size_t, 3>>* out_faces) {
    const size_t n = pts.size();
    if (n < 4) return false;
    // extreme_pair, hull.rs:224-240
    size_t x_min = 0, x_max = 0;
    for (size_t i = 0; i < n; ++i) {
        if (pts[i][0] < pts[x_min][0]) x_min = i;
        if (pts[i][0] > pts[x_max][0]) x_max = i;
    }
    size_t p0 = x_min, p1 = x_max;
    if (x_min == x_max) {
        p0 = 0;
        p1 = n - 1;
    }
    // farthest_from_line, hull.rs:242-263
    const P3 dir = sub3f(pts[p1], pts[p0]);
    const float length_sq = dot3f(dir, dir);
    size_t p2 = 0;
    float best_line = -1.0f;
    for (size_t i = 0; i < n; ++i) {
        const P3 delta = sub3f(pts[i], pts[p0]);
        const float projection = length_sq > 0.0f ? dot3f(delta, dir) / length_sq : 0.0f;
        const P3 closest = {pts[p0][0] + dir[0] * projection, pts[p0][1] + dir[1] * projection,
                            pts[p0][2] + dir[2] * projection};
        const float dist = len3f(sub3f(pts[i], closest));
        if (dist > best_line) {
            best_line = dist;
            p2 = i;
        }
    }
    if (!(best_line > 1e-6f)) return false;  // collinear
    // farthest_from_plane, hull.rs:265-279
    const P3 normal = face_normal(pts, p0, p1, p2);
    const float offset = dot3f(normal, pts[p0]);
    size_t p3 = 0;
    float best_plane = 0.0f;
    for (size_t i = 0; i < n; ++i) {
        const float dist = std::fabs(dot3f(normal, pts[i]) - offset);
        if (dist > best_plane) {
            best_plane = dist;
            p3 = i;
        }
    }
    if (!(best_plane > 1e-6f)) return false;  // coplanar
    // initial_tetra, hull.rs:203-222
    const P3 centroid = {(pts[p0][0] + pts[p1][0] + pts[p2][0] + pts[p3][0]) * 0.25f,
                         (pts[p0][1] + pts[p1][1] + pts[p2][1] + pts[p3][1]) * 0.25f,
                         (pts[p0][2] + pts[p1][2] + pts[p2][2] + pts[p3][2]) * 0.25f};
    std::vector<Face> faces;
    faces.push_back(make_face(p0, p1, p2, pts));
    faces.push_back(make_face(p0, p1, p3, pts));
    faces.push_back(make_face(p0, p2, p3, pts));
    faces.push_back(make_face(p1, p2, p3, pts));
    for (Face& f : faces) {
        if (dot3f(f.normal, centroid) - f.offset > 0.0f) {
            f.normal = {-f.normal[0], -f.normal[1], -f.normal[2]};
            f.offset = -f.offset;
            std::swap(f.b, f.c);
        }
    }
    assign_outside(pts, faces, nullptr);
    std::vector<size_t> indices;
    for (size_t i = 0; i < faces.size(); ++i) indices.push_back(i);
    while (!indices.empty()) {
        std::vector<size_t> live;
        for (size_t i : indices) {
            if (!faces[i].dead && !faces[i].outside.empty()) live.push_back(i);
        }
        if (live.empty()) break;
        const size_t seed = last_max(live, [&](size_t i) { return faces[i].distance(pts[faces[i].outside.front()]); });
        const size_t apex = last_max(faces[seed].outside, [&](size_t p) { return faces[seed].distance(pts[p]); });
        std::vector<size_t> visible;
        for (size_t i = 0; i < faces.size(); ++i) {
            if (!faces[i].dead && i != seed && faces[i].distance(pts[apex]) > 1e-6f) visible.push_back(i);
        }
        visible.push_back(seed);
        auto has_edge = [&](size_t s, size_t e) {
            for (size_t i : visible) {
                const Face& f = faces[i];
                if ((f.a == s && f.b == e) || (f.b == s && f.c == e) || (f.c == s && f.a == e)) return true;
            }
            return false;
        };
        std::vector<std::pair<size_t, size_t>> horizon;
        for (size_t i : visible) {
            const Face& f = faces[i];
            const std::pair<size_t, size_t> edges[3] = {{f.a, f.b}, {f.b, f.c}, {f.c, f.a}};
            for (const auto& e : edges) {
                if (!has_edge(e.second, e.first)) horizon.push_back(e);
            }
        }
        if (horizon.empty()) return false;
        std::vector<size_t> pending;
        for (size_t i : visible) {
            pending.insert(pending.end(), faces[i].outside.begin(), faces[i].outside.end());
            faces[i].dead = true;
        }
        pending.push_back(apex);
        for (const auto& e : horizon) faces.push_back(make_face(e.first, e.second, apex, pts));
        assign_outside(pts, faces, &pending);
        indices.clear();
        for (size_t i = 0; i < faces.size(); ++i) indices.push_back(i);
    }
    out_faces->clear();
    out_vertices->clear();
    for (const Face& f : faces) {
        if (f.dead) continue;
        out_faces->push_back({f.a, f.b, f.c});
    }
    for (const auto& f : *out_faces) {
        for (size_t v : f) {
            bool seen = false;
            for (size_t u : *out_vertices) seen = seen || u == v;
            if (!seen) out_vertices->push_back(v);
        }
    }
    return true;
}

}  // namespace

extern "C" {

int dyn_hull_build(const float* points, uint32_t count, float* vertices, uint32_t* vertex_count,
                   uint32_t* triangles, uint32_t* triangle_count, uint32_t capacity) {
    using dyn::fail;
    if (!points || !vertices || !vertex_count || !triangles || !triangle_count) {
        return fail(DYN_E_INVALID, "dyn_hull_build: null argument");
    }
    std::vector<P3> pts(count);
    for (uint32_t i = 0; i < count; ++i) {
        pts[i] = {points[3 * i], points[3 * i + 1], points[3 * i + 2]};
        if (!std::isfinite(pts[i][0]) || !std::isfinite(pts[i][1]) || !std::isfinite(pts[i][2])) {
            return fail(DYN_E_INVALID, "dyn_hull_build: a hull point must be finite");
        }
    }
    std::vector<size_t> hull_vertices;
    std::vector<std::array<size_t, 3>> hull_faces;
    if (!quickhull(pts, &hull_vertices, &hull_faces)) {
        return fail(DYN_E_INVALID, "dyn_hull_build: hull points must span a volume");
    }
    if (hull_vertices.size() > capacity || hull_faces.size() > capacity) {
        return fail(DYN_E_CAPACITY, "dyn_hull_build: output capacity too small");
    }
    // hull.rs:66-86: vertices in order of first use, triangles remapped
    std::vector<uint32_t> remap(count, 0xFFFFFFFFu);
    uint32_t nv = 0;
    for (size_t v : hull_vertices) {
        if (remap[v] == 0xFFFFFFFFu) {
            remap[v] = nv;
            vertices[3 * nv] = pts[v][0];
            vertices[3 * nv + 1] = pts[v][1];
            vertices[3 * nv + 2] = pts[v][2];
            ++nv;
        }
    }
    for (size_t t = 0; t < hull_faces.size(); ++t) {
        for (int c = 0; c < 3; ++c) triangles[3 * t + c] = remap[hull_faces[t][c]];
    }
    *vertex_count = nv;
    *triangle_count = static_cast<uint32_t>(hull_faces.size());
    return DYN_OK;
}

// solid.rs:62-105
int dyn_hull_solid(const float* vertices, uint32_t vertex_count, const uint32_t* triangles,
                   uint32_t triangle_count, float* volume_out, float* centroid_out,
                   float* second_out) {
    using dyn::fail;
    if (!vertices || !triangles || !volume_out || !centroid_out || !second_out) {
        return fail(DYN_E_INVALID, "dyn_hull_solid: null argument");
    }
    if (vertex_count == 0 || triangle_count == 0) {
        return fail(DYN_E_INVALID, "a hull must contain vertices and triangles");
    }
    double apex[3] = {0.0, 0.0, 0.0};
    for (uint32_t v = 0; v < vertex_count; ++v) {
        for (int a = 0; a < 3; ++a) apex[a] += static_cast<double>(vertices[3 * v + a]);
    }
    const double count = static_cast<double>(vertex_count);
    for (int a = 0; a < 3; ++a) apex[a] = apex[a] / count;
    auto cross = [](const double* a, const double* b, double* r) {
        r[0] = a[1] * b[2] - a[2] * b[1];
        r[1] = a[2] * b[0] - a[0] * b[2];
        r[2] = a[0] * b[1] - a[1] * b[0];
    };
    auto dot = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    auto squared = [](double ax, double bx, double cx) {
        return (ax * ax + bx * bx + cx * cx + ax * bx + ax * cx + bx * cx) / 10.0;
    };
    auto mixed = [](double ax, double bx, double cx, double ay, double by, double cy) {
        return (2.0 * (ax * ay + bx * by + cx * cy) + ax * by + ay * bx + ax * cy + ay * cx + bx * cy +
                by * cx) /
               20.0;
    };
    double volume = 0.0, first[3] = {0.0, 0.0, 0.0}, second[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t t = 0; t < triangle_count; ++t) {
        double p[3][3];
        for (int c = 0; c < 3; ++c) {
            const uint32_t v = triangles[3 * t + c];
            if (v >= vertex_count) return fail(DYN_E_INVALID, "a hull triangle must reference its vertices");
            for (int a = 0; a < 3; ++a) p[c][a] = static_cast<double>(vertices[3 * v + a]) - apex[a];
        }
        double ab[3];
        cross(p[0], p[1], ab);
        const double* A = p[0];
        const double* B = dot(ab, p[2]) < 0.0 ? p[2] : p[1];
        const double* C = dot(ab, p[2]) < 0.0 ? p[1] : p[2];
        double bc[3];
        cross(B, C, bc);
        const double signed_volume = dot(A, bc) / 6.0;
        volume += signed_volume;
        for (int a = 0; a < 3; ++a) first[a] += signed_volume * (A[a] + B[a] + C[a]) / 4.0;
        second[0] += signed_volume * squared(A[0], B[0], C[0]);
        second[3] += signed_volume * squared(A[1], B[1], C[1]);
        second[5] += signed_volume * squared(A[2], B[2], C[2]);
        second[1] += signed_volume * mixed(A[0], B[0], C[0], A[1], B[1], C[1]);
        second[2] += signed_volume * mixed(A[0], B[0], C[0], A[2], B[2], C[2]);
        second[4] += signed_volume * mixed(A[1], B[1], C[1], A[2], B[2], C[2]);
    }
    if (!(volume > 0.0) || !std::isfinite(volume)) {
        return fail(DYN_E_INVALID, "a hull must enclose a positive volume");
    }
    const double centroid[3] = {first[0] / volume, first[1] / volume, first[2] / volume};
    *volume_out = static_cast<float>(volume);
    for (int a = 0; a < 3; ++a) centroid_out[a] = static_cast<float>(apex[a] + centroid[a]);
    const int ii[6] = {0, 0, 0, 1, 1, 2}, jj[6] = {0, 1, 2, 1, 2, 2};
    for (int k = 0; k < 6; ++k) {
        second_out[k] = static_cast<float>((second[k] - volume * centroid[ii[k]] * centroid[jj[k]]) / volume);
    }
    return DYN_OK;
}

}  // extern "C"
