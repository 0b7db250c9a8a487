// Device narrowphase for pairs the analytic generators do not cover: support maps of every convex
// kind (hulls by a linear scan of their vertices), GJK closest points, EPA penetration, the
// direction probe fallback, and manifolds from clipped feature rings. Semantics:
// crates/dynamis-shader/shaders/convex.wgsl:1-1071 and crates/dynamis-rigid/shaders/
// narrowphase.wgsl:479-512. GJK, the probe and the manifolds run one thread per pair; EPA runs one
// wavefront per penetrating pair, with the polytope in the wave's LDS and the lanes scanning a
// hull's vertices and EPA's faces in parallel, reduced to the element the reference's sequential
// scan picks (the first maximum or minimum).
#pragma once

#include "narrow.h"
#include "support_cells.h"

namespace dyn {

constexpr uint32_t kFeatureMax = 8;
constexpr float kCapAlignment = 0.93f;
constexpr uint32_t kCylinderCapPoints = 8;
constexpr uint32_t kCylinderSidePoints = 16;
constexpr uint32_t kCylinderSideFace = 2;
constexpr uint32_t kEmptyRingFace = 1u << 12;  // FEATURE_INDEX_LIMIT
constexpr float kBig = 3.402823466e38f;

// cos / sin of f32(i) * 0.78539816 for the eight cylinder cap points, correctly rounded
__constant__ float kCapCos[8] = {0x1.0p+0f, 0x1.6a09e6p-1f, -0x1.777a5cp-25f, -0x1.6a09e6p-1f,
                                 -0x1.0p+0f, -0x1.6a09e2p-1f, 0x1.99bc5cp-27f, 0x1.6a09eep-1f};
__constant__ float kCapSin[8] = {0x0.0p+0f, 0x1.6a09e6p-1f, 0x1.0p+0f, 0x1.6a09e6p-1f,
                                 -0x1.777a5cp-24f, -0x1.6a09eap-1f, -0x1.0p+0f, -0x1.6a09dep-1f};

__device__ __forceinline__ v3 hull_vertex(const ShapeSet& ss, const dyn_shape_source& src, uint32_t i) {
    const float4 v = ss.vertices[src.vertex_offset + i];
    return mk(v.x, v.y, v.z);
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The hull vertex of largest dot(v, d), first index on ties (the sequential scan's `s >
// best` keeps the first maximum), the kLanes lanes of an aligned lane group scanning every
// kLanes-th vertex and reducing (value, index) pairs. All lanes of the group call it together
// with the same arguments.
template <int kLanes>
__device__ inline v3 hull_support_lanes(const Shape& w, const ShapeSet& ss, v3 d) {
    const dyn_shape_source src = ss.sources[w.source];
    float best = -kBig;
    uint32_t index = 0xFFFFFFFFu;
    for (uint32_t i = lane_id() % kLanes; i < src.vertex_count; i += kLanes) {
        const float s = dot3(hull_vertex(ss, src, i), d);
        if (s > best) {
            best = s;
            index = i;
        }
    }
#pragma unroll
    for (int offset = kLanes / 2; offset > 0; offset >>= 1) {
        const float ob = __shfl_xor(best, offset, kLanes);
        const uint32_t oi = __shfl_xor(index, offset, kLanes);
        if (ob > best || (ob == best && oi < index)) {
            best = ob;
            index = oi;
        }
    }
    return index == 0xFFFFFFFFu ? zero3v() : hull_vertex(ss, src, index);
}

// The same vertex by one thread's sequential scan: of the direction's support-cell candidates
// when the source has them (the same vertex, support_cells.h), else of every vertex.
__device__ inline v3 hull_support_thread(const Shape& w, const ShapeSet& ss, v3 d) {
    const dyn_shape_source src = ss.sources[w.source];
    float best_dot = -kBig;
    v3 best = zero3v();
    uint32_t cell;
    if (ss.cells && support_cell_of(d.x, d.y, d.z, &cell)) {
        const uint2 rec = ss.cells[static_cast<size_t>(w.source) * kSupportCells + cell];
        const uint32_t count = rec.x & 0xFFu;
        if (count != kSupportNoCell) {
            if (count <= kSupportInline) {
                // the unused slots load the first candidate again (not compared)
                const uint32_t i0 = (rec.x >> 8) & 0xFFFu;
                const v3 c0 = hull_vertex(ss, src, i0);
                const v3 c1 = hull_vertex(ss, src, count > 1u ? rec.x >> 20 : i0);
                const v3 c2 = hull_vertex(ss, src, count > 2u ? rec.y & 0xFFFu : i0);
                const v3 c3 = hull_vertex(ss, src, count > 3u ? (rec.y >> 12) & 0xFFFu : i0);
                const auto take = [&](v3 c) {
                    const float s = dot3(c, d);
                    if (s > best_dot) {
                        best_dot = s;
                        best = c;
                    }
                };
                take(c0);
                if (count > 1u) take(c1);
                if (count > 2u) take(c2);
                if (count > 3u) take(c3);
                return best;
            }
            const uint16_t* list = ss.cands + rec.y;
            for (uint32_t k = 0; k < count; ++k) {
                const v3 v = hull_vertex(ss, src, list[k]);
                const float s = dot3(v, d);
                if (s > best_dot) {
                    best_dot = s;
                    best = v;
                }
            }
            return best;
        }
    }
#pragma unroll 8
    for (uint32_t i = 0; i < src.vertex_count; ++i) {
        const v3 v = hull_vertex(ss, src, i);
        const float s = dot3(v, d);
        if (s > best_dot) {
            best_dot = s;
            best = v;
        }
    }
    return best;
}

// convex.wgsl:1-85 (hulls of at most DYN_LINEAR_SUPPORT_VERTICES vertices). With kLanes > 1 an
// aligned group of kLanes lanes calls it together and shares the hull scan.
template <int kLanes = 1>
__device__ inline v3 support_local(const Shape& w, const ShapeSet& ss, v3 d) {
    if (w.kind == DYN_SHAPE_HULL) {
        return kLanes > 1 ? hull_support_lanes<kLanes>(w, ss, d) : hull_support_thread(w, ss, d);
    }
    const v3 n = norm3(d);  // the direction normalized (read by spheres and capsules)
    if (w.kind == DYN_SHAPE_SPHERE) return mul(n, w.radius);
    if (w.kind == DYN_SHAPE_CUBOID) {
        const v3 s = mk(d.x > 0.0f ? 1.0f : -1.0f, d.y > 0.0f ? 1.0f : -1.0f, d.z > 0.0f ? 1.0f : -1.0f);
        return mulv(s, w.half_extents);
    }
    if (w.kind == DYN_SHAPE_CAPSULE) {
        const v3 tip = mul(mul(mk(0.0f, 1.0f, 0.0f), w.half_height), d.y > 0.0f ? 1.0f : -1.0f);
        return add(tip, mul(n, w.radius));
    }
    if (w.kind == DYN_SHAPE_CYLINDER) {
        const v3 base = mul(mul(mk(0.0f, 1.0f, 0.0f), w.half_height), d.y > 0.0f ? 1.0f : -1.0f);
        const v3 radial = sub(d, mk(0.0f, d.y, 0.0f));
        if (len3(radial) < 1e-6f) return base;
        return add(base, mul(norm3(radial), w.radius));
    }
    return zero3v();
}

// convex.wgsl:86-91
template <int kLanes = 1>
__device__ inline v3 support(const Shape& w, const ShapeSet& ss, v3 direction) {
    const v3 sd = mulv(qrot(qconj(w.rotation), direction), w.scale);
    return add(w.center, qrot(w.rotation, mulv(support_local<kLanes>(w, ss, sd), w.scale)));
}

// ---- GJK (convex.wgsl:93-282, 439-521) -------------------------------------------------------

struct SimplexPoint {
    v3 w, a, b;
};

struct SimplexResult {
    v3 v;
    float lambdas[4];
    bool penetrating;
    v3 point_a, point_b;
};

struct Closest {
    float distance;
    v3 point_a, point_b, normal;
    bool penetrating;
};

struct Hit {
    float distance;
    v3 point, normal;
};

struct Epa {
    bool valid;
    v3 normal;
    float depth;
    v3 point;
};

__device__ inline v3 closest_on_triangle(v3 o, v3 a, v3 b, v3 c, v3* weights) {
    const v3 ab = sub(b, a), ac = sub(c, a), ap = sub(o, a);
    const float d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) {
        *weights = mk(1.0f, 0.0f, 0.0f);
        return a;
    }
    const v3 bp = sub(o, b);
    const float d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) {
        *weights = mk(0.0f, 1.0f, 0.0f);
        return b;
    }
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        const float t = d1 / (d1 - d3);
        *weights = mk(1.0f - t, t, 0.0f);
        return add(a, mul(ab, t));
    }
    const v3 cp = sub(o, c);
    const float d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) {
        *weights = mk(0.0f, 0.0f, 1.0f);
        return c;
    }
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        const float t = d2 / (d2 - d6);
        *weights = mk(1.0f - t, 0.0f, t);
        return add(a, mul(ac, t));
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        const float t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        *weights = mk(0.0f, 1.0f - t, t);
        return add(b, mul(sub(c, b), t));
    }
    const float denom = 1.0f / (va + vb + vc);
    const float v = vb * denom, wt = vc * denom;
    *weights = mk(1.0f - v - wt, v, wt);
    return add(add(a, mul(ab, v)), mul(ac, wt));
}

__device__ __forceinline__ void tet_face(uint32_t f, uint32_t* a, uint32_t* b, uint32_t* c) {
    // the four faces of a tetrahedron in the reference's order
    *a = f == 3u ? 1u : 0u;
    *b = f == 0u ? 1u : (f == 1u ? 2u : 3u);
    *c = f == 0u ? 2u : (f == 1u ? 3u : (f == 2u ? 1u : 2u));
}

// |t[i].w - t[j].w| written from the higher slot to the lower (the same bits either way, as
// a - b = -(b - a) exactly), so the faces that share an edge share one square root.
__device__ __forceinline__ float tet_edge(const SimplexPoint* t, uint32_t i, uint32_t j) {
    return len3(sub(t[i > j ? i : j].w, t[i > j ? j : i].w));
}

__device__ inline bool tetra_contains(const SimplexPoint* t, v3 o) {
#pragma unroll
    for (uint32_t f = 0; f < 4; ++f) {
        uint32_t ia, ib, ic;
        tet_face(f, &ia, &ib, &ic);
        const v3 a = t[ia].w, b = t[ib].w, c = t[ic].w;
        const v3 normal = cross3(sub(b, a), sub(c, a));
        const float span = fmx(fmx(tet_edge(t, ib, ia), tet_edge(t, ic, ia)), tet_edge(t, ic, ib));
        if (len3(normal) <= 1e-6f * span * span) return false;
        if (dot3(normal, sub(o, a)) > 0.0f) return false;
    }
    return true;
}

// The simplex lives in four fixed slots that every loop below addresses with indices known at
// compile time (the loops are unrolled over the four slots and guarded by the live count), so
// it stays in registers; the iterations run in the sequential order of the reference.
__device__ __forceinline__ v3 pick3(bool take, v3 a, v3 b) {
    return mk(take ? a.x : b.x, take ? a.y : b.y, take ? a.z : b.z);
}

__device__ __forceinline__ SimplexPoint pick_point(bool take, const SimplexPoint& a, const SimplexPoint& b) {
    SimplexPoint r;
    r.w = pick3(take, a.w, b.w);
    r.a = pick3(take, a.a, b.a);
    r.b = pick3(take, a.b, b.b);
    return r;
}

// s[at] = p, as a select in every slot (a store under a branch is merged into a store through a
// selected pointer, which would keep the simplex in memory).
__device__ __forceinline__ void set_slot(SimplexPoint* s, uint32_t at, const SimplexPoint& p) {
#pragma unroll
    for (uint32_t t = 0; t < 4; ++t) s[t] = pick_point(t == at, p, s[t]);
}

__device__ inline bool simplex_holds(const SimplexPoint* s, uint32_t count, v3 point, float span) {
    const float tolerance = (1e-5f * span) * (1e-5f * span);
    bool held = false;
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        if (i < count) {
            const v3 offset = sub(s[i].w, point);
            held = held || dot3(offset, offset) <= tolerance;
        }
    }
    return held;
}

// Keeps the points whose weight exceeds 1e-6, in order, in the first slots (s[0] alone when none
// does). Slot t takes the t-th kept point: the first kept of four, the second of the last three,
// the third of the last two; the slots past the new count are never read again.
__device__ inline uint32_t simplex_reduce(SimplexPoint* s, const float* lambdas) {
    const bool k0 = lambdas[0] > 1e-6f, k1 = lambdas[1] > 1e-6f, k2 = lambdas[2] > 1e-6f, k3 = lambdas[3] > 1e-6f;
    const uint32_t count = (k0 ? 1u : 0u) + (k1 ? 1u : 0u) + (k2 ? 1u : 0u) + (k3 ? 1u : 0u);
    const uint32_t before2 = (k0 ? 1u : 0u) + (k1 ? 1u : 0u);  // kept among the first two
    const SimplexPoint first = pick_point(k0, s[0], pick_point(k1, s[1], pick_point(k2, s[2], pick_point(k3, s[3], s[0]))));
    const SimplexPoint second = pick_point(before2 == 2u, s[1], pick_point(before2 == 1u && k2, s[2], s[3]));
    const SimplexPoint third = pick_point(before2 == 2u && k2, s[2], s[3]);
    s[0] = first;
    s[1] = second;
    s[2] = third;
    return count == 0u ? 1u : count;
}

__device__ inline SimplexResult simplex_closest(const SimplexPoint* s, uint32_t count) {
    SimplexResult r;
    r.v = mk(kBig, kBig, kBig);
#pragma unroll
    for (int i = 0; i < 4; ++i) r.lambdas[i] = 0.0f;
    r.penetrating = count == 4u && tetra_contains(s, zero3v());
    r.point_a = zero3v();
    r.point_b = zero3v();
    if (r.penetrating) {
        r.v = zero3v();
        return r;
    }
    float best = len3(r.v);  // len3(r.v), kept beside r.v
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        const float length = len3(s[i].w);
        if (i < count && length < best) {
            r.v = s[i].w;
            best = length;
#pragma unroll
            for (int k = 0; k < 4; ++k) r.lambdas[k] = 0.0f;
            r.lambdas[i] = 1.0f;
        }
    }
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
#pragma unroll
        for (uint32_t j = i + 1; j < 4; ++j) {
            if (j >= count) continue;
            const v3 a = s[i].w, b = s[j].w;
            const v3 ab = sub(b, a);
            const float denom = dot3(ab, ab);
            if (denom < 1e-12f) continue;
            const float t = clampf(-dot3(a, ab) / denom, 0.0f, 1.0f);
            const v3 point = add(a, mul(ab, t));
            const float length = len3(point);
            if (length < best) {
                r.v = point;
                best = length;
#pragma unroll
                for (int k = 0; k < 4; ++k) r.lambdas[k] = 0.0f;
                r.lambdas[i] = 1.0f - t;
                r.lambdas[j] = t;
            }
        }
    }
    if (count >= 3u) {
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) {
#pragma unroll
            for (uint32_t j = i + 1; j < 4; ++j) {
#pragma unroll
                for (uint32_t k = j + 1; k < 4; ++k) {
                    if (k >= count) continue;
                    v3 weights;
                    const v3 point = closest_on_triangle(zero3v(), s[i].w, s[j].w, s[k].w, &weights);
                    const float length = len3(point);
                    if (length < best) {
                        r.v = point;
                        best = length;
#pragma unroll
                        for (int q = 0; q < 4; ++q) r.lambdas[q] = 0.0f;
                        r.lambdas[i] = weights.x;
                        r.lambdas[j] = weights.y;
                        r.lambdas[k] = weights.z;
                    }
                }
            }
        }
    }
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        if (i < count) {
            r.point_a = add(r.point_a, smul(r.lambdas[i], s[i].a));
            r.point_b = add(r.point_b, smul(r.lambdas[i], s[i].b));
        }
    }
    return r;
}

template <int kLanes = 1>
__device__ inline SimplexPoint support_pair(const Shape& a, const Shape& b, const ShapeSet& ss, v3 dir) {
    SimplexPoint p;
    p.a = support<kLanes>(a, ss, dir);
    p.b = support<kLanes>(b, ss, neg(dir));
    p.w = sub(p.a, p.b);
    return p;
}

__device__ __forceinline__ float shape_scale(const Shape& w) {
    float scale = w.radius + w.half_height;
    scale = fmx(scale, fmx(fmx(w.half_extents.x, w.half_extents.y), w.half_extents.z));
    return fmx(scale, 1e-4f) * fmx(fmx(w.scale.x, w.scale.y), w.scale.z);
}

__device__ inline v3 simplex_expand_dir(const SimplexPoint* s, uint32_t count) {
    if (count == 1u) return mk(1.0f, 0.0f, 0.0f);
    if (count == 2u) {
        const v3 edge = sub(s[1].w, s[0].w);
        v3 d = cross3(edge, mk(1.0f, 0.0f, 0.0f));
        if (len3(d) < 1e-6f) d = cross3(edge, mk(0.0f, 1.0f, 0.0f));
        return norm3(d);
    }
    const v3 normal = cross3(sub(s[1].w, s[0].w), sub(s[2].w, s[0].w));
    if (len3(normal) > 1e-8f) return norm3(normal);
    return mk(0.0f, 1.0f, 0.0f);
}

// GJK's loop state between iterations: the simplex, the closest point of its hull to the origin,
// and the next iteration's index. A pair can stop after any iteration and go on later (another
// kernel) from this state with exactly the same arithmetic.
struct GjkState {
    SimplexPoint simplex[4];
    uint32_t count;
    SimplexResult closest;
    uint32_t iter;
};

constexpr uint32_t kGjkIterations = 24;  // convex.wgsl:478

// The result convex_closest returns when its loop runs out without an answer (or before it has
// one): the centres and their normalized offset.
__device__ inline Closest closest_default(const Shape& first, const Shape& second) {
    Closest r;
    r.distance = 0.0f;
    r.point_a = first.center;
    r.point_b = second.center;
    r.normal = sign_normalize(sub(second.center, first.center));
    r.penetrating = false;
    return r;
}

// convex_closest (convex.wgsl:464-521), the part before the loop.
template <int kLanes = 1>
__device__ inline void gjk_begin(const Shape& first, const Shape& second, const ShapeSet& ss, GjkState* g) {
    v3 dir = sub(second.center, first.center);
    if (len3(dir) < 1e-8f) dir = mk(1.0f, 0.0f, 0.0f);
#pragma unroll
    for (uint32_t i = 1; i < 4; ++i) g->simplex[i].w = g->simplex[i].a = g->simplex[i].b = zero3v();
    g->simplex[0] = support_pair<kLanes>(first, second, ss, dir);
    g->count = 1;
    g->closest = simplex_closest(g->simplex, g->count);
    g->iter = 0;
}

// The loop of convex_closest from iteration g->iter up to (not including) `limit`. Returns true
// with *r set when the loop has ended (answered, or all kGjkIterations run), false when it
// stopped at `limit` with the state in *g.
template <int kLanes = 1>
__device__ inline bool gjk_run(const Shape& first, const Shape& second, const ShapeSet& ss, GjkState* g,
                               uint32_t limit, Closest* r) {
    *r = closest_default(first, second);
    const float tiny = 1e-4f * fmn(shape_scale(first), shape_scale(second));
    SimplexPoint* simplex = g->simplex;
    uint32_t count = g->count;
    SimplexResult closest = g->closest;
    uint32_t iter = g->iter;
    bool done = false;
    for (; iter < kGjkIterations; ++iter) {
        if (iter >= limit) break;
        if (closest.penetrating || len3(closest.v) < tiny) {
            while (count < 4u) {
                const v3 d = simplex_expand_dir(simplex, count);
                set_slot(simplex, count, support_pair<kLanes>(first, second, ss, d));
                ++count;
            }
            r->penetrating = true;
            r->distance = 0.0f;
            r->point_a = closest.point_a;
            r->point_b = closest.point_b;
            done = true;
            break;
        }
        const v3 nv = neg(closest.v);
        const SimplexPoint sp = support_pair<kLanes>(first, second, ss, nv);
        if (dot3(sp.w, nv) <= dot3(closest.v, nv) + 1e-4f * fmx(1.0f, dot3(closest.v, closest.v))) {
            r->distance = len3(closest.v);
            r->point_a = closest.point_a;
            r->point_b = closest.point_b;
            if (r->distance > 1e-6f) r->normal = sign_normalize(sub(closest.point_b, closest.point_a));
            done = true;
            break;
        }
        count = simplex_reduce(simplex, closest.lambdas);
        if (simplex_holds(simplex, count, sp.w, shape_scale(second))) {
            r->distance = len3(closest.v);
            r->point_a = closest.point_a;
            r->point_b = closest.point_b;
            if (r->distance > 1e-6f) r->normal = sign_normalize(sub(closest.point_b, closest.point_a));
            done = true;
            break;
        }
        set_slot(simplex, count, sp);
        ++count;
        closest = simplex_closest(simplex, count);
    }
    g->count = count;
    g->closest = closest;
    g->iter = iter;
    return done || iter >= kGjkIterations;
}

template <int kLanes = 1>
__device__ inline Closest convex_closest(const Shape& first, const Shape& second, const ShapeSet& ss,
                                         SimplexPoint* simplex, uint32_t* out_count) {
    GjkState g;
    gjk_begin<kLanes>(first, second, ss, &g);
    Closest r;
    gjk_run<kLanes>(first, second, ss, &g, kGjkIterations, &r);
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) simplex[i] = g.simplex[i];
    *out_count = g.count;
    return r;
}

// ---- EPA (convex.wgsl:286-437), one wave per polytope ------------------------------------------

struct EpaFace {
    uint32_t a, b, c, pad;
    float4 normal;
};

// The EPA polytope of one wave in LDS. Only the Minkowski points are kept: the reference also
// stores each point's two support points, but its callers never read them.
struct EpaShared {
    float4 w[64];
    EpaFace faces[128];
    EpaFace fresh[64];
    uint32_t visible[128];
};

__device__ __forceinline__ Epa epa_invalid() {
    Epa r;
    r.valid = false;
    r.normal = mk(0.0f, 1.0f, 0.0f);
    r.depth = 0.0f;
    r.point = zero3v();
    return r;
}

__device__ __forceinline__ v3 lw(const EpaShared& e, uint32_t i) {
    const float4 p = e.w[i];
    return mk(p.x, p.y, p.z);
}

__device__ __forceinline__ EpaFace epa_face(uint32_t a, uint32_t b, uint32_t c, v3 unit, v3 centroid) {
    EpaFace f;
    f.a = a;
    f.b = b;
    f.c = c;
    f.pad = 0u;
    const v3 n = dot3(unit, centroid) > 0.0f ? neg(unit) : unit;
    f.normal = make_float4(n.x, n.y, n.z, 0.0f);
    return f;
}

__device__ __forceinline__ v3 face_normal(const EpaFace& f) { return mk(f.normal.x, f.normal.y, f.normal.z); }

// The face of smallest dot(normal, w[a]) among faces with a unit normal, first index on ties
// (the sequential `dist < best` keeps the first minimum); kBig and face 0 when none qualifies.
__device__ inline void epa_best_face(const EpaShared& e, uint32_t face_count, float* best_dist,
                                     uint32_t* best_face) {
    float best = kBig;
    uint32_t index = 0xFFFFFFFFu;
    for (uint32_t i = lane_id(); i < face_count; i += kWave) {
        const EpaFace f = e.faces[i];
        const float dist = dot3(face_normal(f), lw(e, f.a));
        if (dist < best && len3(face_normal(f)) > 0.5f) {
            best = dist;
            index = i;
        }
    }
#pragma unroll
    for (int offset = kWave / 2; offset > 0; offset >>= 1) {
        const float ob = __shfl_xor(best, offset, kWave);
        const uint32_t oi = __shfl_xor(index, offset, kWave);
        if (ob < best || (ob == best && oi < index)) {
            best = ob;
            index = oi;
        }
    }
    *best_dist = best;
    *best_face = index == 0xFFFFFFFFu ? 0u : index;
}

// `simplex_w` holds the four simplex points' x, y, z (12 floats).
__device__ inline Epa epa_tetrahedron(const Shape& first, const Shape& second, const ShapeSet& ss,
                                      const float* simplex_w, EpaShared& e) {
    const uint32_t lane = lane_id();
    wave_sync();
    if (lane < 4u) e.w[lane] = make_float4(simplex_w[3u * lane], simplex_w[3u * lane + 1u], simplex_w[3u * lane + 2u], 0.0f);
    wave_sync();
    uint32_t vertex_count = 4;
    uint32_t face_count = 0;
    for (uint32_t i = 0; i < 4; ++i) {
        uint32_t a, b, c;
        tet_face(i, &a, &b, &c);
        const v3 normal = cross3(sub(lw(e, b), lw(e, a)), sub(lw(e, c), lw(e, a)));
        const float length_n = len3(normal);
        if (length_n < 1e-10f) return epa_invalid();
        const v3 unit = dv(normal, length_n);
        const v3 centroid = mul(add(add(lw(e, a), lw(e, b)), lw(e, c)), 1.0f / 3.0f);
        if (lane == 0u) e.faces[face_count] = epa_face(a, b, c, unit, centroid);
        ++face_count;
    }
    wave_sync();
    float best_dist = 0.0f;
    uint32_t best_face = 0;
    for (uint32_t iter = 0; iter < 32u; ++iter) {
        epa_best_face(e, face_count, &best_dist, &best_face);
        const v3 bn = face_normal(e.faces[best_face]);
        const v3 probe = sub(support<kWave>(first, ss, bn), support<kWave>(second, ss, neg(bn)));
        const float gain = dot3(probe, bn) - best_dist;
        if (gain < 1e-4f) break;
        if (vertex_count >= 64u) break;
        if (lane == 0u) e.w[vertex_count] = make_float4(probe.x, probe.y, probe.z, 0.0f);
        ++vertex_count;
        wave_sync();
        // faces that see the new point, in face order
        uint32_t visible_count = 0;
        uint64_t seen_masks[2] = {0u, 0u};
        for (uint32_t base = 0; base < face_count; base += kWave) {
            const uint32_t i = base + lane;
            bool vis = false;
            if (i < face_count) {
                const EpaFace f = e.faces[i];
                vis = dot3(face_normal(f), sub(probe, lw(e, f.a))) > 1e-6f;
            }
            const uint64_t mask = __ballot(vis);
            seen_masks[base / kWave] = mask;
            if (vis) e.visible[visible_count + __popcll(mask & ((1ull << lane) - 1ull))] = i;
            visible_count += __popcll(mask);
        }
        wave_sync();
        // horizon edges of the visible faces, in (face, edge) order; the first 64 that make a
        // proper face become new faces
        const uint32_t apex = vertex_count - 1u;
        uint32_t new_count = 0;
        const uint32_t items = visible_count * 3u;
        for (uint32_t base = 0; base < items && new_count < 64u; base += kWave) {
            const uint32_t t = base + lane;
            bool ok = false;
            EpaFace made;
            if (t < items) {
                const uint32_t i = t / 3u, edge = t % 3u;
                const EpaFace face = e.faces[e.visible[i]];
                const uint32_t start = edge == 0u ? face.a : (edge == 1u ? face.b : face.c);
                const uint32_t end = edge == 0u ? face.b : (edge == 1u ? face.c : face.a);
                bool shared = false;
                for (uint32_t j = 0; j < visible_count; ++j) {
                    if (j == i) continue;
                    const EpaFace o = e.faces[e.visible[j]];
                    if ((o.a == end && o.b == start) || (o.b == end && o.c == start) || (o.c == end && o.a == start)) {
                        shared = true;
                    }
                }
                if (!shared) {
                    const v3 normal = cross3(sub(lw(e, end), lw(e, apex)), sub(lw(e, start), lw(e, apex)));
                    const float length_n = len3(normal);
                    if (length_n > 1e-10f) {
                        const v3 unit = dv(normal, length_n);
                        const v3 centroid = mul(add(add(lw(e, apex), lw(e, start)), lw(e, end)), 1.0f / 3.0f);
                        made = epa_face(apex, start, end, unit, centroid);
                        ok = true;
                    }
                }
            }
            const uint64_t mask = __ballot(ok);
            const uint32_t rank = new_count + __popcll(mask & ((1ull << lane) - 1ull));
            if (ok && rank < 64u) e.fresh[rank] = made;
            new_count = min(new_count + static_cast<uint32_t>(__popcll(mask)), 64u);
        }
        wave_sync();
        // drop the visible faces, keeping the order of the rest: a round's kept faces move to
        // slots below the next round's, so each round reads, then writes
        uint32_t write = 0;
        for (uint32_t k = 0; k < 2; ++k) {
            const uint32_t i = k * kWave + lane;
            if (k * kWave >= face_count) break;
            const bool mine = i < face_count && !((seen_masks[k] >> lane) & 1ull);
            const uint64_t mask = __ballot(mine);
            uint4 kept_ids = make_uint4(0u, 0u, 0u, 0u);
            float4 kept_normal = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (mine) {
                kept_ids = make_uint4(e.faces[i].a, e.faces[i].b, e.faces[i].c, e.faces[i].pad);
                kept_normal = e.faces[i].normal;
            }
            wave_sync();
            if (mine) {
                EpaFace& to = e.faces[write + __popcll(mask & ((1ull << lane) - 1ull))];
                to.a = kept_ids.x;
                to.b = kept_ids.y;
                to.c = kept_ids.z;
                to.pad = kept_ids.w;
                to.normal = kept_normal;
            }
            write += __popcll(mask);
            wave_sync();
        }
        face_count = write;
        if (face_count + new_count > 128u) break;
        for (uint32_t i = lane; i < new_count; i += kWave) e.faces[face_count + i] = e.fresh[i];
        face_count += new_count;
        wave_sync();
    }
    epa_best_face(e, face_count, &best_dist, &best_face);
    Epa r;
    r.valid = best_dist < kBig;
    r.normal = face_normal(e.faces[best_face]);
    r.depth = best_dist;
    r.point = zero3v();  // the support points of the polytope are not kept (never read)
    return r;
}

// ---- EPA of small polytopes, kEpaGroup lanes per polytope -------------------------------------
//
// The same expansion as epa_tetrahedron, operation for operation and in the same order, for a
// polytope that stays within kEpaSmallVertices vertices, kEpaSmallFaces faces and
// kEpaSmallFresh new faces per expansion: an aligned group of kEpaGroup lanes shares one
// polytope (kWave / kEpaGroup polytopes per wave), its loops over faces and horizon edges
// striding by the group. A polytope that would outgrow the capacities is given up (the call
// returns false) and is expanded from its start by epa_tetrahedron, so either way the result is
// the sequential expansion's. The support points come from support<1> (the same vertex as the
// wave's shared scan).

constexpr uint32_t kEpaGroup = 8;
constexpr uint32_t kEpaSmallVertices = 16;
constexpr uint32_t kEpaSmallFaces = 24;
constexpr uint32_t kEpaSmallFresh = 16;

struct alignas(16) EpaSmallFace {
    float nx, ny, nz;
    uint32_t abc;  // vertex indices a | b << 8 | c << 16
};

static_assert(sizeof(EpaSmallFace) == 16, "a small face is one 16-byte LDS word");

struct EpaSmallShared {
    float4 w[kEpaSmallVertices];
    EpaSmallFace faces[kEpaSmallFaces];
    EpaSmallFace fresh[kEpaSmallFresh];
    uint8_t visible[kEpaSmallFaces];
};

__device__ __forceinline__ EpaSmallFace epa_small_face(uint32_t a, uint32_t b, uint32_t c, v3 unit, v3 centroid) {
    const v3 n = dot3(unit, centroid) > 0.0f ? neg(unit) : unit;
    EpaSmallFace f;
    f.nx = n.x;
    f.ny = n.y;
    f.nz = n.z;
    f.abc = a | (b << 8) | (c << 16);
    return f;
}

__device__ __forceinline__ v3 small_normal(const EpaSmallFace& f) { return mk(f.nx, f.ny, f.nz); }
__device__ __forceinline__ uint32_t small_a(const EpaSmallFace& f) { return f.abc & 0xFFu; }
__device__ __forceinline__ uint32_t small_b(const EpaSmallFace& f) { return (f.abc >> 8) & 0xFFu; }
__device__ __forceinline__ uint32_t small_c(const EpaSmallFace& f) { return f.abc >> 16; }

__device__ __forceinline__ v3 small_w(const EpaSmallShared& e, uint32_t i) {
    const float4 p = e.w[i];
    return mk(p.x, p.y, p.z);
}

// Lane value of the lane DPP control `kCtrl` names (quad permutes and the half-row mirror stay
// inside an aligned group of 8 lanes).
template <int kCtrl>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), kCtrl, 0xF, 0xF, false));
}
constexpr int kDppQuadSwap1 = 0xB1;     // quad_perm [1, 0, 3, 2]
constexpr int kDppQuadSwap2 = 0x4E;     // quad_perm [2, 3, 0, 1]
constexpr int kDppHalfMirror = 0x141;  // row_half_mirror: lane 7 - l of each 8 lanes

__device__ __forceinline__ void group8_min_first(float* best, uint32_t* index) {
    uint32_t b = __float_as_uint(*best), i = *index;
    const auto step = [&](uint32_t ob, uint32_t oi) {
        const float of = __uint_as_float(ob), bf = __uint_as_float(b);
        if (of < bf || (of == bf && oi < i)) {
            b = ob;
            i = oi;
        }
    };
    step(dpp_u32<kDppQuadSwap1>(b), dpp_u32<kDppQuadSwap1>(i));
    step(dpp_u32<kDppQuadSwap2>(b), dpp_u32<kDppQuadSwap2>(i));
    step(dpp_u32<kDppHalfMirror>(b), dpp_u32<kDppHalfMirror>(i));
    *best = __uint_as_float(b);
    *index = i;
}

// This group's 8 bits of a wave ballot.
__device__ __forceinline__ uint32_t group_ballot(bool p, uint32_t group) {
    return static_cast<uint32_t>(__ballot(p) >> (kEpaGroup * group)) & 0xFFu;
}

__device__ inline void epa_small_best(const EpaSmallShared& e, uint32_t face_count, uint32_t q, float* best_dist,
                                      uint32_t* best_face) {
    float best = kBig;
    uint32_t index = 0xFFFFFFFFu;
    for (uint32_t i = q; i < face_count; i += kEpaGroup) {
        const EpaSmallFace f = e.faces[i];
        const float dist = dot3(small_normal(f), small_w(e, small_a(f)));
        if (dist < best && len3(small_normal(f)) > 0.5f) {
            best = dist;
            index = i;
        }
    }
    group8_min_first(&best, &index);
    *best_dist = best;
    *best_face = index == 0xFFFFFFFFu ? 0u : index;
}

// Returns false when the polytope outgrows the capacities (nothing is returned then), the
// vertices' being min(vertex_limit, kEpaSmallVertices); every lane of the group calls it with the
// same arguments and receives the same result.
// `simplex_w` holds the four simplex points' x, y, z (12 floats).
__device__ inline bool epa_small(const Shape& first, const Shape& second, const ShapeSet& ss,
                                 const float* simplex_w, uint32_t vertex_limit, EpaSmallShared& e, Epa* out) {
    const uint32_t lane = lane_id();
    const uint32_t q = lane % kEpaGroup, group = lane / kEpaGroup;
    const uint32_t below = (1u << q) - 1u;
    wave_sync();
    if (q < 4u) e.w[q] = make_float4(simplex_w[3u * q], simplex_w[3u * q + 1u], simplex_w[3u * q + 2u], 0.0f);
    wave_sync();
    uint32_t vertex_count = 4;
    uint32_t face_count = 0;
    for (uint32_t i = 0; i < 4; ++i) {
        uint32_t a, b, c;
        tet_face(i, &a, &b, &c);
        const v3 normal = cross3(sub(small_w(e, b), small_w(e, a)), sub(small_w(e, c), small_w(e, a)));
        const float length_n = len3(normal);
        if (length_n < 1e-10f) {
            *out = epa_invalid();
            return true;
        }
        const v3 unit = dv(normal, length_n);
        const v3 centroid = mul(add(add(small_w(e, a), small_w(e, b)), small_w(e, c)), 1.0f / 3.0f);
        if (q == 0u) e.faces[face_count] = epa_small_face(a, b, c, unit, centroid);
        ++face_count;
    }
    wave_sync();
    float best_dist = 0.0f;
    uint32_t best_face = 0;
    for (uint32_t iter = 0; iter < 32u; ++iter) {
        epa_small_best(e, face_count, q, &best_dist, &best_face);
        const v3 bn = small_normal(e.faces[best_face]);
        const v3 probe = sub(support<1>(first, ss, bn), support<1>(second, ss, neg(bn)));
        const float gain = dot3(probe, bn) - best_dist;
        if (gain < 1e-4f) break;
        if (vertex_count >= 64u) break;
        if (vertex_count >= vertex_limit || vertex_count >= kEpaSmallVertices) return false;
        if (q == 0u) e.w[vertex_count] = make_float4(probe.x, probe.y, probe.z, 0.0f);
        ++vertex_count;
        wave_sync();
        // faces that see the new point, in face order
        uint32_t visible_count = 0;
        uint32_t seen = 0;
        for (uint32_t base = 0; base < face_count; base += kEpaGroup) {
            const uint32_t i = base + q;
            bool vis = false;
            if (i < face_count) {
                const EpaSmallFace f = e.faces[i];
                vis = dot3(small_normal(f), sub(probe, small_w(e, small_a(f)))) > 1e-6f;
            }
            const uint32_t mask = group_ballot(vis, group);
            seen |= mask << base;
            if (vis) e.visible[visible_count + __popc(mask & below)] = static_cast<uint8_t>(i);
            visible_count += __popc(mask);
        }
        wave_sync();
        // horizon edges of the visible faces, in (face, edge) order
        const uint32_t apex = vertex_count - 1u;
        uint32_t new_count = 0;
        const uint32_t items = visible_count * 3u;
        for (uint32_t base = 0; base < items; base += kEpaGroup) {
            const uint32_t t = base + q;
            bool ok = false;
            EpaSmallFace made;
            if (t < items) {
                const uint32_t i = t / 3u, edge = t % 3u;
                const EpaSmallFace face = e.faces[e.visible[i]];
                const uint32_t fa = small_a(face), fb = small_b(face), fc = small_c(face);
                const uint32_t start = edge == 0u ? fa : (edge == 1u ? fb : fc);
                const uint32_t end = edge == 0u ? fb : (edge == 1u ? fc : fa);
                bool shared = false;
                for (uint32_t j = 0; j < visible_count; ++j) {
                    if (j == i) continue;
                    const EpaSmallFace o = e.faces[e.visible[j]];
                    const uint32_t oa = small_a(o), ob = small_b(o), oc = small_c(o);
                    if ((oa == end && ob == start) || (ob == end && oc == start) || (oc == end && oa == start)) {
                        shared = true;
                    }
                }
                if (!shared) {
                    const v3 normal = cross3(sub(small_w(e, end), small_w(e, apex)), sub(small_w(e, start), small_w(e, apex)));
                    const float length_n = len3(normal);
                    if (length_n > 1e-10f) {
                        const v3 unit = dv(normal, length_n);
                        const v3 centroid = mul(add(add(small_w(e, apex), small_w(e, start)), small_w(e, end)), 1.0f / 3.0f);
                        made = epa_small_face(apex, start, end, unit, centroid);
                        ok = true;
                    }
                }
            }
            const uint32_t mask = group_ballot(ok, group);
            const uint32_t rank = new_count + __popc(mask & below);
            if (ok && rank < kEpaSmallFresh) e.fresh[rank] = made;
            new_count += __popc(mask);
        }
        if (new_count > kEpaSmallFresh) return false;
        wave_sync();
        // drop the visible faces, keeping the order of the rest: a round's kept faces move to
        // slots below the next round's, so each round reads, then writes
        uint32_t write = 0;
        for (uint32_t base = 0; base < face_count; base += kEpaGroup) {
            const uint32_t i = base + q;
            const bool mine = i < face_count && !((seen >> i) & 1u);
            const uint32_t mask = group_ballot(mine, group);
            float4 kept = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (mine) kept = reinterpret_cast<const float4*>(e.faces)[i];
            wave_sync();
            if (mine) reinterpret_cast<float4*>(e.faces)[write + __popc(mask & below)] = kept;
            write += __popc(mask);
            wave_sync();
        }
        face_count = write;
        if (face_count + new_count > 128u) break;
        if (face_count + new_count > kEpaSmallFaces) return false;
        for (uint32_t i = q; i < new_count; i += kEpaGroup) e.faces[face_count + i] = e.fresh[i];
        face_count += new_count;
        wave_sync();
    }
    epa_small_best(e, face_count, q, &best_dist, &best_face);
    Epa r;
    r.valid = best_dist < kBig;
    r.normal = small_normal(e.faces[best_face]);
    r.depth = best_dist;
    r.point = zero3v();  // the support points of the polytope are not kept (never read)
    *out = r;
    return true;
}

// ---- penetration probe and convex_hit (convex.wgsl:523-618) ----------------------------------

template <int kLanes = 1>
__device__ __forceinline__ float support_projection_depth(const Shape& a, const Shape& b, const ShapeSet& ss, v3 dir) {
    const v3 n = norm3(dir);
    return dot3(sub(support<kLanes>(a, ss, n), support<kLanes>(b, ss, neg(n))), n);
}

struct Probe {
    float depth;
    v3 normal;
};

template <int kLanes = 1>
__device__ __forceinline__ void probe_direction(const Shape& first, const Shape& second, const ShapeSet& ss,
                                                v3 direction, Probe* best) {
    const float depth = support_projection_depth<kLanes>(first, second, ss, direction);
    if (depth < best->depth) {
        best->depth = depth;
        best->normal = norm3(direction);
    }
}

// The smallest support-projection overlap over the probe directions, in the reference's order:
// the normals of the simplex's triangles, the centre line, then +-x, +-y, +-z (at most 11 of
// the 16 the reference allows, so its cap never binds). The directions are evaluated as they
// are generated. With `until_separated` the scan stops at the first direction whose overlap is
// not positive: the depth is then not positive either, which is all the GJK half of convex_hit
// reads of it (the normal of such a probe is never used).
template <int kLanes = 1>
__device__ inline Probe penetration_probe(const Shape& first, const Shape& second, const ShapeSet& ss,
                                          const SimplexPoint* s, uint32_t count, bool until_separated) {
    Probe best;
    best.depth = kBig;
    best.normal = mk(0.0f, 1.0f, 0.0f);
    // triangles (i, j, k), i < j < k < count, in loop order: 012, 013, 023, 123
    for (uint32_t t = 0; t < 4; ++t) {
        if ((t == 0u ? 2u : 3u) >= count) break;
        const v3 a = t == 3u ? s[1].w : s[0].w;
        const v3 b = t < 2u ? s[1].w : s[2].w;
        const v3 c = t == 0u ? s[2].w : s[3].w;
        const v3 normal = cross3(sub(b, a), sub(c, a));
        if (len3(normal) > 1e-8f) {
            probe_direction<kLanes>(first, second, ss, normal, &best);
            if (until_separated && !(best.depth > 0.0f)) return best;
        }
    }
    const v3 center_dir = sub(second.center, first.center);
    if (len3(center_dir) > 1e-8f) {
        probe_direction<kLanes>(first, second, ss, center_dir, &best);
        if (until_separated && !(best.depth > 0.0f)) return best;
    }
    for (uint32_t i = 0; i < 6; ++i) {
        const float sgn = (i & 1u) ? -1.0f : 1.0f;
        const v3 axis = i < 2u ? mk(sgn, 0.0f, 0.0f) : (i < 4u ? mk(0.0f, sgn, 0.0f) : mk(0.0f, 0.0f, sgn));
        probe_direction<kLanes>(first, second, ss, axis, &best);
        if (until_separated && !(best.depth > 0.0f)) return best;
    }
    return best;
}

__device__ inline Hit make_hit(float distance, v3 point, v3 normal) {
    Hit h;
    h.distance = distance;
    h.point = point;
    h.normal = normal;
    return h;
}

// convex_hit (convex.wgsl:584-618), GJK half: answers every pair GJK does not find penetrating
// (returns true with the hit); for a penetrating pair returns false with GJK's final simplex,
// which EPA (one wave) and convex_hit_finish take up.
template <int kLanes = 1>
__device__ inline bool convex_hit_after_gjk(const Shape& first, const Shape& second, const ShapeSet& ss,
                                            const Closest& closest, Hit* hit, const SimplexPoint* simplex,
                                            const uint32_t* count) {
    if (closest.penetrating) return false;
    const v3 centers = sub(second.center, first.center);
    const Probe probe = penetration_probe<kLanes>(first, second, ss, simplex, *count, true);
    if (probe.depth > 0.0f) {
        const v3 normal = dot3(probe.normal, centers) < 0.0f ? neg(probe.normal) : probe.normal;
        *hit = make_hit(-probe.depth, mul(add(closest.point_a, closest.point_b), 0.5f), normal);
    } else if (closest.distance > 0.0f) {
        *hit = make_hit(closest.distance, mul(add(closest.point_a, closest.point_b), 0.5f), closest.normal);
    } else {
        *hit = make_hit(kBig, zero3v(), zero3v());  // NO_HIT
    }
    return true;
}

template <int kLanes = 1>
__device__ inline bool convex_hit_gjk(const Shape& first, const Shape& second, const ShapeSet& ss, Hit* hit,
                                      SimplexPoint* simplex, uint32_t* count) {
    const Closest closest = convex_closest<kLanes>(first, second, ss, simplex, count);
    return convex_hit_after_gjk<kLanes>(first, second, ss, closest, hit, simplex, count);
}

// convex_hit, the rest of a penetrating pair once EPA has run on the simplex.
__device__ inline Hit convex_hit_finish(const Shape& first, const Shape& second, const ShapeSet& ss,
                                        const SimplexPoint* simplex, uint32_t count, const Epa& epa) {
    const v3 centers = sub(second.center, first.center);
    Hit result;
    if (epa.valid && epa.depth > 0.0f && epa.depth < kBig && len3(epa.normal) > 0.5f &&
        epa.depth < support_projection_depth(first, second, ss, epa.normal) +
                        0.1f * fmn(shape_scale(first), shape_scale(second))) {
        result = make_hit(-epa.depth,
                          mul(add(support(first, ss, epa.normal), support(second, ss, neg(epa.normal))), 0.5f),
                          epa.normal);
    } else {
        const Probe probe = penetration_probe(first, second, ss, simplex, count, false);
        result = make_hit(-probe.depth,
                          mul(add(support(first, ss, probe.normal), support(second, ss, neg(probe.normal))), 0.5f),
                          probe.normal);
    }
    if (dot3(result.normal, centers) < 0.0f) result.normal = neg(result.normal);
    return result;
}

// ---- feature rings (convex.wgsl:625-902) -----------------------------------------------------

struct Feature {
    v3 p[kFeatureMax];
    uint32_t id[kFeatureMax];
    uint32_t count;
    uint32_t face;
    bool ring;
};

__device__ inline void box_feature_ring(const Shape& w, v3 direction, Feature* f) {
    const v3 axes[3] = {qrot(w.rotation, mk(1.0f, 0.0f, 0.0f)), qrot(w.rotation, mk(0.0f, 1.0f, 0.0f)),
                        qrot(w.rotation, mk(0.0f, 0.0f, 1.0f))};
    uint32_t best = 0;
    for (uint32_t i = 1; i < 3; ++i) {
        if (fabsf(dot3(axes[i], direction)) > fabsf(dot3(axes[best], direction))) best = i;
    }
    const float sign = dot3(axes[best], direction) < 0.0f ? -1.0f : 1.0f;
    const v3 face = mul(axes[best], sign);
    const v3 center = add(w.center, mul(face, comp(w.half_extents, best)));
    const v3 u1 = axes[(best + 1u) % 3u], u2 = axes[(best + 2u) % 3u];
    const float e1 = comp(w.half_extents, (best + 1u) % 3u);
    const float e2 = comp(w.half_extents, (best + 2u) % 3u);
    f->p[0] = add(add(center, mul(u1, e1)), mul(u2, e2));
    f->id[0] = box_corner_id(best, sign, 1.0f, 1.0f);
    f->p[1] = add(sub(center, mul(u1, e1)), mul(u2, e2));
    f->id[1] = box_corner_id(best, sign, -1.0f, 1.0f);
    f->p[2] = sub(sub(center, mul(u1, e1)), mul(u2, e2));
    f->id[2] = box_corner_id(best, sign, -1.0f, -1.0f);
    f->p[3] = sub(add(center, mul(u1, e1)), mul(u2, e2));
    f->id[3] = box_corner_id(best, sign, 1.0f, -1.0f);
    f->face = kFeatureFaceBit | (best * 2u + (sign > 0.0f ? 1u : 0u));
    f->count = 4;
}

// The order of atan2(y, x) of two points decided exactly from the half plane and the sign of a
// cross product (ring_of's angular sort, convex.wgsl:696-711). Groups in increasing angle:
// -pi, (-pi, 0), [0, pi), pi.
__device__ __forceinline__ int angle_group(float x, float y) {
    if (y == 0.0f && x < 0.0f) return signbit(y) ? 0 : 3;
    if (y < 0.0f) return 1;
    return 2;
}

__device__ __forceinline__ bool angle_before(float bx, float by, float ax, float ay) {
    if (bx == 0.0f && by == 0.0f) bx = 1.0f;
    if (ax == 0.0f && ay == 0.0f) ax = 1.0f;
    const int gb = angle_group(bx, by), ga = angle_group(ax, ay);
    if (gb != ga) return gb < ga;
    if (gb == 0 || gb == 3) return false;
    return bx * ay - by * ax > 0.0f;
}

// convex.wgsl:672-721 over a ring of up to 16 gathered points
__device__ inline void ring_of(const Shape& w, const ShapeSet& ss, v3 direction, v3* points, uint32_t* ids,
                               uint32_t count, Feature* f) {
    if (count <= 1u) {
        f->p[0] = support(w, ss, direction);
        f->id[0] = ids[0];
        f->count = 1;
        return;
    }
    const v3 n = sign_normalize(direction);
    v3 center = zero3v();
    for (uint32_t i = 0; i < count; ++i) center = add(center, points[i]);
    center = dv(center, static_cast<float>(count));
    const v3 u = fabsf(n.x) > 0.9f ? norm3(cross3(n, mk(0.0f, 1.0f, 0.0f))) : norm3(cross3(n, mk(1.0f, 0.0f, 0.0f)));
    const v3 v = cross3(n, u);
    for (uint32_t i = 0; i < count; ++i) {
        for (uint32_t j = i + 1; j < count; ++j) {
            const v3 pa = sub(points[i], center), pb = sub(points[j], center);
            if (angle_before(dot3(pb, u), dot3(pb, v), dot3(pa, u), dot3(pa, v))) {
                const v3 tp = points[i];
                points[i] = points[j];
                points[j] = tp;
                const uint32_t ti = ids[i];
                ids[i] = ids[j];
                ids[j] = ti;
            }
        }
    }
    const uint32_t keep = count < kFeatureMax ? count : kFeatureMax;
    for (uint32_t i = 0; i < keep; ++i) {
        f->p[i] = points[i];
        f->id[i] = ids[i];
    }
    f->count = keep;
}

// convex.wgsl:723-793, the linear path
__device__ inline void hull_feature_ring(const Shape& w, const ShapeSet& ss, v3 direction, Feature* f) {
    const dyn_shape_source src = ss.sources[w.source];
    const v3 d = mulv(qrot(qconj(w.rotation), direction), w.scale);
    const float best_dot = dot3(support_local(w, ss, d), d);
    const v3 extent = sub(ld3(src.local_max), ld3(src.local_min));
    const float eps = 1e-3f * fmx(fmx(extent.x, extent.y), extent.z) * fmx(fmx(w.scale.x, w.scale.y), w.scale.z);
    v3 points[16];
    uint32_t ids[16];
    ids[0] = 0u;
    uint32_t count = 0, first_id = kEmptyRingFace;
    for (uint32_t i = 0; i < src.vertex_count && count < 16u; ++i) {
        const v3 v = hull_vertex(ss, src, i);
        if (dot3(v, d) >= best_dot - eps) {
            points[count] = add(w.center, qrot(w.rotation, mulv(v, w.scale)));
            ids[count] = i;
            first_id = first_id < i ? first_id : i;
            ++count;
        }
    }
    f->face = kFeatureFaceBit | first_id;
    ring_of(w, ss, direction, points, ids, count, f);
}

// convex.wgsl:795-819
__device__ inline void capsule_feature(const Shape& w, v3 direction, Feature* f) {
    const v3 ld = mulv(qrot(qconj(w.rotation), direction), w.scale);
    const float cap = ld.y > 0.0f ? 1.0f : -1.0f;
    const v3 radial = mul(sign_normalize(ld), w.radius);
    f->face = kFeatureFaceBit | 0u;
    if (fabsf(ld.y) > kCapAlignment * len3(ld)) {
        const v3 local = add(mk(0.0f, cap * w.half_height, 0.0f), radial);
        f->p[0] = add(w.center, qrot(w.rotation, mulv(local, w.scale)));
        f->id[0] = cap > 0.0f ? 1u : 0u;
        f->count = 1;
        return;
    }
    const v3 upper = add(mk(0.0f, w.half_height, 0.0f), radial);
    const v3 lower = add(mk(0.0f, -w.half_height, 0.0f), radial);
    f->p[0] = add(w.center, qrot(w.rotation, mulv(upper, w.scale)));
    f->p[1] = add(w.center, qrot(w.rotation, mulv(lower, w.scale)));
    f->id[0] = 0u;
    f->id[1] = 1u;
    f->count = 2;
}

// convex.wgsl:821-854
__device__ inline void cylinder_feature(const Shape& w, v3 direction, Feature* f) {
    const v3 axis = qrot(w.rotation, mk(0.0f, 1.0f, 0.0f));
    const float alignment = dot3(axis, norm3(direction));
    if (fabsf(alignment) > kCapAlignment) {
        const float sign = alignment < 0.0f ? -1.0f : 1.0f;
        const uint32_t cap = sign > 0.0f ? 1u : 0u;
        const v3 cap_center = add(w.center, mul(mul(axis, w.half_height), sign));
        v3 u = norm3(cross3(axis, mk(1.0f, 0.0f, 0.0f)));
        if (len3(cross3(axis, mk(1.0f, 0.0f, 0.0f))) < 1e-6f) u = norm3(cross3(axis, mk(0.0f, 0.0f, 1.0f)));
        const v3 v = cross3(axis, u);
        for (uint32_t i = 0; i < kCylinderCapPoints; ++i) {
            f->p[i] = add(cap_center, mul(add(mul(u, kCapCos[i]), mul(v, kCapSin[i])), w.radius));
            f->id[i] = cap * kCylinderCapPoints + i;
        }
        f->face = kFeatureFaceBit | cap;
        f->count = kCylinderCapPoints;
        return;
    }
    const v3 side = sign_normalize(sub(direction, mul(axis, dot3(direction, axis))));
    f->p[0] = add(add(w.center, mul(axis, w.half_height)), mul(side, w.radius));
    f->id[0] = kCylinderSidePoints;
    f->p[1] = add(sub(w.center, mul(axis, w.half_height)), mul(side, w.radius));
    f->id[1] = kCylinderSidePoints + 1u;
    f->face = kFeatureFaceBit | kCylinderSideFace;
    f->count = 2;
}

// convex.wgsl:856-889
__device__ inline void shape_feature(const Shape& w, const ShapeSet& ss, v3 direction, Feature* f) {
    if (w.kind == DYN_SHAPE_CUBOID) {
        box_feature_ring(w, direction, f);
        f->ring = true;
    } else if (w.kind == DYN_SHAPE_HULL) {
        hull_feature_ring(w, ss, direction, f);
        f->ring = f->count > 2u;
    } else if (w.kind == DYN_SHAPE_CAPSULE) {
        capsule_feature(w, direction, f);
        f->ring = false;
    } else if (w.kind == DYN_SHAPE_CYLINDER) {
        cylinder_feature(w, direction, f);
        f->ring = f->count > 2u;
    } else {
        f->p[0] = support(w, ss, direction);
        f->id[0] = 0u;
        f->face = kFeatureFaceBit | 0u;
        f->count = 1;
        f->ring = false;
    }
}

__device__ inline float feature_radius(const Feature& f) {
    v3 center = zero3v();
    for (uint32_t i = 0; i < f.count; ++i) center = add(center, f.p[i]);
    center = dv(center, static_cast<float>(f.count));
    float radius = 0.0f;
    for (uint32_t i = 0; i < f.count; ++i) radius = fmx(radius, len3(sub(f.p[i], center)));
    return radius;
}

// convex.wgsl:910-982: clip a feature against the side planes of a reference ring. Each side
// writes the points it keeps into the other of two buffers (the caller's and a local one), so no
// side copies its output back; the caller's buffer holds the result on return.
__device__ inline uint32_t clip_feature_to_ring(const Feature& inc, const Feature& ref, v3 ref_dir, v3* held,
                                                uint32_t* held_ids) {
    v3 spare[kCandidates];
    uint32_t spare_ids[kCandidates];
    for (uint32_t i = 0; i < inc.count; ++i) {
        held[i] = inc.p[i];
        held_ids[i] = inc.id[i];
    }
    v3* from = held;
    uint32_t* from_ids = held_ids;
    v3* kept = spare;
    uint32_t* kept_ids = spare_ids;
    uint32_t live = inc.count;
    for (uint32_t side = 0; side < ref.count; ++side) {
        const v3 current = ref.p[side], next = ref.p[(side + 1u) % ref.count];
        const v3 plane_normal = norm3(cross3(sub(next, current), ref_dir));
        uint32_t written = 0;
        if (live == 2u) {
            const v3 p0 = from[0], p1 = from[1];
            const float s0 = dot3(sub(p0, current), plane_normal) - kClipMargin;
            const float s1 = dot3(sub(p1, current), plane_normal) - kClipMargin;
            if (s0 <= 0.0f) {
                kept[written] = p0;
                kept_ids[written] = from_ids[0];
                ++written;
            }
            if (s0 * s1 < 0.0f && written < kCandidates) {
                kept[written] = add(p0, mul(sub(p1, p0), s0 / (s0 - s1)));
                kept_ids[written] = kFeatureClip | (side * kCandidates);
                ++written;
            }
            if (s1 <= 0.0f && written < kCandidates) {
                kept[written] = p1;
                kept_ids[written] = from_ids[1];
                ++written;
            }
        } else {
            for (uint32_t i = 0; i < live; ++i) {
                const v3 cp = from[i], np = from[(i + 1u) % live];
                const float cs = dot3(sub(cp, current), plane_normal) - kClipMargin;
                const float ns = dot3(sub(np, current), plane_normal) - kClipMargin;
                if (cs <= 0.0f && written < kCandidates) {
                    kept[written] = cp;
                    kept_ids[written] = from_ids[i];
                    ++written;
                }
                if (cs * ns < 0.0f && written < kCandidates) {
                    kept[written] = add(cp, mul(sub(np, cp), cs / (cs - ns)));
                    kept_ids[written] = kFeatureClip | (side * kCandidates + i);
                    ++written;
                }
            }
        }
        live = written;
        v3* const t = from;
        uint32_t* const t_ids = from_ids;
        from = kept;
        from_ids = kept_ids;
        kept = t;
        kept_ids = t_ids;
        if (live == 0u) break;
    }
    if (from != held) {
        for (uint32_t i = 0; i < live; ++i) {
            held[i] = from[i];
            held_ids[i] = from_ids[i];
        }
    }
    return live;
}

// convex.wgsl:984-1071
__device__ inline bool convex_pair_manifold(const Shape& first, const Shape& second, const ShapeSet& ss,
                                            v3 direction, Manifold& m) {
    Feature fa, fb;
    shape_feature(first, ss, direction, &fa);
    shape_feature(second, ss, neg(direction), &fb);
    if (!fa.ring && !fb.ring) return false;
    bool first_reference = fa.ring;
    if (fa.ring && fb.ring) first_reference = feature_radius(fa) >= feature_radius(fb);
    const Feature& ref = first_reference ? fa : fb;
    const Feature& inc = first_reference ? fb : fa;
    const v3 ref_dir = first_reference ? direction : neg(direction);
    v3 center = zero3v();
    for (uint32_t i = 0; i < ref.count; ++i) center = add(center, ref.p[i]);
    center = dv(center, static_cast<float>(ref.count));
    v3 deepest = zero3v();
    float deepest_depth = -1e30f;
    for (uint32_t i = 0; i < inc.count; ++i) {
        const float depth = dot3(sub(center, inc.p[i]), ref_dir);
        if (depth > deepest_depth) {
            deepest_depth = depth;
            deepest = inc.p[i];
        }
    }
    v3 clipped[kCandidates];
    uint32_t clipped_ids[kCandidates];
    const uint32_t clipped_count = clip_feature_to_ring(inc, ref, ref_dir, clipped, clipped_ids);
    Candidates cand;
    cand.count = 0;
    for (uint32_t i = 0; i < clipped_count && cand.count < kCandidates; ++i) {
        const float depth = dot3(sub(center, clipped[i]), ref_dir);
        if (depth > 0.0f) {
            cand.pos[cand.count] = sub(clipped[i], mul(ref_dir, depth * 0.5f));
            cand.depth[cand.count] = depth;
            cand.feature[cand.count] = first_reference ? feature_pair(kFeatureVertex, ref.face, clipped_ids[i])
                                                       : feature_pair(kFeatureVertex, clipped_ids[i], ref.face);
            ++cand.count;
        }
    }
    if (cand.count == 0u) {
        if (deepest_depth <= 0.0f) return false;
        push(m, sub(deepest, mul(ref_dir, deepest_depth * 0.5f)), deepest_depth, kFeaturePoint);
        return true;
    }
    keep(m, cand);
    return true;
}

// narrowphase.wgsl:479-483
__device__ __forceinline__ void manifold_from_hit(Manifold& m, const Hit& hit, float margin) {
    if (hit.distance <= margin) push(m, hit.point, -hit.distance, kFeaturePoint);
}

// narrowphase.wgsl:485-512
__device__ inline void plane_convex(Manifold& m, const Shape& plane, const Shape& convex, const ShapeSet& ss,
                                    float margin) {
    const v3 n = qrot(plane.rotation, mk(0.0f, 1.0f, 0.0f));
    const float center_side = dot3(sub(convex.center, plane.center), n);
    const v3 facing = center_side < 0.0f ? neg(n) : n;
    emit(m, facing);
    Feature f;
    shape_feature(convex, ss, neg(facing), &f);
    Candidates cand;
    cand.count = 0;
    for (uint32_t i = 0; i < f.count; ++i) {
        const float depth = dot3(sub(plane.center, f.p[i]), facing);
        if (depth <= -margin) continue;
        cand.pos[cand.count] = add(f.p[i], mul(facing, depth * 0.5f));
        cand.depth[cand.count] = depth;
        cand.feature[cand.count] = feature_pair(kFeatureVertex, kFeatureFaceBit | 0u, f.id[i]);
        ++cand.count;
    }
    keep(m, cand);
}

// plane_convex for the shapes k_narrow gives it (sphere, box, capsule): the same feature
// (shape_feature's branches for those kinds), candidates and selection, with every array
// addressed at compile-time indices so that none sits in scratch memory.
__device__ inline void plane_analytic(Manifold& m, const Shape& plane, const Shape& convex, const ShapeSet& ss,
                                      float margin) {
    const v3 n = qrot(plane.rotation, mk(0.0f, 1.0f, 0.0f));
    const float center_side = dot3(sub(convex.center, plane.center), n);
    const v3 facing = center_side < 0.0f ? neg(n) : n;
    emit(m, facing);
    Feature f;
    if (convex.kind == DYN_SHAPE_CUBOID) {
        box_feature_ring(convex, neg(facing), &f);
    } else if (convex.kind == DYN_SHAPE_CAPSULE) {
        capsule_feature(convex, neg(facing), &f);
    } else {
        f.p[0] = support(convex, ss, neg(facing));
        f.id[0] = 0u;
        f.count = 1;
    }
    Cand8 cand;
    cand.count = 0;
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) {
        cand.pos[i] = zero3v();
        cand.depth[i] = 0.0f;
        cand.feature[i] = 0u;
    }
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {  // a box face, a capsule side or one point
        if (i < f.count) {
            const float depth = dot3(sub(plane.center, f.p[i]), facing);
            if (!(depth <= -margin)) {  // plane_convex's test, NaN included
                cand_push(cand, add(f.p[i], mul(facing, depth * 0.5f)), depth,
                          feature_pair(kFeatureVertex, kFeatureFaceBit | 0u, f.id[i]));
            }
        }
    }
    keep8(m, cand);
}

}  // namespace dyn
