// Device contact generation for one candidate pair: analytic sphere / box / capsule pairs,
// box-box SAT with Sutherland-Hodgman clipping, plane against convex, and the manifold ordering
// that keeps at most four spread points. Semantics: crates/dynamis-rigid/shaders/
// narrowphase.wgsl:18-512, crates/dynamis-shader/shaders/convex.wgsl:625-900 and
// core.wgsl:615-760. One thread builds one manifold in registers.
#pragma once

#include "dmath.h"

namespace dyn {

constexpr float kFaceAxisBias = 1e-3f;
constexpr float kClipMargin = 1e-4f;
constexpr uint32_t kCandidates = 16;
constexpr uint32_t kFeaturePoint = 0u;
constexpr uint32_t kFeatureVertex = 1u << 28;
constexpr uint32_t kFeatureEdge = 2u << 28;
constexpr uint32_t kFeatureFace = 3u << 28;
constexpr uint32_t kFeatureKindMask = 0xF0000000u;
constexpr uint32_t kFeatureFieldBits = 14u;
constexpr uint32_t kFeatureFieldMask = (1u << 14) - 1u;
constexpr uint32_t kFeatureClip = 1u << 12;
constexpr uint32_t kFeatureFaceBit = 1u << 13;
constexpr uint32_t kContactAnnounced = 0x80000000u;

struct Body {
    dyn_body_state s;
    dyn_body_desc d;
};

__device__ __forceinline__ bool kind_simulates(const dyn_body_desc& d) {
    return !(d.flags & DYN_BODY_KINEMATIC) && d.inverse_mass > 0.0f;
}
__device__ __forceinline__ bool kind_moves(const dyn_body_desc& d) {
    return (d.flags & DYN_BODY_KINEMATIC) || d.inverse_mass > 0.0f;
}
__device__ __forceinline__ bool body_is_inert(const Body& b) {
    return !kind_simulates(b.d) || b.s.sleeping != 0u;
}
__device__ __forceinline__ void body_frozen(Body& b) {
    b.d.inverse_mass = 0.0f;
    for (int i = 0; i < 6; ++i) b.d.inverse_inertia[i] = 0.0f;
}
__device__ __forceinline__ bool body_is_active(const dyn_body_state& s, const dyn_body_desc& d) {
    return kind_moves(d) && s.sleeping == 0u;
}
__device__ __forceinline__ v3 body_com(const Body& b) {
    return add(ld3(b.s.position), qrot(ld4(b.s.orientation), ld3(b.d.com)));
}
__device__ __forceinline__ v3 apply_inverse_inertia(const Body& b, v3 v) {
    const q4 q = ld4(b.s.orientation);
    return qrot(q, sym_apply(b.d.inverse_inertia, qrot(qconj(q), v)));
}
__device__ __forceinline__ v3 point_velocity(const Body& b, v3 p) {
    return add(ld3(b.s.velocity), cross3(ld3(b.s.angular_velocity), sub(p, body_com(b))));
}
__device__ __forceinline__ v3 relative_velocity(const Body& a, const Body& b, v3 pa, v3 pb) {
    return sub(point_velocity(b, pb), point_velocity(a, pa));
}

struct Shape {
    uint32_t kind;
    float radius, half_height;
    v3 center, half_extents, scale;
    q4 rotation;
    uint32_t source;
};

// Hull shape sources visible to the narrowphase (records.wgsl:256-273).
struct ShapeSet {
    const dyn_shape_source* sources;
    const float4* vertices;
    const uint2* cells;       // support-cell records per source (support_cells.h), or null
    const uint16_t* cands;
};

// core.wgsl:473-484
__device__ __forceinline__ Shape world_collider(const dyn_body_state& s, const dyn_collider& c) {
    Shape w;
    w.kind = c.kind;
    w.radius = c.radius;
    w.half_height = c.half_height;
    const q4 q = ld4(s.orientation);
    w.center = add(ld3(s.position), qrot(q, ld3(c.local_offset)));
    w.half_extents = ld3(c.half_extents);
    w.rotation = qmul(q, ld4(c.local_rotation));
    w.scale = ld3(c.scale);
    w.source = c.source;
    return w;
}

// core.wgsl:490-503: bounds of a shape that reads a shape source
__device__ __forceinline__ void world_shape_bounds(const Shape& w, const ShapeSet& ss, v3* mn, v3* mx) {
    const dyn_shape_source& src = ss.sources[w.source];
    const v3 smin = mulv(ld3(src.local_min), w.scale);
    const v3 smax = mulv(ld3(src.local_max), w.scale);
    const v3 center = add(w.center, qrot(w.rotation, mul(add(smin, smax), 0.5f)));
    const v3 half = mul(sub(smax, smin), 0.5f);
    const v3 extent = add(add(absv(qrot(w.rotation, mk(half.x, 0.0f, 0.0f))),
                              absv(qrot(w.rotation, mk(0.0f, half.y, 0.0f)))),
                          absv(qrot(w.rotation, mk(0.0f, 0.0f, half.z))));
    *mn = sub(center, extent);
    *mx = add(center, extent);
}

// core.wgsl:505-536
__device__ __forceinline__ void world_aabb_of(const Shape& w, const ShapeSet& ss, v3* mn, v3* mx) {
    if (w.kind == DYN_SHAPE_PLANE) {
        *mn = sub(w.center, mk(1e6f, 1e6f, 1e6f));
        *mx = add(w.center, mk(1e6f, 1e6f, 1e6f));
        return;
    }
    if (w.kind == DYN_SHAPE_HULL) {
        world_shape_bounds(w, ss, mn, mx);
        return;
    }
    v3 extent = zero3v();
    if (w.kind == DYN_SHAPE_SPHERE) {
        extent = mk(w.radius, w.radius, w.radius);
    } else if (w.kind == DYN_SHAPE_CUBOID) {
        const v3 e = w.half_extents;
        extent = add(add(absv(qrot(w.rotation, mk(e.x, 0.0f, 0.0f))),
                         absv(qrot(w.rotation, mk(0.0f, e.y, 0.0f)))),
                     absv(qrot(w.rotation, mk(0.0f, 0.0f, e.z))));
    } else if (w.kind == DYN_SHAPE_CAPSULE || w.kind == DYN_SHAPE_CYLINDER) {
        const v3 axis = qrot(w.rotation, mk(0.0f, 1.0f, 0.0f));
        extent = add(absv(mul(axis, w.half_height)), mk(w.radius, w.radius, w.radius));
    }
    const float s = fmx(fmx(w.scale.x, w.scale.y), w.scale.z);
    extent = mul(extent, s);
    *mn = sub(w.center, extent);
    *mx = add(w.center, extent);
}

__device__ __forceinline__ uint32_t feature_pair(uint32_t kind, uint32_t a, uint32_t b) {
    return kind | (a << kFeatureFieldBits) | b;
}
__device__ __forceinline__ uint32_t feature_mirror(uint32_t f) {
    const uint32_t kind = f & kFeatureKindMask;
    if (kind == kFeaturePoint) return f;
    const uint32_t a = (f >> kFeatureFieldBits) & kFeatureFieldMask;
    const uint32_t b = f & kFeatureFieldMask;
    return kind | (b << kFeatureFieldBits) | a;
}

// A manifold under construction: normal and up to four points.
struct Manifold {
    v3 normal;
    uint32_t count;
    v3 pos[4];
    float depth[4];
    uint32_t feature[4];
};

__device__ __forceinline__ void emit(Manifold& m, v3 normal) {
    m.count = 0;
    m.normal = normal;
}

__device__ __forceinline__ v3 sel3(bool take, v3 a, v3 b) {
    return mk(take ? a.x : b.x, take ? a.y : b.y, take ? a.z : b.z);
}

// m's next point, written as a select into each of the four slots (a store through a run-time
// index would keep the manifold in scratch memory).
__device__ __forceinline__ void push(Manifold& m, v3 p, float depth, uint32_t feature) {
    if (m.count >= DYN_CONTACT_MAX_POINTS) return;
#pragma unroll
    for (uint32_t j = 0; j < DYN_CONTACT_MAX_POINTS; ++j) {
        const bool at = j == m.count;
        m.pos[j] = sel3(at, p, m.pos[j]);
        m.depth[j] = at ? depth : m.depth[j];
        m.feature[j] = at ? feature : m.feature[j];
    }
    ++m.count;
}

struct Candidates {
    v3 pos[kCandidates];
    float depth[kCandidates];
    uint32_t feature[kCandidates];
    uint32_t count;
};

__device__ __forceinline__ void swap_candidates(Candidates& c, uint32_t i, uint32_t j) {
    const v3 p = c.pos[i];
    const float d = c.depth[i];
    const uint32_t f = c.feature[i];
    c.pos[i] = c.pos[j];
    c.depth[i] = c.depth[j];
    c.feature[i] = c.feature[j];
    c.pos[j] = p;
    c.depth[j] = d;
    c.feature[j] = f;
}

__device__ __forceinline__ float gap_of(const Candidates& c, uint32_t kept, uint32_t index) {
    float gap = 3.402823466e38f;
    for (uint32_t h = 0; h < kept; ++h) gap = fmn(gap, len3(sub(c.pos[index], c.pos[h])));
    return gap;
}

// core.wgsl:720-760: deepest first (ties by feature), then farthest from the kept ones.
__device__ __forceinline__ void keep(Manifold& m, Candidates& c) {
    const uint32_t n = c.count;
    uint32_t lead = 0;
    for (uint32_t i = 1; i < n; ++i) {
        const float d = c.depth[i], held = c.depth[lead];
        if (d > held || (d == held && c.feature[i] < c.feature[lead])) lead = i;
    }
    if (n > 0) swap_candidates(c, 0, lead);
    // Round k only reorders places k and above, and at most four places are pushed, so the
    // rounds past the fourth place change nothing that is read.
    const uint32_t placed = min(n, static_cast<uint32_t>(DYN_CONTACT_MAX_POINTS));
    for (uint32_t kept = 1; kept < placed; ++kept) {
        uint32_t best = kept;
        float best_gap = gap_of(c, kept, kept);
        for (uint32_t i = kept + 1; i < n; ++i) {
            const float g = gap_of(c, kept, i);
            if (g > best_gap) {
                best_gap = g;
                best = i;
            }
        }
        if (best != kept) swap_candidates(c, kept, best);
    }
    for (uint32_t i = 0; i < n && m.count < DYN_CONTACT_MAX_POINTS; ++i) {
        push(m, c.pos[i], c.depth[i], c.feature[i]);
    }
}

__device__ __forceinline__ v3 box_center(const dyn_body_state& s, const dyn_collider& c) {
    return add(ld3(s.position), qrot(ld4(s.orientation), ld3(c.local_offset)));
}

// narrowphase.wgsl:33-57
__device__ inline void sphere_sphere(Manifold& m, const Body& a, const dyn_collider& ca,
                                     const Body& b, const dyn_collider& cb, float margin) {
    const v3 ac = box_center(a.s, ca), bc = box_center(b.s, cb);
    const v3 delta = sub(bc, ac);
    const float distance = len3(delta);
    v3 normal = sign_normalize(delta);
    if (distance <= 1e-6f) {
        const v3 rel = relative_velocity(a, b, bc, ac);
        if (len3(rel) > 1e-6f) normal = neg(norm3(rel));
    }
    emit(m, normal);
    const float radius_sum = ca.radius + cb.radius;
    if (distance > radius_sum + margin) return;
    const float depth = radius_sum - distance;
    push(m, add(ac, mul(normal, ca.radius - depth * 0.5f)), depth, kFeaturePoint);
}

__device__ __forceinline__ uint32_t smallest_axis(v3 v) {
    const v3 a = absv(v);
    if (a.x < a.y && a.x < a.z) return 0;
    if (a.y < a.z) return 1;
    return 2;
}

// narrowphase.wgsl:85-111 with box_nearest_face (:67-84)
__device__ inline void sphere_box(Manifold& m, const Body& sp, const dyn_collider& sc,
                                  const Body& bx, const dyn_collider& bc, float margin) {
    const v3 center = add(ld3(sp.s.position), qrot(ld4(sp.s.orientation), ld3(sc.local_offset)));
    const q4 q = qmul(ld4(bx.s.orientation), ld4(bc.local_rotation));
    const v3 bcenter = box_center(bx.s, bc);
    const v3 local = qrot(qconj(q), sub(center, bcenter));
    const v3 he = ld3(bc.half_extents);
    const v3 clamped = mk(clampf(local.x, -he.x, he.x), clampf(local.y, -he.y, he.y),
                          clampf(local.z, -he.z, he.z));
    const v3 closest = add(box_center(bx.s, bc), qrot(q, clamped));
    const v3 delta = sub(closest, center);
    const float distance = len3(delta);
    v3 normal = sign_normalize(delta);
    float depth = sc.radius - distance;
    if (distance <= 1e-6f) {
        const v3 l2 = qrot(qconj(q), sub(center, box_center(bx.s, bc)));
        const v3 clearance = sub(he, absv(l2));
        const uint32_t axis = smallest_axis(clearance);
        v3 facing;
        if (axis == 0) facing = mk(l2.x > 0.0f ? -1.0f : 1.0f, 0.0f, 0.0f);
        else if (axis == 1) facing = mk(0.0f, l2.y > 0.0f ? -1.0f : 1.0f, 0.0f);
        else facing = mk(0.0f, 0.0f, l2.z > 0.0f ? -1.0f : 1.0f);
        normal = neg(neg(qrot(q, facing)));
        depth = sc.radius + comp(clearance, axis);
    }
    emit(m, normal);
    if (depth <= -margin) return;
    push(m, sub(closest, mul(normal, depth * 0.5f)), depth, kFeaturePoint);
}

__device__ __forceinline__ void capsule_segment(const Body& b, const dyn_collider& c, v3* s0, v3* s1) {
    const q4 q = qmul(ld4(b.s.orientation), ld4(c.local_rotation));
    const v3 axis = qrot(q, mk(0.0f, 1.0f, 0.0f));
    const v3 center = box_center(b.s, c);
    *s0 = sub(center, mul(axis, c.half_height));
    *s1 = add(center, mul(axis, c.half_height));
}

// core.wgsl:594-602
__device__ __forceinline__ v3 closest_point_segment(v3 p, v3 a, v3 b) {
    const v3 ab = sub(b, a);
    const float denom = dot3(ab, ab);
    if (denom < 1e-10f) return a;
    const float t = clampf(dot3(sub(p, a), ab) / denom, 0.0f, 1.0f);
    return add(a, mul(ab, t));
}

// narrowphase.wgsl:119-140
__device__ inline void sphere_capsule(Manifold& m, const Body& sp, const dyn_collider& sc,
                                      const Body& cp, const dyn_collider& cc, float margin) {
    v3 s0, s1;
    capsule_segment(cp, cc, &s0, &s1);
    const v3 center = add(ld3(sp.s.position), qrot(ld4(sp.s.orientation), ld3(sc.local_offset)));
    const v3 closest = closest_point_segment(center, s0, s1);
    const v3 delta = sub(closest, center);
    const float distance = len3(delta);
    const float radius_sum = sc.radius + cc.radius;
    emit(m, sign_normalize(delta));
    if (distance > radius_sum + margin) return;
    const v3 normal = sign_normalize(delta);
    const float depth = radius_sum - distance;
    push(m, sub(closest, mul(normal, depth * 0.5f)), depth, kFeaturePoint);
}

// narrowphase.wgsl:142-181
__device__ inline void closest_points_segments(v3 a0, v3 a1, v3 b0, v3 b1, v3* pa, v3* pb) {
    const v3 d1 = sub(a1, a0), d2 = sub(b1, b0), r = sub(a0, b0);
    const float a = dot3(d1, d1), e = dot3(d2, d2), f = dot3(d2, r);
    float s = 0.0f, t = 0.0f;
    if (a <= 1e-10f && e <= 1e-10f) {
        s = 0.0f;
        t = 0.0f;
    } else if (a <= 1e-10f) {
        s = 0.0f;
        t = clampf(f / e, 0.0f, 1.0f);
    } else {
        const float c = dot3(d1, r);
        if (e <= 1e-10f) {
            t = 0.0f;
            s = clampf(-c / a, 0.0f, 1.0f);
        } else {
            const float b = dot3(d1, d2);
            const float denom = a * e - b * b;
            s = denom != 0.0f ? clampf((b * f - c * e) / denom, 0.0f, 1.0f) : 0.0f;
            t = (b * s + f) / e;
            if (t < 0.0f) {
                t = 0.0f;
                s = clampf(-c / a, 0.0f, 1.0f);
            } else if (t > 1.0f) {
                t = 1.0f;
                s = clampf((b - c) / a, 0.0f, 1.0f);
            }
        }
    }
    *pa = add(a0, mul(d1, s));
    *pb = add(b0, mul(d2, t));
}

// narrowphase.wgsl:183-204
__device__ inline void capsule_capsule(Manifold& m, const Body& a, const dyn_collider& ca,
                                       const Body& b, const dyn_collider& cb, float margin) {
    v3 a0, a1, b0, b1, pa, pb;
    capsule_segment(a, ca, &a0, &a1);
    capsule_segment(b, cb, &b0, &b1);
    closest_points_segments(a0, a1, b0, b1, &pa, &pb);
    const v3 delta = sub(pb, pa);
    const float distance = len3(delta);
    const float radius_sum = ca.radius + cb.radius;
    emit(m, sign_normalize(delta));
    if (distance > radius_sum + margin) return;
    push(m, mul(add(pa, pb), 0.5f), radius_sum - distance, kFeaturePoint);
}

struct BoxGeom {
    v3 axes[3];
    v3 half;
    v3 center;
};

__device__ __forceinline__ BoxGeom box_geometry(const Body& b, const dyn_collider& c) {
    BoxGeom g;
    const q4 q = qmul(ld4(b.s.orientation), ld4(c.local_rotation));
    g.axes[0] = qrot(q, mk(1.0f, 0.0f, 0.0f));
    g.axes[1] = qrot(q, mk(0.0f, 1.0f, 0.0f));
    g.axes[2] = qrot(q, mk(0.0f, 0.0f, 1.0f));
    g.half = ld3(c.half_extents);
    g.center = box_center(b.s, c);
    return g;
}

__device__ __forceinline__ float support_radius(const BoxGeom& g, v3 axis) {
    return fabsf(dot3(g.axes[0], axis)) * g.half.x + fabsf(dot3(g.axes[1], axis)) * g.half.y +
           fabsf(dot3(g.axes[2], axis)) * g.half.z;
}

__device__ __forceinline__ v3 box_support(const BoxGeom& g, v3 direction) {
    v3 p = g.center;
    for (uint32_t i = 0; i < 3; ++i) {
        const float h = comp(g.half, i);
        p = add(p, mul(g.axes[i], dot3(g.axes[i], direction) > 0.0f ? h : -h));
    }
    return p;
}

// convex.wgsl:625-631
__device__ __forceinline__ uint32_t box_corner_id(uint32_t face_axis, float face_sign, float first,
                                                  float second) {
    uint32_t id = 0;
    id |= face_sign > 0.0f ? (1u << face_axis) : 0u;
    id |= first > 0.0f ? (1u << ((face_axis + 1u) % 3u)) : 0u;
    id |= second > 0.0f ? (1u << ((face_axis + 2u) % 3u)) : 0u;
    return id;
}

// narrowphase.wgsl:232-270
__device__ inline v3 box_face(const BoxGeom& g, v3 face_normal, v3 corners[4], uint32_t ids[4],
                              uint32_t* out_face) {
    uint32_t axis = 0;
    for (uint32_t i = 1; i < 3; ++i) {
        if (fabsf(dot3(face_normal, g.axes[i])) > fabsf(dot3(face_normal, g.axes[axis]))) axis = i;
    }
    const float sign = dot3(face_normal, g.axes[axis]) < 0.0f ? -1.0f : 1.0f;
    const v3 n = mul(g.axes[axis], sign);
    const v3 u = g.axes[(axis + 1u) % 3u], v = g.axes[(axis + 2u) % 3u];
    const float hu = comp(g.half, (axis + 1u) % 3u), hv = comp(g.half, (axis + 2u) % 3u);
    const v3 fc = add(g.center, mul(n, comp(g.half, axis)));
    corners[0] = add(fc, add(mul(u, hu), mul(v, hv)));
    corners[1] = add(fc, add(mul(neg(u), hu), mul(v, hv)));
    corners[2] = add(fc, sub(mul(neg(u), hu), mul(v, hv)));
    corners[3] = add(fc, sub(mul(u, hu), mul(v, hv)));
    ids[0] = box_corner_id(axis, sign, 1.0f, 1.0f);
    ids[1] = box_corner_id(axis, sign, -1.0f, 1.0f);
    ids[2] = box_corner_id(axis, sign, -1.0f, -1.0f);
    ids[3] = box_corner_id(axis, sign, 1.0f, -1.0f);
    *out_face = axis * 2u + (sign > 0.0f ? 1u : 0u);
    return fc;
}

// Eight-slot polygons and candidate lists for the box-box clip, kept in registers: every slot is
// addressed with a compile-time index and a write at a run-time place is a select in each slot.
// The arithmetic and its order are clip_polygon's and keep's.
struct Poly8 {
    v3 p[8];
    uint32_t id[8];
};

__device__ __forceinline__ void poly_set(Poly8& q, uint32_t at, v3 v, uint32_t id) {
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) {
        const bool t = j == at;
        q.p[j] = sel3(t, v, q.p[j]);
        q.id[j] = t ? id : q.id[j];
    }
}

// clip_polygon over a Poly8 (narrowphase.wgsl:272-305)
__device__ inline uint32_t clip_polygon8(const Poly8& in, uint32_t count, v3 pp, v3 pn, uint32_t group,
                                         uint32_t side, Poly8& out) {
    uint32_t n = 0;
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) {
        if (i < count) {
            const v3 cur = in.p[i];
            const v3 nxt = i + 1u < 8u ? sel3(i + 1u < count, in.p[i + 1u < 8u ? i + 1u : 0u], in.p[0]) : in.p[0];
            const float cd = dot3(sub(cur, pp), pn) - kClipMargin;
            const float nd = dot3(sub(nxt, pp), pn) - kClipMargin;
            if (cd <= 0.0f && n < 8u) {
                poly_set(out, n, cur, in.id[i]);
                ++n;
            }
            if (cd * nd < 0.0f) {
                const float t = cd / (cd - nd);
                if (n < 8u) {
                    poly_set(out, n, add(cur, mul(sub(nxt, cur), t)), kFeatureClip | (group * 32u + side * 8u + i));
                    ++n;
                }
            }
        }
    }
    return n;
}

struct Cand8 {
    v3 pos[8];
    float depth[8];
    uint32_t feature[8];
    uint32_t count;
};

__device__ __forceinline__ void cand_push(Cand8& c, v3 p, float depth, uint32_t feature) {
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) {
        const bool at = j == c.count;
        c.pos[j] = sel3(at, p, c.pos[j]);
        c.depth[j] = at ? depth : c.depth[j];
        c.feature[j] = at ? feature : c.feature[j];
    }
    ++c.count;
}

// c[slot] <-> c[other] for a compile-time slot and a run-time other >= slot whose record is
// (op, od, of).
template <uint32_t kSlot>
__device__ __forceinline__ void cand_swap(Cand8& c, uint32_t other, v3 op, float od, uint32_t of) {
    const v3 sp = c.pos[kSlot];
    const float sd = c.depth[kSlot];
    const uint32_t sf = c.feature[kSlot];
#pragma unroll
    for (uint32_t j = kSlot + 1u; j < 8u; ++j) {
        const bool at = j == other;
        c.pos[j] = sel3(at, sp, c.pos[j]);
        c.depth[j] = at ? sd : c.depth[j];
        c.feature[j] = at ? sf : c.feature[j];
    }
    c.pos[kSlot] = op;
    c.depth[kSlot] = od;
    c.feature[kSlot] = of;
}

template <uint32_t kKept>
__device__ __forceinline__ float cand_gap(const Cand8& c, uint32_t index) {
    float gap = 3.402823466e38f;
#pragma unroll
    for (uint32_t h = 0; h < kKept; ++h) gap = fmn(gap, len3(sub(c.pos[index], c.pos[h])));
    return gap;
}

// Round kKept of keep: the farthest remaining candidate from the kept ones moves to slot kKept.
template <uint32_t kKept>
__device__ __forceinline__ void cand_round(Cand8& c, uint32_t n) {
    uint32_t best = kKept;
    float best_gap = cand_gap<kKept>(c, kKept);
    v3 bp = c.pos[kKept];
    float bd = c.depth[kKept];
    uint32_t bf = c.feature[kKept];
#pragma unroll
    for (uint32_t i = kKept + 1u; i < 8u; ++i) {
        if (i < n) {
            const float g = cand_gap<kKept>(c, i);
            if (g > best_gap) {
                best_gap = g;
                best = i;
                bp = c.pos[i];
                bd = c.depth[i];
                bf = c.feature[i];
            }
        }
    }
    if (best != kKept) cand_swap<kKept>(c, best, bp, bd, bf);
}

// keep (core.wgsl:720-760) over a Cand8: the same lead, rounds and pushes.
__device__ inline void keep8(Manifold& m, Cand8& c) {
    const uint32_t n = c.count;
    uint32_t lead = 0;
    v3 lp = c.pos[0];
    float ld = c.depth[0];
    uint32_t lf = c.feature[0];
#pragma unroll
    for (uint32_t i = 1; i < 8; ++i) {
        if (i < n) {
            const float d = c.depth[i];
            if (d > ld || (d == ld && c.feature[i] < lf)) {
                lead = i;
                lp = c.pos[i];
                ld = d;
                lf = c.feature[i];
            }
        }
    }
    if (n > 0) cand_swap<0>(c, lead, lp, ld, lf);
    // rounds past the fourth place change nothing that is pushed
    if (n > 1) cand_round<1>(c, n);
    if (n > 2) cand_round<2>(c, n);
    if (n > 3) cand_round<3>(c, n);
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) {
        if (i < n && m.count < DYN_CONTACT_MAX_POINTS) push(m, c.pos[i], c.depth[i], c.feature[i]);
    }
}

// narrowphase.wgsl:272-305
__device__ inline uint32_t clip_polygon(const v3* pts, const uint32_t* ids, uint32_t count, v3 pp,
                                        v3 pn, uint32_t group, uint32_t side, v3* out_pts,
                                        uint32_t* out_ids) {
    uint32_t out = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const v3 cur = pts[i], nxt = pts[(i + 1u) % count];
        const float cd = dot3(sub(cur, pp), pn) - kClipMargin;
        const float nd = dot3(sub(nxt, pp), pn) - kClipMargin;
        if (cd <= 0.0f && out < 8u) {
            out_pts[out] = cur;
            out_ids[out] = ids[i];
            ++out;
        }
        if (cd * nd < 0.0f) {
            const float t = cd / (cd - nd);
            if (out < 8u) {
                out_pts[out] = add(cur, mul(sub(nxt, cur), t));
                out_ids[out] = kFeatureClip | (group * 32u + side * 8u + i);
                ++out;
            }
        }
    }
    return out;
}

// narrowphase.wgsl:324-473
__device__ inline void box_box_sat(Manifold& m, const Body& a, const dyn_collider& ca,
                                   const Body& b, const dyn_collider& cb, float margin) {
    const BoxGeom left = box_geometry(a, ca), right = box_geometry(b, cb);
    const v3 delta = sub(right.center, left.center);
    float face_depth = 1e30f;
    v3 face_axis = left.axes[0];
    bool face_on_right = false;
    for (uint32_t i = 0; i < 3; ++i) {
        const float overlap =
            comp(left.half, i) + support_radius(right, left.axes[i]) - fabsf(dot3(delta, left.axes[i]));
        if (overlap < face_depth) {
            face_depth = overlap;
            face_axis = left.axes[i];
            face_on_right = false;
        }
        const float other = comp(right.half, i) + support_radius(left, right.axes[i]) -
                            fabsf(dot3(delta, right.axes[i]));
        if (other < face_depth) {
            face_depth = other;
            face_axis = right.axes[i];
            face_on_right = true;
        }
    }
    float edge_depth = 1e30f;
    v3 edge_axis = face_axis;
    uint32_t ef = 0, es = 0;
    for (uint32_t i = 0; i < 3; ++i) {
        for (uint32_t j = 0; j < 3; ++j) {
            const v3 crossed = cross3(left.axes[i], right.axes[j]);
            const float len = len3(crossed);
            if (len < 1e-8f) continue;
            const v3 axis = dv(crossed, len);
            const float overlap = support_radius(left, axis) + support_radius(right, axis) -
                                  fabsf(dot3(delta, axis));
            if (overlap < edge_depth) {
                edge_depth = overlap;
                edge_axis = axis;
                ef = i;
                es = j;
            }
        }
    }
    const float extent = fmx(fmx(left.half.x, left.half.y), left.half.z) +
                         fmx(fmx(right.half.x, right.half.y), right.half.z);
    const bool edge_sep = edge_depth + kFaceAxisBias * extent < face_depth;
    const float depth = edge_sep ? edge_depth : face_depth;
    if (depth <= -margin) {
        m.count = 0;
        m.normal = zero3v();
        return;
    }
    const v3 axis = edge_sep ? edge_axis : face_axis;
    const v3 signed_axis = dot3(axis, delta) < 0.0f ? neg(axis) : axis;
    emit(m, signed_axis);
    if (edge_sep) {
        const v3 p = mul(add(box_support(left, signed_axis), box_support(right, neg(signed_axis))), 0.5f);
        push(m, p, depth, feature_pair(kFeatureEdge, ef, es));
        return;
    }
    const BoxGeom& reference = face_on_right ? right : left;
    const BoxGeom& incident = face_on_right ? left : right;
    const v3 ref_normal = face_on_right ? neg(signed_axis) : signed_axis;
    v3 ref_corners[4];
    uint32_t ref_ids[4], ref_face = 0;
    const v3 ref_center = box_face(reference, ref_normal, ref_corners, ref_ids, &ref_face);
    uint32_t ia = 0;
    for (uint32_t i = 1; i < 3; ++i) {
        if (fabsf(dot3(incident.axes[i], ref_normal)) > fabsf(dot3(incident.axes[ia], ref_normal))) ia = i;
    }
    const v3 inc_normal = mul(incident.axes[ia], dot3(incident.axes[ia], ref_normal) > 0.0f ? -1.0f : 1.0f);
    v3 inc_corners[4];
    uint32_t inc_ids[4], inc_face = 0;
    box_face(incident, inc_normal, inc_corners, inc_ids, &inc_face);
    const uint32_t ref_field = kFeatureFaceBit | ref_face;
    Poly8 pa, pb;
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) {
        pa.p[i] = pb.p[i] = zero3v();
        pa.id[i] = pb.id[i] = 0u;
    }
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        pa.p[i] = inc_corners[i];
        pa.id[i] = inc_ids[i];
    }
    uint32_t pc = 4;
#pragma unroll
    for (uint32_t side = 0; side < 4; ++side) {
        if (pc == 0u) break;
        const v3 cur = ref_corners[side], nxt = ref_corners[(side + 1u) % 4u];
        const v3 plane_normal = norm3(cross3(sub(nxt, cur), ref_normal));
        if (side % 2u == 0u) pc = clip_polygon8(pa, pc, cur, plane_normal, inc_face, side, pb);
        else pc = clip_polygon8(pb, pc, cur, plane_normal, inc_face, side, pa);
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
    for (uint32_t i = 0; i < 8; ++i) {
        if (i < pc) {
            const float pd = dot3(sub(ref_center, pa.p[i]), ref_normal);
            if (pd >= -margin) {
                cand_push(cand, sub(pa.p[i], mul(ref_normal, pd * 0.5f)), pd,
                          face_on_right ? feature_pair(kFeatureFace, pa.id[i], ref_field)
                                        : feature_pair(kFeatureFace, ref_field, pa.id[i]));
            }
        }
    }
    if (cand.count == 0u) {
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) {
            const float pd = dot3(sub(ref_center, inc_corners[i]), ref_normal);
            if (pd >= -margin) {
                cand_push(cand, sub(inc_corners[i], mul(ref_normal, pd * 0.5f)), pd,
                          face_on_right ? feature_pair(kFeatureFace, inc_ids[i], ref_field)
                                        : feature_pair(kFeatureFace, ref_field, inc_ids[i]));
            }
        }
    }
    if (cand.count == 0u) {
        push(m, sub(ref_center, mul(ref_normal, face_depth * 0.5f)), face_depth, kFeaturePoint);
        return;
    }
    keep8(m, cand);
}

}  // namespace dyn
