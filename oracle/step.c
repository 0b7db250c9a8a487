/*
 * step.c - CPU restatement of one rigid-body World::step. TEST INFRASTRUCTURE ONLY (see
 * oracle.h). Pass order: crates/dynamis-world/tests/suite/pipeline.rs:6-49 (apply_commands,
 * prepare, emit_entries, broadphase, narrowphase, build_islands, wake, live, solver_prepare,
 * solve_substeps, sleep, commit, consume_streams). Each function names the WGSL it follows.
 *
 * Modelled: sphere, cuboid and capsule bodies and static planes; analytic pair contacts,
 * box-box SAT with clipping and plane-convex manifolds (narrowphase.wgsl); warm start from the
 * previous step's archive (contact_relay.wgsl) and from the resting archive (contact_begin.wgsl);
 * union-find islands, wake and sleep; the Jacobi mass-splitting contact solver with fixed-point
 * delta accumulation; position correction. Hulls and cylinders take the convex path (GJK, EPA,
 * the direction probe and the clipped feature manifolds of convex.wgsl and narrowphase.wgsl).
 *
 * Built with OpenMP (liboracle_omp.so, the all-core CPU baseline) the per-pair narrowphase and
 * the per-contact and per-row solver loops run in parallel: a pair writes only its own slot, a
 * solver block adds integer delta words (atomic adds; integer sums do not depend on order) and
 * raises the residual maxima and contribution counts atomically, and a row's passes touch only
 * that row, so the OpenMP build steps bit-identically to the serial one.
 *
 * Every collider sits in one grid region per step, its awake flag taken from the owner's activity
 * and wake flag (collider_entries.wgsl:9-35, immovable_entries.wgsl:8-29). The reference keeps
 * sleeping bodies in a separately rebuilt resting region; both give each collider one entry set
 * and the same pair set (a pair is emitted when the boxes overlap and either side is awake).
 * The resting archive is kept as a list of live frozen manifolds searched by pair key.
 */
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "omath.h"
#include "oracle.h"
#include "ostep.h"

#define PI_F 3.141592653589793f
#define FACE_AXIS_BIAS 1e-3f
#define CLIP_MARGIN 1e-4f
#define MANIFOLD_CANDIDATES 16u
#define FEATURE_MAX 8u
#define CONTACT_ANNOUNCED 0x80000000u
#define ISLAND_WAKE 1u
#define NO_SLOT 0xFFFFFFFFu

#define FEATURE_POINT 0u
#define FEATURE_VERTEX (1u << 28)
#define FEATURE_EDGE (2u << 28)
#define FEATURE_FACE (3u << 28)
#define FEATURE_KIND_MASK 0xF0000000u
#define FEATURE_FIELD_BITS 14u
#define FEATURE_FIELD_MASK ((1u << 14) - 1u)
#define FEATURE_CLIP (1u << 12)
#define FEATURE_FACE_BIT (1u << 13)

struct oracle_world {
    dyn_step_params p;
    uint32_t n, d, k;
    dyn_body_state* st;
    dyn_body_desc* de;
    dyn_collider* co;
    uint32_t* owner;
    uint32_t* row_of_body; /* body id -> row, ids < id_capacity */
    uint32_t id_capacity;
    dyn_shape_source* sources; /* hull shape sources and their float4 vertex rows */
    uint32_t source_count;
    float* vertices;
    uint32_t vertex_count;
    uint32_t* wake;
    uint32_t* admitted;
    uint32_t* activity;
    uint32_t* motion;
    uint32_t counters[DYN_COUNTER_BLOCK_WORDS];
    /* broadphase */
    uint32_t entry_cap, pair_cap;
    dyn_aabb* aabbs;
    dyn_aabb* res;
    uint32_t* flags;
    uint32_t* keys;
    uint32_t* order;
    dyn_grid_entry* entries;
    uint32_t* pmaj;
    uint32_t* pmin;
    uint32_t npairs;
    /* contacts */
    uint32_t contact_cap;
    dyn_contact* raw;
    uint32_t* valid;
    dyn_contact* contacts;
    uint32_t* matched;
    uint32_t ncontacts;
    dyn_contact* archive;
    uint32_t narchive;
    dyn_contact* resting;
    uint32_t nresting;
    /* islands and solver */
    uint32_t* parents;
    uint32_t* istate;
    uint32_t* live;
    uint32_t nlive;
    uint32_t* block_counts;
    uint32_t* blocks;
    uint32_t* rows;
    uint32_t nrows;
    float* targets;
    float* resolution;
    uint32_t* vdeltas;
    uint32_t* pdeltas;
    uint32_t* contrib;
    /* body edits queued for the next step (row, edit), in queue order */
    uint32_t* edit_rows;
    dyn_body_edit* edits;
    uint32_t nedits, edit_cap;
    /* capacity planning: each step's (pair, contact) demand, read two steps later */
    uint32_t demand[4][2];
    uint64_t issued;
    uint32_t grows;
    /* the last step's tape: start states, contacts as the solver received them, relay sources */
    dyn_body_state* start;
    dyn_contact* pre;
    uint32_t npre;
    uint32_t* relay_src;
};

/* ---- counters --------------------------------------------------------------------------- */

static uint32_t* ctr(oracle_world* w, uint32_t slot) {
    return &w->counters[slot * DYN_COUNTER_STRIDE_WORDS];
}

/* counter.rs:75-128: the slots whose reset is PerStep (reset_counters.wgsl) */
static const uint32_t kPerStep[] = {
    DYN_COUNTER_ENTRIES, DYN_COUNTER_RESTING_REBUILD, DYN_COUNTER_RESTING_SORT,
    DYN_COUNTER_IMMOVABLE_EMITTED, DYN_COUNTER_PAIRS, DYN_COUNTER_GRID_LEVELS,
    DYN_COUNTER_CONTACTS, DYN_COUNTER_JOINTS, DYN_COUNTER_EVENTS, DYN_COUNTER_REFUSED_PAIRS,
    DYN_COUNTER_REFUSED_EVENTS, DYN_COUNTER_ENTRY_FAULTS, DYN_COUNTER_REFUSED_RESTING,
    DYN_COUNTER_SLEPT, DYN_COUNTER_WOKE, DYN_COUNTER_ACTIVE, DYN_COUNTER_WOKE_DEFERRED,
    DYN_COUNTER_RESTING_GATHER, DYN_COUNTER_ISLANDS, DYN_COUNTER_ISLAND_FAULTS,
    DYN_COUNTER_SOLVE_LINEAR_RESIDUAL, DYN_COUNTER_SOLVE_ANGULAR_RESIDUAL,
    DYN_COUNTER_COARSE_ACTIVE, DYN_COUNTER_GRID_SCALE, DYN_COUNTER_GRID_EXTENT,
    DYN_COUNTER_BREAKS, DYN_COUNTER_PARTICLE_REACH, DYN_COUNTER_COARSE_NEIGHBOURS,
    DYN_COUNTER_LIVE, DYN_COUNTER_LIVE_FAULTS, DYN_COUNTER_ROW_FAULTS, DYN_COUNTER_SOFT_ACTIVE,
    DYN_COUNTER_SOFT_SLEPT, DYN_COUNTER_SOFT_WOKE, DYN_COUNTER_REFUSED_CONTACTS,
    DYN_COUNTER_IMPACTS, DYN_COUNTER_REFUSED_IMPACTS, DYN_COUNTER_SOFT_EVENTS,
    DYN_COUNTER_REFUSED_SOFT_EVENTS};

/* ---- bodies (core.wgsl:188-255) --------------------------------------------------------- */

typedef struct {
    dyn_body_state s;
    dyn_body_desc d;
} body_t;

static int kind_simulates(const dyn_body_desc* d) {
    return !(d->flags & DYN_BODY_KINEMATIC) && d->inverse_mass > 0.0f;
}
static int kind_moves(const dyn_body_desc* d) {
    return (d->flags & DYN_BODY_KINEMATIC) || d->inverse_mass > 0.0f;
}
static int body_is_inert(const body_t* b) { return !kind_simulates(&b->d) || b->s.sleeping != 0u; }
static void body_frozen(body_t* b) {
    b->d.inverse_mass = 0.0f;
    for (int i = 0; i < 6; i++) b->d.inverse_inertia[i] = 0.0f;
}
static int body_is_active(const dyn_body_state* s, const dyn_body_desc* d) {
    return kind_moves(d) && s->sleeping == 0u;
}
static int body_is_moving(const dyn_body_state* s, const dyn_body_desc* d,
                          const dyn_step_params* p) {
    int driven = kind_moves(d) && !kind_simulates(d);
    if (driven && (!vzero(vload(s->velocity)) || !vzero(vload(s->angular_velocity)))) return 1;
    if (!kind_simulates(d) || s->sleeping != 0u) return 0;
    float sv = (d->flags & DYN_OVERRIDE_SLEEP_LINEAR) ? d->sleep_velocity : p->sleep_velocity;
    float sw = (d->flags & DYN_OVERRIDE_SLEEP_ANGULAR) ? d->sleep_angular_velocity
                                                      : p->sleep_angular_velocity;
    return vlen(vload(s->velocity)) > sv || vlen(vload(s->angular_velocity)) > sw;
}
static void freeze_body(dyn_body_state* s) {
    memset(s->velocity, 0, 12);
    memset(s->angular_velocity, 0, 12);
    memset(s->force, 0, 12);
    memset(s->torque, 0, 12);
    memcpy(s->prev_position, s->position, 12);
    memcpy(s->prev_orientation, s->orientation, 16);
    s->sleep_timer = 0.0f;
    s->sleeping = 1u;
}
static v3 body_com(const body_t* b) {
    return vadd(vload(b->s.position), qrot(qload(b->s.orientation), vload(b->d.com)));
}
static v3 apply_inverse_inertia(const body_t* b, v3 v) {
    q4 q = qload(b->s.orientation);
    v3 local = qrot(qconj(q), v);
    return qrot(q, symmetric_apply(b->d.inverse_inertia, local));
}
static v3 point_velocity(const body_t* b, v3 point) {
    return vadd(vload(b->s.velocity), vcross(vload(b->s.angular_velocity), vsub(point, body_com(b))));
}
static v3 relative_velocity(const body_t* a, const body_t* b, v3 pa, v3 pb) {
    return vsub(point_velocity(b, pb), point_velocity(a, pa));
}
static body_t load_body(const oracle_world* w, uint32_t row) {
    body_t b;
    b.s = w->st[row];
    b.d = w->de[row];
    return b;
}

/* ---- shapes (core.wgsl:473-536) --------------------------------------------------------- */

typedef struct {
    uint32_t kind;
    float radius, half_height;
    v3 center, half_extents, scale;
    q4 rotation;
    uint32_t source;
} world_shape;

static world_shape world_collider(const dyn_body_state* s, const dyn_collider* c) {
    world_shape ws;
    ws.kind = c->kind;
    ws.radius = c->radius;
    ws.half_height = c->half_height;
    q4 q = qload(s->orientation);
    ws.center = vadd(vload(s->position), qrot(q, vload(c->local_offset)));
    ws.half_extents = vload(c->half_extents);
    ws.rotation = qmul(q, qload(c->local_rotation));
    ws.scale = vload(c->scale);
    ws.source = c->source;
    return ws;
}

static v3 shape_axis(const world_shape* ws) { return qrot(ws->rotation, V(0.0f, 1.0f, 0.0f)); }

/* core.wgsl:490-503: bounds of a shape that reads a shape source */
static void world_shape_bounds(const oracle_world* w, const world_shape* ws, v3* mn, v3* mx) {
    const dyn_shape_source* src = &w->sources[ws->source];
    v3 scaled_min = vmulv(vload(src->local_min), ws->scale);
    v3 scaled_max = vmulv(vload(src->local_max), ws->scale);
    v3 center = vadd(ws->center, qrot(ws->rotation, vmul(vadd(scaled_min, scaled_max), 0.5f)));
    v3 half = vmul(vsub(scaled_max, scaled_min), 0.5f);
    v3 extent = vadd(vadd(vabs(qrot(ws->rotation, V(half.x, 0.0f, 0.0f))),
                          vabs(qrot(ws->rotation, V(0.0f, half.y, 0.0f)))),
                     vabs(qrot(ws->rotation, V(0.0f, 0.0f, half.z))));
    *mn = vsub(center, extent);
    *mx = vadd(center, extent);
}

/* core.wgsl:505-536 */
static void world_aabb_of(const oracle_world* w, const world_shape* ws, v3* mn, v3* mx) {
    if (ws->kind == DYN_SHAPE_PLANE) {
        *mn = vsub(ws->center, V(1e6f, 1e6f, 1e6f));
        *mx = vadd(ws->center, V(1e6f, 1e6f, 1e6f));
        return;
    }
    if (ws->kind == DYN_SHAPE_HULL) {
        world_shape_bounds(w, ws, mn, mx);
        return;
    }
    v3 extent = V(0.0f, 0.0f, 0.0f);
    if (ws->kind == DYN_SHAPE_SPHERE) {
        extent = V(ws->radius, ws->radius, ws->radius);
    } else if (ws->kind == DYN_SHAPE_CUBOID) {
        v3 e = ws->half_extents;
        extent = vadd(vadd(vabs(qrot(ws->rotation, V(e.x, 0.0f, 0.0f))),
                           vabs(qrot(ws->rotation, V(0.0f, e.y, 0.0f)))),
                      vabs(qrot(ws->rotation, V(0.0f, 0.0f, e.z))));
    } else if (ws->kind == DYN_SHAPE_CAPSULE || ws->kind == DYN_SHAPE_CYLINDER) {
        v3 axis = shape_axis(ws);
        extent = vadd(vabs(vmul(axis, ws->half_height)), V(ws->radius, ws->radius, ws->radius));
    }
    float s = fmaxo(fmaxo(ws->scale.x, ws->scale.y), ws->scale.z);
    extent = vmul(extent, s);
    *mn = vsub(ws->center, extent);
    *mx = vadd(ws->center, extent);
}

/* ---- manifold helpers (core.wgsl:615-760) ----------------------------------------------- */

static uint32_t feature_pair(uint32_t kind, uint32_t a, uint32_t b) {
    return kind | (a << FEATURE_FIELD_BITS) | b;
}
static uint32_t feature_mirror(uint32_t f) {
    uint32_t kind = f & FEATURE_KIND_MASK;
    if (kind == FEATURE_POINT) return f;
    uint32_t a = (f >> FEATURE_FIELD_BITS) & FEATURE_FIELD_MASK;
    uint32_t b = f & FEATURE_FIELD_MASK;
    return kind | (b << FEATURE_FIELD_BITS) | a;
}

typedef struct {
    v3 position;
    float depth;
    uint32_t feature;
} candidate_t;

static void manifold_emit(dyn_contact* c, v3 normal) {
    c->point_count = 0;
    vstore(c->normal, normal);
    c->triangle = 0xFFFFFFFFu;
}

static void manifold_push(dyn_contact* c, v3 point, float depth, uint32_t feature) {
    uint32_t count = c->point_count;
    if (count >= DYN_CONTACT_MAX_POINTS) return;
    dyn_manifold_point* mp = &c->points[count];
    memset(mp, 0, sizeof *mp);
    vstore(mp->position, point);
    mp->depth = depth;
    mp->feature = feature;
    c->point_count = count + 1u;
}

static float manifold_gap(const candidate_t* cand, uint32_t kept, uint32_t index) {
    float gap = 3.402823466e38f;
    for (uint32_t h = 0; h < kept; h++) gap = fmino(gap, vlen(vsub(cand[index].position, cand[h].position)));
    return gap;
}

/* core.wgsl:720-747 */
static void manifold_order(candidate_t* cand, uint32_t count) {
    uint32_t lead = 0;
    for (uint32_t i = 1; i < count; i++) {
        float depth = cand[i].depth, held = cand[lead].depth;
        if (depth > held || (depth == held && cand[i].feature < cand[lead].feature)) lead = i;
    }
    candidate_t t = cand[0];
    cand[0] = cand[lead];
    cand[lead] = t;
    for (uint32_t kept = 1; kept < count; kept++) {
        uint32_t best = kept;
        float best_gap = manifold_gap(cand, kept, kept);
        for (uint32_t i = kept + 1; i < count; i++) {
            float gap = manifold_gap(cand, kept, i);
            if (gap > best_gap) {
                best_gap = gap;
                best = i;
            }
        }
        if (best != kept) {
            t = cand[kept];
            cand[kept] = cand[best];
            cand[best] = t;
        }
    }
}

/* core.wgsl:750-760 */
static void manifold_keep(dyn_contact* c, candidate_t* cand, uint32_t count) {
    manifold_order(cand, count);
    for (uint32_t i = 0; i < count && c->point_count < DYN_CONTACT_MAX_POINTS; i++) {
        manifold_push(c, cand[i].position, cand[i].depth, cand[i].feature);
    }
}

/* ---- analytic pairs (narrowphase.wgsl:33-222) ------------------------------------------- */

static v3 box_center(const dyn_body_state* s, const dyn_collider* c) {
    return vadd(vload(s->position), qrot(qload(s->orientation), vload(c->local_offset)));
}

static float pair_margin(const oracle_world* w, const body_t* a, const body_t* b) {
    if (!(a->d.flags & DYN_BODY_CCD) && !(b->d.flags & DYN_BODY_CCD)) return w->p.contact_margin;
    float closing = vlen(vsub(vload(b->s.velocity), vload(a->s.velocity))) * w->p.dt;
    return w->p.contact_margin + closing;
}

static dyn_contact sphere_sphere(const body_t* a, const dyn_collider* ca, const body_t* b,
                                 const dyn_collider* cb, float margin) {
    dyn_contact c;
    memset(&c, 0, sizeof c);
    v3 ac = box_center(&a->s, ca), bc = box_center(&b->s, cb);
    v3 delta = vsub(bc, ac);
    float distance = vlen(delta);
    v3 normal = sign_normalize(delta);
    if (distance <= 1e-6f) {
        v3 rel = relative_velocity(a, b, bc, ac);
        if (vlen(rel) > 1e-6f) normal = vneg(vnorm(rel));
    }
    manifold_emit(&c, normal);
    float radius_sum = ca->radius + cb->radius;
    if (distance > radius_sum + margin) return c;
    float depth = radius_sum - distance;
    v3 point = vadd(ac, vmul(normal, ca->radius - depth * 0.5f));
    manifold_push(&c, point, depth, FEATURE_POINT);
    return c;
}

static uint32_t smallest_axis(v3 v) {
    v3 av = vabs(v);
    if (av.x < av.y && av.x < av.z) return 0;
    if (av.y < av.z) return 1;
    return 2;
}

static v3 closest_point_box(v3 point, const dyn_body_state* s, const dyn_collider* c) {
    q4 q = qmul(qload(s->orientation), qload(c->local_rotation));
    v3 center = box_center(s, c);
    v3 local = qrot(qconj(q), vsub(point, center));
    v3 he = vload(c->half_extents);
    v3 clamped = V(fclamp(local.x, -he.x, he.x), fclamp(local.y, -he.y, he.y),
                   fclamp(local.z, -he.z, he.z));
    return vadd(box_center(s, c), qrot(q, clamped));
}

static dyn_contact sphere_box(const body_t* sp, const dyn_collider* sc, const body_t* bx,
                              const dyn_collider* bc, float margin) {
    dyn_contact c;
    memset(&c, 0, sizeof c);
    v3 center = vadd(vload(sp->s.position), qrot(qload(sp->s.orientation), vload(sc->local_offset)));
    v3 closest = closest_point_box(center, &bx->s, bc);
    v3 delta = vsub(closest, center);
    float distance = vlen(delta);
    float radius = sc->radius;
    v3 normal = sign_normalize(delta);
    float depth = radius - distance;
    if (distance <= 1e-6f) {
        /* box_nearest_face, narrowphase.wgsl:67-85 */
        q4 q = qmul(qload(bx->s.orientation), qload(bc->local_rotation));
        v3 local = qrot(qconj(q), vsub(center, box_center(&bx->s, bc)));
        v3 clearance = vsub(vload(bc->half_extents), vabs(local));
        uint32_t axis = smallest_axis(clearance);
        v3 facing;
        if (axis == 0) facing = V(local.x > 0.0f ? -1.0f : 1.0f, 0.0f, 0.0f);
        else if (axis == 1) facing = V(0.0f, local.y > 0.0f ? -1.0f : 1.0f, 0.0f);
        else facing = V(0.0f, 0.0f, local.z > 0.0f ? -1.0f : 1.0f);
        v3 face_normal = vneg(qrot(q, facing));
        normal = vneg(face_normal);
        depth = radius + vcomp(clearance, axis);
    }
    manifold_emit(&c, normal);
    if (depth <= -margin) return c;
    v3 point = vsub(closest, vmul(normal, depth * 0.5f));
    manifold_push(&c, point, depth, FEATURE_POINT);
    return c;
}

static void capsule_segment(const body_t* b, const dyn_collider* c, v3* s0, v3* s1) {
    q4 q = qmul(qload(b->s.orientation), qload(c->local_rotation));
    v3 axis = qrot(q, V(0.0f, 1.0f, 0.0f));
    v3 center = box_center(&b->s, c);
    *s0 = vsub(center, vmul(axis, c->half_height));
    *s1 = vadd(center, vmul(axis, c->half_height));
}

static v3 closest_point_segment(v3 point, v3 a, v3 b) {
    v3 ab = vsub(b, a);
    float denom = vdot(ab, ab);
    if (denom < 1e-10f) return a;
    float t = fclamp(vdot(vsub(point, a), ab) / denom, 0.0f, 1.0f);
    return vadd(a, vmul(ab, t));
}

static dyn_contact sphere_capsule(const body_t* sp, const dyn_collider* sc, const body_t* cp,
                                  const dyn_collider* cc, float margin) {
    dyn_contact c;
    memset(&c, 0, sizeof c);
    v3 s0, s1;
    capsule_segment(cp, cc, &s0, &s1);
    v3 center = vadd(vload(sp->s.position), qrot(qload(sp->s.orientation), vload(sc->local_offset)));
    v3 closest = closest_point_segment(center, s0, s1);
    v3 delta = vsub(closest, center);
    float distance = vlen(delta);
    float radius_sum = sc->radius + cc->radius;
    manifold_emit(&c, sign_normalize(delta));
    if (distance > radius_sum + margin) return c;
    v3 normal = sign_normalize(delta);
    float depth = radius_sum - distance;
    v3 point = vsub(closest, vmul(normal, depth * 0.5f));
    manifold_push(&c, point, depth, FEATURE_POINT);
    return c;
}

static void closest_points_segments(v3 a0, v3 a1, v3 b0, v3 b1, v3* pa, v3* pb) {
    v3 d1 = vsub(a1, a0), d2 = vsub(b1, b0), r = vsub(a0, b0);
    float a = vdot(d1, d1), e = vdot(d2, d2), f = vdot(d2, r);
    float s = 0.0f, t = 0.0f;
    if (a <= 1e-10f && e <= 1e-10f) {
        s = 0.0f;
        t = 0.0f;
    } else if (a <= 1e-10f) {
        s = 0.0f;
        t = fclamp(f / e, 0.0f, 1.0f);
    } else {
        float cc = vdot(d1, r);
        if (e <= 1e-10f) {
            t = 0.0f;
            s = fclamp(-cc / a, 0.0f, 1.0f);
        } else {
            float b = vdot(d1, d2);
            float denom = a * e - b * b;
            if (denom != 0.0f) s = fclamp((b * f - cc * e) / denom, 0.0f, 1.0f);
            else s = 0.0f;
            t = (b * s + f) / e;
            if (t < 0.0f) {
                t = 0.0f;
                s = fclamp(-cc / a, 0.0f, 1.0f);
            } else if (t > 1.0f) {
                t = 1.0f;
                s = fclamp((b - cc) / a, 0.0f, 1.0f);
            }
        }
    }
    *pa = vadd(a0, vmul(d1, s));
    *pb = vadd(b0, vmul(d2, t));
}

static dyn_contact capsule_capsule(const body_t* a, const dyn_collider* ca, const body_t* b,
                                   const dyn_collider* cb, float margin) {
    dyn_contact c;
    memset(&c, 0, sizeof c);
    v3 a0, a1, b0, b1, pa, pb;
    capsule_segment(a, ca, &a0, &a1);
    capsule_segment(b, cb, &b0, &b1);
    closest_points_segments(a0, a1, b0, b1, &pa, &pb);
    v3 delta = vsub(pb, pa);
    float distance = vlen(delta);
    float radius_sum = ca->radius + cb->radius;
    manifold_emit(&c, sign_normalize(delta));
    if (distance > radius_sum + margin) return c;
    float depth = radius_sum - distance;
    v3 point = vmul(vadd(pa, pb), 0.5f);
    manifold_push(&c, point, depth, FEATURE_POINT);
    return c;
}

/* ---- box-box SAT + clipping (narrowphase.wgsl:224-473) ---------------------------------- */

typedef struct {
    v3 axes[3];
    v3 half;
    v3 center;
} box_geom;

static box_geom box_geometry(const body_t* b, const dyn_collider* c) {
    box_geom g;
    q4 q = qmul(qload(b->s.orientation), qload(c->local_rotation));
    g.axes[0] = qrot(q, V(1.0f, 0.0f, 0.0f));
    g.axes[1] = qrot(q, V(0.0f, 1.0f, 0.0f));
    g.axes[2] = qrot(q, V(0.0f, 0.0f, 1.0f));
    g.half = vload(c->half_extents);
    g.center = box_center(&b->s, c);
    return g;
}

static float support_radius(const box_geom* g, v3 axis) {
    return fabsf(vdot(g->axes[0], axis)) * g->half.x + fabsf(vdot(g->axes[1], axis)) * g->half.y +
           fabsf(vdot(g->axes[2], axis)) * g->half.z;
}

static float face_overlap(const box_geom* face, const box_geom* other, uint32_t i, v3 delta) {
    v3 axis = face->axes[i];
    return vcomp(face->half, i) + support_radius(other, axis) - fabsf(vdot(delta, axis));
}

static float edge_overlap(const box_geom* l, const box_geom* r, v3 axis, v3 delta) {
    return support_radius(l, axis) + support_radius(r, axis) - fabsf(vdot(delta, axis));
}

static v3 box_support(const box_geom* g, v3 direction) {
    v3 point = g->center;
    for (uint32_t i = 0; i < 3; i++) {
        float h = vcomp(g->half, i);
        point = vadd(point, vmul(g->axes[i], vdot(g->axes[i], direction) > 0.0f ? h : -h));
    }
    return point;
}

/* convex.wgsl:625-631 */
static uint32_t box_corner_id(uint32_t face_axis, float face_sign, float first, float second) {
    uint32_t id = 0;
    id |= face_sign > 0.0f ? (1u << face_axis) : 0u;
    id |= first > 0.0f ? (1u << ((face_axis + 1u) % 3u)) : 0u;
    id |= second > 0.0f ? (1u << ((face_axis + 2u) % 3u)) : 0u;
    return id;
}

/* narrowphase.wgsl:232-270 */
static v3 box_face(const box_geom* g, v3 face_normal, v3 corners[4], uint32_t ids[4],
                   uint32_t* out_face) {
    uint32_t axis = 0;
    for (uint32_t i = 1; i < 3; i++) {
        if (fabsf(vdot(face_normal, g->axes[i])) > fabsf(vdot(face_normal, g->axes[axis]))) axis = i;
    }
    float sign = vdot(face_normal, g->axes[axis]) < 0.0f ? -1.0f : 1.0f;
    v3 n = vmul(g->axes[axis], sign);
    v3 u = g->axes[(axis + 1u) % 3u], v = g->axes[(axis + 2u) % 3u];
    float hu = vcomp(g->half, (axis + 1u) % 3u), hv = vcomp(g->half, (axis + 2u) % 3u);
    v3 fc = vadd(g->center, vmul(n, vcomp(g->half, axis)));
    v3 e0 = vadd(vmul(u, hu), vmul(v, hv));
    v3 e1 = vadd(vmul(vneg(u), hu), vmul(v, hv));
    v3 e2 = vsub(vmul(vneg(u), hu), vmul(v, hv));
    v3 e3 = vsub(vmul(u, hu), vmul(v, hv));
    corners[0] = vadd(fc, e0);
    corners[1] = vadd(fc, e1);
    corners[2] = vadd(fc, e2);
    corners[3] = vadd(fc, e3);
    ids[0] = box_corner_id(axis, sign, 1.0f, 1.0f);
    ids[1] = box_corner_id(axis, sign, -1.0f, 1.0f);
    ids[2] = box_corner_id(axis, sign, -1.0f, -1.0f);
    ids[3] = box_corner_id(axis, sign, 1.0f, -1.0f);
    *out_face = axis * 2u + (sign > 0.0f ? 1u : 0u);
    return fc;
}

/* narrowphase.wgsl:272-305 */
static uint32_t clip_polygon(const v3* pts, const uint32_t* ids, uint32_t count, v3 plane_point,
                             v3 plane_normal, uint32_t group, uint32_t side, v3* out_pts,
                             uint32_t* out_ids) {
    uint32_t out = 0;
    for (uint32_t i = 0; i < count; i++) {
        v3 cur = pts[i], nxt = pts[(i + 1u) % count];
        float cd = vdot(vsub(cur, plane_point), plane_normal) - CLIP_MARGIN;
        float nd = vdot(vsub(nxt, plane_point), plane_normal) - CLIP_MARGIN;
        if (cd <= 0.0f && out < 8u) {
            out_pts[out] = cur;
            out_ids[out] = ids[i];
            out++;
        }
        if (cd * nd < 0.0f) {
            float t = cd / (cd - nd);
            if (out < 8u) {
                out_pts[out] = vadd(cur, vmul(vsub(nxt, cur), t));
                out_ids[out] = FEATURE_CLIP | (group * 32u + side * 8u + i);
                out++;
            }
        }
    }
    return out;
}

static dyn_contact box_box_sat(const body_t* a, const dyn_collider* ca, const body_t* b,
                               const dyn_collider* cb, float margin) {
    dyn_contact c;
    memset(&c, 0, sizeof c);
    box_geom left = box_geometry(a, ca), right = box_geometry(b, cb);
    v3 delta = vsub(right.center, left.center);
    float face_depth = 1e30f;
    v3 face_axis = left.axes[0];
    int face_on_right = 0;
    for (uint32_t i = 0; i < 3; i++) {
        float overlap = face_overlap(&left, &right, i, delta);
        if (overlap < face_depth) {
            face_depth = overlap;
            face_axis = left.axes[i];
            face_on_right = 0;
        }
        float other = face_overlap(&right, &left, i, delta);
        if (other < face_depth) {
            face_depth = other;
            face_axis = right.axes[i];
            face_on_right = 1;
        }
    }
    float edge_depth = 1e30f;
    v3 edge_axis = face_axis;
    uint32_t ef = 0, es = 0;
    for (uint32_t i = 0; i < 3; i++) {
        for (uint32_t j = 0; j < 3; j++) {
            v3 crossed = vcross(left.axes[i], right.axes[j]);
            float len = vlen(crossed);
            if (len < 1e-8f) continue;
            v3 axis = vdiv(crossed, len);
            float overlap = edge_overlap(&left, &right, axis, delta);
            if (overlap < edge_depth) {
                edge_depth = overlap;
                edge_axis = axis;
                ef = i;
                es = j;
            }
        }
    }
    float extent = fmaxo(fmaxo(left.half.x, left.half.y), left.half.z) +
                   fmaxo(fmaxo(right.half.x, right.half.y), right.half.z);
    int edge_sep = edge_depth + FACE_AXIS_BIAS * extent < face_depth;
    float depth = edge_sep ? edge_depth : face_depth;
    if (depth <= -margin) return c;
    v3 axis = edge_sep ? edge_axis : face_axis;
    v3 signed_axis = vdot(axis, delta) < 0.0f ? vneg(axis) : axis;
    manifold_emit(&c, signed_axis);
    if (edge_sep) {
        v3 point = vmul(vadd(box_support(&left, signed_axis), box_support(&right, vneg(signed_axis))), 0.5f);
        manifold_push(&c, point, depth, feature_pair(FEATURE_EDGE, ef, es));
        return c;
    }
    const box_geom* reference = face_on_right ? &right : &left;
    const box_geom* incident = face_on_right ? &left : &right;
    v3 ref_normal = face_on_right ? vneg(signed_axis) : signed_axis;
    v3 ref_corners[4];
    uint32_t ref_ids[4], ref_face = 0;
    v3 ref_center = box_face(reference, ref_normal, ref_corners, ref_ids, &ref_face);
    uint32_t ia = 0;
    for (uint32_t i = 1; i < 3; i++) {
        if (fabsf(vdot(incident->axes[i], ref_normal)) > fabsf(vdot(incident->axes[ia], ref_normal))) ia = i;
    }
    v3 inc_normal = vmul(incident->axes[ia], vdot(incident->axes[ia], ref_normal) > 0.0f ? -1.0f : 1.0f);
    v3 inc_corners[4];
    uint32_t inc_ids[4], inc_face = 0;
    box_face(incident, inc_normal, inc_corners, inc_ids, &inc_face);
    uint32_t ref_field = FEATURE_FACE_BIT | ref_face;
    v3 pa[8], pb[8];
    uint32_t ida[8], idb[8];
    for (uint32_t i = 0; i < 4; i++) {
        pa[i] = inc_corners[i];
        ida[i] = inc_ids[i];
    }
    uint32_t pc = 4;
    for (uint32_t side = 0; side < 4; side++) {
        v3 cur = ref_corners[side], nxt = ref_corners[(side + 1u) % 4u];
        v3 plane_normal = vnorm(vcross(vsub(nxt, cur), ref_normal));
        if (side % 2u == 0u) pc = clip_polygon(pa, ida, pc, cur, plane_normal, inc_face, side, pb, idb);
        else pc = clip_polygon(pb, idb, pc, cur, plane_normal, inc_face, side, pa, ida);
        if (pc == 0u) break;
    }
    candidate_t cand[MANIFOLD_CANDIDATES];
    uint32_t cc = 0;
    for (uint32_t i = 0; i < pc && cc < MANIFOLD_CANDIDATES; i++) {
        float pd = vdot(vsub(ref_center, pa[i]), ref_normal);
        if (pd >= -margin) {
            cand[cc].position = vsub(pa[i], vmul(ref_normal, pd * 0.5f));
            cand[cc].depth = pd;
            cand[cc].feature = face_on_right ? feature_pair(FEATURE_FACE, ida[i], ref_field)
                                             : feature_pair(FEATURE_FACE, ref_field, ida[i]);
            cc++;
        }
    }
    if (cc == 0u) {
        for (uint32_t i = 0; i < 4 && cc < MANIFOLD_CANDIDATES; i++) {
            float pd = vdot(vsub(ref_center, inc_corners[i]), ref_normal);
            if (pd >= -margin) {
                cand[cc].position = vsub(inc_corners[i], vmul(ref_normal, pd * 0.5f));
                cand[cc].depth = pd;
                cand[cc].feature = face_on_right ? feature_pair(FEATURE_FACE, inc_ids[i], ref_field)
                                                 : feature_pair(FEATURE_FACE, ref_field, inc_ids[i]);
                cc++;
            }
        }
    }
    if (cc == 0u) {
        manifold_push(&c, vsub(ref_center, vmul(ref_normal, face_depth * 0.5f)), face_depth, FEATURE_POINT);
        return c;
    }
    manifold_keep(&c, cand, cc);
    return c;
}

/* ---- shape sources (records.wgsl:256-273) and support maps (convex.wgsl:1-91) --------------- */

#define LINEAR_SUPPORT_VERTICES 4096u
#define CAP_ALIGNMENT 0.93f
#define CYLINDER_CAP_POINTS 8u
#define CYLINDER_SIDE_POINTS 16u
#define CYLINDER_SIDE_FACE 2u
#define FEATURE_INDEX_LIMIT (1u << 12)
#define EMPTY_RING_FACE FEATURE_INDEX_LIMIT
#define NO_TRIANGLE 0xFFFFFFFFu
#define FLT_BIG 3.402823466e38f

/* cos / sin of f32(i) * 0.78539816 for the eight cylinder cap points (convex.wgsl:836-840),
 * correctly rounded to FP32 */
static const float kCapCos[8] = {0x1.0p+0f, 0x1.6a09e6p-1f, -0x1.777a5cp-25f, -0x1.6a09e6p-1f,
                                 -0x1.0p+0f, -0x1.6a09e2p-1f, 0x1.99bc5cp-27f, 0x1.6a09eep-1f};
static const float kCapSin[8] = {0x0.0p+0f, 0x1.6a09e6p-1f, 0x1.0p+0f, 0x1.6a09e6p-1f,
                                 -0x1.777a5cp-24f, -0x1.6a09eap-1f, -0x1.0p+0f, -0x1.6a09dep-1f};

static v3 hull_vertex(const oracle_world* w, const dyn_shape_source* src, uint32_t i) {
    const float* v = &w->vertices[4u * (size_t)(src->vertex_offset + i)];
    return V(v[0], v[1], v[2]);
}

static v3 support_local(const oracle_world* w, const world_shape* ws, v3 d) {
    v3 n = vnorm(d);
    if (ws->kind == DYN_SHAPE_SPHERE) return vmul(n, ws->radius);
    if (ws->kind == DYN_SHAPE_CUBOID) {
        v3 s = V(d.x > 0.0f ? 1.0f : -1.0f, d.y > 0.0f ? 1.0f : -1.0f, d.z > 0.0f ? 1.0f : -1.0f);
        return vmulv(s, ws->half_extents);
    }
    if (ws->kind == DYN_SHAPE_CAPSULE) {
        v3 tip = vmul(vmul(V(0.0f, 1.0f, 0.0f), ws->half_height), d.y > 0.0f ? 1.0f : -1.0f);
        return vadd(tip, vmul(n, ws->radius));
    }
    if (ws->kind == DYN_SHAPE_CYLINDER) {
        v3 base = vmul(vmul(V(0.0f, 1.0f, 0.0f), ws->half_height), d.y > 0.0f ? 1.0f : -1.0f);
        v3 radial = vsub(d, V(0.0f, d.y, 0.0f));
        if (vlen(radial) < 1e-6f) return base;
        return vadd(base, vmul(vnorm(radial), ws->radius));
    }
    if (ws->kind != DYN_SHAPE_HULL || ws->source >= w->source_count) return V(0.0f, 0.0f, 0.0f);
    /* the linear support scan (hulls above LINEAR_SUPPORT_VERTICES are refused at upload) */
    const dyn_shape_source* src = &w->sources[ws->source];
    float best_dot = -FLT_BIG;
    v3 best = V(0.0f, 0.0f, 0.0f);
    for (uint32_t i = 0; i < src->vertex_count; i++) {
        v3 v = hull_vertex(w, src, i);
        float s = vdot(v, d);
        if (s > best_dot) {
            best_dot = s;
            best = v;
        }
    }
    return best;
}

static v3 support(const oracle_world* w, const world_shape* ws, v3 direction) {
    v3 ld = qrot(qconj(ws->rotation), direction);
    v3 sd = vmulv(ld, ws->scale);
    v3 local = support_local(w, ws, sd);
    return vadd(ws->center, qrot(ws->rotation, vmulv(local, ws->scale)));
}

/* ---- GJK (convex.wgsl:93-282, 439-521) ----------------------------------------------------- */

typedef struct { v3 w, a, b; } simplex_point;

typedef struct {
    v3 v;
    float lambdas[4];
    int penetrating;
    v3 point_a, point_b;
} simplex_result;

typedef struct {
    float distance;
    v3 point_a, point_b, normal;
    int penetrating;
} convex_closest_t;

typedef struct {
    float distance;
    v3 point, normal;
    uint32_t triangle;
} shape_hit;

typedef struct {
    int valid;
    v3 normal;
    float depth;
    v3 point;
} epa_result;

static v3 closest_on_triangle(v3 origin, v3 a, v3 b, v3 c, v3* weights) {
    v3 ab = vsub(b, a), ac = vsub(c, a), ap = vsub(origin, a);
    float d1 = vdot(ab, ap), d2 = vdot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) {
        *weights = V(1.0f, 0.0f, 0.0f);
        return a;
    }
    v3 bp = vsub(origin, b);
    float d3 = vdot(ab, bp), d4 = vdot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) {
        *weights = V(0.0f, 1.0f, 0.0f);
        return b;
    }
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        float t = d1 / (d1 - d3);
        *weights = V(1.0f - t, t, 0.0f);
        return vadd(a, vmul(ab, t));
    }
    v3 cp = vsub(origin, c);
    float d5 = vdot(ab, cp), d6 = vdot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) {
        *weights = V(0.0f, 0.0f, 1.0f);
        return c;
    }
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        float t = d2 / (d2 - d6);
        *weights = V(1.0f - t, 0.0f, t);
        return vadd(a, vmul(ac, t));
    }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        float t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        *weights = V(0.0f, 1.0f - t, t);
        return vadd(b, vmul(vsub(c, b), t));
    }
    float denom = 1.0f / (va + vb + vc);
    float v = vb * denom, wgt = vc * denom;
    *weights = V(1.0f - v - wgt, v, wgt);
    return vadd(vadd(a, vmul(ab, v)), vmul(ac, wgt));
}

static float face_span(v3 a, v3 b, v3 c) {
    return fmaxo(fmaxo(vlen(vsub(b, a)), vlen(vsub(c, a))), vlen(vsub(c, b)));
}

/* the four faces of a tetrahedron, in the reference's order */
static const uint32_t kTetFaces[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {1, 3, 2}};

static int tetra_contains(const simplex_point* tet, v3 origin) {
    for (uint32_t f = 0; f < 4; f++) {
        v3 a = tet[kTetFaces[f][0]].w, b = tet[kTetFaces[f][1]].w, c = tet[kTetFaces[f][2]].w;
        v3 normal = vcross(vsub(b, a), vsub(c, a));
        float span = face_span(a, b, c);
        if (vlen(normal) <= 1e-6f * span * span) return 0;
        if (vdot(normal, vsub(origin, a)) > 0.0f) return 0;
    }
    return 1;
}

static int simplex_holds(const simplex_point* s, uint32_t count, v3 point, float span) {
    float tolerance = (1e-5f * span) * (1e-5f * span);
    for (uint32_t i = 0; i < count; i++) {
        v3 offset = vsub(s[i].w, point);
        if (vdot(offset, offset) <= tolerance) return 1;
    }
    return 0;
}

static uint32_t simplex_reduce(simplex_point* s, const float* lambdas) {
    simplex_point kept[4];
    uint32_t count = 0;
    for (uint32_t i = 0; i < 4; i++) {
        if (lambdas[i] > 1e-6f) kept[count++] = s[i];
    }
    if (count == 0u) {
        kept[0] = s[0];
        count = 1;
    }
    for (uint32_t i = 0; i < count; i++) s[i] = kept[i];
    return count;
}

static simplex_result simplex_closest(const simplex_point* s, uint32_t count) {
    simplex_result r;
    r.v = V(FLT_BIG, FLT_BIG, FLT_BIG);
    for (int i = 0; i < 4; i++) r.lambdas[i] = 0.0f;
    r.penetrating = count == 4u && tetra_contains(s, V(0.0f, 0.0f, 0.0f));
    r.point_a = V(0.0f, 0.0f, 0.0f);
    r.point_b = V(0.0f, 0.0f, 0.0f);
    if (r.penetrating) {
        r.v = V(0.0f, 0.0f, 0.0f);
        return r;
    }
    for (uint32_t i = 0; i < count; i++) {
        float d = vlen(s[i].w);
        if (d < vlen(r.v)) {
            r.v = s[i].w;
            for (int k = 0; k < 4; k++) r.lambdas[k] = 0.0f;
            r.lambdas[i] = 1.0f;
        }
    }
    for (uint32_t i = 0; i < count; i++) {
        for (uint32_t j = i + 1; j < count; j++) {
            v3 a = s[i].w, b = s[j].w;
            v3 ab = vsub(b, a);
            float denom = vdot(ab, ab);
            if (denom < 1e-12f) continue;
            float t = fclamp(-vdot(a, ab) / denom, 0.0f, 1.0f);
            v3 point = vadd(a, vmul(ab, t));
            if (vlen(point) < vlen(r.v)) {
                r.v = point;
                for (int k = 0; k < 4; k++) r.lambdas[k] = 0.0f;
                r.lambdas[i] = 1.0f - t;
                r.lambdas[j] = t;
            }
        }
    }
    if (count >= 3u) {
        for (uint32_t i = 0; i < count; i++) {
            for (uint32_t j = i + 1; j < count; j++) {
                for (uint32_t k = j + 1; k < count; k++) {
                    v3 weights;
                    v3 point = closest_on_triangle(V(0.0f, 0.0f, 0.0f), s[i].w, s[j].w, s[k].w, &weights);
                    if (vlen(point) < vlen(r.v)) {
                        r.v = point;
                        for (int q = 0; q < 4; q++) r.lambdas[q] = 0.0f;
                        r.lambdas[i] = weights.x;
                        r.lambdas[j] = weights.y;
                        r.lambdas[k] = weights.z;
                    }
                }
            }
        }
    }
    for (uint32_t i = 0; i < count; i++) {
        r.point_a = vadd(r.point_a, smul(r.lambdas[i], s[i].a));
        r.point_b = vadd(r.point_b, smul(r.lambdas[i], s[i].b));
    }
    return r;
}

static simplex_point support_pair(const oracle_world* w, const world_shape* a, const world_shape* b, v3 dir) {
    simplex_point p;
    p.a = support(w, a, dir);
    p.b = support(w, b, vneg(dir));
    p.w = vsub(p.a, p.b);
    return p;
}

static float shape_scale(const world_shape* ws) {
    float scale = ws->radius + ws->half_height;
    scale = fmaxo(scale, fmaxo(fmaxo(ws->half_extents.x, ws->half_extents.y), ws->half_extents.z));
    return fmaxo(scale, 1e-4f) * fmaxo(fmaxo(ws->scale.x, ws->scale.y), ws->scale.z);
}

static v3 simplex_expand_dir(const simplex_point* s, uint32_t count) {
    if (count == 1u) return V(1.0f, 0.0f, 0.0f);
    if (count == 2u) {
        v3 edge = vsub(s[1].w, s[0].w);
        v3 d = vcross(edge, V(1.0f, 0.0f, 0.0f));
        if (vlen(d) < 1e-6f) d = vcross(edge, V(0.0f, 1.0f, 0.0f));
        return vnorm(d);
    }
    v3 normal = vcross(vsub(s[1].w, s[0].w), vsub(s[2].w, s[0].w));
    if (vlen(normal) > 1e-8f) return vnorm(normal);
    return V(0.0f, 1.0f, 0.0f);
}

static convex_closest_t convex_closest(const oracle_world* w, const world_shape* first,
                                       const world_shape* second, simplex_point* simplex,
                                       uint32_t* out_count) {
    uint32_t count = 0;
    convex_closest_t r;
    r.distance = 0.0f;
    r.point_a = first->center;
    r.point_b = second->center;
    r.normal = sign_normalize(vsub(second->center, first->center));
    r.penetrating = 0;
    v3 dir = vsub(second->center, first->center);
    if (vlen(dir) < 1e-8f) dir = V(1.0f, 0.0f, 0.0f);
    simplex[0] = support_pair(w, first, second, dir);
    count = 1;
    simplex_result closest = simplex_closest(simplex, count);
    float tiny = 1e-4f * fmino(shape_scale(first), shape_scale(second));
    for (uint32_t iter = 0; iter < 24u; iter++) {
        if (closest.penetrating || vlen(closest.v) < tiny) {
            while (count < 4u) {
                v3 d = simplex_expand_dir(simplex, count);
                simplex[count] = support_pair(w, first, second, d);
                count++;
            }
            r.penetrating = 1;
            r.distance = 0.0f;
            r.point_a = closest.point_a;
            r.point_b = closest.point_b;
            break;
        }
        simplex_point sp = support_pair(w, first, second, vneg(closest.v));
        v3 nv = vneg(closest.v);
        if (vdot(sp.w, nv) <= vdot(closest.v, nv) + 1e-4f * fmaxo(1.0f, vdot(closest.v, closest.v))) {
            r.distance = vlen(closest.v);
            r.point_a = closest.point_a;
            r.point_b = closest.point_b;
            if (r.distance > 1e-6f) r.normal = sign_normalize(vsub(closest.point_b, closest.point_a));
            break;
        }
        count = simplex_reduce(simplex, closest.lambdas);
        if (simplex_holds(simplex, count, sp.w, shape_scale(second))) {
            r.distance = vlen(closest.v);
            r.point_a = closest.point_a;
            r.point_b = closest.point_b;
            if (r.distance > 1e-6f) r.normal = sign_normalize(vsub(closest.point_b, closest.point_a));
            break;
        }
        simplex[count] = sp;
        count++;
        closest = simplex_closest(simplex, count);
    }
    *out_count = count;
    return r;
}

/* ---- EPA (convex.wgsl:286-437) -------------------------------------------------------------- */

typedef struct { uint32_t a, b, c; v3 normal; } epa_face;

static epa_result epa_invalid(void) {
    epa_result r;
    r.valid = 0;
    r.normal = V(0.0f, 1.0f, 0.0f);
    r.depth = 0.0f;
    r.point = V(0.0f, 0.0f, 0.0f);
    return r;
}

static epa_face epa_make_face(uint32_t a, uint32_t b, uint32_t c, v3 unit, v3 centroid) {
    epa_face f;
    f.a = a;
    f.b = b;
    f.c = c;
    f.normal = vdot(unit, centroid) > 0.0f ? vneg(unit) : unit;
    return f;
}

static epa_result epa_tetrahedron(const oracle_world* w, const world_shape* first,
                                  const world_shape* second, const simplex_point* simplex,
                                  uint32_t count) {
    simplex_point polytope[64];
    epa_face new_faces[64];
    epa_face faces[128];
    uint32_t visible[128];
    uint32_t vertex_count = 0;
    for (uint32_t i = 0; i < count; i++) polytope[vertex_count++] = simplex[i];
    uint32_t face_count = 0;
    if (vertex_count != 4u) return epa_invalid();
    for (uint32_t i = 0; i < 4; i++) {
        uint32_t a = kTetFaces[i][0], b = kTetFaces[i][1], c = kTetFaces[i][2];
        v3 normal = vcross(vsub(polytope[b].w, polytope[a].w), vsub(polytope[c].w, polytope[a].w));
        float length_n = vlen(normal);
        if (length_n < 1e-10f) return epa_invalid();
        v3 unit = vdiv(normal, length_n);
        v3 centroid = vmul(vadd(vadd(polytope[a].w, polytope[b].w), polytope[c].w), 1.0f / 3.0f);
        faces[face_count++] = epa_make_face(a, b, c, unit, centroid);
    }
    float best_dist = 0.0f;
    uint32_t best_face = 0;
    for (uint32_t iter = 0; iter < 32u; iter++) {
        best_dist = FLT_BIG;
        best_face = 0;
        for (uint32_t i = 0; i < face_count; i++) {
            float dist = vdot(faces[i].normal, polytope[faces[i].a].w);
            if (dist < best_dist && vlen(faces[i].normal) > 0.5f) {
                best_dist = dist;
                best_face = i;
            }
        }
        v3 probe = support_pair(w, first, second, faces[best_face].normal).w;
        float gain = vdot(probe, faces[best_face].normal) - best_dist;
        if (gain < 1e-4f) break;
        if (vertex_count >= 64u) break;
        polytope[vertex_count].w = probe;
        polytope[vertex_count].a = support(w, first, faces[best_face].normal);
        polytope[vertex_count].b = support(w, second, vneg(faces[best_face].normal));
        vertex_count++;
        uint32_t visible_count = 0;
        for (uint32_t i = 0; i < face_count; i++) {
            v3 vert = polytope[faces[i].a].w;
            if (vdot(faces[i].normal, vsub(probe, vert)) > 1e-6f) visible[visible_count++] = i;
        }
        uint32_t new_count = 0;
        for (uint32_t i = 0; i < visible_count; i++) {
            epa_face face = faces[visible[i]];
            uint32_t edges[3][2] = {{face.a, face.b}, {face.b, face.c}, {face.c, face.a}};
            for (uint32_t e = 0; e < 3; e++) {
                uint32_t start = edges[e][0], end = edges[e][1];
                int shared = 0;
                for (uint32_t j = 0; j < visible_count; j++) {
                    if (j == i) continue;
                    epa_face other = faces[visible[j]];
                    if ((other.a == end && other.b == start) || (other.b == end && other.c == start) ||
                        (other.c == end && other.a == start)) {
                        shared = 1;
                    }
                }
                if (!shared && new_count < 64u) {
                    uint32_t apex = vertex_count - 1u;
                    v3 normal = vcross(vsub(polytope[end].w, polytope[apex].w),
                                       vsub(polytope[start].w, polytope[apex].w));
                    float length_n = vlen(normal);
                    if (length_n > 1e-10f) {
                        v3 unit = vdiv(normal, length_n);
                        v3 centroid = vmul(vadd(vadd(polytope[apex].w, polytope[start].w), polytope[end].w),
                                           1.0f / 3.0f);
                        new_faces[new_count++] = epa_make_face(apex, start, end, unit, centroid);
                    }
                }
            }
        }
        uint32_t write = 0, seen = 0;
        for (uint32_t i = 0; i < face_count; i++) {
            if (seen < visible_count && visible[seen] == i) {
                seen++;
                continue;
            }
            faces[write++] = faces[i];
        }
        face_count = write;
        if (face_count + new_count > 128u) break;
        for (uint32_t i = 0; i < new_count; i++) faces[face_count++] = new_faces[i];
    }
    best_dist = FLT_BIG;
    for (uint32_t i = 0; i < face_count; i++) {
        float dist = vdot(faces[i].normal, polytope[faces[i].a].w);
        if (dist < best_dist && vlen(faces[i].normal) > 0.5f) {
            best_dist = dist;
            best_face = i;
        }
    }
    epa_result r;
    r.valid = best_dist < FLT_BIG;
    r.normal = faces[best_face].normal;
    r.depth = best_dist;
    r.point = polytope[faces[best_face].a].a;
    return r;
}

/* ---- penetration fallbacks and convex_hit (convex.wgsl:523-618) ------------------------------ */

static float support_projection_depth(const oracle_world* w, const world_shape* a, const world_shape* b, v3 dir) {
    v3 n = vnorm(dir);
    return vdot(vsub(support(w, a, n), support(w, b, vneg(n))), n);
}

static epa_result convex_penetration_probe(const oracle_world* w, const world_shape* first,
                                           const world_shape* second, const simplex_point* s,
                                           uint32_t count) {
    epa_result best;
    best.valid = 0;
    best.normal = V(0.0f, 1.0f, 0.0f);
    best.depth = FLT_BIG;
    best.point = vmul(vadd(first->center, second->center), 0.5f);
    v3 probed[16];
    uint32_t probe_count = 0;
    for (uint32_t i = 0; i < count; i++) {
        for (uint32_t j = i + 1; j < count; j++) {
            for (uint32_t k = j + 1; k < count; k++) {
                v3 normal = vcross(vsub(s[j].w, s[i].w), vsub(s[k].w, s[i].w));
                if (vlen(normal) > 1e-8f && probe_count < 16u) probed[probe_count++] = normal;
            }
        }
    }
    v3 center_dir = vsub(second->center, first->center);
    if (vlen(center_dir) > 1e-8f && probe_count < 16u) probed[probe_count++] = center_dir;
    const v3 axes[6] = {V(1.0f, 0.0f, 0.0f), V(-1.0f, 0.0f, 0.0f), V(0.0f, 1.0f, 0.0f),
                        V(0.0f, -1.0f, 0.0f), V(0.0f, 0.0f, 1.0f), V(0.0f, 0.0f, -1.0f)};
    for (uint32_t i = 0; i < 6; i++) {
        if (probe_count < 16u) probed[probe_count++] = axes[i];
    }
    for (uint32_t i = 0; i < probe_count; i++) {
        float depth = support_projection_depth(w, first, second, probed[i]);
        if (depth < best.depth) {
            best.depth = depth;
            best.normal = vnorm(probed[i]);
        }
    }
    best.valid = best.depth < FLT_BIG;
    v3 pa = support(w, first, best.normal), pb = support(w, second, vneg(best.normal));
    best.point = vmul(vadd(pa, pb), 0.5f);
    return best;
}

static shape_hit make_hit(float distance, v3 point, v3 normal) {
    shape_hit h;
    h.distance = distance;
    h.point = point;
    h.normal = normal;
    h.triangle = NO_TRIANGLE;
    return h;
}

static shape_hit convex_hit(const oracle_world* w, const world_shape* first, const world_shape* second) {
    simplex_point simplex[4];
    uint32_t count = 0;
    convex_closest_t closest = convex_closest(w, first, second, simplex, &count);
    v3 centers = vsub(second->center, first->center);
    if (!closest.penetrating) {
        epa_result probe = convex_penetration_probe(w, first, second, simplex, count);
        if (probe.depth > 0.0f) {
            v3 normal = probe.normal;
            if (vdot(normal, centers) < 0.0f) normal = vneg(normal);
            return make_hit(-probe.depth, vmul(vadd(closest.point_a, closest.point_b), 0.5f), normal);
        }
        if (closest.distance > 0.0f) {
            return make_hit(closest.distance, vmul(vadd(closest.point_a, closest.point_b), 0.5f), closest.normal);
        }
        return make_hit(FLT_BIG, V(0.0f, 0.0f, 0.0f), V(0.0f, 0.0f, 0.0f)); /* NO_HIT */
    }
    epa_result epa = epa_tetrahedron(w, first, second, simplex, count);
    shape_hit result;
    if (epa.valid && epa.depth > 0.0f && epa.depth < FLT_BIG && vlen(epa.normal) > 0.5f &&
        epa.depth < support_projection_depth(w, first, second, epa.normal) +
                        0.1f * fmino(shape_scale(first), shape_scale(second))) {
        v3 pa = support(w, first, epa.normal), pb = support(w, second, vneg(epa.normal));
        result = make_hit(-epa.depth, vmul(vadd(pa, pb), 0.5f), epa.normal);
    } else {
        epa_result probe = convex_penetration_probe(w, first, second, simplex, count);
        v3 pa = support(w, first, probe.normal), pb = support(w, second, vneg(probe.normal));
        result = make_hit(-probe.depth, vmul(vadd(pa, pb), 0.5f), probe.normal);
    }
    if (vdot(result.normal, centers) < 0.0f) result.normal = vneg(result.normal);
    return result;
}

/* ---- feature rings (convex.wgsl:625-900) ----------------------------------------------------- */

static uint32_t box_feature_ring(const world_shape* ws, v3 direction, v3* out, uint32_t* ids,
                                 uint32_t* out_face) {
    v3 axes[3] = {qrot(ws->rotation, V(1.0f, 0.0f, 0.0f)), qrot(ws->rotation, V(0.0f, 1.0f, 0.0f)),
                  qrot(ws->rotation, V(0.0f, 0.0f, 1.0f))};
    uint32_t best = 0;
    for (uint32_t i = 1; i < 3; i++) {
        if (fabsf(vdot(axes[i], direction)) > fabsf(vdot(axes[best], direction))) best = i;
    }
    float sign = vdot(axes[best], direction) < 0.0f ? -1.0f : 1.0f;
    v3 face = vmul(axes[best], sign);
    v3 center = vadd(ws->center, vmul(face, vcomp(ws->half_extents, best)));
    v3 u1 = axes[(best + 1u) % 3u], u2 = axes[(best + 2u) % 3u];
    float e1 = vcomp(ws->half_extents, (best + 1u) % 3u);
    float e2 = vcomp(ws->half_extents, (best + 2u) % 3u);
    out[0] = vadd(vadd(center, vmul(u1, e1)), vmul(u2, e2));
    ids[0] = box_corner_id(best, sign, 1.0f, 1.0f);
    out[1] = vadd(vsub(center, vmul(u1, e1)), vmul(u2, e2));
    ids[1] = box_corner_id(best, sign, -1.0f, 1.0f);
    out[2] = vsub(vsub(center, vmul(u1, e1)), vmul(u2, e2));
    ids[2] = box_corner_id(best, sign, -1.0f, -1.0f);
    out[3] = vsub(vadd(center, vmul(u1, e1)), vmul(u2, e2));
    ids[3] = box_corner_id(best, sign, 1.0f, -1.0f);
    *out_face = FEATURE_FACE_BIT | (best * 2u + (sign > 0.0f ? 1u : 0u));
    return 4;
}

/* Angular order about the ring normal: atan2(y, x) of the two points compared, decided exactly
 * from the half plane and the sign of a cross product instead of two rounded arctangents
 * (ring_of, convex.wgsl:696-711). Groups in increasing angle: -pi, (-pi, 0), [0, pi), pi. */
static int angle_group(float x, float y) {
    if (y == 0.0f && x < 0.0f) return signbit(y) ? 0 : 3;
    if (y < 0.0f) return 1;
    return 2;
}

static int angle_before(float bx, float by, float ax, float ay) {
    if (bx == 0.0f && by == 0.0f) bx = 1.0f;
    if (ax == 0.0f && ay == 0.0f) ax = 1.0f;
    int gb = angle_group(bx, by), ga = angle_group(ax, ay);
    if (gb != ga) return gb < ga;
    if (gb == 0 || gb == 3) return 0;
    return bx * ay - by * ax > 0.0f;
}

static uint32_t ring_of(const oracle_world* w, const world_shape* ws, v3 direction, v3* points,
                        uint32_t* ids, uint32_t count, v3* out, uint32_t* out_ids) {
    if (count <= 1u) {
        out[0] = support(w, ws, direction);
        out_ids[0] = ids[0];
        return 1;
    }
    v3 n = sign_normalize(direction);
    v3 center = V(0.0f, 0.0f, 0.0f);
    for (uint32_t i = 0; i < count; i++) center = vadd(center, points[i]);
    center = vdiv(center, (float)count);
    v3 u = fabsf(n.x) > 0.9f ? vnorm(vcross(n, V(0.0f, 1.0f, 0.0f))) : vnorm(vcross(n, V(1.0f, 0.0f, 0.0f)));
    v3 v = vcross(n, u);
    for (uint32_t i = 0; i < count; i++) {
        for (uint32_t j = i + 1; j < count; j++) {
            v3 pa = vsub(points[i], center), pb = vsub(points[j], center);
            if (angle_before(vdot(pb, u), vdot(pb, v), vdot(pa, u), vdot(pa, v))) {
                v3 tp = points[i];
                points[i] = points[j];
                points[j] = tp;
                uint32_t ti = ids[i];
                ids[i] = ids[j];
                ids[j] = ti;
            }
        }
    }
    uint32_t keep = count < FEATURE_MAX ? count : FEATURE_MAX;
    for (uint32_t i = 0; i < keep; i++) {
        out[i] = points[i];
        out_ids[i] = ids[i];
    }
    return keep;
}

/* convex.wgsl:723-793, the linear path */
static uint32_t hull_feature_ring(const oracle_world* w, const world_shape* ws, v3 direction, v3* out,
                                  uint32_t* out_ids, uint32_t* out_face) {
    const dyn_shape_source* src = &w->sources[ws->source];
    v3 d = vmulv(qrot(qconj(ws->rotation), direction), ws->scale);
    float best_dot = vdot(support_local(w, ws, d), d);
    v3 extent = vsub(vload(src->local_max), vload(src->local_min));
    float eps = 1e-3f * fmaxo(fmaxo(extent.x, extent.y), extent.z) *
                fmaxo(fmaxo(ws->scale.x, ws->scale.y), ws->scale.z);
    v3 points[16];
    uint32_t ids[16];
    ids[0] = 0u;
    uint32_t count = 0, first_id = EMPTY_RING_FACE;
    for (uint32_t i = 0; i < src->vertex_count && count < 16u; i++) {
        v3 v = hull_vertex(w, src, i);
        if (vdot(v, d) >= best_dot - eps) {
            points[count] = vadd(ws->center, qrot(ws->rotation, vmulv(v, ws->scale)));
            ids[count] = i;
            first_id = first_id < i ? first_id : i;
            count++;
        }
    }
    *out_face = FEATURE_FACE_BIT | first_id;
    return ring_of(w, ws, direction, points, ids, count, out, out_ids);
}

/* convex.wgsl:795-819 */
static uint32_t capsule_feature(const world_shape* ws, v3 direction, v3* out, uint32_t* ids, uint32_t* out_face) {
    v3 ld = vmulv(qrot(qconj(ws->rotation), direction), ws->scale);
    float cap = ld.y > 0.0f ? 1.0f : -1.0f;
    v3 radial = vmul(sign_normalize(ld), ws->radius);
    *out_face = FEATURE_FACE_BIT | 0u;
    if (fabsf(ld.y) > CAP_ALIGNMENT * vlen(ld)) {
        v3 local = vadd(V(0.0f, cap * ws->half_height, 0.0f), radial);
        out[0] = vadd(ws->center, qrot(ws->rotation, vmulv(local, ws->scale)));
        ids[0] = cap > 0.0f ? 1u : 0u;
        return 1;
    }
    v3 upper = vadd(V(0.0f, ws->half_height, 0.0f), radial);
    v3 lower = vadd(V(0.0f, -ws->half_height, 0.0f), radial);
    out[0] = vadd(ws->center, qrot(ws->rotation, vmulv(upper, ws->scale)));
    out[1] = vadd(ws->center, qrot(ws->rotation, vmulv(lower, ws->scale)));
    ids[0] = 0u;
    ids[1] = 1u;
    return 2;
}

/* convex.wgsl:821-854 */
static uint32_t cylinder_feature(const world_shape* ws, v3 direction, v3* out, uint32_t* ids, uint32_t* out_face) {
    v3 axis = shape_axis(ws);
    float alignment = vdot(axis, vnorm(direction));
    if (fabsf(alignment) > CAP_ALIGNMENT) {
        float sign = alignment < 0.0f ? -1.0f : 1.0f;
        uint32_t cap = sign > 0.0f ? 1u : 0u;
        v3 cap_center = vadd(ws->center, vmul(vmul(axis, ws->half_height), sign));
        v3 u = vnorm(vcross(axis, V(1.0f, 0.0f, 0.0f)));
        if (vlen(vcross(axis, V(1.0f, 0.0f, 0.0f))) < 1e-6f) u = vnorm(vcross(axis, V(0.0f, 0.0f, 1.0f)));
        v3 v = vcross(axis, u);
        for (uint32_t i = 0; i < CYLINDER_CAP_POINTS; i++) {
            out[i] = vadd(cap_center, vmul(vadd(vmul(u, kCapCos[i]), vmul(v, kCapSin[i])), ws->radius));
            ids[i] = cap * CYLINDER_CAP_POINTS + i;
        }
        *out_face = FEATURE_FACE_BIT | cap;
        return CYLINDER_CAP_POINTS;
    }
    v3 side = sign_normalize(vsub(direction, vmul(axis, vdot(direction, axis))));
    out[0] = vadd(vadd(ws->center, vmul(axis, ws->half_height)), vmul(side, ws->radius));
    ids[0] = CYLINDER_SIDE_POINTS;
    out[1] = vadd(vsub(ws->center, vmul(axis, ws->half_height)), vmul(side, ws->radius));
    ids[1] = CYLINDER_SIDE_POINTS + 1u;
    *out_face = FEATURE_FACE_BIT | CYLINDER_SIDE_FACE;
    return 2;
}

/* convex.wgsl:856-889 */
static uint32_t shape_feature(const oracle_world* w, const world_shape* ws, v3 direction, v3* out,
                              uint32_t* ids, uint32_t* out_face, int* ring) {
    if (ws->kind == DYN_SHAPE_CUBOID) {
        *ring = 1;
        return box_feature_ring(ws, direction, out, ids, out_face);
    }
    if (ws->kind == DYN_SHAPE_HULL) {
        uint32_t count = hull_feature_ring(w, ws, direction, out, ids, out_face);
        *ring = count > 2u;
        return count;
    }
    if (ws->kind == DYN_SHAPE_CAPSULE) {
        *ring = 0;
        return capsule_feature(ws, direction, out, ids, out_face);
    }
    if (ws->kind == DYN_SHAPE_CYLINDER) {
        uint32_t count = cylinder_feature(ws, direction, out, ids, out_face);
        *ring = count > 2u;
        return count;
    }
    *ring = 0;
    out[0] = support(w, ws, direction);
    ids[0] = 0u;
    *out_face = FEATURE_FACE_BIT | 0u;
    return 1;
}

/* convex.wgsl:891-902 */
static float feature_radius(const v3* points, uint32_t count) {
    v3 center = V(0.0f, 0.0f, 0.0f);
    for (uint32_t i = 0; i < count; i++) center = vadd(center, points[i]);
    center = vdiv(center, (float)count);
    float radius = 0.0f;
    for (uint32_t i = 0; i < count; i++) radius = fmaxo(radius, vlen(vsub(points[i], center)));
    return radius;
}

/* convex.wgsl:910-982: Sutherland-Hodgman clip of a feature against the side planes of a ring */
static uint32_t clip_feature_to_ring(const v3* points, const uint32_t* ids, uint32_t count,
                                     const v3* ring, uint32_t ring_count, v3 ref_dir, v3* held,
                                     uint32_t* held_ids) {
    v3 kept[MANIFOLD_CANDIDATES];
    uint32_t kept_ids[MANIFOLD_CANDIDATES];
    for (uint32_t i = 0; i < count; i++) {
        held[i] = points[i];
        held_ids[i] = ids[i];
    }
    uint32_t live = count;
    for (uint32_t side = 0; side < ring_count; side++) {
        v3 current = ring[side], next = ring[(side + 1u) % ring_count];
        v3 plane_normal = vnorm(vcross(vsub(next, current), ref_dir));
        uint32_t written = 0;
        if (live == 2u) {
            v3 p0 = held[0], p1 = held[1];
            float s0 = vdot(vsub(p0, current), plane_normal) - CLIP_MARGIN;
            float s1 = vdot(vsub(p1, current), plane_normal) - CLIP_MARGIN;
            if (s0 <= 0.0f) {
                kept[written] = p0;
                kept_ids[written] = held_ids[0];
                written++;
            }
            if (s0 * s1 < 0.0f && written < MANIFOLD_CANDIDATES) {
                kept[written] = vadd(p0, vmul(vsub(p1, p0), s0 / (s0 - s1)));
                kept_ids[written] = FEATURE_CLIP | (side * MANIFOLD_CANDIDATES);
                written++;
            }
            if (s1 <= 0.0f && written < MANIFOLD_CANDIDATES) {
                kept[written] = p1;
                kept_ids[written] = held_ids[1];
                written++;
            }
        } else {
            for (uint32_t i = 0; i < live; i++) {
                v3 cp = held[i], np = held[(i + 1u) % live];
                float cs = vdot(vsub(cp, current), plane_normal) - CLIP_MARGIN;
                float ns = vdot(vsub(np, current), plane_normal) - CLIP_MARGIN;
                if (cs <= 0.0f && written < MANIFOLD_CANDIDATES) {
                    kept[written] = cp;
                    kept_ids[written] = held_ids[i];
                    written++;
                }
                if (cs * ns < 0.0f && written < MANIFOLD_CANDIDATES) {
                    kept[written] = vadd(cp, vmul(vsub(np, cp), cs / (cs - ns)));
                    kept_ids[written] = FEATURE_CLIP | (side * MANIFOLD_CANDIDATES + i);
                    written++;
                }
            }
        }
        live = written;
        for (uint32_t i = 0; i < written; i++) {
            held[i] = kept[i];
            held_ids[i] = kept_ids[i];
        }
        if (live == 0u) break;
    }
    return live;
}

/* convex.wgsl:984-1071 */
static int convex_pair_manifold(const oracle_world* w, const world_shape* first, const world_shape* second,
                                v3 direction, dyn_contact* c) {
    v3 fp[FEATURE_MAX], sp[FEATURE_MAX];
    uint32_t fi[FEATURE_MAX], si[FEATURE_MAX];
    uint32_t first_face = 0, second_face = 0;
    int first_ring = 0, second_ring = 0;
    uint32_t first_count = shape_feature(w, first, direction, fp, fi, &first_face, &first_ring);
    uint32_t second_count = shape_feature(w, second, vneg(direction), sp, si, &second_face, &second_ring);
    if (!first_ring && !second_ring) return 0;
    int first_reference = first_ring;
    if (first_ring && second_ring) {
        first_reference = feature_radius(fp, first_count) >= feature_radius(sp, second_count);
    }
    const v3 *reference, *incident;
    const uint32_t* incident_ids;
    uint32_t reference_count, incident_count, reference_face;
    v3 ref_dir;
    if (first_reference) {
        reference = fp;
        reference_count = first_count;
        reference_face = first_face;
        incident = sp;
        incident_ids = si;
        incident_count = second_count;
        ref_dir = direction;
    } else {
        reference = sp;
        reference_count = second_count;
        reference_face = second_face;
        incident = fp;
        incident_ids = fi;
        incident_count = first_count;
        ref_dir = vneg(direction);
    }
    v3 center = V(0.0f, 0.0f, 0.0f);
    for (uint32_t i = 0; i < reference_count; i++) center = vadd(center, reference[i]);
    center = vdiv(center, (float)reference_count);
    v3 deepest = V(0.0f, 0.0f, 0.0f);
    float deepest_depth = -1e30f;
    for (uint32_t i = 0; i < incident_count; i++) {
        float depth = vdot(vsub(center, incident[i]), ref_dir);
        if (depth > deepest_depth) {
            deepest_depth = depth;
            deepest = incident[i];
        }
    }
    v3 clipped[MANIFOLD_CANDIDATES];
    uint32_t clipped_ids[MANIFOLD_CANDIDATES];
    uint32_t clipped_count = clip_feature_to_ring(incident, incident_ids, incident_count, reference,
                                                  reference_count, ref_dir, clipped, clipped_ids);
    candidate_t cand[MANIFOLD_CANDIDATES];
    uint32_t cc = 0;
    for (uint32_t i = 0; i < clipped_count && cc < MANIFOLD_CANDIDATES; i++) {
        float depth = vdot(vsub(center, clipped[i]), ref_dir);
        if (depth > 0.0f) {
            cand[cc].position = vsub(clipped[i], vmul(ref_dir, depth * 0.5f));
            cand[cc].depth = depth;
            cand[cc].feature = first_reference ? feature_pair(FEATURE_VERTEX, reference_face, clipped_ids[i])
                                               : feature_pair(FEATURE_VERTEX, clipped_ids[i], reference_face);
            cc++;
        }
    }
    if (cc == 0u) {
        if (deepest_depth <= 0.0f) return 0;
        manifold_push(c, vsub(deepest, vmul(ref_dir, deepest_depth * 0.5f)), deepest_depth, FEATURE_POINT);
        return 1;
    }
    manifold_keep(c, cand, cc);
    return 1;
}

/* narrowphase.wgsl:479-483 */
static void manifold_from_hit(dyn_contact* c, shape_hit hit, float margin) {
    if (hit.distance <= margin) manifold_push(c, hit.point, -hit.distance, FEATURE_POINT);
}

/* narrowphase.wgsl:485-512 */
static dyn_contact plane_convex(const oracle_world* w, const world_shape* plane, const world_shape* convex,
                               float margin) {
    dyn_contact c;
    memset(&c, 0, sizeof c);
    v3 n = qrot(plane->rotation, V(0.0f, 1.0f, 0.0f));
    float center_side = vdot(vsub(convex->center, plane->center), n);
    v3 facing = center_side < 0.0f ? vneg(n) : n;
    manifold_emit(&c, facing);
    v3 pts[FEATURE_MAX];
    uint32_t ids[FEATURE_MAX];
    uint32_t feature = 0;
    int ring = 0;
    uint32_t count = shape_feature(w, convex, vneg(facing), pts, ids, &feature, &ring);
    candidate_t cand[MANIFOLD_CANDIDATES];
    uint32_t cc = 0;
    for (uint32_t i = 0; i < count; i++) {
        float depth = vdot(vsub(plane->center, pts[i]), facing);
        if (depth <= -margin) continue;
        cand[cc].position = vadd(pts[i], vmul(facing, depth * 0.5f));
        cand[cc].depth = depth;
        cand[cc].feature = feature_pair(FEATURE_VERTEX, FEATURE_FACE_BIT | 0u, ids[i]);
        cc++;
    }
    manifold_keep(&c, cand, cc);
    return c;
}

/* ---- narrowphase work (narrowphase.wgsl:518-685) ---------------------------------------- */

static int shape_analytic(uint32_t kind) {
    return kind == DYN_SHAPE_SPHERE || kind == DYN_SHAPE_CUBOID || kind == DYN_SHAPE_CAPSULE;
}
static int scaled(const dyn_collider* c) {
    return c->scale[0] != 1.0f || c->scale[1] != 1.0f || c->scale[2] != 1.0f;
}
static float material_combine(float a, float b, uint32_t mode) {
    if (mode == 1u) return fmino(a, b);
    if (mode == 2u) return fmaxo(a, b);
    if (mode == 3u) return (a + b) * 0.5f;
    return sqrtf(a * b);
}
static int filters_intersect(uint32_t g1, uint32_t m1, uint32_t g2, uint32_t m2) {
    return (g1 & m2) != 0u && (g2 & m1) != 0u;
}

static void narrow_work(oracle_world* w, uint32_t index) {
    w->valid[index] = 0;
    uint32_t fs = w->pmaj[index], ss = w->pmin[index];
    uint32_t fb = w->owner[fs], sb = w->owner[ss];
    if (fb == DYN_NO_BODY || sb == DYN_NO_BODY || fb == sb) return;
    body_t first = load_body(w, fb), second = load_body(w, sb);
    if (!kind_moves(&first.d) && !kind_moves(&second.d)) return;
    const dyn_collider* fc = &w->co[fs];
    const dyn_collider* sc = &w->co[ss];
    if (fc->kind == DYN_SHAPE_NONE || sc->kind == DYN_SHAPE_NONE) return;
    uint32_t g1 = fc->collision_group != DYN_NO_COLLISION_FILTER ? fc->collision_group : first.d.collision_group;
    uint32_t m1 = fc->collision_mask != DYN_NO_COLLISION_FILTER ? fc->collision_mask : first.d.collision_mask;
    uint32_t g2 = sc->collision_group != DYN_NO_COLLISION_FILTER ? sc->collision_group : second.d.collision_group;
    uint32_t m2 = sc->collision_mask != DYN_NO_COLLISION_FILTER ? sc->collision_mask : second.d.collision_mask;
    if (!filters_intersect(g1, m1, g2, m2)) return;
    dyn_contact c;
    memset(&c, 0, sizeof c);
    int generated = 0;
    int sensor = (fc->flags & DYN_COLLIDER_SENSOR) || (sc->flags & DYN_COLLIDER_SENSOR);
    int fw = fc->kind == DYN_SHAPE_PLANE, sw = sc->kind == DYN_SHAPE_PLANE;
    if (fw && sw) return;
    float margin = pair_margin(w, &first, &second);
    world_shape wf = world_collider(&first.s, fc), wsd = world_collider(&second.s, sc);
    if (fw) {
        c = plane_convex(w, &wf, &wsd, margin);
        generated = c.point_count > 0u;
    } else if (sw) {
        c = plane_convex(w, &wsd, &wf, margin);
        c.normal[0] = -c.normal[0];
        c.normal[1] = -c.normal[1];
        c.normal[2] = -c.normal[2];
        for (uint32_t i = 0; i < c.point_count; i++) c.points[i].feature = feature_mirror(c.points[i].feature);
        generated = c.point_count > 0u;
    } else if (scaled(fc) || scaled(sc) || !shape_analytic(fc->kind) || !shape_analytic(sc->kind) ||
               (fc->kind == DYN_SHAPE_CAPSULE && sc->kind == DYN_SHAPE_CUBOID) ||
               (fc->kind == DYN_SHAPE_CUBOID && sc->kind == DYN_SHAPE_CAPSULE)) {
        /* GJK / EPA with a clipped feature manifold (narrowphase.wgsl:594-628, 654-664) */
        shape_hit hit = convex_hit(w, &wf, &wsd);
        if (hit.distance <= margin) {
            manifold_emit(&c, hit.normal);
            generated = 1;
            if (!convex_pair_manifold(w, &wf, &wsd, hit.normal, &c)) manifold_from_hit(&c, hit, margin);
        }
    } else {
        uint32_t a = fc->kind, b = sc->kind;
        if (a == DYN_SHAPE_SPHERE && b == DYN_SHAPE_SPHERE) c = sphere_sphere(&first, fc, &second, sc, margin);
        else if (a == DYN_SHAPE_SPHERE && b == DYN_SHAPE_CUBOID) c = sphere_box(&first, fc, &second, sc, margin);
        else if (a == DYN_SHAPE_SPHERE && b == DYN_SHAPE_CAPSULE) c = sphere_capsule(&first, fc, &second, sc, margin);
        else if (a == DYN_SHAPE_CUBOID && b == DYN_SHAPE_SPHERE) {
            c = sphere_box(&second, sc, &first, fc, margin);
            c.normal[0] = -c.normal[0];
            c.normal[1] = -c.normal[1];
            c.normal[2] = -c.normal[2];
        } else if (a == DYN_SHAPE_CUBOID && b == DYN_SHAPE_CUBOID) c = box_box_sat(&first, fc, &second, sc, margin);
        else if (a == DYN_SHAPE_CAPSULE && b == DYN_SHAPE_SPHERE) {
            c = sphere_capsule(&second, sc, &first, fc, margin);
            c.normal[0] = -c.normal[0];
            c.normal[1] = -c.normal[1];
            c.normal[2] = -c.normal[2];
        } else if (a == DYN_SHAPE_CAPSULE && b == DYN_SHAPE_CAPSULE) c = capsule_capsule(&first, fc, &second, sc, margin);
        generated = 1;
    }
    if (!generated) return;
    c.a = fs;
    c.b = ss;
    c.sensor = sensor ? 1u : 0u;
    c.first_body_id = first.s.body_id;
    c.second_body_id = second.s.body_id;
    c.first_generation = first.s.generation;
    c.second_generation = second.s.generation;
    c.surface = 0xFFFFFFFFu; /* NO_SURFACE: no triangle scene */
    c.friction = material_combine(fc->friction, sc->friction, w->p.friction_combine);
    c.restitution = material_combine(fc->restitution, sc->restitution, w->p.restitution_combine);
    c.rolling_friction = fmaxo(fc->rolling_friction, sc->rolling_friction);
    c.spin_friction = fmaxo(fc->spin_friction, sc->spin_friction);
    c.relaxation = fc->relaxation + sc->relaxation;
    c.damping_ratio = fmaxo(fc->damping_ratio, sc->damping_ratio);
    c.events = (fc->flags & sc->flags) & (2u | 4u);
    if (c.point_count > 0u) {
        /* manifold_anchor, core.wgsl:675-681 */
        for (uint32_t i = 0; i < c.point_count; i++) {
            v3 point = vload(c.points[i].position);
            vstore(c.points[i].local_a, qrot(qconj(qload(first.s.orientation)), vsub(point, vload(first.s.position))));
            vstore(c.points[i].local_b, qrot(qconj(qload(second.s.orientation)), vsub(point, vload(second.s.position))));
        }
        w->raw[index] = c;
        w->valid[index] = 1u;
    }
}

/* ---- identity and warm start (identity.wgsl, contact_relay.wgsl, contact_begin.wgsl) ---- */

static int contact_touches(const dyn_contact* c, float threshold) {
    for (uint32_t i = 0; i < c->point_count; i++) {
        if (c->points[i].depth >= -threshold) return 1;
    }
    return 0;
}
static int same_roles(const dyn_contact* h, const dyn_contact* c) {
    return h->first_body_id == c->first_body_id && h->first_generation == c->first_generation &&
           h->second_body_id == c->second_body_id && h->second_generation == c->second_generation &&
           h->a == c->a && h->b == c->b;
}
static int same_pair(const dyn_contact* h, const dyn_contact* c) {
    int flipped = h->first_body_id == c->second_body_id && h->first_generation == c->second_generation &&
                  h->second_body_id == c->first_body_id && h->second_generation == c->first_generation &&
                  h->a == c->b && h->b == c->a;
    return same_roles(h, c) || flipped;
}
static float normal_impulse(const dyn_contact* c) {
    float t = 0.0f;
    for (uint32_t i = 0; i < c->point_count; i++) t = t + fmaxo(c->points[i].accumulated_normal, 0.0f);
    return t;
}
static float tangent_impulse(const dyn_contact* c) {
    float t = 0.0f;
    for (uint32_t i = 0; i < c->point_count; i++) {
        const dyn_manifold_point* p = &c->points[i];
        t = t + sqrtf(p->accumulated_tangent_1 * p->accumulated_tangent_1 +
                      p->accumulated_tangent_2 * p->accumulated_tangent_2);
    }
    return t;
}
static dyn_contact contact_carried(dyn_contact c) {
    c.carried_normal = normal_impulse(&c);
    c.carried_tangent = tangent_impulse(&c);
    return c;
}
static float carried_distance(const dyn_manifold_point* h, const dyn_manifold_point* c) {
    return vlen(vsub(vload(h->local_a), vload(c->local_a))) + vlen(vsub(vload(h->local_b), vload(c->local_b)));
}
/* identity.wgsl:85-117 */
static dyn_contact relay_impulses(const dyn_contact* current, const dyn_contact* held, uint32_t* picks) {
    dyn_contact r = *current;
    int taken[4] = {0, 0, 0, 0};
    for (uint32_t pi = 0; pi < current->point_count; pi++) {
        const dyn_manifold_point* point = &r.points[pi];
        uint32_t picked = NO_SLOT;
        float picked_distance = 0.0f;
        for (uint32_t hi = 0; hi < held->point_count; hi++) {
            if (taken[hi]) continue;
            const dyn_manifold_point* hp = &held->points[hi];
            if (!(hp->feature == point->feature || hp->feature == feature_mirror(point->feature))) continue;
            float d = carried_distance(hp, point);
            if (picked == NO_SLOT || d < picked_distance) {
                picked = hi;
                picked_distance = d;
            }
        }
        if (picked == NO_SLOT) continue;
        taken[picked] = 1;
        if (picks) picks[pi] = picked;
        r.points[pi].accumulated_normal = held->points[picked].accumulated_normal;
        r.points[pi].accumulated_tangent_1 = held->points[picked].accumulated_tangent_1;
        r.points[pi].accumulated_tangent_2 = held->points[picked].accumulated_tangent_2;
    }
    return r;
}
/* The relay of identity.wgsl:85-117 alone (test hook): `current` with the impulses of `held`
 * relayed point by point; picks[i] = the held point relayed into point i, or ~0. */
void oracle_relay_impulses(const dyn_contact* current, const dyn_contact* held, dyn_contact* out, uint32_t* picks) {
    uint32_t local[4] = {NO_SLOT, NO_SLOT, NO_SLOT, NO_SLOT};
    *out = relay_impulses(current, held, local);
    if (picks) memcpy(picks, local, sizeof(local));
}
/* body_row.wgsl:1-13 */
static uint32_t resolve_row(const oracle_world* w, uint32_t id, uint32_t gen) {
    if (id >= w->id_capacity) return DYN_NO_BODY;
    uint32_t row = w->row_of_body[id];
    if (row >= w->n) return DYN_NO_BODY;
    const dyn_body_state* s = &w->st[row];
    if (s->body_id != id || s->generation != gen) return DYN_NO_BODY;
    return row;
}
static uint32_t current_slot(const oracle_world* w, const dyn_contact* c) {
    uint32_t lo = 0, hi = w->ncontacts;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2u;
        const dyn_contact* m = &w->contacts[mid];
        if (m->a < c->a || (m->a == c->a && m->b < c->b)) lo = mid + 1u; else hi = mid;
    }
    if (lo < w->ncontacts && w->contacts[lo].a == c->a && w->contacts[lo].b == c->b) return lo;
    return NO_SLOT;
}

static void contact_relay(oracle_world* w) {
    for (uint32_t i = 0; i < w->narchive; i++) {
        const dyn_contact* held = &w->archive[i];
        uint32_t r1 = resolve_row(w, held->first_body_id, held->first_generation);
        uint32_t r2 = resolve_row(w, held->second_body_id, held->second_generation);
        if (r1 == DYN_NO_BODY || r2 == DYN_NO_BODY) continue;
        uint32_t slot = current_slot(w, held);
        if (slot == NO_SLOT) continue;
        dyn_contact* cur = &w->contacts[slot];
        if (!same_pair(held, cur)) continue;
        if (held->events & CONTACT_ANNOUNCED) {
            w->matched[slot] = 1u;
            dyn_contact relayed = *cur;
            relayed.events |= CONTACT_ANNOUNCED;
            /* contact_carries_over == contact_same_pair, which holds here */
            uint32_t* picks = w->relay_src + 4u * slot;
            for (int q = 0; q < 4; q++) picks[q] = NO_SLOT;
            relayed = contact_carried(relay_impulses(cur, held, picks));
            for (int q = 0; q < 4; q++) {
                if (picks[q] != NO_SLOT) picks[q] = 4u * i + picks[q];
            }
            relayed.events |= CONTACT_ANNOUNCED;
            *cur = relayed;
        }
    }
}

static uint32_t resting_slot(const oracle_world* w, const dyn_contact* c) {
    for (uint32_t i = 0; i < w->nresting; i++) {
        if (same_pair(&w->resting[i], c)) return i;
    }
    return NO_SLOT;
}

static void contact_begin(oracle_world* w) {
    for (uint32_t i = 0; i < w->ncontacts; i++) {
        if (w->matched[i]) continue;
        dyn_contact* c = &w->contacts[i];
        uint32_t slot = resting_slot(w, c);
        if (slot != NO_SLOT) {
            dyn_contact revived = *c;
            if (w->resting[slot].events & CONTACT_ANNOUNCED) {
                revived.events |= CONTACT_ANNOUNCED;
                revived = contact_carried(relay_impulses(&revived, &w->resting[slot], NULL));
                revived.events |= CONTACT_ANNOUNCED;
            } else if (contact_touches(c, w->p.slop)) {
                revived.events |= CONTACT_ANNOUNCED;
            }
            *c = revived;
            continue;
        }
        if (contact_touches(c, w->p.slop)) c->events |= CONTACT_ANNOUNCED;
    }
}

/* ---- islands (island_*.wgsl) ------------------------------------------------------------ */

static uint32_t island_find(const oracle_world* w, uint32_t node) {
    uint32_t cur = node, parent = w->parents[cur];
    while (parent != cur) {
        cur = parent;
        parent = w->parents[cur];
    }
    return cur;
}
static void island_union(oracle_world* w, uint32_t a, uint32_t b) {
    if (a >= w->d || b >= w->d || a == b) return;
    uint32_t ra = island_find(w, a), rb = island_find(w, b);
    if (ra == rb) return;
    uint32_t lower = ra < rb ? ra : rb, higher = ra < rb ? rb : ra;
    w->parents[higher] = lower;
}
static void carry_static_wake(oracle_world* w, uint32_t movable, uint32_t partner) {
    if (partner < w->d) return;
    if (w->wake[partner]) w->wake[movable] |= 1u;
}
static void island_couple(oracle_world* w, uint32_t a, uint32_t b) {
    carry_static_wake(w, a, b);
    carry_static_wake(w, b, a);
    island_union(w, a, b);
}
static uint32_t rounds_for(uint32_t bodies) {
    uint32_t n = bodies < 2u ? 2u : bodies, r = 0;
    while (n > 1u) { n >>= 1; r++; }
    return r + 1u;
}

static void build_islands(oracle_world* w) {
    for (uint32_t i = 0; i < w->d; i++) {
        w->parents[i] = i;
        w->istate[i] = 0;
    }
    for (uint32_t i = 0; i < w->ncontacts; i++) {
        const dyn_contact* c = &w->contacts[i];
        if (c->point_count == 0u || c->sensor == 1u || !contact_touches(c, w->p.slop)) continue;
        island_couple(w, w->owner[c->a], w->owner[c->b]);
    }
    for (uint32_t i = 0; i < w->nresting; i++) {
        const dyn_contact* c = &w->resting[i];
        uint32_t r1 = resolve_row(w, c->first_body_id, c->first_generation);
        uint32_t r2 = resolve_row(w, c->second_body_id, c->second_generation);
        if (r1 == DYN_NO_BODY || r2 == DYN_NO_BODY) continue;
        island_couple(w, r1, r2);
    }
    uint32_t rounds = rounds_for(w->d);
    uint32_t* next = (uint32_t*)malloc((w->d ? w->d : 1u) * 4u);
    for (uint32_t r = 0; r < rounds; r++) {
        /* island_flatten.wgsl as one data-parallel round */
        for (uint32_t i = 0; i < w->d; i++) {
            uint32_t parent = w->parents[i], gp = w->parents[parent];
            next[i] = parent < gp ? parent : gp;
        }
        memcpy(w->parents, next, w->d * 4u);
    }
    free(next);
    for (uint32_t i = 0; i < w->d; i++) {
        if (w->wake[i] == 0u && w->motion[i] == 0u) continue;
        uint32_t root = w->parents[i];
        if (w->istate[root] < ISLAND_WAKE) w->istate[root] = ISLAND_WAKE;
    }
    for (uint32_t i = 0; i < w->d; i++) {
        if (w->parents[i] == i) (*ctr(w, DYN_COUNTER_ISLANDS))++;
    }
}

static void island_wake(oracle_world* w) {
    for (uint32_t i = 0; i < w->d; i++) {
        uint32_t root = w->parents[i];
        dyn_body_state* s = &w->st[i];
        if (s->sleeping != 0u && (w->istate[root] & ISLAND_WAKE)) {
            (*ctr(w, DYN_COUNTER_WOKE))++;
            (*ctr(w, DYN_COUNTER_WOKE_DEFERRED))++;
            w->admitted[i] = 0;
            s->sleeping = 0;
            s->sleep_timer = 0.0f;
        }
        w->wake[i] = 0;
    }
}

static void island_sleep(oracle_world* w) {
    for (uint32_t i = 0; i < w->d; i++) {
        dyn_body_state* s = &w->st[i];
        if (s->sleeping != 0u || !kind_simulates(&w->de[i])) continue;
        uint32_t root = w->parents[i];
        int disturbed = (w->istate[root] & ISLAND_WAKE) != 0u || w->wake[i] != 0u;
        if (disturbed) {
            s->sleep_timer = 0.0f;
        } else {
            s->sleep_timer = s->sleep_timer + w->p.dt;
            if (s->sleep_timer >= w->p.sleep_time) {
                freeze_body(s);
                (*ctr(w, DYN_COUNTER_SLEPT))++;
                w->admitted[i] = 0;
            }
        }
    }
}

/* ---- solver (solver_*.wgsl, substep_*.wgsl, position_*.wgsl) --------------------------- */

static void split_body(body_t* b, float scale) {
    b->d.inverse_mass = b->d.inverse_mass * scale;
    for (int i = 0; i < 6; i++) b->d.inverse_inertia[i] = b->d.inverse_inertia[i] * scale;
}

static void make_tangents(v3 n, v3* t1, v3* t2) {
    v3 t = fabsf(n.x) > 0.9f ? V(0.0f, 1.0f, 0.0f) : V(1.0f, 0.0f, 0.0f);
    *t1 = vnorm(vcross(n, t));
    *t2 = vcross(n, *t1);
}

static float point_momentum_mass(const body_t* a, const body_t* b, v3 pa, v3 pb, v3 axis) {
    v3 ra = vsub(pa, body_com(a)), rb = vsub(pb, body_com(b));
    v3 rax = vcross(ra, axis), rbx = vcross(rb, axis);
    return a->d.inverse_mass + b->d.inverse_mass + vdot(rax, apply_inverse_inertia(a, rax)) +
           vdot(rbx, apply_inverse_inertia(b, rbx));
}

static void apply_pair_impulse(body_t* a, body_t* b, v3 pa, v3 pb, v3 impulse) {
    vstore(a->s.velocity, vsub(vload(a->s.velocity), vmul(impulse, a->d.inverse_mass)));
    vstore(a->s.angular_velocity, vsub(vload(a->s.angular_velocity),
                                       apply_inverse_inertia(a, vcross(vsub(pa, body_com(a)), impulse))));
    vstore(b->s.velocity, vadd(vload(b->s.velocity), vmul(impulse, b->d.inverse_mass)));
    vstore(b->s.angular_velocity, vadd(vload(b->s.angular_velocity),
                                       apply_inverse_inertia(b, vcross(vsub(pb, body_com(b)), impulse))));
}

static v3 manifold_arm(const body_t* b, const dyn_manifold_point* p, int first) {
    return vadd(qrot(qload(b->s.orientation), vload(first ? p->local_a : p->local_b)), vload(b->s.position));
}

static int block_resolves(const dyn_contact* c) { return c->point_count > 0u && c->sensor == 0u; }

uint32_t oracle_threads(void) {
#ifdef _OPENMP
    return (uint32_t)omp_get_max_threads();
#else
    return 1u;
#endif
}

/* Integer additions and maxima shared by parallel solver blocks (plain in the serial build). */
static void add_word(uint32_t* p, uint32_t v) {
#ifdef _OPENMP
    __atomic_fetch_add(p, v, __ATOMIC_RELAXED);
#else
    *p += v;
#endif
}

static void max_word(uint32_t* p, uint32_t v) {
#ifdef _OPENMP
    uint32_t seen = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (v > seen && !__atomic_compare_exchange_n(p, &seen, v, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
    }
#else
    if (v > *p) *p = v;
#endif
}

static void accumulate(uint32_t* deltas, uint32_t row, v3 lin, v3 ang, float scale) {
    uint32_t* d = deltas + row * 8u;
    add_word(&d[0], solver_word(lin.x, scale));
    add_word(&d[1], solver_word(lin.y, scale));
    add_word(&d[2], solver_word(lin.z, scale));
    add_word(&d[3], solver_word(ang.x, scale));
    add_word(&d[4], solver_word(ang.y, scale));
    add_word(&d[5], solver_word(ang.z, scale));
}

static void solver_blocks(oracle_world* w) {
    uint32_t blocks = *ctr(w, DYN_COUNTER_BLOCKS);
    for (uint32_t i = 0; i < blocks; i++) {
        const dyn_contact* c = &w->contacts[i];
        uint32_t fb = w->owner[c->a], sb = w->owner[c->b];
        w->blocks[2 * i] = fb;
        w->blocks[2 * i + 1] = sb;
        if (!block_resolves(c)) continue;
        body_t first = load_body(w, fb), second = load_body(w, sb);
        for (uint32_t pi = 0; pi < c->point_count; pi++) {
            const dyn_manifold_point* p = &c->points[pi];
            v3 pos = vload(p->position);
            float speed = vdot(relative_velocity(&first, &second, pos, pos), vload(c->normal));
            float target = 0.0f;
            if (p->depth < 0.0f) target = p->depth / w->p.dt;
            else if (speed < -w->p.restitution_threshold) target = -speed * c->restitution;
            w->targets[i * 4u + pi] = target;
        }
        uint32_t both[2] = {fb, sb};
        for (int k = 0; k < 2; k++) {
            if (w->block_counts[both[k]]++ == 0u) w->rows[w->nrows++] = both[k];
        }
    }
    *ctr(w, DYN_COUNTER_SOLVER_ROWS) = w->nrows;
}

typedef struct {
    body_t first, second, split_first, split_second;
} block_pair;

static block_pair block_bodies(const oracle_world* w, uint32_t fb, uint32_t sb) {
    block_pair bp;
    bp.first = load_body(w, fb);
    bp.second = load_body(w, sb);
    bp.split_first = bp.first;
    bp.split_second = bp.second;
    uint32_t c1 = w->block_counts[fb], c2 = w->block_counts[sb];
    split_body(&bp.split_first, (float)(c1 > 1u ? c1 : 1u));
    split_body(&bp.split_second, (float)(c2 > 1u ? c2 : 1u));
    if (body_is_inert(&bp.first)) {
        body_frozen(&bp.first);
        body_frozen(&bp.split_first);
    }
    if (body_is_inert(&bp.second)) {
        body_frozen(&bp.second);
        body_frozen(&bp.split_second);
    }
    return bp;
}

static float row_mass(const body_t* a, const body_t* b, float count, v3 reach_a, v3 reach_b,
                      v3 axis, v3 anchor_a, v3 anchor_b) {
    v3 arm_a = vcross(vsub(anchor_a, body_com(a)), axis);
    v3 arm_b = vcross(vsub(anchor_b, body_com(b)), axis);
    return count * (a->d.inverse_mass + b->d.inverse_mass) +
           vdot(apply_inverse_inertia(a, vcross(reach_a, axis)), arm_a) +
           vdot(apply_inverse_inertia(b, vcross(reach_b, axis)), arm_b);
}

static float normal_delta(const dyn_contact* c, float mass, float truth, float step, float goal,
                          float depth, float speed, float accumulated) {
    if (c->relaxation <= 0.0f) return (goal - speed) / mass;
    float bias = fmaxo(depth / step, goal);
    float drag = 2.0f * c->damping_ratio * c->relaxation / step;
    float softness = c->relaxation * c->relaxation * truth / (step * step);
    float prior = speed - accumulated * truth;
    float wanted = (bias - (1.0f + drag) * prior) / (softness + (1.0f + drag) * truth);
    return truth * (wanted - accumulated) / mass;
}

static void commit_block(oracle_world* w, uint32_t fb, uint32_t sb, v3 da, v3 sa, v3 db, v3 sbv,
                         int residual) {
    accumulate(w->vdeltas, fb, da, sa, DYN_SOLVER_VELOCITY_SCALE);
    accumulate(w->vdeltas, sb, db, sbv, DYN_SOLVER_VELOCITY_SCALE);
    if (!residual) return;
    uint32_t lw = solver_word(fmaxo(vlen(da), vlen(db)), DYN_SOLVER_VELOCITY_SCALE);
    uint32_t aw = solver_word(fmaxo(vlen(sa), vlen(sbv)), DYN_SOLVER_VELOCITY_SCALE);
    max_word(ctr(w, DYN_COUNTER_SOLVE_LINEAR_RESIDUAL), lw);
    max_word(ctr(w, DYN_COUNTER_SOLVE_ANGULAR_RESIDUAL), aw);
}

static const v3 kZero = {0.0f, 0.0f, 0.0f};

/* solver_contact_block.wgsl:31-193 */
static void solve_block(oracle_world* w, uint32_t ci, int residual) {
    dyn_contact contact = w->contacts[ci];
    uint32_t fb = w->blocks[2 * ci], sb = w->blocks[2 * ci + 1];
    if (!block_resolves(&contact)) {
        commit_block(w, fb, sb, kZero, kZero, kZero, kZero, residual);
        return;
    }
    block_pair pair = block_bodies(w, fb, sb);
    if (body_is_inert(&pair.first) && body_is_inert(&pair.second)) {
        commit_block(w, fb, sb, kZero, kZero, kZero, kZero, residual);
        return;
    }
    body_t first = pair.first, second = pair.second;
    v3 normal = vload(contact.normal), t1, t2;
    make_tangents(normal, &t1, &t2);
    dyn_contact updated = contact;
    v3 reach_a = kZero, reach_b = kZero;
    for (uint32_t pi = 0; pi < contact.point_count; pi++) {
        const dyn_manifold_point* p = &contact.points[pi];
        reach_a = vadd(reach_a, vsub(manifold_arm(&pair.first, p, 1), body_com(&pair.first)));
        reach_b = vadd(reach_b, vsub(manifold_arm(&pair.second, p, 0), body_com(&pair.second)));
    }
    float count = (float)(contact.point_count > 1u ? contact.point_count : 1u);
    float normal_total = 0.0f;
    for (uint32_t pi = 0; pi < contact.point_count; pi++) {
        normal_total = normal_total + fmaxo(contact.points[pi].accumulated_normal, 0.0f);
    }
    v3 impulses[4] = {kZero, kZero, kZero, kZero};
    for (uint32_t pi = 0; pi < contact.point_count; pi++) {
        const dyn_manifold_point* p = &contact.points[pi];
        v3 aa = manifold_arm(&pair.first, p, 1), ab = manifold_arm(&pair.second, p, 0);
        float an = p->accumulated_normal, at1 = p->accumulated_tangent_1, at2 = p->accumulated_tangent_2;
        float nspeed = vdot(relative_velocity(&pair.first, &pair.second, aa, ab), normal);
        float nmass = row_mass(&pair.split_first, &pair.split_second, count, reach_a, reach_b, normal, aa, ab);
        float truth = point_momentum_mass(&pair.first, &pair.second, aa, ab, normal);
        float delta = normal_delta(&contact, nmass, truth, w->p.dt, w->targets[ci * 4u + pi], p->depth,
                                   nspeed, an);
        float next_n = fmaxo(0.0f, an + delta);
        impulses[pi] = vadd(impulses[pi], vmul(normal, next_n - an));
        an = next_n;
        float limit = contact.friction * an;
        float m1 = point_momentum_mass(&pair.split_first, &pair.split_second, aa, ab, t1) * count;
        float s1 = vdot(relative_velocity(&pair.first, &pair.second, aa, ab), t1);
        float next_1 = fclamp(at1 - s1 / m1, -limit, limit);
        impulses[pi] = vadd(impulses[pi], vmul(t1, next_1 - at1));
        at1 = next_1;
        float m2 = point_momentum_mass(&pair.split_first, &pair.split_second, aa, ab, t2) * count;
        float s2 = vdot(relative_velocity(&pair.first, &pair.second, aa, ab), t2);
        float remaining = sqrtf(fmaxo(limit * limit - at1 * at1, 0.0f));
        float next_2 = fclamp(at2 - s2 / m2, -remaining, remaining);
        impulses[pi] = vadd(impulses[pi], vmul(t2, next_2 - at2));
        at2 = next_2;
        updated.points[pi].accumulated_normal = an;
        updated.points[pi].accumulated_tangent_1 = at1;
        updated.points[pi].accumulated_tangent_2 = at2;
    }
    for (uint32_t pi = 0; pi < contact.point_count; pi++) {
        const dyn_manifold_point* p = &contact.points[pi];
        v3 aa = manifold_arm(&pair.first, p, 1), ab = manifold_arm(&pair.second, p, 0);
        apply_pair_impulse(&first, &second, aa, ab, impulses[pi]);
    }
    {
        v3 rel_spin = vsub(vload(second.s.angular_velocity), vload(first.s.angular_velocity));
        v3 planar = vsub(rel_spin, vmul(normal, vdot(rel_spin, normal)));
        float planar_len = vlen(planar);
        if (planar_len > 1e-6f) {
            v3 axis = vdiv(planar, planar_len);
            float mass = vdot(axis, apply_inverse_inertia(&pair.split_first, axis)) +
                         vdot(axis, apply_inverse_inertia(&pair.split_second, axis));
            if (mass > 0.0f) {
                float lim = contact.rolling_friction * normal_total;
                float applied = fclamp(-planar_len / mass, -lim, lim);
                vstore(first.s.angular_velocity, vsub(vload(first.s.angular_velocity),
                                                      apply_inverse_inertia(&first, vmul(axis, applied))));
                vstore(second.s.angular_velocity, vadd(vload(second.s.angular_velocity),
                                                       apply_inverse_inertia(&second, vmul(axis, applied))));
            }
        }
        float twist = vdot(rel_spin, normal);
        if (fabsf(twist) > 1e-6f) {
            float mass = vdot(normal, apply_inverse_inertia(&pair.split_first, normal)) +
                         vdot(normal, apply_inverse_inertia(&pair.split_second, normal));
            if (mass > 0.0f) {
                float lim = contact.spin_friction * normal_total;
                float applied = fclamp(-twist / mass, -lim, lim);
                vstore(first.s.angular_velocity, vsub(vload(first.s.angular_velocity),
                                                      apply_inverse_inertia(&first, vmul(normal, applied))));
                vstore(second.s.angular_velocity, vadd(vload(second.s.angular_velocity),
                                                       apply_inverse_inertia(&second, vmul(normal, applied))));
            }
        }
        w->contacts[ci] = updated;
    }
    commit_block(w, fb, sb, vsub(vload(first.s.velocity), vload(pair.first.s.velocity)),
                 vsub(vload(first.s.angular_velocity), vload(pair.first.s.angular_velocity)),
                 vsub(vload(second.s.velocity), vload(pair.second.s.velocity)),
                 vsub(vload(second.s.angular_velocity), vload(pair.second.s.angular_velocity)), residual);
}

/* solver_contact_block.wgsl:195-233 */
static void warm_block(oracle_world* w, uint32_t ci) {
    const dyn_contact* contact = &w->contacts[ci];
    uint32_t fb = w->blocks[2 * ci], sb = w->blocks[2 * ci + 1];
    if (!block_resolves(contact)) {
        commit_block(w, fb, sb, kZero, kZero, kZero, kZero, 0);
        return;
    }
    body_t fl = load_body(w, fb), sl = load_body(w, sb);
    body_t first = fl, second = sl;
    if (body_is_inert(&fl)) body_frozen(&first);
    if (body_is_inert(&sl)) body_frozen(&second);
    v3 normal = vload(contact->normal), t1, t2;
    make_tangents(normal, &t1, &t2);
    for (uint32_t pi = 0; pi < contact->point_count; pi++) {
        const dyn_manifold_point* p = &contact->points[pi];
        v3 aa = manifold_arm(&first, p, 1), ab = manifold_arm(&second, p, 0);
        v3 impulse = vadd(vadd(vmul(normal, p->accumulated_normal), vmul(t1, p->accumulated_tangent_1)),
                          vmul(t2, p->accumulated_tangent_2));
        apply_pair_impulse(&first, &second, aa, ab, impulse);
    }
    commit_block(w, fb, sb, vsub(vload(first.s.velocity), vload(fl.s.velocity)),
                 vsub(vload(first.s.angular_velocity), vload(fl.s.angular_velocity)),
                 vsub(vload(second.s.velocity), vload(sl.s.velocity)),
                 vsub(vload(second.s.angular_velocity), vload(sl.s.angular_velocity)), 0);
}

/* solver_block_apply.wgsl */
static void block_apply(oracle_world* w) {
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < w->nrows; i++) {
        uint32_t row = w->rows[i];
        uint32_t* d = w->vdeltas + row * 8u;
        v3 lin = V(solver_value(d[0], DYN_SOLVER_VELOCITY_SCALE), solver_value(d[1], DYN_SOLVER_VELOCITY_SCALE),
                   solver_value(d[2], DYN_SOLVER_VELOCITY_SCALE));
        v3 ang = V(solver_value(d[3], DYN_SOLVER_VELOCITY_SCALE), solver_value(d[4], DYN_SOLVER_VELOCITY_SCALE),
                   solver_value(d[5], DYN_SOLVER_VELOCITY_SCALE));
        memset(d, 0, 6 * 4u);
        if (vzero(lin) && vzero(ang)) continue;
        dyn_body_state* s = &w->st[row];
        vstore(s->velocity, vadd(vload(s->velocity), lin));
        vstore(s->angular_velocity, vadd(vload(s->angular_velocity), ang));
    }
}

/* substep_integrate.wgsl:15-76 (no fields) */
static void integrate(oracle_world* w) {
    float h = w->p.substep_dt;
#pragma omp parallel for schedule(static)
    for (uint32_t li = 0; li < w->nlive; li++) {
        uint32_t row = w->live[li];
        const dyn_body_desc* d = &w->de[row];
        if (!kind_simulates(d)) continue;
        dyn_body_state* s = &w->st[row];
        q4 q = qload(s->orientation);
        v3 acc = vadd(vmul(V(w->p.gravity[0], w->p.gravity[1], w->p.gravity[2]), d->gravity_scale),
                      vmul(vload(s->force), d->inverse_mass));
        v3 spin = vload(s->angular_velocity);
        vstore(s->velocity, vadd(vload(s->velocity), vmul(acc, h)));
        /* gyroscopic_spin */
        v3 gspin = spin;
        int iso = d->inertia[1] == 0.0f && d->inertia[2] == 0.0f && d->inertia[4] == 0.0f &&
                  d->inertia[0] == d->inertia[3] && d->inertia[3] == d->inertia[5];
        if (!iso) {
            v3 body_spin = qrot(qconj(q), spin);
            v3 momentum = symmetric_apply(d->inertia, body_spin);
            if (!vzero(vcross(body_spin, momentum))) {
                v3 solved = momentum;
                for (int it = 0; it < 4; it++) {
                    v3 mid = vmul(vadd(momentum, solved), 0.5f);
                    solved = rotate_about(vmul(vneg(symmetric_apply(d->inverse_inertia, mid)), h), momentum);
                }
                gspin = qrot(q, symmetric_apply(d->inverse_inertia, solved));
            }
        }
        v3 torque_term = qrot(q, symmetric_apply(d->inverse_inertia, qrot(qconj(q), vmul(vload(s->torque), h))));
        vstore(s->angular_velocity, vadd(gspin, torque_term));
        float speed = vlen(vload(s->velocity));
        if (speed > w->p.max_velocity) vstore(s->velocity, vmul(vload(s->velocity), w->p.max_velocity / speed));
        float sspeed = vlen(vload(s->angular_velocity));
        if (sspeed > w->p.max_angular_velocity)
            vstore(s->angular_velocity, vmul(vload(s->angular_velocity), w->p.max_angular_velocity / sspeed));
        vstore(s->velocity, vdiv(vload(s->velocity), 1.0f + d->linear_damping * h));
        vstore(s->angular_velocity, vdiv(vload(s->angular_velocity), 1.0f + d->angular_damping * h));
    }
}

/* substep_advance.wgsl */
static void advance(oracle_world* w) {
    float h = w->p.substep_dt;
#pragma omp parallel for schedule(static)
    for (uint32_t li = 0; li < w->nlive; li++) {
        dyn_body_state* s = &w->st[w->live[li]];
        vstore(s->position, vadd(vload(s->position), vmul(vload(s->velocity), h)));
        q4 spin = {s->angular_velocity[0], s->angular_velocity[1], s->angular_velocity[2], 0.0f};
        q4 q = qload(s->orientation);
        q4 dq = qmul(spin, q);
        float k = 0.5f * h;
        q4 sum = {q.x + k * dq.x, q.y + k * dq.y, q.z + k * dq.z, q.w + k * dq.w};
        qstore(s->orientation, qnorm(sum));
    }
}

/* position_correction.wgsl:5-46 + position_block.wgsl */
static void position_block(oracle_world* w) {
    uint32_t blocks = *ctr(w, DYN_COUNTER_BLOCKS);
#pragma omp parallel for schedule(static)
    for (uint32_t ci = 0; ci < blocks; ci++) {
        const dyn_contact* c = &w->contacts[ci];
        if (!block_resolves(c) || c->relaxation > 0.0f) continue;
        uint32_t fr = w->blocks[2 * ci], sr = w->blocks[2 * ci + 1];
        body_t fl = load_body(w, fr), sl = load_body(w, sr);
        body_t first = fl, second = sl;
        if (body_is_inert(&fl)) body_frozen(&first);
        if (body_is_inert(&sl)) body_frozen(&second);
        v3 normal = vload(c->normal);
        float resolved = vdot(normal, vsub(vload(&w->resolution[4 * sr]), vload(&w->resolution[4 * fr])));
        v3 total = kZero;
        uint32_t contributing = 0;
        for (uint32_t pi = 0; pi < c->point_count; pi++) {
            float error = fmaxo(c->points[pi].depth - w->p.slop - resolved, 0.0f);
            if (error <= 0.0f) continue;
            float k = first.d.inverse_mass + second.d.inverse_mass;
            if (k == 0.0f) continue;
            total = vadd(total, vmul(normal, w->p.relaxation * error / k));
            contributing++;
        }
        if (contributing == 0u) continue;
        v3 correction = vdiv(total, (float)contributing);
        v3 la = vmul(vneg(correction), first.d.inverse_mass);
        v3 lb = vmul(correction, second.d.inverse_mass);
        if (!(vdot(la, la) > 0.0f || vdot(lb, lb) > 0.0f)) continue; /* pair_holds */
        add_word(&w->contrib[fr], 1u);
        add_word(&w->contrib[sr], 1u);
        accumulate(w->pdeltas, fr, la, kZero, DYN_SOLVER_POSITION_SCALE);
        accumulate(w->pdeltas, sr, lb, kZero, DYN_SOLVER_POSITION_SCALE);
    }
}

/* position_apply.wgsl */
static void position_apply(oracle_world* w) {
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < w->nrows; i++) {
        uint32_t row = w->rows[i];
        uint32_t contributing = w->contrib[row];
        w->contrib[row] = 0;
        uint32_t* d = w->pdeltas + row * 8u;
        v3 lin = V(solver_value(d[0], DYN_SOLVER_POSITION_SCALE), solver_value(d[1], DYN_SOLVER_POSITION_SCALE),
                   solver_value(d[2], DYN_SOLVER_POSITION_SCALE));
        v3 ang = V(solver_value(d[3], DYN_SOLVER_POSITION_SCALE), solver_value(d[4], DYN_SOLVER_POSITION_SCALE),
                   solver_value(d[5], DYN_SOLVER_POSITION_SCALE));
        memset(d, 0, 6 * 4u);
        if (contributing == 0u) continue;
        v3 al = vdiv(lin, (float)contributing), aa = vdiv(ang, (float)contributing);
        dyn_body_state* s = &w->st[row];
        vstore(s->position, vadd(vload(s->position), al));
        q4 dq = {aa.x * 0.5f, aa.y * 0.5f, aa.z * 0.5f, 1.0f};
        qstore(s->orientation, qnorm(qmul(dq, qload(s->orientation))));
        vstore(&w->resolution[4 * row], vadd(vload(&w->resolution[4 * row]), al));
    }
}

/* ---- the step --------------------------------------------------------------------------- */

static void prepare_and_index(oracle_world* w) {
    /* begin_step.wgsl (dynamic rows) */
    for (uint32_t i = 0; i < w->d; i++) {
        dyn_body_state* s = &w->st[i];
        if (s->sleeping != 0u) continue;
        memcpy(s->prev_position, s->position, 12);
        memcpy(s->prev_orientation, s->orientation, 16);
    }
    /* broadphase_aabb.wgsl:52-62 */
    for (uint32_t i = 0; i < w->k; i++) {
        uint32_t ow = w->owner[i];
        const dyn_body_state* s = &w->st[ow];
        const dyn_body_desc* d = &w->de[ow];
        const dyn_collider* c = &w->co[i];
        world_shape ws = world_collider(s, c);
        v3 tmn, tmx;
        world_aabb_of(w, &ws, &tmn, &tmx);
        world_shape up = ws;
        up.rotation.x = up.rotation.y = up.rotation.z = 0.0f;
        up.rotation.w = 1.0f;
        v3 rmn, rmx;
        world_aabb_of(w, &up, &rmn, &rmx);
        vstore(w->res[i].min, rmn);
        vstore(w->res[i].max, rmx);
        v3 smn = tmn, smx = tmx;
        if (kind_moves(d)) {
            v3 travel = vabs(vmul(vload(s->velocity), w->p.dt));
            float spin = vlen(vmul(vload(s->angular_velocity), w->p.dt));
            v3 com = vadd(vload(s->position), qrot(qload(s->orientation), vload(d->com)));
            v3 center = vadd(vload(s->position), qrot(qload(s->orientation), vload(c->local_offset)));
            float radius = vlen(vmul(vsub(tmx, tmn), 0.5f)) + vlen(vsub(center, com));
            float reach_r = spin >= PI_F ? 2.0f * radius : 2.0f * radius * sin_rn(0.5f * spin);
            float r = w->p.contact_margin + reach_r;
            v3 reach = V(r, r, r);
            smn = vsub(vsub(tmn, travel), reach);
            smx = vadd(vadd(tmx, travel), reach);
        }
        vstore(w->aabbs[i].min, smn);
        vstore(w->aabbs[i].max, smx);
        w->aabbs[i]._pad0 = w->aabbs[i]._pad1 = 0.0f;
        /* entry flags: collider_emit.wgsl:31-63 with collider_awake (collider_entries.wgsl:12) */
        uint32_t f = 0;
        if (kind_moves(d)) f |= DYN_BP_MOBILE;
        if (w->activity[ow] != 0u || w->wake[ow] != 0u) f |= DYN_BP_AWAKE;
        w->flags[i] = f;
    }
    uint32_t bp_counters[DYN_COUNTER_BLOCK_WORDS];
    oracle_broadphase(w->aabbs, w->owner, w->flags, w->res, w->k, w->keys, w->order, w->entries,
                      w->entry_cap, w->pmaj, w->pmin, w->pair_cap, bp_counters, 1u);
    const uint32_t carried[] = {DYN_COUNTER_ENTRIES, DYN_COUNTER_PAIRS, DYN_COUNTER_REFUSED_PAIRS,
                                DYN_COUNTER_ENTRY_FAULTS, DYN_COUNTER_GRID_LEVELS,
                                DYN_COUNTER_COARSE_ACTIVE, DYN_COUNTER_GRID_SCALE,
                                DYN_COUNTER_GRID_EXTENT};
    for (size_t i = 0; i < sizeof carried / sizeof carried[0]; i++) {
        *ctr(w, carried[i]) = bp_counters[carried[i] * DYN_COUNTER_STRIDE_WORDS];
    }
    uint32_t pairs = *ctr(w, DYN_COUNTER_PAIRS);
    w->npairs = pairs < w->pair_cap ? pairs : w->pair_cap;
}

/* body_edits.wgsl:10-18 */
static void mark_disturbed(oracle_world* w, dyn_body_state* s, uint32_t row, int disturbed) {
    if (s->sleeping != 0u) {
        (*ctr(w, DYN_COUNTER_WOKE))++;
        w->wake[row] |= 1u;
        w->admitted[row] = 0;
    } else if (disturbed) {
        w->wake[row] |= 1u;
    }
    s->sleeping = 0u;
    s->sleep_timer = 0.0f;
}

/* body_edits.wgsl:34-101, one edit on its row */
static void apply_edit(oracle_world* w, uint32_t row, const dyn_body_edit* e) {
    dyn_body_state* s = &w->st[row];
    const dyn_body_desc* d = &w->de[row];
    const dyn_body_state* es = &e->state;
    body_t b;
    b.d = *d;
    switch (e->kind) {
        case DYN_EDIT_PATCH: {
            if (e->mask & DYN_PATCH_POSITION) {
                memcpy(s->position, es->position, 12);
                memcpy(s->prev_position, es->position, 12);
            }
            if (e->mask & DYN_PATCH_ORIENTATION) {
                memcpy(s->orientation, es->orientation, 16);
                memcpy(s->prev_orientation, es->orientation, 16);
            }
            if (e->mask & DYN_PATCH_VELOCITY) memcpy(s->velocity, es->velocity, 12);
            if (e->mask & DYN_PATCH_ANGULAR_VELOCITY) memcpy(s->angular_velocity, es->angular_velocity, 12);
            int teleport = (e->mask & (DYN_PATCH_POSITION | DYN_PATCH_ORIENTATION)) != 0u;
            mark_disturbed(w, s, row, teleport || !kind_simulates(d));
            break;
        }
        case DYN_EDIT_FORCE:
            vstore(s->force, vadd(vload(s->force), vload(es->force)));
            mark_disturbed(w, s, row, 0);
            break;
        case DYN_EDIT_FORCE_AT_POINT:
            b.s = *s;
            vstore(s->force, vadd(vload(s->force), vload(es->force)));
            vstore(s->torque, vadd(vload(s->torque), vcross(vsub(vload(es->position), body_com(&b)), vload(es->force))));
            mark_disturbed(w, s, row, 0);
            break;
        case DYN_EDIT_TORQUE:
            vstore(s->torque, vadd(vload(s->torque), vload(es->torque)));
            mark_disturbed(w, s, row, 0);
            break;
        case DYN_EDIT_IMPULSE:
            vstore(s->velocity, vadd(vload(s->velocity), vmul(vload(es->velocity), d->inverse_mass)));
            mark_disturbed(w, s, row, 0);
            break;
        case DYN_EDIT_IMPULSE_AT_POINT: {
            v3 impulse = vload(es->velocity);
            vstore(s->velocity, vadd(vload(s->velocity), vmul(impulse, d->inverse_mass)));
            b.s = *s;
            v3 lever = vsub(vload(es->position), body_com(&b));
            vstore(s->angular_velocity, vadd(vload(s->angular_velocity), apply_inverse_inertia(&b, vcross(lever, impulse))));
            mark_disturbed(w, s, row, 0);
            break;
        }
        case DYN_EDIT_ANGULAR_IMPULSE:
            b.s = *s;
            vstore(s->angular_velocity, vadd(vload(s->angular_velocity), apply_inverse_inertia(&b, vload(es->angular_velocity))));
            mark_disturbed(w, s, row, 0);
            break;
        case DYN_EDIT_SLEEP:
            if (s->sleeping == 0u) {
                (*ctr(w, DYN_COUNTER_SLEPT))++;
                w->admitted[row] = 0;
            }
            freeze_body(s);
            w->wake[row] = 0u;
            break;
        case DYN_EDIT_WAKE:
            if (s->sleeping != 0u) {
                (*ctr(w, DYN_COUNTER_WOKE))++;
                w->admitted[row] = 0;
            }
            s->sleep_timer = 0.0f;
            s->sleeping = 0u;
            w->wake[row] |= 1u;
            break;
        default:
            break;
    }
}

/* Queues body edits for the next step (dyn_world_edit); -1 when an id names no body or a kind is
 * unknown (nothing is queued then). */
int oracle_world_edit(oracle_world* w, const uint32_t* body_ids, const dyn_body_edit* edits, uint32_t count) {
    for (uint32_t i = 0; i < count; i++) {
        uint32_t id = body_ids[i];
        if (id >= w->id_capacity || w->row_of_body[id] >= w->n) return -1;
        if (edits[i].kind > DYN_EDIT_ANGULAR_IMPULSE) return -1;
    }
    if (w->nedits + count > w->edit_cap) {
        uint32_t cap = w->edit_cap ? w->edit_cap : 16u;
        while (cap < w->nedits + count) cap *= 2u;
        w->edits = realloc(w->edits, (size_t)cap * sizeof *w->edits);
        w->edit_rows = realloc(w->edit_rows, (size_t)cap * sizeof *w->edit_rows);
        w->edit_cap = cap;
    }
    for (uint32_t i = 0; i < count; i++) {
        w->edit_rows[w->nedits] = w->row_of_body[body_ids[i]];
        w->edits[w->nedits] = edits[i];
        w->nedits++;
    }
    return 0;
}

/* Capacity planning (World::apply_plan, crates/dynamis-world/src/backend/mod.rs:194-230; stream
 * plans crates/dynamis-broadphase/src/capacity.rs:27-80, crates/dynamis-rigid/src/capacity.rs:
 * 50-140): the demand of the step two steps back (its uncapped pair total; the larger of its
 * uncapped contact and resting totals) grows a stream it outgrew to max(demand, 2 x capacity)
 * (crates/dynamis-domain/src/capacity.rs:15-23 grown). The library plans with the same lag, so the
 * two engines refuse and grow at the same steps. Per-step arrays grow zero-filled; the archives
 * keep their contents. */
#define PLAN_LAG 2u
#define PLAN_RING 4u

static void* grow_array(void* ptr, size_t old_count, size_t new_count, size_t size) {
    char* p = realloc(ptr, (new_count ? new_count : 1u) * size);
    if (!p) abort();
    if (new_count > old_count) memset(p + old_count * size, 0, (new_count - old_count) * size);
    return p;
}

static uint32_t grown_capacity(uint32_t capacity, uint32_t required) {
    if (required <= capacity) return capacity;
    uint64_t twice = 2ull * capacity;
    uint64_t next = required > twice ? required : twice;
    return next > (1u << 30) ? (1u << 30) : (uint32_t)next;
}

static void plan_capacity(oracle_world* w) {
    if (w->issued < PLAN_LAG) return;
    const uint32_t* d = w->demand[(w->issued - PLAN_LAG) % PLAN_RING];
    uint32_t pairs = grown_capacity(w->pair_cap, d[0]);
    uint32_t contacts = grown_capacity(w->contact_cap, d[1]);
    if (pairs != w->pair_cap || contacts != w->contact_cap) w->grows++;
    if (pairs != w->pair_cap) {
        size_t o = w->pair_cap, n = pairs;
        w->pmaj = grow_array(w->pmaj, o, n, sizeof *w->pmaj);
        w->pmin = grow_array(w->pmin, o, n, sizeof *w->pmin);
        w->raw = grow_array(w->raw, o, n, sizeof *w->raw);
        w->valid = grow_array(w->valid, o, n, sizeof *w->valid);
        w->pair_cap = pairs;
    }
    if (contacts != w->contact_cap) {
        size_t o = w->contact_cap, n = contacts;
        w->contacts = grow_array(w->contacts, o, n, sizeof *w->contacts);
        w->matched = grow_array(w->matched, o, n, sizeof *w->matched);
        w->archive = grow_array(w->archive, o, n, sizeof *w->archive);
        w->resting = grow_array(w->resting, o, n, sizeof *w->resting);
        w->blocks = grow_array(w->blocks, 2u * o, 2u * n, sizeof *w->blocks);
        w->targets = grow_array(w->targets, 4u * o, 4u * n, sizeof *w->targets);
        w->pre = grow_array(w->pre, o, n, sizeof *w->pre);
        w->relay_src = grow_array(w->relay_src, 4u * o, 4u * n, sizeof *w->relay_src);
        w->contact_cap = contacts;
    }
}

static void record_demand(oracle_world* w) {
    uint32_t* d = w->demand[w->issued % PLAN_RING];
    d[0] = *ctr(w, DYN_COUNTER_PAIRS);
    uint64_t resting = (uint64_t)*ctr(w, DYN_COUNTER_RESTING) + *ctr(w, DYN_COUNTER_REFUSED_RESTING);
    uint32_t contacts = *ctr(w, DYN_COUNTER_CONTACTS);
    d[1] = resting > contacts ? (uint32_t)(resting > 0xFFFFFFFFu ? 0xFFFFFFFFu : resting) : contacts;
    w->issued++;
}

void oracle_world_capacity(const oracle_world* w, uint32_t* out) {
    out[0] = w->pair_cap;
    out[1] = w->contact_cap;
    out[2] = w->grows;
}

void oracle_world_step(oracle_world* w, float dt) {
    plan_capacity(w);
    w->p.dt = dt;
    w->p.substep_dt = dt / (float)w->p.substeps;
    for (size_t i = 0; i < sizeof kPerStep / sizeof kPerStep[0]; i++) *ctr(w, kPerStep[i]) = 0;
    /* apply_commands: clear_inputs, activity (all rows) */
    for (uint32_t i = 0; i < w->n; i++) {
        memset(w->st[i].force, 0, 12);
        memset(w->st[i].torque, 0, 12);
    }
    /* body_edits.wgsl: the queued edits, each body's in queue order (edits of different bodies
     * commute) */
    for (uint32_t i = 0; i < w->nedits; i++) apply_edit(w, w->edit_rows[i], &w->edits[i]);
    w->nedits = 0;
    memcpy(w->start, w->st, (size_t)w->n * sizeof *w->st);
    for (uint32_t i = 0; i < w->n; i++) {
        int awake = body_is_active(&w->st[i], &w->de[i]);
        w->activity[i] = awake ? 1u : 0u;
        w->motion[i] = body_is_moving(&w->st[i], &w->de[i], &w->p) ? 1u : 0u;
        if (awake) *ctr(w, DYN_COUNTER_ACTIVE) = 1u;
    }
    prepare_and_index(w);

    /* narrowphase + compaction (compact_*.wgsl) */
#pragma omp parallel for schedule(dynamic, 256)
    for (uint32_t i = 0; i < w->npairs; i++) narrow_work(w, i);
    w->ncontacts = 0;
    for (uint32_t i = 0; i < w->npairs; i++) {
        if (!w->valid[i]) continue;
        if (w->ncontacts >= w->contact_cap) {
            (*ctr(w, DYN_COUNTER_REFUSED_CONTACTS))++;
            continue;
        }
        dyn_contact c = w->raw[i];
        c.carried_normal = 0.0f;
        c.carried_tangent = 0.0f;
        w->contacts[w->ncontacts] = c;
        w->matched[w->ncontacts] = 0;
        w->ncontacts++;
    }
    uint32_t total_valid = 0;
    for (uint32_t i = 0; i < w->npairs; i++) total_valid += w->valid[i];
    *ctr(w, DYN_COUNTER_CONTACTS) = total_valid;

    /* build_islands */
    for (uint32_t i = 0; i < 4u * w->ncontacts; i++) w->relay_src[i] = NO_SLOT;
    contact_relay(w);
    contact_begin(w);
    memcpy(w->pre, w->contacts, (size_t)w->ncontacts * sizeof *w->contacts);
    w->npre = w->ncontacts;
    build_islands(w);
    /* wake, live */
    island_wake(w);
    w->nlive = 0;
    for (uint32_t i = 0; i < w->d; i++) {
        if (body_is_active(&w->st[i], &w->de[i])) w->live[w->nlive++] = i;
    }
    *ctr(w, DYN_COUNTER_LIVE) = w->nlive;
    /* solver_prepare */
    for (uint32_t i = 0; i < w->n; i++) {
        w->block_counts[i] = 0;
        memset(&w->resolution[4 * i], 0, 16);
    }
    *ctr(w, DYN_COUNTER_BLOCKS) = w->ncontacts;
    w->nrows = 0;
    /* solve_substeps */
    solver_blocks(w);
    for (uint32_t sub = 0; sub < w->p.substeps; sub++) {
        integrate(w);
        if (sub == 0) {
#pragma omp parallel for schedule(static)
            for (uint32_t ci = 0; ci < w->ncontacts; ci++) warm_block(w, ci);
            block_apply(w);
        }
        for (uint32_t it = 0; it < w->p.solve_iterations; it++) {
            /* solver_block_solve.wgsl:85-87 with the round counter of solver_block_apply.wgsl:
               only the warm-started first substep reaches rounds == solve_iterations */
            int residual = sub == 0 && it + 1u == w->p.solve_iterations;
#pragma omp parallel for schedule(static)
            for (uint32_t ci = 0; ci < w->ncontacts; ci++) solve_block(w, ci, residual);
            block_apply(w);
        }
        advance(w);
        for (uint32_t it = 0; it < w->p.position_iterations; it++) {
            position_block(w);
            position_apply(w);
        }
    }
    /* sleep */
    island_sleep(w);
    /* commit: thaw, archive, static wake clear, freeze */
    uint32_t kept = 0;
    for (uint32_t i = 0; i < w->nresting; i++) {
        const dyn_contact* c = &w->resting[i];
        uint32_t r1 = resolve_row(w, c->first_body_id, c->first_generation);
        uint32_t r2 = resolve_row(w, c->second_body_id, c->second_generation);
        if (r1 == DYN_NO_BODY || r2 == DYN_NO_BODY) continue;
        if (w->activity[r1] == 0u && w->activity[r2] == 0u) w->resting[kept++] = *c;
    }
    w->nresting = kept;
    memcpy(w->archive, w->contacts, (size_t)w->ncontacts * sizeof(dyn_contact));
    w->narchive = w->ncontacts;
    *ctr(w, DYN_COUNTER_ARCHIVED) = w->narchive;
    for (uint32_t i = w->d; i < w->n; i++) w->wake[i] = 0;
    for (uint32_t i = 0; i < w->ncontacts; i++) {
        const dyn_contact* c = &w->contacts[i];
        uint32_t f = w->owner[c->a], s = w->owner[c->b];
        if (body_is_active(&w->st[f], &w->de[f]) || body_is_active(&w->st[s], &w->de[s])) continue;
        if (w->nresting >= w->contact_cap) {
            (*ctr(w, DYN_COUNTER_REFUSED_RESTING))++;
            continue;
        }
        w->resting[w->nresting++] = *c;
    }
    *ctr(w, DYN_COUNTER_RESTING) = w->nresting;
    (*ctr(w, DYN_COUNTER_STEP))++;
    record_demand(w);
}

/* ---- lifetime and accessors ------------------------------------------------------------- */

#define ALLOC(ptr, count) ((ptr) = calloc(((count) != 0u) ? (count) : 1u, sizeof *(ptr)))

oracle_world* oracle_world_create(const dyn_step_params* params, const dyn_body_state* states,
                                  const dyn_body_desc* descs, uint32_t body_count,
                                  uint32_t dynamic_count, const dyn_collider* colliders,
                                  const uint32_t* owners, uint32_t collider_count,
                                  uint32_t contact_capacity, uint32_t pair_capacity) {
    oracle_world* w = calloc(1, sizeof *w);
    w->p = *params;
    w->n = body_count;
    w->d = dynamic_count;
    w->k = collider_count;
    ALLOC(w->st, body_count);
    ALLOC(w->de, body_count);
    ALLOC(w->co, collider_count);
    ALLOC(w->owner, collider_count);
    memcpy(w->st, states, (size_t)body_count * sizeof *states);
    memcpy(w->de, descs, (size_t)body_count * sizeof *descs);
    memcpy(w->co, colliders, (size_t)collider_count * sizeof *colliders);
    memcpy(w->owner, owners, (size_t)collider_count * 4u);
    /* the id -> row table the host keeps beside the rows (pool.rs rows) */
    w->id_capacity = 0;
    for (uint32_t i = 0; i < body_count; i++) {
        if (states[i].body_id + 1u > w->id_capacity) w->id_capacity = states[i].body_id + 1u;
    }
    ALLOC(w->row_of_body, w->id_capacity);
    for (uint32_t i = 0; i < w->id_capacity; i++) w->row_of_body[i] = DYN_NO_BODY;
    for (uint32_t i = 0; i < body_count; i++) w->row_of_body[states[i].body_id] = i;
    ALLOC(w->wake, body_count);
    ALLOC(w->admitted, body_count);
    ALLOC(w->activity, body_count);
    ALLOC(w->motion, body_count);
    w->entry_cap = 8u * (collider_count ? collider_count : 1u);
    w->pair_cap = pair_capacity;
    w->contact_cap = contact_capacity;
    ALLOC(w->aabbs, collider_count);
    ALLOC(w->res, collider_count);
    ALLOC(w->flags, collider_count);
    ALLOC(w->keys, w->entry_cap);
    ALLOC(w->order, w->entry_cap);
    ALLOC(w->entries, w->entry_cap);
    ALLOC(w->pmaj, w->pair_cap);
    ALLOC(w->pmin, w->pair_cap);
    ALLOC(w->raw, w->pair_cap);
    ALLOC(w->valid, w->pair_cap);
    ALLOC(w->contacts, w->contact_cap);
    ALLOC(w->matched, w->contact_cap);
    ALLOC(w->archive, w->contact_cap);
    ALLOC(w->resting, w->contact_cap);
    ALLOC(w->parents, body_count);
    ALLOC(w->istate, body_count);
    ALLOC(w->live, body_count);
    ALLOC(w->block_counts, body_count);
    ALLOC(w->blocks, 2u * w->contact_cap);
    ALLOC(w->rows, body_count);
    ALLOC(w->targets, 4u * w->contact_cap);
    ALLOC(w->resolution, 4u * body_count);
    ALLOC(w->vdeltas, 8u * body_count);
    ALLOC(w->pdeltas, 8u * body_count);
    ALLOC(w->contrib, body_count);
    ALLOC(w->start, body_count);
    ALLOC(w->pre, w->contact_cap);
    ALLOC(w->relay_src, 4u * w->contact_cap);
    return w;
}

int oracle_world_shapes(oracle_world* w, const dyn_shape_source* sources, uint32_t source_count,
                        const float* vertices, uint32_t vertex_count) {
    for (uint32_t i = 0; i < source_count; i++) {
        if (sources[i].vertex_count > LINEAR_SUPPORT_VERTICES ||
            (uint64_t)sources[i].vertex_offset + sources[i].vertex_count > vertex_count) {
            return -1;
        }
    }
    free(w->sources);
    free(w->vertices);
    ALLOC(w->sources, source_count);
    ALLOC(w->vertices, 4u * (size_t)vertex_count);
    memcpy(w->sources, sources, (size_t)source_count * sizeof *sources);
    memcpy(w->vertices, vertices, 16u * (size_t)vertex_count);
    w->source_count = source_count;
    w->vertex_count = vertex_count;
    return 0;
}

/* Test hook: GJK closest points and convex_hit of two colliders at given poses, over the given
 * shape sources. out = {distance, point_a xyz, point_b xyz, normal xyz, penetrating,
 * hit distance, hit normal xyz}. */
int oracle_convex_query(const dyn_shape_source* sources, uint32_t source_count, const float* vertices,
                        uint32_t vertex_count, const dyn_collider* ca, const dyn_body_state* sa,
                        const dyn_collider* cb, const dyn_body_state* sb, float* out) {
    oracle_world w;
    memset(&w, 0, sizeof w);
    w.sources = (dyn_shape_source*)sources;
    w.source_count = source_count;
    w.vertices = (float*)vertices;
    w.vertex_count = vertex_count;
    world_shape a = world_collider(sa, ca), b = world_collider(sb, cb);
    simplex_point simplex[4];
    uint32_t count = 0;
    convex_closest_t c = convex_closest(&w, &a, &b, simplex, &count);
    shape_hit hit = convex_hit(&w, &a, &b);
    out[0] = c.distance;
    vstore(out + 1, c.point_a);
    vstore(out + 4, c.point_b);
    vstore(out + 7, c.normal);
    out[10] = c.penetrating ? 1.0f : 0.0f;
    out[11] = hit.distance;
    vstore(out + 12, hit.normal);
    return 0;
}

void oracle_world_destroy(oracle_world* w) {
    if (!w) return;
    void* ptrs[] = {w->st, w->de, w->co, w->owner, w->row_of_body, w->sources, w->vertices, w->wake, w->admitted, w->activity, w->motion,
                    w->aabbs, w->res, w->flags, w->keys, w->order, w->entries, w->pmaj, w->pmin,
                    w->raw, w->valid, w->contacts, w->matched, w->archive, w->resting,
                    w->parents, w->istate, w->live, w->block_counts, w->blocks, w->rows,
                    w->targets, w->resolution, w->vdeltas, w->pdeltas, w->contrib,
                    w->start, w->pre, w->relay_src, w->edits, w->edit_rows};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; i++) free(ptrs[i]);
    free(w);
}

uint32_t oracle_world_tape(const oracle_world* w, dyn_body_state* start, dyn_contact* pre, uint32_t* relay_src,
                           uint32_t capacity) {
    memcpy(start, w->start, (size_t)w->n * sizeof *start);
    uint32_t count = w->npre < capacity ? w->npre : capacity;
    memcpy(pre, w->pre, (size_t)count * sizeof *pre);
    memcpy(relay_src, w->relay_src, (size_t)count * 16u);
    return w->npre;
}

void oracle_world_states(const oracle_world* w, dyn_body_state* out) {
    memcpy(out, w->st, (size_t)w->n * sizeof *out);
}

uint32_t oracle_world_contacts(const oracle_world* w, dyn_contact* out, uint32_t capacity) {
    uint32_t n = w->ncontacts < capacity ? w->ncontacts : capacity;
    memcpy(out, w->contacts, (size_t)n * sizeof *out);
    return w->ncontacts;
}

uint32_t oracle_world_resting(const oracle_world* w, dyn_contact* out, uint32_t capacity) {
    uint32_t n = w->nresting < capacity ? w->nresting : capacity;
    memcpy(out, w->resting, (size_t)n * sizeof *out);
    return w->nresting;
}

void oracle_world_counters(const oracle_world* w, uint32_t* out) {
    memcpy(out, w->counters, sizeof w->counters);
}

/* The solver alone (solver.rs:89-101 SolverReset, solver.rs:304-334 SolveSubsteps) over the
 * given contacts and the world's current records, with the live rows of the step (the simulated
 * rows that are awake): the same statements as the step between its island pass and its sleep
 * pass. The contacts are read (their accumulated impulses warm-start the first substep) and
 * written back with the new impulses; count <= the world's contact capacity. */
int oracle_world_solve(oracle_world* w, dyn_contact* contacts, uint32_t count, float dt) {
    if (count > w->contact_cap) return -1;
    for (uint32_t i = 0; i < count; i++) {
        if (contacts[i].a >= w->k || contacts[i].b >= w->k) return -1;
        if (w->owner[contacts[i].a] >= w->n || w->owner[contacts[i].b] >= w->n) return -1;
    }
    w->p.dt = dt;
    w->p.substep_dt = dt / (float)w->p.substeps;
    memcpy(w->contacts, contacts, (size_t)count * sizeof *contacts);
    w->ncontacts = count;
    w->nlive = 0;
    for (uint32_t i = 0; i < w->d; i++) {
        if (body_is_active(&w->st[i], &w->de[i])) w->live[w->nlive++] = i;
    }
    *ctr(w, DYN_COUNTER_LIVE) = w->nlive;
    for (uint32_t i = 0; i < w->n; i++) {
        w->block_counts[i] = 0;
        memset(&w->resolution[4 * i], 0, 16);
    }
    *ctr(w, DYN_COUNTER_BLOCKS) = w->ncontacts;
    w->nrows = 0;
    solver_blocks(w);
    for (uint32_t sub = 0; sub < w->p.substeps; sub++) {
        integrate(w);
        if (sub == 0) {
#pragma omp parallel for schedule(static)
            for (uint32_t ci = 0; ci < w->ncontacts; ci++) warm_block(w, ci);
            block_apply(w);
        }
        for (uint32_t it = 0; it < w->p.solve_iterations; it++) {
            int residual = sub == 0 && it + 1u == w->p.solve_iterations;
#pragma omp parallel for schedule(static)
            for (uint32_t ci = 0; ci < w->ncontacts; ci++) solve_block(w, ci, residual);
            block_apply(w);
        }
        advance(w);
        for (uint32_t it = 0; it < w->p.position_iterations; it++) {
            position_block(w);
            position_apply(w);
        }
    }
    memcpy(contacts, w->contacts, (size_t)count * sizeof *contacts);
    return 0;
}

/* The narrowphase pass alone (narrowphase.wgsl:518-685, dispatched at narrowphase.rs:126) over
 * the given candidate pairs and the world's current records: contacts_raw and contact_valid for
 * pairs [0, count), count <= the world's pair capacity. The world's own pair list is replaced. */
int oracle_world_narrowphase(oracle_world* w, const uint32_t* major, const uint32_t* minor, uint32_t count,
                             dyn_contact* raw, uint32_t* valid) {
    if (count > w->pair_cap) return -1;
    for (uint32_t i = 0; i < count; i++) {
        if (major[i] >= w->k || minor[i] >= w->k) return -1;
    }
    memcpy(w->pmaj, major, (size_t)count * 4u);
    memcpy(w->pmin, minor, (size_t)count * 4u);
    w->npairs = count;
#pragma omp parallel for schedule(dynamic, 256)
    for (uint32_t i = 0; i < count; i++) narrow_work(w, i);
    for (uint32_t i = 0; i < count; i++) {
        valid[i] = w->valid[i];
        if (w->valid[i]) raw[i] = w->raw[i];
    }
    return 0;
}

uint32_t oracle_world_pairs(const oracle_world* w, uint32_t* major, uint32_t* minor,
                            uint32_t capacity) {
    uint32_t n = w->npairs < capacity ? w->npairs : capacity;
    memcpy(major, w->pmaj, (size_t)n * 4u);
    memcpy(minor, w->pmin, (size_t)n * 4u);
    return w->npairs;
}

/* crates/dynamis-model/src/mass.rs:259-280 */
void oracle_inertia_inverse(const float* in, float* out) {
    float xx = in[0], xy = in[1], xz = in[2], yy = in[3], yz = in[4], zz = in[5];
    float cxx = yy * zz - yz * yz;
    float cxy = xz * yz - xy * zz;
    float cxz = xy * yz - xz * yy;
    float cyy = xx * zz - xz * xz;
    float cyz = xy * xz - xx * yz;
    float czz = xx * yy - xy * xy;
    float det = xx * cxx + xy * cxy + xz * cxz;
    out[0] = cxx / det;
    out[1] = cxy / det;
    out[2] = cxz / det;
    out[3] = cyy / det;
    out[4] = cyz / det;
    out[5] = czz / det;
}
