// One rigid-body World::step on gfx950: the pass sequence of the reference
// (crates/dynamis-world/tests/suite/pipeline.rs:6-49) for worlds without joints, CCD, soft bodies,
// characters, vehicles, fields or command edits, as a fixed sequence of HIP kernels on the
// world's stream.
//
// Device choices that keep the reference's results:
//  - every collider is emitted into one grid region per step with its owner's current awake
//    flag; the pair set equals the reference's (overlapping boxes with an awake side);
//  - per-step appends (contacts, resting manifolds) use prefix sums, never a shared atomic
//    counter, so their order is the canonical one;
//  - solver rows are all rows a resolving contact touches; a row whose accumulated words are
//    zero is left alone, exactly as the reference's apply does. Inert (static, kinematic or
//    sleeping) bodies receive exactly zero deltas, so their words are never touched, and their
//    claim / contribution marks are plain stores rather than counted atomics (their counts are
//    never read: the mass split of an inert body is frozen to zero);
//  - the resting archive is kept sorted by (min body id, max body id) for the contact_begin
//    lookup (the reference sorts an index over a free-list archive, commit.rs:278-295).
#include <algorithm>
#include <vector>

#include "convex.h"
#include "sort.h"

namespace dyn {

namespace {

constexpr int kThreads = 256;
// Row kernels that count into a counter word stride over their rows with at most this many
// workgroups, each adding its workgroup's count with one atomic: atomics on one word run one
// after another, and one per wave over 1.5 M rows cost more than the rows' own work.
constexpr uint32_t kCountBlocks = 1024;
constexpr uint32_t kIslandWake = 1u;
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;
constexpr uint32_t kFlattenFlags = 8;   // per flatten pass: a word per round (at most 6 rounds)
constexpr uint32_t kHookRounds = 2;     // hooking rounds (each with its flatten pass) before the link
constexpr uint32_t kFlattenPasses = kHookRounds + 2;
constexpr uint32_t kFlattenHops = 63;   // hops a flatten walk takes past a row's parent
constexpr float kHalfTurn = 3.141592653589793f;

__device__ __forceinline__ uint32_t* cw(uint32_t* counters, uint32_t slot) {
    return counters + slot * DYN_COUNTER_STRIDE_WORDS;
}

// counter.rs:75-128: slots cleared every step (reset_counters.wgsl)
__constant__ uint32_t kPerStepSlots[] = {
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
constexpr uint32_t kPerStepCount = sizeof(kPerStepSlots) / sizeof(kPerStepSlots[0]);

// Device view of a world's buffers (passed by value to kernels).
struct W {
    dyn_step_params p;
    uint32_t n, d, k;
    dyn_body_state* st;
    dyn_body_desc* de;
    dyn_collider* co;
    uint32_t* owner;
    const uint32_t* worlds;  // world of each collider in a batch, or null for one world
    uint32_t* row_of_body;  // body id -> row, ids < id_cap
    uint32_t id_cap;
    const dyn_shape_source* sources;  // hull shape sources and their float4 vertex rows
    const float4* vertices;
    const uint2* support_cells;      // per source and cube-map cell: a two-word record (support_cells.h)
    const uint16_t* support_cands;   // candidate vertex indices, ascending within a cell
    uint32_t source_count;
    uint32_t* wake;
    uint32_t* activity;
    uint32_t* motion;
    uint32_t* counters;
    dyn_aabb* aabbs;
    dyn_aabb* res;
    uint32_t* flags;
    uint32_t* pmaj;
    uint32_t* pmin;
    uint32_t pair_cap;
    dyn_contact* raw;
    uint32_t* valid;
    uint8_t* pkind;        // per convex pair: the unordered class of its two shape kinds
    uint32_t* worklist;    // convex pairs grouped by kind class (k_class_scatter)
    uint32_t* class_hist;  // per kind class and block of pairs: count, then offset in the class;
                           // then the classes' totals
    uint32_t* work_total;  // device words: entries on the worklist of modes 0-2, the hull-class GJK pairs,
                           // the resumed GJK pairs (mode 4), the EPA polytopes given a wave, the
                           // candidate pairs (mode 6)
    uint32_t* work_minor;  // the sort's unused minor lane
    uint32_t* dest;
    dyn_contact* contacts;
    uint2* ckeys;          // per contact: its collider pair (a, b), the key the relay searches
    uint32_t* matched;
    uint32_t contact_cap;
    uint32_t* ncontacts;  // device word: contacts this step
    dyn_contact* archive;
    uint32_t* narchive;   // device word
    dyn_contact* resting;
    uint32_t* nresting;   // device word
    dyn_contact* rest_tmp;
    uint32_t* rest_major;
    uint32_t* rest_minor;
    uint32_t* rest_slot;
    uint32_t* rest_flags;
    uint32_t* parents;
    uint32_t* istate;
    uint32_t* flat_flags;  // kFlattenFlags words per flatten pass: set when a round left a row short of its root
    uint2* iedges;         // per contact: the rows it links in the island pass, or kNoSlot twice (null: read the contacts)
    uint32_t* block_counts;
    uint32_t* blocks;
    float* targets;
    float4* prows;  // per contact block: the position phase's three float4 (k_blocks)
    float4* srows;    // per contact block, per substep: kSolveRow float4 for the velocity iterations,
                      // section-major (solve_section)
    float4* sbodies;  // per row, per substep: orientation, frozen inverse inertia and mass, split count, velocities
    float4* sposes;   // per row, per substep: position and inert flag, local centre of mass
    float4* ppose;    // per claimed row, during the position iterations: position, orientation
    float4* sanchor;  // per contact point, per step (first k_solve_prep): local anchors and depth,
                      // point-major like srows
    float4* bbodies;  // per row, per step (k_solver_reset): world centre of mass and inverse mass,
                      // velocity and inert flag, angular velocity, for k_blocks
    float4* simp;     // per contact block, for the step's iterations: accumulated normal, tangent 1, tangent 2
    float* resolution;
    // Solver deltas without atomics: every block writes its fixed-point words to its own slots
    // and every row sums the words of the blocks on its list. The words are integers, so the
    // sum does not depend on the order of the list.
    // A row's list is a run of entries from row_start, one per block on it. The velocity words
    // stay with their block (k_solve writes a block's 12 words together) and the row gathers
    // them through adj; a block writes its position words straight into its sides' entries, so
    // the position pass reads a row's run in order.
    uint32_t* slots;        // per block and side: k_blocks' place in the row's list, then
                            // (k_fill_lists) the entry's index; kNoSlot for an inert side
    uint32_t* list_counts;  // per row: blocks on its list (0 for inert rows)
    uint32_t* row_start;    // per row: its first list entry
    uint32_t* adj;          // per list entry: block << 1 | side
    uint32_t* vwords;       // per block: 12 velocity words (first lin, ang; second lin, ang)
    uint4* pwords;          // per list entry: {contributed, lin words} of its block's last
                            // position pass
    uint32_t* contrib;      // inert rows: marked when a block of theirs contributed
    uint32_t* scan_blocks;   // per-block sums of compaction scans
    uint32_t* scan_total;    // two words: totals of the last scans
    uint32_t* relay_src;     // while a tape is recorded: per contact point, the archive point relayed into it
    // Body edits queued by dyn_world_edit for the next step: runs (row, first edit, edit count)
    // sorted by row over the edits in queue order; *edit_runs_count runs (0: no edits).
    const dyn_body_edit* edits;
    const uint4* edit_runs;
    uint32_t* edit_runs_count;
};

__device__ __forceinline__ Body load_body(const W& w, uint32_t row) {
    Body b;
    b.s = w.st[row];
    b.d = w.de[row];
    return b;
}

// ---- counters, inputs, activity, begin_step -------------------------------------------------

__global__ void k_reset_counters(W w) {
    const uint32_t t = threadIdx.x;
    if (t < kPerStepCount) *cw(w.counters, kPerStepSlots[t]) = 0u;
}

// body_edits.wgsl:10-18 mark_disturbed
__device__ __forceinline__ void mark_disturbed(const W& w, dyn_body_state& s, uint32_t row, bool disturbed) {
    if (s.sleeping != 0u) {
        atomicAdd(cw(w.counters, DYN_COUNTER_WOKE), 1u);
        atomicOr(&w.wake[row], 1u);
    } else if (disturbed) {
        atomicOr(&w.wake[row], 1u);
    }
    s.sleeping = 0u;
    s.sleep_timer = 0.0f;
}

// body_edits.wgsl:34-101: the row's queued edits, in queue order, on its state. (The reference also
// clears the row's admission to the resting grid region, body_admitted; this build emits every
// collider into one region each step, so there is no admission to clear.)
__device__ inline void apply_edits(const W& w, uint32_t row, dyn_body_state& s, const dyn_body_desc& d) {
    const uint32_t runs = *w.edit_runs_count;
    if (runs == 0u) return;
    uint32_t lo = 0u, hi = runs;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (w.edit_runs[mid].x < row) lo = mid + 1u;
        else hi = mid;
    }
    if (lo >= runs || w.edit_runs[lo].x != row) return;
    const uint4 run = w.edit_runs[lo];
    Body b;
    b.d = d;
    for (uint32_t e = 0; e < run.z; ++e) {
        const dyn_body_edit& edit = w.edits[run.y + e];
        const dyn_body_state& es = edit.state;
        if (edit.kind == DYN_EDIT_PATCH) {
            if (edit.mask & DYN_PATCH_POSITION) {
                for (int c = 0; c < 3; ++c) s.position[c] = s.prev_position[c] = es.position[c];
            }
            if (edit.mask & DYN_PATCH_ORIENTATION) {
                for (int c = 0; c < 4; ++c) s.orientation[c] = s.prev_orientation[c] = es.orientation[c];
            }
            if (edit.mask & DYN_PATCH_VELOCITY) {
                for (int c = 0; c < 3; ++c) s.velocity[c] = es.velocity[c];
            }
            if (edit.mask & DYN_PATCH_ANGULAR_VELOCITY) {
                for (int c = 0; c < 3; ++c) s.angular_velocity[c] = es.angular_velocity[c];
            }
            const bool teleport = (edit.mask & (DYN_PATCH_POSITION | DYN_PATCH_ORIENTATION)) != 0u;
            mark_disturbed(w, s, row, teleport || !kind_simulates(d));
        } else if (edit.kind == DYN_EDIT_FORCE) {
            st3(s.force, add(ld3(s.force), ld3(es.force)));
            mark_disturbed(w, s, row, false);
        } else if (edit.kind == DYN_EDIT_FORCE_AT_POINT) {
            b.s = s;
            st3(s.force, add(ld3(s.force), ld3(es.force)));
            st3(s.torque, add(ld3(s.torque), cross3(sub(ld3(es.position), body_com(b)), ld3(es.force))));
            mark_disturbed(w, s, row, false);
        } else if (edit.kind == DYN_EDIT_TORQUE) {
            st3(s.torque, add(ld3(s.torque), ld3(es.torque)));
            mark_disturbed(w, s, row, false);
        } else if (edit.kind == DYN_EDIT_IMPULSE) {
            st3(s.velocity, add(ld3(s.velocity), mul(ld3(es.velocity), d.inverse_mass)));
            mark_disturbed(w, s, row, false);
        } else if (edit.kind == DYN_EDIT_IMPULSE_AT_POINT) {
            const v3 impulse = ld3(es.velocity);
            st3(s.velocity, add(ld3(s.velocity), mul(impulse, d.inverse_mass)));
            b.s = s;
            const v3 lever = sub(ld3(es.position), body_com(b));
            st3(s.angular_velocity, add(ld3(s.angular_velocity), apply_inverse_inertia(b, cross3(lever, impulse))));
            mark_disturbed(w, s, row, false);
        } else if (edit.kind == DYN_EDIT_ANGULAR_IMPULSE) {
            b.s = s;
            st3(s.angular_velocity, add(ld3(s.angular_velocity), apply_inverse_inertia(b, ld3(es.angular_velocity))));
            mark_disturbed(w, s, row, false);
        } else if (edit.kind == DYN_EDIT_SLEEP) {
            if (s.sleeping == 0u) atomicAdd(cw(w.counters, DYN_COUNTER_SLEPT), 1u);
            for (int c = 0; c < 3; ++c) {  // core.wgsl:218-227 freeze_body
                s.velocity[c] = s.angular_velocity[c] = s.force[c] = s.torque[c] = 0.0f;
                s.prev_position[c] = s.position[c];
            }
            for (int c = 0; c < 4; ++c) s.prev_orientation[c] = s.orientation[c];
            s.sleep_timer = 0.0f;
            s.sleeping = 1u;
            atomic_store_agent(&w.wake[row], 0u);
        } else if (edit.kind == DYN_EDIT_WAKE) {
            if (s.sleeping != 0u) atomicAdd(cw(w.counters, DYN_COUNTER_WOKE), 1u);
            s.sleep_timer = 0.0f;
            s.sleeping = 0u;
            atomicOr(&w.wake[row], 1u);
        }
    }
}

// clear_inputs.wgsl, body_edits.wgsl, activity.wgsl (all rows), begin_step.wgsl (dynamic rows)
__global__ void __launch_bounds__(kThreads) k_begin(W w) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool awake = false;
    if (i < w.n) {
        dyn_body_state s = w.st[i];
        const dyn_body_desc d = w.de[i];
        s.force[0] = s.force[1] = s.force[2] = 0.0f;
        s.torque[0] = s.torque[1] = s.torque[2] = 0.0f;
        apply_edits(w, i, s, d);
        awake = body_is_active(s, d);
        w.activity[i] = awake ? 1u : 0u;
        bool moving;
        const bool driven = kind_moves(d) && !kind_simulates(d);
        if (driven && (!zero3(ld3(s.velocity)) || !zero3(ld3(s.angular_velocity)))) {
            moving = true;
        } else if (!kind_simulates(d) || s.sleeping != 0u) {
            moving = false;
        } else {
            const float sv = (d.flags & DYN_OVERRIDE_SLEEP_LINEAR) ? d.sleep_velocity : w.p.sleep_velocity;
            const float sw = (d.flags & DYN_OVERRIDE_SLEEP_ANGULAR) ? d.sleep_angular_velocity
                                                                   : w.p.sleep_angular_velocity;
            moving = len3(ld3(s.velocity)) > sv || len3(ld3(s.angular_velocity)) > sw;
        }
        w.motion[i] = moving ? 1u : 0u;
        if (i < w.d && s.sleeping == 0u) {
            s.prev_position[0] = s.position[0];
            s.prev_position[1] = s.position[1];
            s.prev_position[2] = s.position[2];
            for (int c = 0; c < 4; ++c) s.prev_orientation[c] = s.orientation[c];
        }
        w.st[i] = s;
    }
    if (__any(awake) && lane_id() == 0) *cw(w.counters, DYN_COUNTER_ACTIVE) = 1u;
}

// broadphase_aabb.wgsl:26-62, plus the entry flags of collider_emit.wgsl:31-63
__global__ void __launch_bounds__(kThreads) k_aabb(W w) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w.k) return;
    const uint32_t ow = w.owner[i];
    const dyn_body_state s = w.st[ow];
    const dyn_body_desc d = w.de[ow];
    const dyn_collider c = w.co[i];
    const Shape ws = world_collider(s, c);
    const ShapeSet ss{w.sources, w.vertices, w.support_cells, w.support_cands};
    v3 tmn, tmx;
    world_aabb_of(ws, ss, &tmn, &tmx);
    Shape up = ws;
    up.rotation = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    v3 rmn, rmx;
    world_aabb_of(up, ss, &rmn, &rmx);
    v3 smn = tmn, smx = tmx;
    if (kind_moves(d)) {
        const v3 travel = absv(mul(ld3(s.velocity), w.p.dt));
        const float spin = len3(mul(ld3(s.angular_velocity), w.p.dt));
        const v3 com = add(ld3(s.position), qrot(ld4(s.orientation), ld3(d.com)));
        const v3 center = add(ld3(s.position), qrot(ld4(s.orientation), ld3(c.local_offset)));
        const float radius = len3(mul(sub(tmx, tmn), 0.5f)) + len3(sub(center, com));
        const float reach_r = spin >= kHalfTurn ? 2.0f * radius : 2.0f * radius * sin_rn(0.5f * spin);
        const float r = w.p.contact_margin + reach_r;
        const v3 reach = mk(r, r, r);
        smn = sub(sub(tmn, travel), reach);
        smx = add(add(tmx, travel), reach);
    }
    float4* a = reinterpret_cast<float4*>(w.aabbs + i);
    a[0] = make_float4(smn.x, smn.y, smn.z, 0.0f);
    a[1] = make_float4(smx.x, smx.y, smx.z, 0.0f);
    float4* r = reinterpret_cast<float4*>(w.res + i);
    r[0] = make_float4(rmn.x, rmn.y, rmn.z, 0.0f);
    r[1] = make_float4(rmx.x, rmx.y, rmx.z, 0.0f);
    uint32_t f = 0;
    if (kind_moves(d)) f |= DYN_BP_MOBILE;
    if (w.activity[ow] != 0u || w.wake[ow] != 0u) f |= DYN_BP_AWAKE;
    w.flags[i] = f;
}

// ---- narrowphase (narrowphase.wgsl:518-685) -------------------------------------------------

__device__ __forceinline__ float material_combine(float a, float b, uint32_t mode) {
    if (mode == 1u) return fmn(a, b);
    if (mode == 2u) return fmx(a, b);
    if (mode == 3u) return (a + b) * 0.5f;
    return sqrtf(a * b);
}

__device__ __forceinline__ bool analytic(uint32_t kind) {
    return kind == DYN_SHAPE_SPHERE || kind == DYN_SHAPE_CUBOID || kind == DYN_SHAPE_CAPSULE;
}

__device__ __forceinline__ bool scaled(const dyn_collider& c) {
    return c.scale[0] != 1.0f || c.scale[1] != 1.0f || c.scale[2] != 1.0f;
}

__device__ __forceinline__ void mirror(Manifold& m) {
    m.normal = neg(m.normal);
}

// Whether a pair leaves the analytic generators for GJK/EPA and clipped features
// (narrowphase.wgsl:551-664): planes against hulls and cylinders, scaled shapes, any hull or
// cylinder, and capsule-box pairs.
__device__ __forceinline__ bool convex_path(const dyn_collider& fc, const dyn_collider& sc) {
    const bool fw = fc.kind == DYN_SHAPE_PLANE, sw = sc.kind == DYN_SHAPE_PLANE;
    if (fw) return !analytic(sc.kind);
    if (sw) return !analytic(fc.kind);
    if (scaled(fc) || scaled(sc) || !analytic(fc.kind) || !analytic(sc.kind)) return true;
    return (fc.kind == DYN_SHAPE_CAPSULE && sc.kind == DYN_SHAPE_CUBOID) ||
           (fc.kind == DYN_SHAPE_CUBOID && sc.kind == DYN_SHAPE_CAPSULE);
}

// The pair filter of narrowphase.wgsl:518-550; false when the pair makes no contact.
__device__ __forceinline__ bool pair_admitted(const W& w, uint32_t fs, uint32_t ss, uint32_t* fb_out,
                                              uint32_t* sb_out) {
    // indices outside the records (possible only for pairs handed to dyn_narrowphase) make no
    // contact
    if (fs >= w.k || ss >= w.k) return false;
    const uint32_t fb = w.owner[fs], sb = w.owner[ss];
    if (fb == DYN_NO_BODY || sb == DYN_NO_BODY || fb == sb || fb >= w.n || sb >= w.n) return false;
    const dyn_body_desc& fd = w.de[fb];
    const dyn_body_desc& sd = w.de[sb];
    if (!kind_moves(fd) && !kind_moves(sd)) return false;
    const dyn_collider& fc = w.co[fs];
    const dyn_collider& sc = w.co[ss];
    if (fc.kind == DYN_SHAPE_NONE || sc.kind == DYN_SHAPE_NONE) return false;
    if ((fc.kind == DYN_SHAPE_HULL && fc.source >= w.source_count) ||
        (sc.kind == DYN_SHAPE_HULL && sc.source >= w.source_count)) {
        return false;
    }
    const uint32_t g1 = fc.collision_group != DYN_NO_COLLISION_FILTER ? fc.collision_group : fd.collision_group;
    const uint32_t m1 = fc.collision_mask != DYN_NO_COLLISION_FILTER ? fc.collision_mask : fd.collision_mask;
    const uint32_t g2 = sc.collision_group != DYN_NO_COLLISION_FILTER ? sc.collision_group : sd.collision_group;
    const uint32_t m2 = sc.collision_mask != DYN_NO_COLLISION_FILTER ? sc.collision_mask : sd.collision_mask;
    if (!((g1 & m2) != 0u && (g2 & m1) != 0u)) return false;
    if (fc.kind == DYN_SHAPE_PLANE && sc.kind == DYN_SHAPE_PLANE) return false;
    *fb_out = fb;
    *sb_out = sb;
    return true;
}

__device__ __forceinline__ float pair_margin(const W& w, const Body& first, const Body& second) {
    if ((first.d.flags & DYN_BODY_CCD) || (second.d.flags & DYN_BODY_CCD)) {
        return w.p.contact_margin + len3(sub(ld3(second.s.velocity), ld3(first.s.velocity))) * w.p.dt;
    }
    return w.p.contact_margin;
}

// What a candidate pair's `valid` word says between the narrowphase kernels. Only kPairContact
// survives them: every other mark is cleared by the kernel that takes the pair up.
constexpr uint32_t kPairNone = 0;
constexpr uint32_t kPairContact = 1;  // its contact is written in the pair's raw slot
constexpr uint32_t kPairGjk = 2;      // convex pair for k_narrow_convex
constexpr uint32_t kPairEpa = 3;      // GJK found it penetrating; simplex parked for k_narrow_epa
constexpr uint32_t kPairEpaDone = 4;  // EPA result parked for k_narrow_finish
constexpr uint32_t kPairHit = 5;      // GJK hit within the margin parked for k_narrow_finish
constexpr uint32_t kPairPlane = 6;    // plane against a hull or cylinder, for k_narrow_finish
constexpr uint32_t kPairGjkResume = 7;  // GJK's state parked for k_narrow_convex_resume
constexpr uint32_t kPairEpaWave = 8;    // EPA result of the wave kernel parked for k_narrow_finish_tail

// Material, identity and anchors of a generated manifold (narrowphase.wgsl:665-684).
__device__ inline void write_contact(const W& w, uint32_t index, uint32_t fs, uint32_t ss, const Body& first,
                                     const Body& second, const dyn_collider& fc, const dyn_collider& sc,
                                     const Manifold& m, bool mirrored_features) {
    dyn_contact c;
    c.a = fs;
    c.b = ss;
    c.point_count = m.count;
    c.sensor = ((fc.flags & DYN_COLLIDER_SENSOR) || (sc.flags & DYN_COLLIDER_SENSOR)) ? 1u : 0u;
    c.first_body_id = first.s.body_id;
    c.second_body_id = second.s.body_id;
    c.first_generation = first.s.generation;
    c.second_generation = second.s.generation;
    st3(c.normal, m.normal);
    c.events = (fc.flags & sc.flags) & (2u | 4u);
    c.surface = 0xFFFFFFFFu;
    c.triangle = 0xFFFFFFFFu;
    c.friction = material_combine(fc.friction, sc.friction, w.p.friction_combine);
    c.restitution = material_combine(fc.restitution, sc.restitution, w.p.restitution_combine);
    c.rolling_friction = fmx(fc.rolling_friction, sc.rolling_friction);
    c.spin_friction = fmx(fc.spin_friction, sc.spin_friction);
    c.carried_normal = 0.0f;
    c.carried_tangent = 0.0f;
    c.relaxation = fc.relaxation + sc.relaxation;
    c.damping_ratio = fmx(fc.damping_ratio, sc.damping_ratio);
    c._pad_head[0] = c._pad_head[1] = 0u;
    const q4 qa = qconj(ld4(first.s.orientation)), qb = qconj(ld4(second.s.orientation));
    const v3 pa = ld3(first.s.position), pb = ld3(second.s.position);
    for (uint32_t i = 0; i < 4; ++i) {
        dyn_manifold_point& mp = c.points[i];
        if (i < m.count) {
            st3(mp.position, m.pos[i]);
            mp.depth = m.depth[i];
            st3(mp.local_a, qrot(qa, sub(m.pos[i], pa)));
            st3(mp.local_b, qrot(qb, sub(m.pos[i], pb)));
            mp.feature = mirrored_features ? feature_mirror(m.feature[i]) : m.feature[i];
        } else {
            mp.position[0] = mp.position[1] = mp.position[2] = 0.0f;
            mp.depth = 0.0f;
            mp.local_a[0] = mp.local_a[1] = mp.local_a[2] = 0.0f;
            mp.local_b[0] = mp.local_b[1] = mp.local_b[2] = 0.0f;
            mp.feature = 0u;
        }
        mp._pad_a = 0.0f;
        mp._pad_tail = 0u;
        mp.accumulated_normal = 0.0f;
        mp.accumulated_tangent_1 = 0.0f;
        mp.accumulated_tangent_2 = 0.0f;
    }
    w.raw[index] = c;
    w.valid[index] = kPairContact;
}

// Whether two shapes' unswept bounds are apart by more than the margin along some axis. Then
// the reference makes no contact: the axis probe of convex_hit sees the gap, and GJK's distance
// is at least the gap (convex.wgsl:584-604); the 1 mm slack covers the rounding of the bounds.
__device__ __forceinline__ bool bounds_apart(const Shape& a, const Shape& b, const ShapeSet& ss, float margin) {
    v3 amn, amx, bmn, bmx;
    world_aabb_of(a, ss, &amn, &amx);
    world_aabb_of(b, ss, &bmn, &bmx);
    const float reach = margin + 1e-3f;
    return bmn.x - amx.x > reach || amn.x - bmx.x > reach || bmn.y - amx.y > reach ||
           amn.y - bmx.y > reach || bmn.z - amx.z > reach || amn.z - bmx.z > reach;
}

// ---- convex worklists: pairs grouped by the kinds of their shapes ---------------------------

constexpr uint32_t kClasses = 32;       // unordered pairs of 7 kind slots (28) and spare
constexpr uint32_t kHullClasses = 21;   // the classes of pairs with a hull come last, from here
constexpr uint32_t kClassItems = 16;
constexpr uint32_t kClassBlock = kThreads * kClassItems;  // 4096 pairs per block
constexpr uint32_t kNoClass = 0xFFu;

__device__ __forceinline__ uint32_t kind_slot(uint32_t kind) {
    switch (kind) {
        case DYN_SHAPE_SPHERE: return 0u;
        case DYN_SHAPE_CUBOID: return 1u;
        case DYN_SHAPE_CAPSULE: return 2u;
        case DYN_SHAPE_CYLINDER: return 3u;
        case DYN_SHAPE_PLANE: return 4u;
        case DYN_SHAPE_HULL: return 6u;
        default: return 5u;
    }
}

// Pairs without a hull: the row-major upper triangle of slots 0-5 (classes 0-20); pairs with a
// hull: kHullClasses + the other shape's slot (classes 21-27).
__device__ __forceinline__ uint8_t kind_class(uint32_t first_kind, uint32_t second_kind) {
    uint32_t a = kind_slot(first_kind), b = kind_slot(second_kind);
    if (a > b) {
        const uint32_t t = a;
        a = b;
        b = t;
    }
    if (b == 6u) return static_cast<uint8_t>(kHullClasses + a);
    return static_cast<uint8_t>(a * 6u - (a * (a - 1u)) / 2u + (b - a));
}

// The pairs a narrowphase kernel takes: kPairGjk (mode 0), the pairs k_narrow_finish completes
// (mode 1), kPairEpa (mode 2), kPairGjkResume (mode 4), every candidate pair (mode 6, for
// k_narrow; its marks are not written yet).
constexpr uint32_t kWorkAllPairs = 6;
__device__ __forceinline__ uint32_t work_class(const W& w, uint32_t index, uint32_t mode) {
    if (mode == kWorkAllPairs) return static_cast<uint32_t>(w.pkind[index]);
    const uint32_t v = w.valid[index];
    const bool take = mode == 0u   ? v == kPairGjk
                      : mode == 1u ? (v == kPairEpaDone || v == kPairHit || v == kPairPlane)
                      : mode == 4u ? v == kPairGjkResume
                                   : v == kPairEpa;
    return take ? static_cast<uint32_t>(w.pkind[index]) : kNoClass;
}

__global__ void __launch_bounds__(kThreads) k_class_count(W w, uint32_t mode) {
    __shared__ uint32_t hist[kClasses];
    if (threadIdx.x < kClasses) hist[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t live = min(*cw(w.counters, DYN_COUNTER_PAIRS), w.pair_cap);
    const uint32_t base = blockIdx.x * kClassBlock;
    for (uint32_t j = 0; j < kClassItems; ++j) {
        const uint32_t index = base + j * kThreads + threadIdx.x;
        if (index >= live) break;
        const uint32_t c = work_class(w, index, mode);
        if (c != kNoClass) atomicAdd(&hist[c], 1u);
    }
    __syncthreads();
    if (threadIdx.x < kClasses) w.class_hist[threadIdx.x * gridDim.x + blockIdx.x] = hist[threadIdx.x];
}

// One workgroup per kind class: the class's block counts become offsets within the class, and
// the class's total goes to class_hist[kClasses * blocks + class].
__global__ void __launch_bounds__(1024) k_class_scan(uint32_t* class_hist, uint32_t blocks) {
    __shared__ uint32_t s_scan[1024 / kWave];
    __shared__ uint32_t s_carry;
    uint32_t* row = class_hist + static_cast<size_t>(blockIdx.x) * blocks;
    if (threadIdx.x == 0) s_carry = 0;
    for (uint32_t first = 0; first < blocks; first += 1024) {
        const uint32_t b = first + threadIdx.x;
        const uint32_t v = b < blocks ? row[b] : 0u;
        uint32_t total = 0;
        const uint32_t excl = block_exclusive_sum<1024>(v, s_scan, &total);
        if (b < blocks) row[b] = s_carry + excl;
        __syncthreads();
        if (threadIdx.x == 0) s_carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) class_hist[static_cast<size_t>(kClasses) * blocks + blockIdx.x] = s_carry;
}

// After k_class_scan: each block places its pairs of every class at the class's start (the
// totals of the classes before it) plus its offset within the class (in any order within the
// block: every pair's result goes to its own slot); block 0 writes the worklist's length.
__global__ void __launch_bounds__(kThreads) k_class_scatter(W w, uint32_t mode) {
    static_assert(kClasses <= kWave, "one wave sums the class totals");
    __shared__ uint32_t cursor[kClasses];
    if (threadIdx.x < kWave) {
        const uint32_t c = threadIdx.x;
        const uint32_t total = c < kClasses ? w.class_hist[static_cast<size_t>(kClasses) * gridDim.x + c] : 0u;
        const uint32_t inclusive = wave_inclusive_sum(total);
        if (c < kClasses) cursor[c] = inclusive - total + w.class_hist[c * gridDim.x + blockIdx.x];
        if (blockIdx.x == 0u && c == kWave - 1u) w.work_total[mode] = inclusive;
    }
    __syncthreads();
    const uint32_t live = min(*cw(w.counters, DYN_COUNTER_PAIRS), w.pair_cap);
    const uint32_t base = blockIdx.x * kClassBlock;
    for (uint32_t j = 0; j < kClassItems; ++j) {
        const uint32_t index = base + j * kThreads + threadIdx.x;
        if (index >= live) break;
        const uint32_t c = work_class(w, index, mode);
        if (c != kNoClass) w.worklist[atomicAdd(&cursor[c], 1u)] = index;
    }
}

// Analytic pairs; convex-path pairs are marked for the kernels below.
// Every candidate pair's kind class (pairs naming colliders outside the records: class 0), so
// k_narrow's worklist groups its pairs by kind and a wave runs one generator.
__global__ void __launch_bounds__(kThreads) k_pair_class(W w) {
    const uint32_t index = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t live = min(*cw(w.counters, DYN_COUNTER_PAIRS), w.pair_cap);
    if (index >= live) return;
    const uint32_t fs = w.pmaj[index], ss = w.pmin[index];
    w.pkind[index] = (fs < w.k && ss < w.k) ? kind_class(w.co[fs].kind, w.co[ss].kind) : 0u;
}

// One thread per candidate pair of the kind-grouped worklist (every pair's result goes to its
// own slot, so the order changes nothing).
__global__ void __launch_bounds__(kThreads) k_narrow(W w) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= w.work_total[kWorkAllPairs]) return;
    const uint32_t index = w.worklist[t];
    w.valid[index] = kPairNone;
    const uint32_t fs = w.pmaj[index], ss = w.pmin[index];
    uint32_t fb, sb;
    if (!pair_admitted(w, fs, ss, &fb, &sb)) return;
    const dyn_collider fc = w.co[fs], sc = w.co[ss];
    const Body first = load_body(w, fb), second = load_body(w, sb);
    const float margin = pair_margin(w, first, second);
    const ShapeSet shapes{w.sources, w.vertices, w.support_cells, w.support_cands};
    if (convex_path(fc, sc)) {  // pkind holds the pair's class (k_pair_class)
        if (fc.kind == DYN_SHAPE_PLANE || sc.kind == DYN_SHAPE_PLANE) {
            w.valid[index] = kPairPlane;
        } else if (!bounds_apart(world_collider(first.s, fc), world_collider(second.s, sc), shapes, margin)) {
            w.valid[index] = kPairGjk;
        }
        return;
    }
    Manifold m;
    m.count = 0;
    m.normal = zero3v();
    bool mirrored_features = false;
    if (fc.kind == DYN_SHAPE_PLANE) {  // the other shape is a sphere, box or capsule (convex_path)
        plane_analytic(m, world_collider(first.s, fc), world_collider(second.s, sc), shapes, margin);
    } else if (sc.kind == DYN_SHAPE_PLANE) {
        plane_analytic(m, world_collider(second.s, sc), world_collider(first.s, fc), shapes, margin);
        mirror(m);
        mirrored_features = true;
    } else {
        const uint32_t a = fc.kind, b = sc.kind;
        if (a == DYN_SHAPE_SPHERE && b == DYN_SHAPE_SPHERE) sphere_sphere(m, first, fc, second, sc, margin);
        else if (a == DYN_SHAPE_SPHERE && b == DYN_SHAPE_CUBOID) sphere_box(m, first, fc, second, sc, margin);
        else if (a == DYN_SHAPE_SPHERE && b == DYN_SHAPE_CAPSULE) sphere_capsule(m, first, fc, second, sc, margin);
        else if (a == DYN_SHAPE_CUBOID && b == DYN_SHAPE_SPHERE) {
            sphere_box(m, second, sc, first, fc, margin);
            mirror(m);
        } else if (a == DYN_SHAPE_CUBOID && b == DYN_SHAPE_CUBOID) box_box_sat(m, first, fc, second, sc, margin);
        else if (a == DYN_SHAPE_CAPSULE && b == DYN_SHAPE_SPHERE) {
            sphere_capsule(m, second, sc, first, fc, margin);
            mirror(m);
        } else if (a == DYN_SHAPE_CAPSULE && b == DYN_SHAPE_CAPSULE) capsule_capsule(m, first, fc, second, sc, margin);
    }
    if (m.count == 0u) return;
    write_contact(w, index, fs, ss, first, second, fc, sc, m, mirrored_features);
}

// Intermediate results of a convex pair, parked in the pair's raw contact slot (which only its
// final contact overwrites) between k_narrow_convex, k_narrow_epa and k_narrow_finish. A
// penetrating pair: floats 0-11 GJK's four Minkowski points, 12 EPA's validity, 13-15 its
// normal, 16 its depth. A hit within the margin: 0 its distance, 1-3 its point, 4-6 its normal.
__device__ __forceinline__ float* parked(const W& w, uint32_t index) {
    return reinterpret_cast<float*>(w.raw + index);
}

// GJK's loop state parked in the pair's raw slot between k_narrow_convex and
// k_narrow_convex_resume: the simplex (w, a, b of four points), the closest result (v, lambdas,
// the penetration flag, both closest points), the simplex size and the next iteration.
constexpr uint32_t kParkedGjkFloats = 36u + 3u + 4u + 1u + 3u + 3u + 2u;
static_assert(kParkedGjkFloats * 4u <= sizeof(dyn_contact), "GJK's state fits the pair's raw slot");

__device__ inline void park_gjk(float* park, const GjkState& g) {
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        const v3 p[3] = {g.simplex[i].w, g.simplex[i].a, g.simplex[i].b};
#pragma unroll
        for (uint32_t k = 0; k < 3; ++k) {
            park[9 * i + 3 * k] = p[k].x;
            park[9 * i + 3 * k + 1] = p[k].y;
            park[9 * i + 3 * k + 2] = p[k].z;
        }
    }
    const SimplexResult& c = g.closest;
    const float tail[kParkedGjkFloats - 36u] = {c.v.x, c.v.y, c.v.z, c.lambdas[0], c.lambdas[1], c.lambdas[2],
                                                 c.lambdas[3], c.penetrating ? 1.0f : 0.0f, c.point_a.x,
                                                 c.point_a.y, c.point_a.z, c.point_b.x, c.point_b.y, c.point_b.z,
                                                 __uint_as_float(g.count), __uint_as_float(g.iter)};
#pragma unroll
    for (uint32_t k = 0; k < kParkedGjkFloats - 36u; ++k) park[36u + k] = tail[k];
}

__device__ inline void unpark_gjk(const float* park, GjkState* g) {
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        g->simplex[i].w = mk(park[9 * i], park[9 * i + 1], park[9 * i + 2]);
        g->simplex[i].a = mk(park[9 * i + 3], park[9 * i + 4], park[9 * i + 5]);
        g->simplex[i].b = mk(park[9 * i + 6], park[9 * i + 7], park[9 * i + 8]);
    }
    const float* t = park + 36;
    g->closest.v = mk(t[0], t[1], t[2]);
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) g->closest.lambdas[k] = t[3 + k];
    g->closest.penetrating = t[7] != 0.0f;
    g->closest.point_a = mk(t[8], t[9], t[10]);
    g->closest.point_b = mk(t[11], t[12], t[13]);
    g->count = __float_as_uint(t[14]);
    g->iter = __float_as_uint(t[15]);
}

// GJK iterations every convex pair runs in k_narrow_convex; pairs that need more go on in
// k_narrow_convex_resume from their parked state, packed together, so the waves of the first
// kernel do not wait on the few pairs that run to the 24-iteration cap.
constexpr uint32_t kGjkFirstIterations = 8;

struct ConvexPair {
    uint32_t index;
    Shape first, second;
    float margin;
};

__device__ inline ConvexPair load_convex_pair(const W& w, uint32_t index) {
    ConvexPair c;
    c.index = index;
    const uint32_t fs = w.pmaj[index], ss = w.pmin[index];
    const uint32_t fb = w.owner[fs], sb = w.owner[ss];
    const dyn_collider fc = w.co[fs], sc = w.co[ss];
    const dyn_body_state fst = w.st[fb], sst = w.st[sb];
    c.margin = w.p.contact_margin;
    if ((w.de[fb].flags & DYN_BODY_CCD) || (w.de[sb].flags & DYN_BODY_CCD)) {
        c.margin += len3(sub(ld3(sst.velocity), ld3(fst.velocity))) * w.p.dt;  // pair_margin
    }
    c.first = world_collider(fst, fc);
    c.second = world_collider(sst, sc);
    return c;
}

// The rest of convex_hit once GJK has answered: a penetrating pair's simplex is parked for
// k_narrow_epa, a hit within the margin for k_narrow_finish; separated pairs end here.
__device__ inline void convex_pair_after_gjk(const W& w, const ConvexPair& c, const ShapeSet& shapes,
                                             const Closest& closest, const GjkState& g) {
    Hit hit;
    float* park = parked(w, c.index);
    if (!convex_hit_after_gjk<1>(c.first, c.second, shapes, closest, &hit, g.simplex, &g.count)) {
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) {
            park[3 * i] = g.simplex[i].w.x;
            park[3 * i + 1] = g.simplex[i].w.y;
            park[3 * i + 2] = g.simplex[i].w.z;
        }
        w.valid[c.index] = kPairEpa;
        return;
    }
    if (!(hit.distance <= c.margin)) return;
    park[0] = hit.distance;
    park[1] = hit.point.x;
    park[2] = hit.point.y;
    park[3] = hit.point.z;
    park[4] = hit.normal.x;
    park[5] = hit.normal.y;
    park[6] = hit.normal.z;
    w.valid[c.index] = kPairHit;
}

// GJK of the pairs marked kPairGjk (convex_hit, convex.wgsl:584-618), one thread per pair of the
// worklist, for its first kGjkFirstIterations iterations: a pair GJK answers by then goes on to
// the probe and is parked for EPA or the manifold kernel (or ends); a pair that needs more parks
// its GJK state (kPairGjkResume) for k_narrow_convex_resume.
// Three waves a SIMD (168 registers, a few spilled): the loop is vector-ALU bound and the third
// wave hides more of its latency than the spills cost (k_narrow_convex 708 -> 627 us and the
// resume kernel 338 -> 277 us on the settled 100k pile; four waves, 128 registers, were slower).
__global__ void __launch_bounds__(kThreads, 3) k_narrow_convex(W w, uint32_t class_blocks) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0u) {
        uint32_t hull = 0;  // GJK pairs of the hull classes
        for (uint32_t c = kHullClasses; c < kClasses; ++c) hull += w.class_hist[kClasses * class_blocks + c];
        w.work_total[3] = hull;
        w.work_total[5] = 0u;  // EPA polytopes k_narrow_epa_small gives up, counted there
    }
    if (t >= w.work_total[0]) return;
    const uint32_t index = w.worklist[t];
    w.valid[index] = kPairNone;
    const ConvexPair c = load_convex_pair(w, index);
    const ShapeSet shapes{w.sources, w.vertices, w.support_cells, w.support_cands};
    GjkState g;
    gjk_begin<1>(c.first, c.second, shapes, &g);
    Closest closest;
    if (!gjk_run<1>(c.first, c.second, shapes, &g, kGjkFirstIterations, &closest)) {
        park_gjk(parked(w, index), g);
        w.valid[index] = kPairGjkResume;
        return;
    }
    convex_pair_after_gjk(w, c, shapes, closest, g);
}

// The remaining GJK iterations of the pairs k_narrow_convex parked (kPairGjkResume), from their
// parked state, then the rest of convex_hit as in k_narrow_convex.
__global__ void __launch_bounds__(kThreads, 3) k_narrow_convex_resume(W w) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= w.work_total[4]) return;
    const uint32_t index = w.worklist[t];
    w.valid[index] = kPairNone;
    const ConvexPair c = load_convex_pair(w, index);
    const ShapeSet shapes{w.sources, w.vertices, w.support_cells, w.support_cands};
    GjkState g;
    unpark_gjk(parked(w, index), &g);
    Closest closest;
    gjk_run<1>(c.first, c.second, shapes, &g, kGjkIterations, &closest);
    convex_pair_after_gjk(w, c, shapes, closest, g);
}

constexpr uint32_t kEpaWaves = 4;  // waves per workgroup of k_narrow_epa and k_narrow_epa_small

// The vertex capacity k_narrow_epa_small runs with (dyn_set_epa_small_vertices).
uint32_t g_epa_small_vertices = kEpaSmallVertices;

__device__ __forceinline__ void park_epa(float* park, const Epa& r) {
    park[12] = r.valid ? 1.0f : 0.0f;
    park[13] = r.normal.x;
    park[14] = r.normal.y;
    park[15] = r.normal.z;
    park[16] = r.depth;
}

// EPA of the pairs marked kPairEpa, kEpaGroup lanes per pair, the groups striding over the
// worklist. A pair whose polytope outgrows the small capacities keeps its mark and is listed for
// k_narrow_epa (work_total[5] entries of work_minor); the others are parked (lane 0 of the group)
// and marked kPairEpaDone.
__global__ void __launch_bounds__(kEpaWaves * kWave) k_narrow_epa_small(W w, uint32_t vertex_limit) {
    constexpr uint32_t kGroupsPerBlock = kEpaWaves * kWave / kEpaGroup;
    __shared__ EpaSmallShared epa[kGroupsPerBlock];
    const uint32_t slot = threadIdx.x / kEpaGroup;
    const uint32_t q = threadIdx.x % kEpaGroup;
    const uint32_t total = w.work_total[2];
    const uint32_t groups = gridDim.x * kGroupsPerBlock;
    const ShapeSet shapes{w.sources, w.vertices, w.support_cells, w.support_cands};
    for (uint32_t t = blockIdx.x * kGroupsPerBlock + slot; t < total; t += groups) {
        const uint32_t index = w.worklist[t];
        const uint32_t fs = w.pmaj[index], ss = w.pmin[index];
        const Shape wf = world_collider(w.st[w.owner[fs]], w.co[fs]);
        const Shape wsd = world_collider(w.st[w.owner[ss]], w.co[ss]);
        float* park = parked(w, index);
        Epa r;
        if (!epa_small(wf, wsd, shapes, park, vertex_limit, epa[slot], &r)) {
            if (q == 0u) w.work_minor[atomicAdd(&w.work_total[5], 1u)] = index;
            continue;
        }
        if (q == 0u) {
            park_epa(park, r);
            w.valid[index] = kPairEpaDone;
        }
    }
}

// EPA of the pairs k_narrow_epa_small gave up (listed in work_minor), one wave per pair, the
// waves striding over that list. Lane 0 parks the result.
__global__ void __launch_bounds__(kEpaWaves * kWave) k_narrow_epa(W w) {
    __shared__ EpaShared epa[kEpaWaves];
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t total = w.work_total[5];
    const uint32_t waves = gridDim.x * kEpaWaves;
    const ShapeSet shapes{w.sources, w.vertices, w.support_cells, w.support_cands};
    for (uint32_t t = blockIdx.x * kEpaWaves + wave; t < total; t += waves) {
        const uint32_t index = w.work_minor[t];
        const uint32_t fs = w.pmaj[index], ss = w.pmin[index];
        const Shape wf = world_collider(w.st[w.owner[fs]], w.co[fs]);
        const Shape wsd = world_collider(w.st[w.owner[ss]], w.co[ss]);
        float* park = parked(w, index);
        const Epa r = epa_tetrahedron(wf, wsd, shapes, park, epa[wave]);
        if (lane == 0u) {
            park_epa(park, r);
            w.valid[index] = kPairEpaWave;
        }
    }
}

// The manifolds of the convex path (narrowphase.wgsl:551-664): planes against hulls and
// cylinders, GJK hits within the margin, and penetrating pairs after EPA (the rest of
// convex_hit first).
// Four waves a SIMD: its clip polygons live in scratch either way (315 -> 298 us on the pile).
__device__ __forceinline__ void finish_pair(const W& w, uint32_t index, uint32_t mark) {
    w.valid[index] = kPairNone;
    const uint32_t fs = w.pmaj[index], ss = w.pmin[index];
    const uint32_t fb = w.owner[fs], sb = w.owner[ss];
    const dyn_collider fc = w.co[fs], sc = w.co[ss];
    const Body first = load_body(w, fb), second = load_body(w, sb);
    const float margin = pair_margin(w, first, second);
    const ShapeSet shapes{w.sources, w.vertices, w.support_cells, w.support_cands};
    const Shape wf = world_collider(first.s, fc), wsd = world_collider(second.s, sc);
    const float* park = parked(w, index);
    Manifold m;
    m.count = 0;
    m.normal = zero3v();
    bool mirrored_features = false;
    if (mark == kPairPlane) {
        if (fc.kind == DYN_SHAPE_PLANE) {
            plane_convex(m, wf, wsd, shapes, margin);
        } else {
            plane_convex(m, wsd, wf, shapes, margin);
            mirror(m);
            mirrored_features = true;
        }
    } else {
        Hit hit;
        if (mark == kPairHit) {
            hit = make_hit(park[0], mk(park[1], park[2], park[3]), mk(park[4], park[5], park[6]));
        } else {
            SimplexPoint simplex[4];
#pragma unroll
            for (uint32_t i = 0; i < 4; ++i) simplex[i].w = mk(park[3 * i], park[3 * i + 1], park[3 * i + 2]);
            Epa epa;
            epa.valid = park[12] != 0.0f;
            epa.normal = mk(park[13], park[14], park[15]);
            epa.depth = park[16];
            epa.point = zero3v();
            hit = convex_hit_finish(wf, wsd, shapes, simplex, 4u, epa);
            if (!(hit.distance <= margin)) return;
        }
        emit(m, hit.normal);
        if (!convex_pair_manifold(wf, wsd, shapes, hit.normal, m)) manifold_from_hit(m, hit, margin);
    }
    if (m.count == 0u) return;
    write_contact(w, index, fs, ss, first, second, fc, sc, m, mirrored_features);
}

__global__ void __launch_bounds__(kThreads, 4) k_narrow_finish(W w) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= w.work_total[1]) return;
    const uint32_t index = w.worklist[t];
    const uint32_t mark = w.valid[index];
    if (mark != kPairEpaDone && mark != kPairHit && mark != kPairPlane) return;
    finish_pair(w, index, mark);
}

// The manifolds of the polytopes k_narrow_epa expanded (listed in work_minor). Their marks keep
// them off k_narrow_finish's worklist, so the two can run side by side on two streams.
__global__ void __launch_bounds__(kThreads, 4) k_narrow_finish_tail(W w) {
    const uint32_t total = w.work_total[5];
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const uint32_t index = w.work_minor[t];
        if (w.valid[index] == kPairEpaWave) finish_pair(w, index, kPairEpaDone);
    }
}

// ---- stable flag compaction: per-block counts, one-workgroup scan, destination map ---------

constexpr int kScanItems = 4;
constexpr int kScanBlock = kThreads * kScanItems;  // 1024 flags per block

// flags[i] for i < *count (or < limit when count is null); with kCounts the values are counts
// and the scan sums them instead of the set flags.
template <bool kCounts>
__global__ void __launch_bounds__(kThreads) k_scan_count(const uint32_t* __restrict__ flags,
                                                         const uint32_t* count, uint32_t limit,
                                                         uint32_t* block_sums) {
    __shared__ uint32_t s_scan[kThreads / kWave];
    const uint32_t n = count ? min(*count, limit) : limit;
    const uint32_t first = blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const uint32_t i = first + j;
        if (i < n) sum += kCounts ? flags[i] : (flags[i] ? 1u : 0u);
    }
    uint32_t total = 0;
    block_exclusive_sum<kThreads>(sum, s_scan, &total);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(1024) k_scan_blocks(uint32_t* block_sums, uint32_t blocks,
                                                      uint32_t* total_out) {
    __shared__ uint32_t s_scan[1024 / kWave];
    __shared__ uint32_t s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    for (uint32_t first = 0; first < blocks; first += 1024) {
        const uint32_t b = first + threadIdx.x;
        const uint32_t v = b < blocks ? block_sums[b] : 0u;
        uint32_t total = 0;
        const uint32_t excl = block_exclusive_sum<1024>(v, s_scan, &total);
        if (b < blocks) block_sums[b] = s_carry + excl;
        __syncthreads();
        if (threadIdx.x == 0) s_carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total_out = s_carry;
}

// dest[i] = rank of flag i among set flags (kNoSlot when clear), or with kCounts the sum of
// the counts before i; the same partition as k_scan_count, whose block sums k_scan_blocks
// turned into offsets.
template <bool kCounts>
__global__ void __launch_bounds__(kThreads) k_scan_dest(const uint32_t* __restrict__ flags,
                                                        const uint32_t* count, uint32_t limit,
                                                        const uint32_t* __restrict__ block_offsets,
                                                        uint32_t* dest) {
    __shared__ uint32_t s_scan[kThreads / kWave];
    const uint32_t n = count ? min(*count, limit) : limit;
    const uint32_t first = blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
    uint32_t f[kScanItems];
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const uint32_t i = first + j;
        f[j] = i < n ? (kCounts ? flags[i] : (flags[i] ? 1u : 0u)) : 0u;
        sum += f[j];
    }
    uint32_t at = block_offsets[blockIdx.x] + block_exclusive_sum<kThreads>(sum, s_scan, nullptr);
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const uint32_t i = first + j;
        if (i < n) {
            dest[i] = (kCounts || f[j]) ? at : kNoSlot;
            at += f[j];
        }
    }
}

// compact_scatter.wgsl: valid raw contacts in pair order; carried impulses cleared (the raw
// record already has them at zero).
// The workgroup lists its pairs' (raw slot, destination) and copies the records together, its
// lanes on consecutive 16-byte pieces, instead of each lane walking its own 352-byte record.
static_assert(sizeof(dyn_contact) % sizeof(float4) == 0, "contacts copy in 16-byte pieces");
__global__ void __launch_bounds__(kThreads) k_contacts_scatter(W w) {
    constexpr uint32_t kPieces = sizeof(dyn_contact) / sizeof(float4);
    __shared__ uint2 job[kThreads];
    __shared__ uint32_t jobs;
    if (threadIdx.x == 0) jobs = 0u;
    __syncthreads();
    const uint32_t index = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t live = min(*cw(w.counters, DYN_COUNTER_PAIRS), w.pair_cap);
    if (index < live && w.valid[index] == kPairContact) {
        const uint32_t at = w.dest[index];
        if (at < w.contact_cap) job[atomicAdd(&jobs, 1u)] = make_uint2(index, at);
    }
    __syncthreads();
    const uint32_t n = jobs;
    for (uint32_t item = threadIdx.x; item < n * kPieces; item += blockDim.x) {
        const uint32_t k = item / kPieces, piece = item - k * kPieces;
        const uint2 jb = job[k];
        reinterpret_cast<float4*>(w.contacts + jb.y)[piece] = reinterpret_cast<const float4*>(w.raw + jb.x)[piece];
    }
    if (threadIdx.x < n) {
        const uint2 jb = job[threadIdx.x];
        const dyn_contact& c = w.raw[jb.x];
        w.ckeys[jb.y] = make_uint2(c.a, c.b);
        w.matched[jb.y] = 0u;
    }
}

__global__ void k_contacts_total(W w) {
    if (threadIdx.x != 0) return;
    const uint32_t total = w.scan_total[0];
    *cw(w.counters, DYN_COUNTER_CONTACTS) = total;
    const uint32_t kept = min(total, w.contact_cap);
    *cw(w.counters, DYN_COUNTER_REFUSED_CONTACTS) = total - kept;
    *w.ncontacts = kept;
    *cw(w.counters, DYN_COUNTER_BLOCKS) = kept;
}

// ---- warm start and identity (identity.wgsl, contact_relay.wgsl, contact_begin.wgsl) ------

__device__ __forceinline__ bool touches(const dyn_contact& c, float threshold) {
    for (uint32_t i = 0; i < c.point_count; ++i) {
        if (c.points[i].depth >= -threshold) return true;
    }
    return false;
}

__device__ __forceinline__ bool same_pair(const dyn_contact& h, const dyn_contact& c) {
    const bool roles = h.first_body_id == c.first_body_id && h.first_generation == c.first_generation &&
                       h.second_body_id == c.second_body_id && h.second_generation == c.second_generation &&
                       h.a == c.a && h.b == c.b;
    const bool flipped = h.first_body_id == c.second_body_id && h.first_generation == c.second_generation &&
                         h.second_body_id == c.first_body_id && h.second_generation == c.first_generation &&
                         h.a == c.b && h.b == c.a;
    return roles || flipped;
}

__device__ __forceinline__ void carry(dyn_contact& c) {
    float n = 0.0f, t = 0.0f;
    for (uint32_t i = 0; i < c.point_count; ++i) n = n + fmx(c.points[i].accumulated_normal, 0.0f);
    for (uint32_t i = 0; i < c.point_count; ++i) {
        const dyn_manifold_point& p = c.points[i];
        t = t + sqrtf(p.accumulated_tangent_1 * p.accumulated_tangent_1 +
                      p.accumulated_tangent_2 * p.accumulated_tangent_2);
    }
    c.carried_normal = n;
    c.carried_tangent = t;
}

// identity.wgsl:85-117
__device__ inline void relay_impulses(dyn_contact& r, const dyn_contact& held, uint32_t* picks = nullptr) {
    bool taken[4] = {false, false, false, false};
    for (uint32_t pi = 0; pi < r.point_count; ++pi) {
        const dyn_manifold_point& point = r.points[pi];
        uint32_t picked = kNoSlot;
        float picked_distance = 0.0f;
        for (uint32_t hi = 0; hi < held.point_count; ++hi) {
            if (taken[hi]) continue;
            const dyn_manifold_point& hp = held.points[hi];
            if (!(hp.feature == point.feature || hp.feature == feature_mirror(point.feature))) continue;
            const float d = len3(sub(ld3(hp.local_a), ld3(point.local_a))) +
                            len3(sub(ld3(hp.local_b), ld3(point.local_b)));
            if (picked == kNoSlot || d < picked_distance) {
                picked = hi;
                picked_distance = d;
            }
        }
        if (picked == kNoSlot) continue;
        taken[picked] = true;
        if (picks) picks[pi] = picked;
        r.points[pi].accumulated_normal = held.points[picked].accumulated_normal;
        r.points[pi].accumulated_tangent_1 = held.points[picked].accumulated_tangent_1;
        r.points[pi].accumulated_tangent_2 = held.points[picked].accumulated_tangent_2;
    }
}

// An archived contact's points, read once into registers (all four slots; those past the count
// are never used).
struct HeldPoints {
    uint32_t count;
    uint32_t feature[4];
    v3 la[4], lb[4];
    float n[4], t1[4], t2[4];
};

__device__ __forceinline__ HeldPoints held_points(const dyn_contact& held) {
    HeldPoints h;
    h.count = held.point_count;
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        const dyn_manifold_point& p = held.points[i];
        h.feature[i] = p.feature;
        h.la[i] = ld3(p.local_a);
        h.lb[i] = ld3(p.local_b);
        h.n[i] = p.accumulated_normal;
        h.t1[i] = p.accumulated_tangent_1;
        h.t2[i] = p.accumulated_tangent_2;
    }
    return h;
}

// relay_impulses from the registers above: the same matches, in the same order, with the same
// distances.
__device__ inline void relay_held(dyn_contact& r, const HeldPoints& h, uint32_t* picks) {
    uint32_t taken = 0;
    for (uint32_t pi = 0; pi < r.point_count; ++pi) {
        const dyn_manifold_point& point = r.points[pi];
        const uint32_t f = point.feature, fm = feature_mirror(point.feature);
        const v3 pa = ld3(point.local_a), pb = ld3(point.local_b);
        uint32_t picked = kNoSlot;
        float picked_distance = 0.0f;
#pragma unroll
        for (uint32_t hi = 0; hi < 4; ++hi) {
            if (hi >= h.count || ((taken >> hi) & 1u)) continue;
            if (!(h.feature[hi] == f || h.feature[hi] == fm)) continue;
            const float d = len3(sub(h.la[hi], pa)) + len3(sub(h.lb[hi], pb));
            if (picked == kNoSlot || d < picked_distance) {
                picked = hi;
                picked_distance = d;
            }
        }
        if (picked == kNoSlot) continue;
        taken |= 1u << picked;
        picks[pi] = picked;
        float n = 0.0f, t1 = 0.0f, t2 = 0.0f;
#pragma unroll
        for (uint32_t hi = 0; hi < 4; ++hi) {
            if (hi == picked) {
                n = h.n[hi];
                t1 = h.t1[hi];
                t2 = h.t2[hi];
            }
        }
        r.points[pi].accumulated_normal = n;
        r.points[pi].accumulated_tangent_1 = t1;
        r.points[pi].accumulated_tangent_2 = t2;
    }
}

// body_row.wgsl:1-13
__device__ __forceinline__ uint32_t resolve_row(const W& w, uint32_t id, uint32_t gen) {
    if (id >= w.id_cap) return DYN_NO_BODY;
    const uint32_t row = w.row_of_body[id];
    if (row >= w.n) return DYN_NO_BODY;
    const dyn_body_state& s = w.st[row];
    if (s.body_id != id || s.generation != gen) return DYN_NO_BODY;
    return row;
}

// The first index of the (a, b)-sorted keys[0, count) whose key is not below (a, b): bracketed by
// steps of 1, 2, 4, ... out from `hint`, then bisected, which returns the same index as a bisection
// of the whole range in a few probes when the key sits near the hint. The archive and the current
// contacts are both in canonical pair order and mostly hold the same pairs, so an archived
// contact's match sits near its own scaled position.
__device__ __forceinline__ uint32_t lower_bound_near(const uint2* __restrict__ keys, uint32_t count, uint32_t hint,
                                                     uint32_t a, uint32_t b) {
    if (count == 0u) return 0u;
    const auto below = [&](uint32_t i) {
        const uint2 m = keys[i];
        return m.x < a || (m.x == a && m.y < b);
    };
    if (hint >= count) hint = count - 1u;
    uint32_t lo, hi;  // the answer lies in [lo, hi]
    if (below(hint)) {
        lo = hint + 1u;
        hi = count;
        for (uint32_t step = 1u;; step <<= 1) {
            if (step >= count - hint) break;  // hint + step would pass the end
            const uint32_t probe = hint + step;
            if (!below(probe)) {
                hi = probe;
                break;
            }
            lo = probe + 1u;
        }
    } else {
        lo = 0u;
        hi = hint;
        for (uint32_t step = 1u;; step <<= 1) {
            if (step > hint) break;
            const uint32_t probe = hint - step;
            if (below(probe)) {
                lo = probe + 1u;
                break;
            }
            hi = probe;
        }
    }
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2u;
        if (below(mid)) lo = mid + 1u; else hi = mid;
    }
    return lo;
}

// contact_relay.wgsl:34-71 (events are not modelled: no collider here requests them)
__global__ void __launch_bounds__(kThreads) k_relay(W w) {
    const uint32_t index = blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= *w.narchive) return;
    // records are read in place and only what the relay changes is written back: the points'
    // impulses, the carried totals and the event bits (not the whole 352-byte record)
    const dyn_contact& held = w.archive[index];
    const HeldPoints points = held_points(held);  // in flight through the checks and the search
    if (resolve_row(w, held.first_body_id, held.first_generation) == DYN_NO_BODY ||
        resolve_row(w, held.second_body_id, held.second_generation) == DYN_NO_BODY) return;
    const uint32_t count = *w.ncontacts;
    // lower bound of (a, b) over the dense keys (8 bytes a contact, not its 352-byte record),
    // galloping out from where the archived contact would sit in the current list
    const uint32_t hint = static_cast<uint32_t>(static_cast<uint64_t>(index) * count / max(*w.narchive, 1u));
    const uint32_t lo = lower_bound_near(w.ckeys, count, hint, held.a, held.b);
    if (lo >= count) return;
    const uint2 at = w.ckeys[lo];
    if (at.x != held.a || at.y != held.b) return;
    dyn_contact& cur = w.contacts[lo];
    if (!same_pair(held, cur)) return;
    if (!(held.events & kContactAnnounced)) return;
    w.matched[lo] = 1u;
    uint32_t picks[4] = {kNoSlot, kNoSlot, kNoSlot, kNoSlot};
    relay_held(cur, points, picks);
    if (w.relay_src) {
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) w.relay_src[4u * lo + q] = picks[q] == kNoSlot ? kNoSlot : 4u * index + picks[q];
    }
    carry(cur);
    cur.events |= kContactAnnounced;
}

// contact_begin.wgsl:41-71 over the resting archive sorted by (min id, max id)
__global__ void __launch_bounds__(kThreads) k_contact_begin(W w, float slop) {
    const uint32_t index = blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= *w.ncontacts || w.matched[index]) return;
    // changed in place: the fields below are the only ones a copy written back would change
    dyn_contact& c = w.contacts[index];
    const uint32_t kmaj = min(c.first_body_id, c.second_body_id);
    const uint32_t kmin = max(c.first_body_id, c.second_body_id);
    const uint32_t count = *w.nresting;
    uint32_t lo = 0, hi = count;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2u;
        const uint32_t ma = w.rest_major[mid];
        if (ma < kmaj || (ma == kmaj && w.rest_minor[mid] < kmin)) lo = mid + 1u; else hi = mid;
    }
    uint32_t slot = kNoSlot;
    for (; lo < count && w.rest_major[lo] == kmaj && w.rest_minor[lo] == kmin; ++lo) {
        if (same_pair(w.resting[lo], c)) {
            slot = lo;
            break;
        }
    }
    if (slot != kNoSlot) {
        const dyn_contact& rest = w.resting[slot];
        if (rest.events & kContactAnnounced) {
            c.events |= kContactAnnounced;
            relay_impulses(c, rest);
            carry(c);
        } else if (touches(c, slop)) {
            c.events |= kContactAnnounced;
        }
        return;
    }
    if (touches(c, slop)) c.events |= kContactAnnounced;
}

// ---- the differentiable rollout's tape ---------------------------------------------------------
//
// While a tape is recorded, each step writes the states it starts from and its contacts as the
// solver receives them (after the warm-start relay) into the layouts of dynamis_amd/csrc/adjoint.h:
// per row 16 floats (position, orientation, velocity, spin), per contact 76 floats (rows, flags,
// colliders, normal, combined friction and its partials, restitution; per point the anchors,
// depth, world position, relay source and warm-start impulses).

__global__ void __launch_bounds__(kThreads) k_tape_sources(W w) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 4u * w.contact_cap) w.relay_src[i] = kNoSlot;
}

__device__ __forceinline__ void friction_partials(float a, float b, uint32_t mode, float* da, float* db) {
    if (mode == 1u) {  // min
        *da = a < b ? 1.0f : 0.0f;
        *db = a < b ? 0.0f : 1.0f;
    } else if (mode == 2u) {  // max
        *da = a > b ? 1.0f : 0.0f;
        *db = a > b ? 0.0f : 1.0f;
    } else if (mode == 3u) {  // average
        *da = *db = 0.5f;
    } else {  // geometric mean
        const float c = sqrtf(a * b);
        *da = c > 0.0f ? b / (2.0f * c) : 0.0f;
        *db = c > 0.0f ? a / (2.0f * c) : 0.0f;
    }
}

// error bits: 1 a resolving contact is soft or has rolling or spin friction, 2 a body sleeps
__global__ void __launch_bounds__(kThreads) k_tape_record(W w, float* states, float* records, uint32_t* count,
                                                          uint32_t* error) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *count = *w.ncontacts;
    if (i < w.n) {
        const dyn_body_state& s = w.st[i];
        float4* o = reinterpret_cast<float4*>(states + 16u * static_cast<size_t>(i));
        o[0] = make_float4(s.position[0], s.position[1], s.position[2], 0.0f);
        o[1] = make_float4(s.orientation[0], s.orientation[1], s.orientation[2], s.orientation[3]);
        o[2] = make_float4(s.velocity[0], s.velocity[1], s.velocity[2], 0.0f);
        o[3] = make_float4(s.angular_velocity[0], s.angular_velocity[1], s.angular_velocity[2], 0.0f);
        if (s.sleeping != 0u && i < w.d) atomicOr(error, 2u);
    }
    if (i >= *w.ncontacts) return;
    const dyn_contact& c = w.contacts[i];
    const uint32_t fb = w.owner[c.a], sb = w.owner[c.b];
    const bool fi = body_is_inert(load_body(w, fb)), si = body_is_inert(load_body(w, sb));
    const bool resolves = c.point_count > 0u && c.sensor == 0u;  // block_resolves
    if (resolves && (c.relaxation > 0.0f || c.rolling_friction != 0.0f || c.spin_friction != 0.0f)) atomicOr(error, 1u);
    const uint32_t flags = c.point_count | (fi ? 8u : 0u) | (si ? 16u : 0u) | (resolves ? 32u : 0u);
    float da, db;
    friction_partials(w.co[c.a].friction, w.co[c.b].friction, w.p.friction_combine, &da, &db);
    float4* o = reinterpret_cast<float4*>(records + 76u * static_cast<size_t>(i));
    o[0] = make_float4(__uint_as_float(fb), __uint_as_float(sb), __uint_as_float(flags), __uint_as_float(c.a));
    o[1] = make_float4(c.normal[0], c.normal[1], c.normal[2], c.friction);
    o[2] = make_float4(da, db, c.restitution, __uint_as_float(c.b));
    for (uint32_t p = 0; p < 4; ++p) {
        const dyn_manifold_point& m = c.points[p];
        const bool has = p < c.point_count;
        const uint32_t src = has ? w.relay_src[4u * i + p] : kNoSlot;
        float4* q = o + 3u + 4u * p;
        q[0] = has ? make_float4(m.local_a[0], m.local_a[1], m.local_a[2], m.depth) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        q[1] = has ? make_float4(m.local_b[0], m.local_b[1], m.local_b[2], 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        q[2] = make_float4(has ? m.position[0] : 0.0f, has ? m.position[1] : 0.0f, has ? m.position[2] : 0.0f,
                           __uint_as_float(src));
        q[3] = has ? make_float4(m.accumulated_normal, m.accumulated_tangent_1, m.accumulated_tangent_2, 0.0f)
                   : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// ---- islands (island_*.wgsl) ------------------------------------------------------------------

__global__ void __launch_bounds__(kThreads) k_island_init(W w) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kFlattenPasses * kFlattenFlags) w.flat_flags[i] = 0u;  // launched with at least kThreads lanes
    if (i >= w.d) return;
    w.parents[i] = i;
    w.istate[i] = 0u;
}

// The root of a node by a read-only walk. Halving the path on the way (pointing each visited
// node at its grandparent) was measured: on the settled 100k pile its stores cost more than the
// hops they save, and the flatten rounds compress the trees after the hooks anyway. Parents only
// ever move to an ancestor and a root only to a lower root, so the components and their
// minimum-row roots are the same either way.
__device__ __forceinline__ uint32_t island_find(const W& w, uint32_t node) {
    uint32_t cur = node;
    uint32_t parent = atomic_load_agent(&w.parents[cur]);
    while (parent != cur) {
        cur = parent;
        parent = atomic_load_agent(&w.parents[cur]);
    }
    return cur;
}

// island_link.wgsl:13-45: the higher root is hooked under the lower one. The hook is an
// atomic minimum on the higher root's parent word, which never fails: when that node is no
// longer a root, its parent has become min(its parent, lower) and the node it pointed to is
// linked with `lower` next, so no connection is lost. Parents only ever point to lower rows,
// so every component ends as one tree under its minimum row, as with the reference's
// compare-and-swap retries (which count a fault after 256 failures; this loop counts one if
// 256 hooks do not finish it).
__device__ inline void island_couple(const W& w, uint32_t a, uint32_t b) {
    if (b >= w.d && atomic_load_agent(&w.wake[b]) != 0u) atomicOr(&w.wake[a], 1u);
    if (a >= w.d && atomic_load_agent(&w.wake[a]) != 0u) atomicOr(&w.wake[b], 1u);
    if (a >= w.d || b >= w.d || a == b) return;
    // rows with one parent are joined already (most pairs on a seeded pile)
    if (atomic_load_agent(&w.parents[a]) == atomic_load_agent(&w.parents[b])) return;
    uint32_t ra = island_find(w, a), rb = island_find(w, b);
    for (uint32_t attempt = 0; attempt < 256u; ++attempt) {
        if (ra == rb) return;
        const uint32_t lower = min(ra, rb), higher = max(ra, rb);
        const uint32_t was = atomicMin(&w.parents[higher], lower);
        if (was == higher) return;  // it was a root
        ra = island_find(w, lower);
        rb = island_find(w, was);
    }
    atomicAdd(cw(w.counters, DYN_COUNTER_ISLAND_FAULTS), 1u);
}

// The rows a contact links in the island pass (island_link.wgsl:47-60): a touching, non-sensor
// contact between colliders and bodies inside the records; kNoSlot twice otherwise.
__device__ __forceinline__ uint2 island_edge(const W& w, uint32_t i, float slop) {
    const uint2 none = make_uint2(kNoSlot, kNoSlot);
    const dyn_contact& c = w.contacts[i];
    if (c.point_count == 0u || c.sensor == 1u || !touches(c, slop)) return none;
    // colliders or bodies outside the records (possible only for contacts handed to dyn_islands)
    // link nothing
    if (c.a >= w.k || c.b >= w.k) return none;
    const uint32_t fb = w.owner[c.a], sb = w.owner[c.b];
    if (fb >= w.n || sb >= w.n) return none;
    return make_uint2(fb, sb);
}

__device__ __forceinline__ uint2 island_edge_at(const W& w, uint32_t i, float slop) {
    return w.iedges ? w.iedges[i] : island_edge(w, i, slop);
}

// Completes the components over every linking contact with the retrying hooks above; after the
// seeding and hooking rounds most contacts' rows share a parent and return at once.
__global__ void __launch_bounds__(kThreads) k_island_link(W w, float slop) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= min(*w.ncontacts, w.contact_cap)) return;
    const uint2 e = island_edge_at(w, i, slop);
    if (e.x == kNoSlot) return;
    island_couple(w, e.x, e.y);
}

// A hooking round over the linking contacts between simulated rows, run on a flattened forest
// (every parent a root): a contact whose rows' parents differ points the higher parent at the
// lower one with one atomic minimum and no retries. A hook landing on a row that is no longer a
// root moves that row to a lower row of its component; a connection it drops is restored by
// k_island_link, which visits every contact. Parents still point only at lower rows of the same
// component.
// For the first kMergedKeys distinct `key`s of the wave's lanes with `take`, one atomic minimum of
// the smallest of their `value`s on words[key]; the lanes with other keys send their own. The
// words end as with one atomic minimum per lane. Every lane of the wave calls it.
constexpr uint32_t kMergedKeys = 8;
__device__ inline void wave_min_per_key(uint32_t* words, bool take, uint32_t key, uint32_t value) {
    uint64_t pending = __ballot(take);
    for (uint32_t merged = 0; pending != 0u && merged < kMergedKeys; ++merged) {
        const uint32_t leader = static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(pending))) - 1u;
        const uint32_t target = __shfl(key, leader, kWave);
        const bool mine = take && key == target;
        uint32_t least = mine ? value : 0xFFFFFFFFu;
#pragma unroll
        for (int offset = kWave / 2; offset > 0; offset >>= 1) least = min(least, __shfl_xor(least, offset, kWave));
        if (lane_id() == leader) atomicMin(&words[target], least);
        pending &= ~__ballot(mine);
    }
    if ((pending >> lane_id()) & 1u) atomicMin(&words[key], value);
}

__global__ void __launch_bounds__(kThreads) k_island_hook(W w, float slop) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool take = false;
    uint32_t hi = 0, lo = 0;
    if (i < min(*w.ncontacts, w.contact_cap)) {
        const uint2 e = island_edge_at(w, i, slop);
        if (e.x < w.d && e.y < w.d) {
            const uint32_t ra = atomic_load_agent(&w.parents[e.x]), rb = atomic_load_agent(&w.parents[e.y]);
            take = ra != rb;
            hi = max(ra, rb);
            lo = min(ra, rb);
        }
    }
    // a contact's neighbours in the list mostly join the same rows, so the wave's hooks on one
    // root go out as one atomic
    wave_min_per_key(w.parents, take, hi, lo);
}

// dyn_solve's set-up: the inert rows' contribution marks at zero (a world keeps them zero
// between position passes: zeroed at allocation, cleared by k_position_apply), the solver's
// counters and the check word at zero, the contact count clamped to the capacity (into the
// stage's own count word, words[0]).
__global__ void __launch_bounds__(kThreads) k_solve_stage_reset(W w, const uint32_t* count, uint32_t* words) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < w.n) w.contrib[i] = 0u;
    if (i != 0u) return;
    *cw(w.counters, DYN_COUNTER_LIVE) = 0u;
    *cw(w.counters, DYN_COUNTER_SOLVER_ROWS) = 0u;
    *cw(w.counters, DYN_COUNTER_SOLVE_LINEAR_RESIDUAL) = 0u;
    *cw(w.counters, DYN_COUNTER_SOLVE_ANGULAR_RESIDUAL) = 0u;
    words[0] = min(*count, w.contact_cap);
    words[1] = 0u;
}

// dyn_solve's check: words[1] set when a contact names a collider or a body outside the records.
__global__ void __launch_bounds__(kThreads) k_solve_stage_check(W w, uint32_t* words) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= words[0]) return;
    const dyn_contact& c = w.contacts[i];
    if (c.a >= w.k || c.b >= w.k || w.owner[c.a] >= w.n || w.owner[c.b] >= w.n) atomicOr(&words[1], 1u);
}

// The island pass's first kernel: each contact's linking rows go to iedges (when present), and
// every linking contact between two simulated rows points the higher row's parent at the lower
// row (an atomic minimum, so each row ends at its lowest linked neighbour below it, or itself).
// Parents still point only at lower rows of the same component, so the hooks and the flatten
// rounds end at the same minimum-row roots; most rows then start a few hops from their root.
__global__ void __launch_bounds__(kThreads) k_island_seed(W w, float slop) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool take = false;
    uint32_t hi = 0, lo = 0;
    if (i < min(*w.ncontacts, w.contact_cap)) {
        const uint2 e = island_edge(w, i, slop);
        if (w.iedges) w.iedges[i] = e;
        take = e.x < w.d && e.y < w.d && e.x != e.y;
        hi = max(e.x, e.y);
        lo = min(e.x, e.y);
    }
    // a row's contacts sit together in the list but mostly as its lower side: one atomic a lane
    if (take) atomicMin(&w.parents[hi], lo);
}

// dyn_islands' set-up: no wake or motion flags, the island counters at zero.
__global__ void __launch_bounds__(kThreads) k_islands_stage_reset(W w) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < w.n) w.wake[i] = 0u;
    if (i < w.d) w.motion[i] = 0u;
    if (i == 0u) {
        *cw(w.counters, DYN_COUNTER_ISLANDS) = 0u;
        *cw(w.counters, DYN_COUNTER_ISLAND_FAULTS) = 0u;
    }
}

__global__ void __launch_bounds__(kThreads) k_island_link_resting(W w) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *w.nresting) return;
    const dyn_contact& c = w.resting[i];
    const uint32_t r1 = resolve_row(w, c.first_body_id, c.first_generation);
    const uint32_t r2 = resolve_row(w, c.second_body_id, c.second_generation);
    if (r1 == DYN_NO_BODY || r2 == DYN_NO_BODY) return;
    island_couple(w, r1, r2);
}

// island_flatten.wgsl, with walks of up to kFlattenHops hops in place of one: round `round` of a
// flatten pass whose per-round flags are `flags`. Each row walks up from its parent to its root
// or kFlattenHops hops further and points there, so a round divides every depth by at least
// kFlattenHops + 1 (rows moved up by other rows in the same round only shorten the walks), and
// flatten_rounds(d) rounds leave every row pointing at its root. No hooks run during a pass, so
// a round in which every walk reached its root leaves the forest flat: the rounds after it
// return at once.
__global__ void __launch_bounds__(kThreads) k_island_flatten(W w, uint32_t* flags, uint32_t round) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0u) flags[round + 1u] = 0u;  // read by round + 2; no round writes it before round + 1
    if (round > 0u && flags[round - 1u] == 0u) return;
    if (i >= w.d) return;
    const uint32_t parent = atomic_load_agent(&w.parents[i]);
    uint32_t at = parent;
    uint32_t up = atomic_load_agent(&w.parents[at]);
    for (uint32_t hop = 0; up != at && hop < kFlattenHops; ++hop) {
        at = up;
        up = atomic_load_agent(&w.parents[at]);
    }
    if (at != parent) atomic_store_agent(&w.parents[i], at);
    if (up != at) flags[round] = 1u;  // stopped short of the root
}

// All the flatten rounds of a small world in one workgroup, a barrier after each round (within a
// round, as across the launched rounds, a row may read a parent another row has already moved
// up, which only shortens its walk); it stops after the first round whose walks all reached
// their roots.
constexpr uint32_t kFlattenGroupRows = 8192;
__global__ void __launch_bounds__(1024) k_island_flatten_rounds(W w, uint32_t rounds) {
    for (uint32_t r = 0; r < rounds; ++r) {
        int short_of_root = 0;
        for (uint32_t i = threadIdx.x; i < w.d; i += 1024u) {
            const uint32_t parent = atomic_load_agent(&w.parents[i]);
            uint32_t at = parent;
            uint32_t up = atomic_load_agent(&w.parents[at]);
            for (uint32_t hop = 0; up != at && hop < kFlattenHops; ++hop) {
                at = up;
                up = atomic_load_agent(&w.parents[at]);
            }
            if (at != parent) atomic_store_agent(&w.parents[i], at);
            if (up != at) short_of_root = 1;
        }
        if (!__syncthreads_or(short_of_root)) break;
    }
}

// v summed over the workgroup (kThreads threads, every one of them calling), added to *counter
// by thread 0 with one atomic when it is not zero.
__device__ inline void block_count(uint32_t* counter, uint32_t v) {
    __shared__ uint32_t partial[kThreads / kWave];
    v = wave_sum(v);
    if (lane_id() == 0) partial[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
#pragma unroll
        for (uint32_t k = 0; k < kThreads / kWave; ++k) total += partial[k];
        if (total) atomicAdd(counter, total);
    }
    __syncthreads();  // the array is reused by the next call
}

// island_aggregate.wgsl + island_count.wgsl
__global__ void __launch_bounds__(kThreads) k_island_aggregate(W w) {
    uint32_t roots = 0;
    for (uint32_t base = blockIdx.x * blockDim.x; base < w.d; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        uint32_t root = kNoSlot;
        bool raise = false;
        if (i < w.d) {
            root = w.parents[i];
            raise = w.wake[i] != 0u || w.motion[i] != 0u;
            roots += root == i ? 1u : 0u;
        }
        // One island holds many neighbouring rows: one lane per distinct root of the wave raises
        // its flag, and not at all once the flag is there.
        // (a batch's small worlds put many roots in one wave: past kMergedKeys each lane raises
        // its own)
        uint64_t pending = __ballot(raise);
        for (uint32_t merged = 0; pending != 0u && merged < kMergedKeys; ++merged) {
            const uint32_t leader = static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(pending))) - 1u;
            const uint32_t target = __shfl(root, leader, kWave);
            const uint64_t same = __ballot(raise && root == target);
            if (lane_id() == leader && !(atomic_load_agent(&w.istate[target]) & kIslandWake)) {
                atomicMax(&w.istate[target], kIslandWake);
            }
            pending &= ~same;
        }
        if (((pending >> lane_id()) & 1u) && !(atomic_load_agent(&w.istate[root]) & kIslandWake)) {
            atomicMax(&w.istate[root], kIslandWake);
        }
    }
    block_count(cw(w.counters, DYN_COUNTER_ISLANDS), roots);
}

// island_wake.wgsl
__global__ void __launch_bounds__(kThreads) k_island_wake(W w) {
    uint32_t woke = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < w.d; i += gridDim.x * blockDim.x) {
        const uint32_t root = w.parents[i];
        if (w.st[i].sleeping != 0u && (w.istate[root] & kIslandWake)) {
            ++woke;
            w.st[i].sleeping = 0u;
            w.st[i].sleep_timer = 0.0f;
        }
        w.wake[i] = 0u;
    }
    block_count(cw(w.counters, DYN_COUNTER_WOKE), woke);
    block_count(cw(w.counters, DYN_COUNTER_WOKE_DEFERRED), woke);
}

// live.wgsl count, solver_reset.wgsl
__global__ void __launch_bounds__(kThreads) k_solver_reset(W w) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *cw(w.counters, DYN_COUNTER_SOLVER_ROWS) = 0u;  // maintained: k_rows_count refills it
    }
    uint32_t live = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < w.n; i += gridDim.x * blockDim.x) {
        w.block_counts[i] = 0u;
        reinterpret_cast<float4*>(w.resolution)[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const Body b = load_body(w, i);
        if (i < w.d && body_is_active(b.s, b.d)) ++live;
        // what k_blocks reads of a block's body, in one 48-byte record
        const v3 com = body_com(b);
        float4* r = w.bbodies + 3u * i;
        r[0] = make_float4(com.x, com.y, com.z, b.d.inverse_mass);
        r[1] = make_float4(b.s.velocity[0], b.s.velocity[1], b.s.velocity[2], body_is_inert(b) ? 1.0f : 0.0f);
        r[2] = make_float4(b.s.angular_velocity[0], b.s.angular_velocity[1], b.s.angular_velocity[2], 0.0f);
    }
    block_count(cw(w.counters, DYN_COUNTER_LIVE), live);
}

// ---- solver ------------------------------------------------------------------------------------

__device__ __forceinline__ bool block_resolves(const dyn_contact& c) {
    return c.point_count > 0u && c.sensor == 0u;
}

// solver_blocks.wgsl:24-48
__global__ void __launch_bounds__(kThreads) k_blocks(W w) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *w.ncontacts) return;
    const dyn_contact& c = w.contacts[i];
    const uint32_t fb = w.owner[c.a], sb = w.owner[c.b];
    w.blocks[2 * i] = fb;
    w.blocks[2 * i + 1] = sb;
    float4* prow = w.prows + 3u * i;
    if (!block_resolves(c)) {
        prow[0] = make_float4(__uint_as_float(fb), __uint_as_float(sb), __uint_as_float(0u), 0.0f);
        w.slots[2 * i] = w.slots[2 * i + 1] = kNoSlot;
        return;
    }
    // the bodies as k_solver_reset recorded them: centre of mass, inverse mass, velocities, inert
    const float4* rf = w.bbodies + 3u * fb;
    const float4* rs = w.bbodies + 3u * sb;
    const float4 f0 = rf[0], f1 = rf[1], f2 = rf[2], s0 = rs[0], s1 = rs[1], s2 = rs[2];
    const bool fi = f1.w != 0.0f, si = s1.w != 0.0f;
    const v3 normal = ld3(c.normal);
    {
        // position_correction.wgsl:5-46 reads only what stays fixed through the substeps
        const uint32_t flags = c.relaxation > 0.0f ? 0u : (c.point_count | (fi ? 8u : 0u) | (si ? 16u : 0u));
        prow[0] = make_float4(__uint_as_float(fb), __uint_as_float(sb), __uint_as_float(flags), fi ? 0.0f : f0.w);
        prow[1] = make_float4(si ? 0.0f : s0.w, normal.x, normal.y, normal.z);
        prow[2] = make_float4(c.points[0].depth, c.point_count > 1u ? c.points[1].depth : 0.0f,
                              c.point_count > 2u ? c.points[2].depth : 0.0f,
                              c.point_count > 3u ? c.points[3].depth : 0.0f);
    }
    const v3 com_f = mk(f0.x, f0.y, f0.z), v_f = mk(f1.x, f1.y, f1.z), w_f = mk(f2.x, f2.y, f2.z);
    const v3 com_s = mk(s0.x, s0.y, s0.z), v_s = mk(s1.x, s1.y, s1.z), w_s = mk(s2.x, s2.y, s2.z);
    for (uint32_t pi = 0; pi < c.point_count; ++pi) {
        const v3 pos = ld3(c.points[pi].position);
        // relative_velocity: each body's point velocity v + w x (p - centre of mass)
        const v3 rel = sub(add(v_s, cross3(w_s, sub(pos, com_s))), add(v_f, cross3(w_f, sub(pos, com_f))));
        const float speed = dot3(rel, normal);
        float target = 0.0f;
        if (c.points[pi].depth < 0.0f) target = c.points[pi].depth / w.p.dt;
        else if (speed < -w.p.restitution_threshold) target = -speed * c.restitution;
        w.targets[i * 4u + pi] = target;
    }
    // a non-inert row's count is the length of its list; the count before this block's
    // increment is the block's place on it
    uint32_t slot_f = kNoSlot, slot_s = kNoSlot;
    if (fi) w.block_counts[fb] = 1u; else slot_f = atomicAdd(&w.block_counts[fb], 1u);
    if (si) w.block_counts[sb] = 1u; else slot_s = atomicAdd(&w.block_counts[sb], 1u);
    w.slots[2 * i] = slot_f;
    w.slots[2 * i + 1] = slot_s;
}

__global__ void __launch_bounds__(kThreads) k_rows_count(W w) {
    uint32_t claimed = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < w.n; i += gridDim.x * blockDim.x) {
        const uint32_t blocks = w.block_counts[i];
        const bool listed = blocks != 0u && !(!kind_simulates(w.de[i]) || w.st[i].sleeping != 0u);
        w.list_counts[i] = listed ? blocks : 0u;
        claimed += blocks != 0u ? 1u : 0u;
    }
    block_count(cw(w.counters, DYN_COUNTER_SOLVER_ROWS), claimed);
}

// The list entry of every block's non-inert sides: the row's first entry plus the block's place.
__global__ void __launch_bounds__(kThreads) k_fill_lists(W w) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *w.ncontacts) return;
    uint2& slot = reinterpret_cast<uint2*>(w.slots)[i];
    const uint2 place = slot;
    const uint2 entry = make_uint2(place.x != kNoSlot ? w.row_start[w.blocks[2 * i]] + place.x : kNoSlot,
                                   place.y != kNoSlot ? w.row_start[w.blocks[2 * i + 1]] + place.y : kNoSlot);
    slot = entry;
    if (entry.x != kNoSlot) w.adj[entry.x] = i << 1;
    if (entry.y != kNoSlot) w.adj[entry.y] = (i << 1) | 1u;
}

// One side's six velocity words of block ci.
__device__ __forceinline__ void store_words(const W& w, uint32_t ci, uint32_t side, v3 lin, v3 ang) {
    uint2* d = reinterpret_cast<uint2*>(w.vwords + ci * 12u + side * 6u);
    const float scale = DYN_SOLVER_VELOCITY_SCALE;
    d[0] = make_uint2(solver_word(lin.x, scale), solver_word(lin.y, scale));
    d[1] = make_uint2(solver_word(lin.z, scale), solver_word(ang.x, scale));
    d[2] = make_uint2(solver_word(ang.y, scale), solver_word(ang.z, scale));
}

__device__ __forceinline__ void make_tangents(v3 n, v3* t1, v3* t2) {
    const v3 t = fabsf(n.x) > 0.9f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
    *t1 = norm3(cross3(n, t));
    *t2 = cross3(n, *t1);
}

__device__ __forceinline__ float momentum_mass(const Body& a, const Body& b, v3 pa, v3 pb, v3 axis) {
    const v3 rax = cross3(sub(pa, body_com(a)), axis), rbx = cross3(sub(pb, body_com(b)), axis);
    return a.d.inverse_mass + b.d.inverse_mass + dot3(rax, apply_inverse_inertia(a, rax)) +
           dot3(rbx, apply_inverse_inertia(b, rbx));
}

__device__ __forceinline__ void apply_pair_impulse(Body& a, Body& b, v3 pa, v3 pb, v3 impulse) {
    st3(a.s.velocity, sub(ld3(a.s.velocity), mul(impulse, a.d.inverse_mass)));
    st3(a.s.angular_velocity,
        sub(ld3(a.s.angular_velocity), apply_inverse_inertia(a, cross3(sub(pa, body_com(a)), impulse))));
    st3(b.s.velocity, add(ld3(b.s.velocity), mul(impulse, b.d.inverse_mass)));
    st3(b.s.angular_velocity,
        add(ld3(b.s.angular_velocity), apply_inverse_inertia(b, cross3(sub(pb, body_com(b)), impulse))));
}

__device__ __forceinline__ v3 arm_at(const Body& b, v3 local) {
    return add(qrot(ld4(b.s.orientation), local), ld3(b.s.position));
}

__device__ __forceinline__ v3 arm(const Body& b, const dyn_manifold_point& p, bool first) {
    return arm_at(b, ld3(first ? p.local_a : p.local_b));
}

// solver_block_solve.wgsl:56-83: inert bodies receive exactly zero words (skipped).
__device__ __forceinline__ void commit_block(const W& w, uint32_t ci, bool fa_inert, bool sb_inert, v3 da,
                                             v3 sa, v3 db, v3 sbv, bool residual) {
    if (!fa_inert) store_words(w, ci, 0u, da, sa);
    if (!sb_inert) store_words(w, ci, 1u, db, sbv);
    if (!residual) return;
    atomicMax(cw(w.counters, DYN_COUNTER_SOLVE_LINEAR_RESIDUAL),
              solver_word(fmx(len3(da), len3(db)), DYN_SOLVER_VELOCITY_SCALE));
    atomicMax(cw(w.counters, DYN_COUNTER_SOLVE_ANGULAR_RESIDUAL),
              solver_word(fmx(len3(sa), len3(sbv)), DYN_SOLVER_VELOCITY_SCALE));
}

__device__ __forceinline__ float normal_delta(float relaxation, float damping_ratio, float mass, float truth,
                                              float step, float goal, float depth, float speed,
                                              float accumulated) {
    if (relaxation <= 0.0f) return (goal - speed) / mass;
    const float bias = fmx(depth / step, goal);
    const float drag = 2.0f * damping_ratio * relaxation / step;
    const float softness = relaxation * relaxation * truth / (step * step);
    const float prior = speed - accumulated * truth;
    const float wanted = (bias - (1.0f + drag) * prior) / (softness + (1.0f + drag) * truth);
    return truth * (wanted - accumulated) / mass;
}

// ---- velocity iterations over per-substep rows -------------------------------------------------
//
// Within one substep's iterations (solver_block_solve.wgsl:56-101) only the velocities and the
// accumulated impulses change: the bodies' poses, the masses, the split counts and every
// quantity of a block built from them stay put. k_solver_bodies and k_solve_prep evaluate those
// once per substep with exactly the expressions solver_contact_block.wgsl evaluates in every
// iteration, and k_solve reads them back, so the iterations produce the same bits.

constexpr uint32_t kSolveRow = 15;    // float4 per block: 3 of header, 3 per manifold point

// A block's solver row in sections of 3 float4: the header (section 0) and each manifold point's
// (section 1 + point), laid out section-major over the contact capacity, so the lanes of a wave
// read one section of neighbouring blocks as one contiguous run and the sections of points a
// block does not have are never read.
__device__ __forceinline__ float4* solve_section(const W& w, uint32_t ci, uint32_t section) {
    return w.srows + (static_cast<size_t>(section) * w.contact_cap + ci) * 3u;
}
constexpr uint32_t kSolverBody = 5;   // float4 per row: q, inertia 0-3, inertia 4-5 mass count, v, w
constexpr uint32_t kRowActive = 1u << 5;

__device__ __forceinline__ v3 inverse_inertia_at(float4 q, const float* inertia, v3 v) {
    return qrot(q, sym_apply(inertia, qrot(qconj(q), v)));  // apply_inverse_inertia
}

struct SolverBody {
    float4 q;
    float inertia[6];  // frozen (zero) for inert bodies
    float inverse_mass;
    float count;       // max(block count, 1), the mass split
    v3 v, w;           // the row's velocities (k_apply keeps them equal to the body state's)
};

__device__ __forceinline__ SolverBody load_solver_body(const W& w, uint32_t row) {
    const float4* p = w.sbodies + kSolverBody * row;
    const float4 a = p[0], b = p[1], c = p[2], d = p[3], e = p[4];
    SolverBody r;
    r.q = a;
    r.inertia[0] = b.x;
    r.inertia[1] = b.y;
    r.inertia[2] = b.z;
    r.inertia[3] = b.w;
    r.inertia[4] = c.x;
    r.inertia[5] = c.y;
    r.inverse_mass = c.z;
    r.count = c.w;
    r.v = mk(d.x, d.y, d.z);
    r.w = mk(e.x, e.y, e.z);
    return r;
}

// solver_block_solve.wgsl:24-54 for every claimed row: the frozen or unsplit body of a block.
__device__ __forceinline__ void solver_bodies_at(const W& w, uint32_t row) {
    if (row >= w.n || w.block_counts[row] == 0u) return;
    Body b = load_body(w, row);
    const bool inert = body_is_inert(b);
    if (inert) body_frozen(b);
    float4* p = w.sbodies + kSolverBody * row;
    p[0] = ld4(b.s.orientation);
    p[1] = make_float4(b.d.inverse_inertia[0], b.d.inverse_inertia[1], b.d.inverse_inertia[2],
                       b.d.inverse_inertia[3]);
    p[2] = make_float4(b.d.inverse_inertia[4], b.d.inverse_inertia[5], b.d.inverse_mass,
                       static_cast<float>(max(w.block_counts[row], 1u)));
    p[3] = make_float4(b.s.velocity[0], b.s.velocity[1], b.s.velocity[2], 0.0f);
    p[4] = make_float4(b.s.angular_velocity[0], b.s.angular_velocity[1], b.s.angular_velocity[2], 0.0f);
    float4* o = w.sposes + 2u * row;
    o[0] = make_float4(b.s.position[0], b.s.position[1], b.s.position[2], inert ? 1.0f : 0.0f);
    o[1] = make_float4(b.d.com[0], b.d.com[1], b.d.com[2], 0.0f);
}

// A claimed row's body as the substep's blocks see it, from the row's two copies above: pose,
// local centre of mass, velocities, and inverse mass and inertia (zero when inert). The body
// records' other fields are not read by the per-block passes and stay zero.
__device__ __forceinline__ Body posed_body(const W& w, uint32_t row, bool* inert) {
    const float4* p = w.sbodies + kSolverBody * row;
    const float4* o = w.sposes + 2u * row;
    const float4 q = p[0], i0 = p[1], i1 = p[2], v = p[3], av = p[4], at = o[0], com = o[1];
    Body b{};
    st4(b.s.orientation, q);
    b.s.position[0] = at.x;
    b.s.position[1] = at.y;
    b.s.position[2] = at.z;
    b.s.velocity[0] = v.x;
    b.s.velocity[1] = v.y;
    b.s.velocity[2] = v.z;
    b.s.angular_velocity[0] = av.x;
    b.s.angular_velocity[1] = av.y;
    b.s.angular_velocity[2] = av.z;
    b.d.com[0] = com.x;
    b.d.com[1] = com.y;
    b.d.com[2] = com.z;
    b.d.inverse_inertia[0] = i0.x;
    b.d.inverse_inertia[1] = i0.y;
    b.d.inverse_inertia[2] = i0.z;
    b.d.inverse_inertia[3] = i0.w;
    b.d.inverse_inertia[4] = i1.x;
    b.d.inverse_inertia[5] = i1.y;
    b.d.inverse_mass = i1.z;
    *inert = at.w != 0.0f;
    return b;
}

__global__ void __launch_bounds__(kThreads) k_solver_bodies(W w) {
    solver_bodies_at(w, blockIdx.x * blockDim.x + threadIdx.x);
}

// solver_contact_block.wgsl:195-233
__device__ __forceinline__ void warm_at(const W& w, uint32_t ci) {
    if (ci >= *w.ncontacts) return;
    const dyn_contact& contact = w.contacts[ci];
    if (!block_resolves(contact)) return;
    const uint32_t fb = w.blocks[2 * ci], sb = w.blocks[2 * ci + 1];
    bool fi, si;
    const Body fl = posed_body(w, fb, &fi), sl = posed_body(w, sb, &si);  // inert ones frozen
    Body first = fl, second = sl;
    const v3 normal = ld3(contact.normal);
    v3 t1, t2;
    make_tangents(normal, &t1, &t2);
    for (uint32_t pi = 0; pi < contact.point_count; ++pi) {
        const dyn_manifold_point& p = contact.points[pi];
        const v3 aa = arm(first, p, true), ab = arm(second, p, false);
        const v3 impulse = add(add(mul(normal, p.accumulated_normal), mul(t1, p.accumulated_tangent_1)),
                               mul(t2, p.accumulated_tangent_2));
        apply_pair_impulse(first, second, aa, ab, impulse);
    }
    commit_block(w, ci, fi, si, sub(ld3(first.s.velocity), ld3(fl.s.velocity)),
                 sub(ld3(first.s.angular_velocity), ld3(fl.s.angular_velocity)),
                 sub(ld3(second.s.velocity), ld3(sl.s.velocity)),
                 sub(ld3(second.s.angular_velocity), ld3(sl.s.angular_velocity)), false);
}

__global__ void __launch_bounds__(kThreads) k_warm(W w) {
    warm_at(w, blockIdx.x * blockDim.x + threadIdx.x);
}

// A block's point pi: its local anchors and depth, from the contact record or from the copy the
// step's first k_solve_prep made.
__device__ __forceinline__ void prep_anchor(const W& w, const dyn_contact& contact, uint32_t ci, uint32_t pi,
                                            bool from_contact, v3* la, v3* lb, float* depth) {
    if (from_contact) {
        const dyn_manifold_point& p = contact.points[pi];
        *la = ld3(p.local_a);
        *lb = ld3(p.local_b);
        *depth = p.depth;
        return;
    }
    const float4* anchor = w.sanchor + (static_cast<size_t>(pi) * w.contact_cap + ci) * 2u;
    const float4 a = anchor[0], b = anchor[1];
    *la = mk(a.x, a.y, a.z);
    *lb = mk(b.x, b.y, b.z);
    *depth = a.w;
}

// solver_contact_block.wgsl:31-110, the parts of a block that depend on poses and masses only:
// each point's arms from the centres of mass, its row masses, target speed and depth.
__device__ __forceinline__ void solve_prep_at(const W& w, uint32_t ci, uint32_t load_impulses) {
    if (ci >= *w.ncontacts) return;
    float4* row = solve_section(w, ci, 0u);
    const dyn_contact& contact = w.contacts[ci];
    const uint32_t point_count = contact.point_count;
    const uint32_t fb = w.blocks[2 * ci], sb = w.blocks[2 * ci + 1];
    if (point_count == 0u || contact.sensor != 0u) {  // block_resolves
        row[0] = make_float4(__uint_as_float(fb), __uint_as_float(sb), __uint_as_float(0u), 0.0f);
        return;
    }
    if (load_impulses) {  // the step's iterations keep the impulses in simp (k_impulses_store)
        float an[4] = {0.0f, 0.0f, 0.0f, 0.0f}, a1[4] = {0.0f, 0.0f, 0.0f, 0.0f}, a2[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (uint32_t pi = 0; pi < 4; ++pi) {
            if (pi < point_count) {
                an[pi] = contact.points[pi].accumulated_normal;
                a1[pi] = contact.points[pi].accumulated_tangent_1;
                a2[pi] = contact.points[pi].accumulated_tangent_2;
            }
        }
        float4* imp = w.simp + 3u * ci;
        imp[0] = make_float4(an[0], an[1], an[2], an[3]);
        imp[1] = make_float4(a1[0], a1[1], a1[2], a1[3]);
        imp[2] = make_float4(a2[0], a2[1], a2[2], a2[3]);
    }
    // the split bodies scale the frozen copies: an inert body's zeros stay zero
    bool fi, si;
    const Body pf = posed_body(w, fb, &fi), ps = posed_body(w, sb, &si);
    Body split_f = pf, split_s = ps;
    const float c1 = w.sbodies[kSolverBody * fb + 2u].w, c2 = w.sbodies[kSolverBody * sb + 2u].w;
    split_f.d.inverse_mass = split_f.d.inverse_mass * c1;
    split_s.d.inverse_mass = split_s.d.inverse_mass * c2;
    for (int k = 0; k < 6; ++k) {
        split_f.d.inverse_inertia[k] = split_f.d.inverse_inertia[k] * c1;
        split_s.d.inverse_inertia[k] = split_s.d.inverse_inertia[k] * c2;
    }
    const uint32_t flags = point_count | (fi ? 8u : 0u) | (si ? 16u : 0u) | ((fi && si) ? 0u : kRowActive);
    const v3 normal = ld3(contact.normal);
    row[0] = make_float4(__uint_as_float(fb), __uint_as_float(sb), __uint_as_float(flags), contact.friction);
    row[1] = make_float4(normal.x, normal.y, normal.z, contact.relaxation);
    row[2] = make_float4(contact.damping_ratio, contact.rolling_friction, contact.spin_friction, 0.0f);
    if (fi && si) return;  // both inert: zero words (commit_block of zeros)
    v3 t1, t2;
    make_tangents(normal, &t1, &t2);
    // the points' local anchors and depths stay fixed through the step: the first substep copies
    // them from the contact's 64-byte point records into 32 bytes a point (point-major), which
    // the later substeps read instead
    v3 reach_a = zero3v(), reach_b = zero3v();
    const v3 com_f = body_com(pf), com_s = body_com(ps);
    for (uint32_t pi = 0; pi < point_count; ++pi) {
        v3 la, lb;
        float depth;
        prep_anchor(w, contact, ci, pi, load_impulses != 0u, &la, &lb, &depth);
        if (load_impulses) {
            float4* anchor = w.sanchor + (static_cast<size_t>(pi) * w.contact_cap + ci) * 2u;
            anchor[0] = make_float4(la.x, la.y, la.z, depth);
            anchor[1] = make_float4(lb.x, lb.y, lb.z, 0.0f);
        }
        reach_a = add(reach_a, sub(arm_at(pf, la), com_f));
        reach_b = add(reach_b, sub(arm_at(ps, lb), com_s));
    }
    const float count = static_cast<float>(max(point_count, 1u));
    for (uint32_t pi = 0; pi < point_count; ++pi) {
        v3 la, lb;
        float depth;
        prep_anchor(w, contact, ci, pi, load_impulses != 0u, &la, &lb, &depth);
        const v3 aa = arm_at(pf, la), ab = arm_at(ps, lb);
        const v3 ra = sub(aa, body_com(pf)), rb = sub(ab, body_com(ps));  // point_velocity's arms
        const v3 arm_a = cross3(sub(aa, body_com(split_f)), normal);
        const v3 arm_b = cross3(sub(ab, body_com(split_s)), normal);
        const float nmass = count * (split_f.d.inverse_mass + split_s.d.inverse_mass) +
                            dot3(apply_inverse_inertia(split_f, cross3(reach_a, normal)), arm_a) +
                            dot3(apply_inverse_inertia(split_s, cross3(reach_b, normal)), arm_b);
        const float truth = momentum_mass(pf, ps, aa, ab, normal);
        const float m1 = momentum_mass(split_f, split_s, aa, ab, t1) * count;
        const float m2 = momentum_mass(split_f, split_s, aa, ab, t2) * count;
        float4* q = solve_section(w, ci, 1u + pi);
        q[0] = make_float4(ra.x, ra.y, ra.z, nmass);
        q[1] = make_float4(rb.x, rb.y, rb.z, truth);
        q[2] = make_float4(m1, m2, w.targets[ci * 4u + pi], depth);
    }
}

__global__ void __launch_bounds__(kThreads) k_solve_prep(W w, uint32_t load_impulses) {
    solve_prep_at(w, blockIdx.x * blockDim.x + threadIdx.x, load_impulses);
}

// solver_contact_block.wgsl:31-193 over the block's row: the points' impulses are applied as
// each point is solved (every point reads only the velocities the block started from, so this
// adds the same impulses in the same order as applying them all after the loop).
__device__ __forceinline__ void solve_at(const W& w, uint32_t ci, uint32_t residual) {
    if (ci >= *w.ncontacts) return;
    const float4* row = solve_section(w, ci, 0u);
    const float4 h0 = row[0];
    const uint32_t flags = __float_as_uint(h0.z);
    if (!(flags & kRowActive)) return;  // not resolved, or both bodies inert
    const uint32_t fb = __float_as_uint(h0.x), sb = __float_as_uint(h0.y);
    const bool fi = (flags & 8u) != 0u, si = (flags & 16u) != 0u;
    const uint32_t point_count = flags & 7u;
    const float friction = h0.w;
    const float4 h1 = row[1], h2 = row[2];
    const v3 normal = mk(h1.x, h1.y, h1.z);
    const float relaxation = h1.w, damping_ratio = h2.x;
    const SolverBody bf = load_solver_body(w, fb), bs = load_solver_body(w, sb);
    float split_f[6], split_s[6];
    for (int k = 0; k < 6; ++k) {
        split_f[k] = bf.inertia[k] * bf.count;
        split_s[k] = bs.inertia[k] * bs.count;
    }
    const v3 vf = bf.v, wf = bf.w, vs = bs.v, ws = bs.w;
    v3 first_v = vf, first_w = wf, second_v = vs, second_w = ws;
    v3 t1, t2;
    make_tangents(normal, &t1, &t2);
    float4* imp = w.simp + 3u * ci;
    const float4 n4 = imp[0], t14 = imp[1], t24 = imp[2];
    float acc_n[4] = {n4.x, n4.y, n4.z, n4.w}, acc_1[4] = {t14.x, t14.y, t14.z, t14.w};
    float acc_2[4] = {t24.x, t24.y, t24.z, t24.w};
    float normal_total = 0.0f;
#pragma unroll
    for (uint32_t pi = 0; pi < 4; ++pi) {
        if (pi < point_count) normal_total = normal_total + fmx(acc_n[pi], 0.0f);
    }
#pragma unroll
    for (uint32_t pi = 0; pi < 4; ++pi) {
        if (pi >= point_count) continue;
        const float4* q = solve_section(w, ci, 1u + pi);
        const float4 qa = q[0], qb = q[1], qc = q[2];
        const v3 ra = mk(qa.x, qa.y, qa.z), rb = mk(qb.x, qb.y, qb.z);
        const float nmass = qa.w, truth = qb.w, m1 = qc.x, m2 = qc.y, target = qc.z, depth = qc.w;
        float an = acc_n[pi], at1 = acc_1[pi], at2 = acc_2[pi];
        // relative_velocity of the velocities the block started from
        const v3 rel = sub(add(vs, cross3(ws, rb)), add(vf, cross3(wf, ra)));
        v3 impulse = zero3v();
        const float nspeed = dot3(rel, normal);
        const float delta = normal_delta(relaxation, damping_ratio, nmass, truth, w.p.dt, target, depth,
                                         nspeed, an);
        const float next_n = fmx(0.0f, an + delta);
        impulse = add(impulse, mul(normal, next_n - an));
        an = next_n;
        const float limit = friction * an;
        const float s1 = dot3(rel, t1);
        const float next_1 = clampf(at1 - s1 / m1, -limit, limit);
        impulse = add(impulse, mul(t1, next_1 - at1));
        at1 = next_1;
        const float s2 = dot3(rel, t2);
        const float remaining = sqrtf(fmx(limit * limit - at1 * at1, 0.0f));
        const float next_2 = clampf(at2 - s2 / m2, -remaining, remaining);
        impulse = add(impulse, mul(t2, next_2 - at2));
        at2 = next_2;
        acc_n[pi] = an;
        acc_1[pi] = at1;
        acc_2[pi] = at2;
        // apply_pair_impulse on the unsplit (frozen when inert) bodies
        first_v = sub(first_v, mul(impulse, bf.inverse_mass));
        first_w = sub(first_w, inverse_inertia_at(bf.q, bf.inertia, cross3(ra, impulse)));
        second_v = add(second_v, mul(impulse, bs.inverse_mass));
        second_w = add(second_w, inverse_inertia_at(bs.q, bs.inertia, cross3(rb, impulse)));
    }
    const v3 rel_spin = sub(second_w, first_w);
    const v3 planar = sub(rel_spin, mul(normal, dot3(rel_spin, normal)));
    const float planar_len = len3(planar);
    if (planar_len > 1e-6f) {
        const v3 axis = dv(planar, planar_len);
        const float mass = dot3(axis, inverse_inertia_at(bf.q, split_f, axis)) +
                           dot3(axis, inverse_inertia_at(bs.q, split_s, axis));
        if (mass > 0.0f) {
            const float lim = h2.y * normal_total;
            const float applied = clampf(-planar_len / mass, -lim, lim);
            first_w = sub(first_w, inverse_inertia_at(bf.q, bf.inertia, mul(axis, applied)));
            second_w = add(second_w, inverse_inertia_at(bs.q, bs.inertia, mul(axis, applied)));
        }
    }
    const float twist = dot3(rel_spin, normal);
    if (fabsf(twist) > 1e-6f) {
        const float mass = dot3(normal, inverse_inertia_at(bf.q, split_f, normal)) +
                           dot3(normal, inverse_inertia_at(bs.q, split_s, normal));
        if (mass > 0.0f) {
            const float lim = h2.z * normal_total;
            const float applied = clampf(-twist / mass, -lim, lim);
            first_w = sub(first_w, inverse_inertia_at(bf.q, bf.inertia, mul(normal, applied)));
            second_w = add(second_w, inverse_inertia_at(bs.q, bs.inertia, mul(normal, applied)));
        }
    }
    imp[0] = make_float4(acc_n[0], acc_n[1], acc_n[2], acc_n[3]);
    imp[1] = make_float4(acc_1[0], acc_1[1], acc_1[2], acc_1[3]);
    imp[2] = make_float4(acc_2[0], acc_2[1], acc_2[2], acc_2[3]);
    commit_block(w, ci, fi, si, sub(first_v, vf), sub(first_w, wf), sub(second_v, vs), sub(second_w, ws),
                 residual != 0u);
}

__global__ void __launch_bounds__(kThreads) k_solve(W w, uint32_t residual) {
    solve_at(w, blockIdx.x * blockDim.x + threadIdx.x, residual);
}

// The iterations' impulses back into the contact records, for the archive and the next step.
__device__ __forceinline__ void impulses_store_at(const W& w, uint32_t ci) {
    if (ci >= *w.ncontacts) return;
    dyn_contact& contact = w.contacts[ci];
    const uint32_t point_count = contact.point_count;
    if (point_count == 0u || contact.sensor != 0u) return;
    const float4* imp = w.simp + 3u * ci;
    const float4 n4 = imp[0], t14 = imp[1], t24 = imp[2];
    const float an[4] = {n4.x, n4.y, n4.z, n4.w}, a1[4] = {t14.x, t14.y, t14.z, t14.w};
    const float a2[4] = {t24.x, t24.y, t24.z, t24.w};
#pragma unroll
    for (uint32_t pi = 0; pi < 4; ++pi) {
        if (pi < point_count) {
            contact.points[pi].accumulated_normal = an[pi];
            contact.points[pi].accumulated_tangent_1 = a1[pi];
            contact.points[pi].accumulated_tangent_2 = a2[pi];
        }
    }
}

__global__ void __launch_bounds__(kThreads) k_impulses_store(W w) {
    impulses_store_at(w, blockIdx.x * blockDim.x + threadIdx.x);
}

// solver_block_apply.wgsl over every claimed row
// List entries whose loads k_apply and k_position_apply keep in flight together: a row of up to
// kApplyBatch blocks reads its entries (and, in k_apply, then their blocks' words) in one round
// (integer sums: any order gives the same words).
constexpr uint32_t kApplyBatch = 8;

// The velocities live in the row's solver record (k_solver_bodies copies them from the body
// state at the substep's start) while the iterations run: only the substep's last apply
// (`commit`) writes them to the body state, for every claimed row, as the same bits an update of
// both after every apply would leave there.
// Thread t serves row t >> lanes_log2 with its 2^lanes_log2 neighbours (one wave): each lane sums
// every 2^lanes_log2-th entry of the row's list and the lanes add their sums with shuffles
// (integer sums: the same words in any order); the first lane finishes the row. A small world
// gets several lanes a row, so the rows' dependent loads overlap more.
__device__ __forceinline__ uint32_t row_lanes_sum(uint32_t v, uint32_t lanes_log2) {
    for (uint32_t o = 1; o < (1u << lanes_log2); o <<= 1) v += __shfl_xor(v, static_cast<int>(o), kWave);
    return v;
}

__device__ __forceinline__ void apply_at(const W& w, uint32_t t, uint32_t commit, uint32_t lanes_log2) {
    const uint32_t lanes = 1u << lanes_log2;
    const uint32_t row = t >> lanes_log2, sub = t & (lanes - 1u);
    const bool live = row < w.n;
    // list_counts is 0 for rows without blocks and for inert rows, which receive no words
    const uint32_t count = live ? w.list_counts[row] : 0u;
    const uint32_t start = live ? w.row_start[row] : 0u;
    float4* p = w.sbodies + kSolverBody * (live ? row : 0u);  // the solver body's copy
    float4 pv = make_float4(0.0f, 0.0f, 0.0f, 0.0f), pw = pv;
    if (sub == 0u && count != 0u) {  // in flight with the list
        pv = p[3];
        pw = p[4];
    }
    const v3 v0 = mk(pv.x, pv.y, pv.z), av0 = mk(pw.x, pw.y, pw.z);
    const uint32_t* list = w.adj + start;
    uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0;
    for (uint32_t e = sub; e < count; e += kApplyBatch * lanes) {
        uint32_t entry[kApplyBatch];
#pragma unroll
        for (uint32_t j = 0; j < kApplyBatch; ++j) entry[j] = e + j * lanes < count ? list[e + j * lanes] : kNoSlot;
        uint2 a[kApplyBatch], b[kApplyBatch], c[kApplyBatch];
#pragma unroll
        for (uint32_t j = 0; j < kApplyBatch; ++j) {
            a[j] = b[j] = c[j] = make_uint2(0u, 0u);
            if (entry[j] != kNoSlot) {
                const uint2* d =
                    reinterpret_cast<const uint2*>(w.vwords + (entry[j] >> 1) * 12u + (entry[j] & 1u) * 6u);
                a[j] = d[0];
                b[j] = d[1];
                c[j] = d[2];
            }
        }
#pragma unroll
        for (uint32_t j = 0; j < kApplyBatch; ++j) {
            s0 += a[j].x;
            s1 += a[j].y;
            s2 += b[j].x;
            s3 += b[j].y;
            s4 += c[j].x;
            s5 += c[j].y;
        }
    }
    if (lanes_log2 != 0u) {
        s0 = row_lanes_sum(s0, lanes_log2);
        s1 = row_lanes_sum(s1, lanes_log2);
        s2 = row_lanes_sum(s2, lanes_log2);
        s3 = row_lanes_sum(s3, lanes_log2);
        s4 = row_lanes_sum(s4, lanes_log2);
        s5 = row_lanes_sum(s5, lanes_log2);
    }
    if (sub != 0u || count == 0u) return;
    v3 v = v0, av = av0;
    if ((s0 | s1 | s2 | s3 | s4 | s5) != 0u) {
        const v3 lin = mk(solver_value(s0, DYN_SOLVER_VELOCITY_SCALE), solver_value(s1, DYN_SOLVER_VELOCITY_SCALE),
                          solver_value(s2, DYN_SOLVER_VELOCITY_SCALE));
        const v3 ang = mk(solver_value(s3, DYN_SOLVER_VELOCITY_SCALE), solver_value(s4, DYN_SOLVER_VELOCITY_SCALE),
                          solver_value(s5, DYN_SOLVER_VELOCITY_SCALE));
        if (!(zero3(lin) && zero3(ang))) {
            v = add(v0, lin);
            av = add(av0, ang);
            p[3] = make_float4(v.x, v.y, v.z, 0.0f);
            p[4] = make_float4(av.x, av.y, av.z, 0.0f);
        }
    }
    if (commit) {
        dyn_body_state& s = w.st[row];
        st3(s.velocity, v);
        st3(s.angular_velocity, av);
    }
}

__global__ void __launch_bounds__(kThreads) k_apply(W w, uint32_t commit, uint32_t lanes_log2) {
    apply_at(w, blockIdx.x * blockDim.x + threadIdx.x, commit, lanes_log2);
}

// substep_integrate.wgsl:15-76 over the awake dynamic rows (no fields)
__device__ __forceinline__ void integrate_at(const W& w, uint32_t row) {
    if (row >= w.d) return;
    const dyn_body_desc d = w.de[row];
    dyn_body_state s = w.st[row];
    if (!body_is_active(s, d) || !kind_simulates(d)) return;
    const float h = w.p.substep_dt;
    const q4 q = ld4(s.orientation);
    const v3 acc = add(mul(mk(w.p.gravity[0], w.p.gravity[1], w.p.gravity[2]), d.gravity_scale),
                       mul(ld3(s.force), d.inverse_mass));
    const v3 spin = ld3(s.angular_velocity);
    st3(s.velocity, add(ld3(s.velocity), mul(acc, h)));
    v3 gspin = spin;
    const bool iso = d.inertia[1] == 0.0f && d.inertia[2] == 0.0f && d.inertia[4] == 0.0f &&
                     d.inertia[0] == d.inertia[3] && d.inertia[3] == d.inertia[5];
    if (!iso) {
        const v3 body_spin = qrot(qconj(q), spin);
        const v3 momentum = sym_apply(d.inertia, body_spin);
        if (!zero3(cross3(body_spin, momentum))) {
            v3 solved = momentum;
            for (int it = 0; it < 4; ++it) {
                const v3 mid = mul(add(momentum, solved), 0.5f);
                solved = rotate_about(mul(neg(sym_apply(d.inverse_inertia, mid)), h), momentum);
            }
            gspin = qrot(q, sym_apply(d.inverse_inertia, solved));
        }
    }
    const v3 torque_term = qrot(q, sym_apply(d.inverse_inertia, qrot(qconj(q), mul(ld3(s.torque), h))));
    st3(s.angular_velocity, add(gspin, torque_term));
    const float speed = len3(ld3(s.velocity));
    if (speed > w.p.max_velocity) st3(s.velocity, mul(ld3(s.velocity), w.p.max_velocity / speed));
    const float sspeed = len3(ld3(s.angular_velocity));
    if (sspeed > w.p.max_angular_velocity)
        st3(s.angular_velocity, mul(ld3(s.angular_velocity), w.p.max_angular_velocity / sspeed));
    st3(s.velocity, dv(ld3(s.velocity), 1.0f + d.linear_damping * h));
    st3(s.angular_velocity, dv(ld3(s.angular_velocity), 1.0f + d.angular_damping * h));
    w.st[row] = s;
}

__global__ void __launch_bounds__(kThreads) k_integrate(W w) {
    integrate_at(w, blockIdx.x * blockDim.x + threadIdx.x);
}

// substep_advance.wgsl
__device__ __forceinline__ void advance_at(const W& w, uint32_t row) {
    if (row >= w.d) return;
    dyn_body_state s = w.st[row];
    if (!body_is_active(s, w.de[row])) return;
    const float h = w.p.substep_dt;
    st3(s.position, add(ld3(s.position), mul(ld3(s.velocity), h)));
    const q4 spin = make_float4(s.angular_velocity[0], s.angular_velocity[1], s.angular_velocity[2], 0.0f);
    const q4 q = ld4(s.orientation);
    const q4 dq = qmul(spin, q);
    const float k = 0.5f * h;
    st4(s.orientation, qnorm(make_float4(q.x + k * dq.x, q.y + k * dq.y, q.z + k * dq.z, q.w + k * dq.w)));
    w.st[row] = s;
}

__global__ void __launch_bounds__(kThreads) k_advance(W w) {
    advance_at(w, blockIdx.x * blockDim.x + threadIdx.x);
}

// position_correction.wgsl:5-46, position_block.wgsl, from the block's position row
// position_correction.wgsl:5-46 for one block of `count` points: false when it contributes
// nothing, else its two sides' corrections (inert rows only get their "contributed" mark).
__device__ __forceinline__ bool position_correction(const W& w, float4 h0, const float4* prow, uint32_t flags,
                                                    uint32_t count, v3* la_out, v3* lb_out) {
    const uint32_t fr = __float_as_uint(h0.x), sr = __float_as_uint(h0.y);
    const bool fi = (flags & 8u) != 0u, si = (flags & 16u) != 0u;
    const float imf = h0.w;
    const float4 h1 = prow[1];
    const float ims = h1.x;
    const v3 normal = mk(h1.y, h1.z, h1.w);
    const float4 depths = prow[2];
    const float4 rs = reinterpret_cast<const float4*>(w.resolution)[sr];
    const float4 rf = reinterpret_cast<const float4*>(w.resolution)[fr];
    const float resolved = dot3(normal, sub(mk(rs.x, rs.y, rs.z), mk(rf.x, rf.y, rf.z)));
    v3 total = zero3v();
    uint32_t contributing = 0;
#pragma unroll
    for (uint32_t pi = 0; pi < 4; ++pi) {
        if (pi >= count) continue;
        const float depth = pi == 0 ? depths.x : (pi == 1 ? depths.y : (pi == 2 ? depths.z : depths.w));
        const float error = fmx(depth - w.p.slop - resolved, 0.0f);
        if (error <= 0.0f) continue;
        const float k = imf + ims;
        if (k == 0.0f) continue;
        total = add(total, mul(normal, w.p.relaxation * error / k));
        ++contributing;
    }
    if (contributing == 0u) return false;
    const v3 correction = dv(total, static_cast<float>(contributing));
    const v3 la = mul(neg(correction), imf), lb = mul(correction, ims);
    if (!(dot3(la, la) > 0.0f || dot3(lb, lb) > 0.0f)) return false;
    // Inert rows get exactly zero words; only their "contributed" mark matters.
    if (fi) w.contrib[fr] = 1u;
    if (si) w.contrib[sr] = 1u;
    *la_out = la;
    *lb_out = lb;
    return true;
}

__device__ __forceinline__ void position_block_at(const W& w, uint32_t ci) {
    if (ci >= *w.ncontacts) return;
    const float4* prow = w.prows + 3u * ci;
    const float4 h0 = prow[0];
    const uint32_t flags = __float_as_uint(h0.z);
    const uint32_t count = flags & 7u;  // 0: not resolved, or relaxed by compliance
    const uint2 entries = reinterpret_cast<const uint2*>(w.slots)[ci];
    v3 la, lb;
    // every entry of the block is written once per pass: "no contribution" unless corrected
    uint4 wa = make_uint4(0u, 0u, 0u, 0u), wb = wa;
    if (count != 0u && position_correction(w, h0, prow, flags, count, &la, &lb)) {
        const float scale = DYN_SOLVER_POSITION_SCALE;
        wa = make_uint4(1u, solver_word(la.x, scale), solver_word(la.y, scale), solver_word(la.z, scale));
        wb = make_uint4(1u, solver_word(lb.x, scale), solver_word(lb.y, scale), solver_word(lb.z, scale));
    }
    if (entries.x != kNoSlot) w.pwords[entries.x] = wa;
    if (entries.y != kNoSlot) w.pwords[entries.y] = wb;
}

__global__ void __launch_bounds__(kThreads) k_position_block(W w) {
    position_block_at(w, blockIdx.x * blockDim.x + threadIdx.x);
}

// position_apply.wgsl over every claimed row
// A claimed row (count > 0) keeps its pose in ppose through the iterations: the first reads the
// body state and writes ppose (moved or not), the last writes the body state (every claimed row,
// as the same bits per-iteration updates of the state would leave). An inert row with a
// contributing block, which differs from iteration to iteration, reads and writes the state.
__device__ __forceinline__ void position_apply_at(const W& w, uint32_t t, uint32_t first, uint32_t last,
                                                  uint32_t lanes_log2) {
    const uint32_t lanes = 1u << lanes_log2;
    const uint32_t row = t >> lanes_log2, sub = t & (lanes - 1u);
    const bool live = row < w.n;
    // list_counts is 0 for rows without blocks and for inert rows; contrib, zero outside a
    // position pass, is marked only for inert rows with a contributing block
    const uint32_t count = live ? w.list_counts[row] : 0u;
    const uint32_t start = live ? w.row_start[row] : 0u;
    const uint32_t at = live ? row : 0u;
    dyn_body_state& s = w.st[at];
    float4& r = reinterpret_cast<float4*>(w.resolution)[at];
    float4* pose = w.ppose + 2u * at;
    const bool in_state = count == 0u || first;
    v3 p0 = zero3v();
    float4 q0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), r0 = q0;
    if (live && sub == 0u) {  // in flight with the list
        if (in_state) {
            p0 = ld3(s.position);
            q0 = ld4(s.orientation);
        } else {
            const float4 a = pose[0];
            p0 = mk(a.x, a.y, a.z);
            q0 = pose[1];
        }
        r0 = r;
    }
    uint32_t contributing = 0;
    uint4 lo = make_uint4(0u, 0u, 0u, 0u);  // linear words, then the (zero) angular x
    uint2 hi = make_uint2(0u, 0u);          // angular y, z: the position blocks send none
    const uint4* run = w.pwords + start;
    for (uint32_t e = sub; e < count; e += kApplyBatch * lanes) {
        uint4 word[kApplyBatch];
#pragma unroll
        for (uint32_t j = 0; j < kApplyBatch; ++j) {
            word[j] = e + j * lanes < count ? run[e + j * lanes] : make_uint4(0u, 0u, 0u, 0u);
        }
#pragma unroll
        for (uint32_t j = 0; j < kApplyBatch; ++j) {
            if (word[j].x == 0u) continue;  // the block did not contribute
            ++contributing;
            lo.x += word[j].y;
            lo.y += word[j].z;
            lo.z += word[j].w;
        }
    }
    if (lanes_log2 != 0u) {
        contributing = row_lanes_sum(contributing, lanes_log2);
        lo.x = row_lanes_sum(lo.x, lanes_log2);
        lo.y = row_lanes_sum(lo.y, lanes_log2);
        lo.z = row_lanes_sum(lo.z, lanes_log2);
    }
    if (!live || sub != 0u) return;
    if (count == 0u) {
        contributing = w.contrib[row];  // inert row: only the mark
        if (contributing == 0u) return;
        w.contrib[row] = 0u;
    } else {
        if (contributing == 0u) {
            if (first && !last) {
                pose[0] = make_float4(p0.x, p0.y, p0.z, 0.0f);
                pose[1] = q0;
            } else if (last && !first) {
                st3(s.position, p0);
                st4(s.orientation, q0);
            }
            return;
        }
    }
    const v3 lin = mk(solver_value(lo.x, DYN_SOLVER_POSITION_SCALE), solver_value(lo.y, DYN_SOLVER_POSITION_SCALE),
                      solver_value(lo.z, DYN_SOLVER_POSITION_SCALE));
    const v3 ang = mk(solver_value(lo.w, DYN_SOLVER_POSITION_SCALE), solver_value(hi.x, DYN_SOLVER_POSITION_SCALE),
                      solver_value(hi.y, DYN_SOLVER_POSITION_SCALE));
    const float cf = static_cast<float>(contributing);
    const v3 al = dv(lin, cf), aa = dv(ang, cf);
    const v3 p1 = add(p0, al);
    const float4 q1 = qnorm(qmul(make_float4(aa.x * 0.5f, aa.y * 0.5f, aa.z * 0.5f, 1.0f), q0));
    if (count == 0u || last) {
        st3(s.position, p1);
        st4(s.orientation, q1);
    } else {
        pose[0] = make_float4(p1.x, p1.y, p1.z, 0.0f);
        pose[1] = q1;
    }
    const v3 nr = add(mk(r0.x, r0.y, r0.z), al);
    r = make_float4(nr.x, nr.y, nr.z, 0.0f);
}

__global__ void __launch_bounds__(kThreads) k_position_apply(W w, uint32_t first, uint32_t last, uint32_t lanes_log2) {
    position_apply_at(w, blockIdx.x * blockDim.x + threadIdx.x, first, last, lanes_log2);
}

// island_sleep.wgsl
__global__ void __launch_bounds__(kThreads) k_sleep(W w) {
    uint32_t slept = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < w.d; i += gridDim.x * blockDim.x) {
        dyn_body_state& s = w.st[i];
        if (s.sleeping != 0u || !kind_simulates(w.de[i])) continue;
        const uint32_t root = w.parents[i];
        const bool disturbed = (w.istate[root] & kIslandWake) != 0u || w.wake[i] != 0u;
        const float timer = disturbed ? 0.0f : s.sleep_timer + w.p.dt;
        if (disturbed || !(timer >= w.p.sleep_time)) {  // (a NaN timer stays awake)
            s.sleep_timer = timer;  // the only field an awake row changes here
            continue;
        }
        for (int c = 0; c < 3; ++c) {
            s.velocity[c] = 0.0f;
            s.angular_velocity[c] = 0.0f;
            s.force[c] = 0.0f;
            s.torque[c] = 0.0f;
            s.prev_position[c] = s.position[c];
        }
        for (int c = 0; c < 4; ++c) s.prev_orientation[c] = s.orientation[c];
        s.sleep_timer = 0.0f;
        s.sleeping = 1u;
        ++slept;
    }
    block_count(cw(w.counters, DYN_COUNTER_SLEPT), slept);
}

// ---- commit: thaw, archive, static wake clear, freeze (commit.rs:187-209) ---------------

// rest_flags[i] = keep resting manifold i (both bodies inactive at the step's start)
__global__ void __launch_bounds__(kThreads) k_thaw(W w) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *w.nresting) return;
    const dyn_contact& c = w.resting[i];
    const uint32_t r1 = resolve_row(w, c.first_body_id, c.first_generation);
    const uint32_t r2 = resolve_row(w, c.second_body_id, c.second_generation);
    const bool keep = r1 != DYN_NO_BODY && r2 != DYN_NO_BODY && w.activity[r1] == 0u && w.activity[r2] == 0u;
    w.rest_flags[i] = keep ? 1u : 0u;
}

// contact_archive.wgsl + freeze_contacts.wgsl flags (freeze flags after the resting flags)
// contact_archive.wgsl keeps the step's contacts for the next step's relay: here the contact and
// archive buffers trade places at the end of the step (swap_contact_buffers) instead of copying,
// so this pass only marks the contacts freeze_contacts.wgsl:13-45 moves to the resting archive.
__global__ void __launch_bounds__(kThreads) k_archive(W w, uint32_t* freeze_flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *w.ncontacts) return;
    const uint2 ab = w.ckeys[i];
    const uint32_t f = w.owner[ab.x], s = w.owner[ab.y];
    const bool frozen = !body_is_active(w.st[f], w.de[f]) && !body_is_active(w.st[s], w.de[s]);
    freeze_flags[i] = frozen ? 1u : 0u;
}

__global__ void __launch_bounds__(kThreads) k_archive_count(W w) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *w.narchive = *w.ncontacts;
        *cw(w.counters, DYN_COUNTER_ARCHIVED) = *w.ncontacts;
    }
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w.d && i < w.n) w.wake[i] = 0u;  // static_wake_clear.wgsl
}

// New resting list = kept manifolds then frozen manifolds, keyed by (min id, max id).
__global__ void __launch_bounds__(kThreads) k_rest_gather(W w, const uint32_t* keep_dest,
                                                          const uint32_t* keep_total,
                                                          const uint32_t* freeze_dest,
                                                          uint32_t* out_count) {
    // the workgroup lists its records (source, destination) and copies them together, its lanes
    // on consecutive 16-byte pieces
    constexpr uint32_t kPieces = sizeof(dyn_contact) / sizeof(float4);
    __shared__ const dyn_contact* from[2 * kThreads];
    __shared__ uint32_t to[2 * kThreads];
    __shared__ uint32_t jobs;
    if (threadIdx.x == 0) jobs = 0u;
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t kept = *keep_total;
    const uint32_t old = *w.nresting;
    const uint32_t ncont = *w.ncontacts;
    if (i == 0) {
        // freeze_dest totals are written after the map by k_scan_blocks into scan_total[1]
        const uint32_t total = kept + w.scan_total[1];
        const uint32_t capped = min(total, w.contact_cap);
        *out_count = capped;
        *cw(w.counters, DYN_COUNTER_REFUSED_RESTING) = total - capped;
        *cw(w.counters, DYN_COUNTER_RESTING) = capped;
    }
    if (i < old && keep_dest[i] != kNoSlot) {
        const uint32_t at = keep_dest[i];
        if (at < w.contact_cap) {
            const dyn_contact& c = w.resting[i];
            const uint32_t k = atomicAdd(&jobs, 1u);
            from[k] = &c;
            to[k] = at;
            w.rest_major[at] = min(c.first_body_id, c.second_body_id);
            w.rest_minor[at] = max(c.first_body_id, c.second_body_id);
            w.rest_slot[at] = at;
        }
    }
    if (i < ncont && freeze_dest[i] != kNoSlot) {
        const uint32_t at = kept + freeze_dest[i];
        if (at < w.contact_cap) {
            const dyn_contact& c = w.contacts[i];
            const uint32_t k = atomicAdd(&jobs, 1u);
            from[k] = &c;
            to[k] = at;
            w.rest_major[at] = min(c.first_body_id, c.second_body_id);
            w.rest_minor[at] = max(c.first_body_id, c.second_body_id);
            w.rest_slot[at] = at;
        }
    }
    __syncthreads();
    const uint32_t n = jobs;
    for (uint32_t item = threadIdx.x; item < n * kPieces; item += blockDim.x) {
        const uint32_t k = item / kPieces, piece = item - k * kPieces;
        reinterpret_cast<float4*>(w.rest_tmp + to[k])[piece] = reinterpret_cast<const float4*>(from[k])[piece];
    }
}

// resting[i] = rest_tmp[rest_slot[i]], a workgroup's records copied together (lanes on
// consecutive 16-byte pieces).
__global__ void __launch_bounds__(kThreads) k_rest_place(W w, const uint32_t* count) {
    constexpr uint32_t kPieces = sizeof(dyn_contact) / sizeof(float4);
    __shared__ uint32_t from[kThreads];
    const uint32_t base = blockIdx.x * blockDim.x;
    const uint32_t i = base + threadIdx.x;
    const uint32_t n = *count;
    if (i == 0) *w.nresting = n;
    if (i < n) from[threadIdx.x] = w.rest_slot[i];
    __syncthreads();
    const uint32_t m = base < n ? min(n - base, static_cast<uint32_t>(kThreads)) : 0u;
    for (uint32_t item = threadIdx.x; item < m * kPieces; item += blockDim.x) {
        const uint32_t k = item / kPieces, piece = item - k * kPieces;
        reinterpret_cast<float4*>(w.resting + base + k)[piece] =
            reinterpret_cast<const float4*>(w.rest_tmp + from[k])[piece];
    }
}

__global__ void k_consume(W w) {
    if (threadIdx.x == 0) {
        *cw(w.counters, DYN_COUNTER_STEP) += 1u;
        *w.edit_runs_count = 0u;  // the step applied the queued edits
    }
}

}  // namespace

}  // namespace dyn

// ---- host side: the World object behind the C ABI ---------------------------------------------

// A second stream and the events that fork it from and join it to a world's stream.
struct SideStream {
    hipStream_t stream;
    hipEvent_t fork, join;
};

struct dyn_world {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    dyn::W w{};
    uint32_t entry_cap = 0;
    uint32_t* entry_keys = nullptr;
    uint32_t* entry_order = nullptr;
    dyn_grid_entry* entries = nullptr;
    uint32_t* words = nullptr;        // small device words (counts)
    uint32_t* freeze_flags = nullptr;
    uint32_t* keep_dest = nullptr;
    uint32_t* freeze_dest = nullptr;
    void* bp_scratch = nullptr;
    size_t bp_scratch_bytes = 0;
    void* sort_scratch = nullptr;
    size_t sort_scratch_bytes = 0;
    std::vector<void*> owned;
    void* shape_sources = nullptr;   // owned separately: replaced by each shape upload
    void* shape_vertices = nullptr;
    void* support_cells = nullptr;   // the hull sources' support cells (owned like the sources)
    void* support_cands = nullptr;
    uint32_t collider_count = 0;
    std::vector<uint32_t> host_kinds;    // collider kinds and sources, for upload validation
    std::vector<uint32_t> host_sources;
    uint32_t hull_colliders = 0;
    uint32_t convex_colliders = 0;  // colliders some pair may send to the convex path
    // The whole step captured once as a HIP graph and replayed: every launch of a step has
    // its grid from the capacities and its arguments from the world, so only dt (and the
    // shape pointers, replaced by an upload) can change what a step launches.
    bool graphs = true;
    // The step captured once per parity: the contact and archive buffers trade places every step,
    // so even and odd steps launch the same kernels over swapped pointers.
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    float graph_dt[2] = {0.0f, 0.0f};
    uint32_t parity = 0;
    // body edits queued since the last step (dyn_world_edit): row and edit in queue order, and the
    // device buffers they are uploaded to (capacities in edits and runs)
    std::vector<uint32_t> host_row_of_body;
    std::vector<std::pair<uint32_t, dyn_body_edit>> pending_edits;
    std::vector<dyn_body_edit> edit_upload;
    std::vector<uint4> run_upload;
    dyn_body_edit* edit_buf = nullptr;
    uint4* run_buf = nullptr;
    uint32_t edit_cap = 0, run_cap = 0;
    // Capacity planning (crates/dynamis-world/src/backend/mod.rs:194-230 apply_plan;
    // crates/dynamis-broadphase/src/capacity.rs, crates/dynamis-rigid/src/capacity.rs): after
    // every step its demand (the uncapped pair, contact and resting totals) is copied to pinned
    // host memory; before step s the demand of step s - kPlanLag is read (waiting for that step
    // if it has not ended), and a demand above a stream's capacity grows the stream to
    // max(demand, 2 x capacity) (dynamis-domain capacity.rs grown). A fixed lag keeps the
    // trajectory independent of host timing. The per-step streams (pairs, contacts and their
    // solver and compaction arrays) live in one group that growth reallocates; the durable ones
    // (archive, resting archive and its index) are copied over.
    std::vector<void*> streams;       // the group's allocations
    uint32_t* snap_header = nullptr;  // device: snapshot magic, contact capacity, rows, colliders
    uint32_t* plan_host = nullptr;    // pinned: kPlanRing demand records of kPlanWords words
    hipEvent_t plan_events[4] = {nullptr, nullptr, nullptr, nullptr};
    SideStream side{nullptr, nullptr, nullptr};  // the narrowphase's second stream (default priority)
    uint64_t issued = 0;              // dyn_world_step calls since create
    uint32_t growth_locks = 0;        // holders (differentiable rollouts) whose buffers fix the capacity
    uint32_t grows = 0;
    bool tuning = false;              // steps of dyn_world_tune_solver: no demand is recorded
    float last_dt = 0.0f;             // the dt of the last step
    uint32_t solver_threads = dyn::kThreads;  // workgroup size of the solver substeps' kernels
    // the tape a recorded step writes (dyn::world_record); null when not recording
    float* tape_states = nullptr;
    float* tape_contacts = nullptr;
    uint32_t* tape_count = nullptr;
    uint32_t* tape_error = nullptr;
};

namespace {

template <typename T>
int alloc_in(std::vector<void*>& owned, T** ptr, size_t count) {
    void* p = nullptr;
    DYN_HIP(hipMalloc(&p, (count ? count : 1) * sizeof(T)));
    owned.push_back(p);
    DYN_HIP(hipMemset(p, 0, (count ? count : 1) * sizeof(T)));
    *ptr = static_cast<T*>(p);
    return DYN_OK;
}

template <typename T>
int alloc(dyn_world* world, T** ptr, size_t count) {
    return alloc_in(world->owned, ptr, count);
}

constexpr uint32_t kPlanLag = 2;    // steps between a step and the plan that reads its demand
constexpr uint32_t kPlanRing = 4;   // demand records in flight (> kPlanLag)
// The counter words a demand record copies: slots PAIRS (10) through REFUSED_RESTING (20).
constexpr uint32_t kPlanFirst = DYN_COUNTER_PAIRS;
constexpr uint32_t kPlanWords = (DYN_COUNTER_REFUSED_RESTING - kPlanFirst) * DYN_COUNTER_STRIDE_WORDS + 1u;
static_assert(DYN_COUNTER_CONTACTS > kPlanFirst && DYN_COUNTER_RESTING > kPlanFirst &&
                  DYN_COUNTER_REFUSED_RESTING > DYN_COUNTER_RESTING,
              "the demand record spans the pair, contact and resting counters");

// The per-step streams sized by the pair and contact capacities, allocated as one group (growth
// frees and reallocates the whole group). Durable contents are the caller's to copy.
int alloc_streams(dyn_world* world, uint32_t pair_capacity, uint32_t contact_capacity) {
    using namespace dyn;
    W& w = world->w;
    std::vector<void*>& g = world->streams;
    const uint32_t body_count = w.n;
    int rc = DYN_OK;
#define A(ptr, count) if (rc == DYN_OK) rc = alloc_in(g, &(ptr), (count))
    A(w.pmaj, pair_capacity);
    A(w.pmin, pair_capacity);
    A(w.raw, pair_capacity);
    A(w.valid, pair_capacity);
    A(w.pkind, pair_capacity);
    A(w.worklist, pair_capacity);
    A(w.class_hist, kClasses * (static_cast<size_t>(ceil_div(pair_capacity ? pair_capacity : 1u, kClassBlock)) + 1u));
    A(w.work_minor, pair_capacity);
    A(w.dest, pair_capacity);
    A(w.contacts, contact_capacity);
    A(w.ckeys, contact_capacity);
    A(w.matched, contact_capacity);
    A(w.archive, contact_capacity);
    A(w.resting, contact_capacity);
    A(w.rest_tmp, contact_capacity);
    A(w.rest_major, contact_capacity);
    A(w.rest_minor, contact_capacity);
    A(w.rest_slot, contact_capacity);
    A(w.rest_flags, contact_capacity);
    A(w.blocks, 2u * static_cast<size_t>(contact_capacity));
    A(w.targets, 4u * static_cast<size_t>(contact_capacity));
    A(w.prows, 3u * static_cast<size_t>(contact_capacity));
    A(w.srows, static_cast<size_t>(kSolveRow) * contact_capacity);
    A(w.simp, 3u * static_cast<size_t>(contact_capacity));
    A(w.sanchor, 8u * static_cast<size_t>(contact_capacity));
    A(w.slots, 2u * static_cast<size_t>(contact_capacity));
    A(w.adj, 2u * static_cast<size_t>(contact_capacity));
    A(w.vwords, 12u * static_cast<size_t>(contact_capacity));
    A(w.pwords, 2u * static_cast<size_t>(contact_capacity));
    A(w.iedges, contact_capacity);
    uint32_t scan_limit = pair_capacity > contact_capacity ? pair_capacity : contact_capacity;
    if (body_count > scan_limit) scan_limit = body_count;  // the rows' list offsets
    A(w.scan_blocks, ceil_div(scan_limit ? scan_limit : 1u, kScanBlock) + 1u);
    A(world->freeze_flags, contact_capacity);
    A(world->keep_dest, contact_capacity);
    A(world->freeze_dest, contact_capacity);
#undef A
    if (rc != DYN_OK) return rc;
    w.pair_cap = pair_capacity;
    w.contact_cap = contact_capacity;
    world->bp_scratch_bytes = dyn_broadphase_scratch_bytes(w.k, world->entry_cap, pair_capacity);
    world->sort_scratch_bytes = dyn_sort_scratch_bytes(contact_capacity, 2u * dyn_sort_key_words(w.id_cap ? w.id_cap : 1u));
    for (auto [ptr, bytes] : {std::pair<void**, size_t>{&world->bp_scratch, world->bp_scratch_bytes},
                              {&world->sort_scratch, world->sort_scratch_bytes}}) {
        *ptr = nullptr;
        DYN_HIP(hipMalloc(ptr, bytes ? bytes : 1u));
        g.push_back(*ptr);
    }
    return DYN_OK;
}

void free_streams(std::vector<void*>& owned) {
    for (void* p : owned) (void)hipFree(p);
    owned.clear();
}

void drop_graph(dyn_world* world);

constexpr uint32_t kSnapMagicWord = 0x44594E53u;  // "DYNS", the first word of a snapshot

int write_snap_header(dyn_world* world) {
    const uint32_t header[4] = {kSnapMagicWord, world->w.contact_cap, world->w.n, world->w.k};
    DYN_HIP(hipMemcpy(world->snap_header, header, sizeof header, hipMemcpyHostToDevice));
    return DYN_OK;
}

// Grows the stream group to the given capacities (never below the current ones): a new group,
// the durable streams copied into it, the old group freed, the captured graphs dropped (they hold
// the old pointers and launch shapes).
int grow_streams(dyn_world* world, uint32_t pair_capacity, uint32_t contact_capacity) {
    using namespace dyn;
    W& w = world->w;
    pair_capacity = std::max(pair_capacity, w.pair_cap);
    contact_capacity = std::max(contact_capacity, w.contact_cap);
    if (pair_capacity == w.pair_cap && contact_capacity == w.contact_cap) return DYN_OK;
    if (contact_capacity > (1u << 30) || pair_capacity > (1u << 30)) {
        return fail(DYN_E_INVALID, "capacity growth past 2^30 pairs or contacts");
    }
    DYN_HIP(hipStreamSynchronize(world->stream));
    const W old = w;
    std::vector<void*> old_group;
    old_group.swap(world->streams);
    void* old_scratch[2] = {world->bp_scratch, world->sort_scratch};
    (void)old_scratch;
    const int rc = alloc_streams(world, pair_capacity, contact_capacity);
    if (rc != DYN_OK) {
        free_streams(world->streams);
        world->streams.swap(old_group);  // the world keeps its old streams
        w = old;
        world->bp_scratch_bytes = dyn_broadphase_scratch_bytes(w.k, world->entry_cap, w.pair_cap);
        world->sort_scratch_bytes = dyn_sort_scratch_bytes(w.contact_cap, 2u * dyn_sort_key_words(w.id_cap ? w.id_cap : 1u));
        world->bp_scratch = old_scratch[0];
        world->sort_scratch = old_scratch[1];
        return rc;
    }
    const size_t cap = old.contact_cap;
    for (auto [dst, src, bytes] : {std::tuple<void*, const void*, size_t>{w.archive, old.archive, sizeof(dyn_contact) * cap},
                                   {w.resting, old.resting, sizeof(dyn_contact) * cap},
                                   {w.rest_major, old.rest_major, 4u * cap},
                                   {w.rest_minor, old.rest_minor, 4u * cap},
                                   {w.rest_slot, old.rest_slot, 4u * cap}}) {
        if (bytes) DYN_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, world->stream));
    }
    DYN_HIP(hipStreamSynchronize(world->stream));
    free_streams(old_group);
    drop_graph(world);
    ++world->grows;
    return write_snap_header(world);
}

uint32_t blocks_for(uint32_t n) { return dyn::ceil_div(n ? n : 1u, dyn::kThreads); }
// The grid of a row kernel that counts into a counter word (kCountBlocks).
uint32_t count_blocks(uint32_t n) { return std::min(blocks_for(n), dyn::kCountBlocks); }

int scan(const dyn::W& w, hipStream_t st, const uint32_t* flags, const uint32_t* count, uint32_t limit,
         uint32_t* dest, uint32_t* total, bool counts = false) {
    using namespace dyn;
    const uint32_t blocks = ceil_div(limit ? limit : 1u, kScanBlock);
    {
        ProfileScope scope("k_scan_count", st);
        if (counts) k_scan_count<true><<<blocks, kThreads, 0, st>>>(flags, count, limit, w.scan_blocks);
        else k_scan_count<false><<<blocks, kThreads, 0, st>>>(flags, count, limit, w.scan_blocks);
    }
    DYN_LAUNCHED("k_scan_count");
    {
        ProfileScope scope("k_scan_blocks", st);
        k_scan_blocks<<<1, 1024, 0, st>>>(w.scan_blocks, blocks, total);
    }
    DYN_LAUNCHED("k_scan_blocks");
    {
        ProfileScope scope("k_scan_dest", st);
        if (counts) k_scan_dest<true><<<blocks, kThreads, 0, st>>>(flags, count, limit, w.scan_blocks, dest);
        else k_scan_dest<false><<<blocks, kThreads, 0, st>>>(flags, count, limit, w.scan_blocks, dest);
    }
    DYN_LAUNCHED("k_scan_dest");
    return DYN_OK;
}

int scan(dyn_world* world, const uint32_t* flags, const uint32_t* count, uint32_t limit, uint32_t* dest,
         uint32_t* total, bool counts = false) {
    return scan(world->w, world->stream, flags, count, limit, dest, total, counts);
}

// The worklist of convex pairs of one mode, grouped by kind class: block histograms, their
// per-class scans, the scatter (total in work_total[mode]).
int worklist(const dyn::W& w, hipStream_t st, uint32_t mode) {
    using namespace dyn;
    const uint32_t blocks = ceil_div(w.pair_cap ? w.pair_cap : 1u, kClassBlock);
    {
        ProfileScope scope("k_class_count", st);
        k_class_count<<<blocks, kThreads, 0, st>>>(w, mode);
    }
    DYN_LAUNCHED("k_class_count");
    {
        ProfileScope scope("k_class_scan", st);
        k_class_scan<<<kClasses, 1024, 0, st>>>(w.class_hist, blocks);
    }
    DYN_LAUNCHED("k_class_scan");
    {
        ProfileScope scope("k_class_scatter", st);
        k_class_scatter<<<blocks, kThreads, 0, st>>>(w, mode);
    }
    DYN_LAUNCHED("k_class_scatter");
    return DYN_OK;
}

#define LAUNCH_ON(stream, name, grid, ...)                                 \
    do {                                                                   \
        {                                                                  \
            ::dyn::ProfileScope scope_(#name, (stream));                   \
            name<<<(grid), ::dyn::kThreads, 0, (stream)>>>(__VA_ARGS__);   \
        }                                                                  \
        DYN_LAUNCHED(#name);                                               \
    } while (0)
#define LAUNCH(name, grid, ...) LAUNCH_ON(world->stream, name, grid, __VA_ARGS__)

// The narrowphase of the candidate pairs on pmaj/pmin (narrowphase.wgsl:518-685): k_narrow
// writes the analytic manifolds and marks the convex-path pairs, which (when `convex`) GJK, EPA
// and the convex manifolds take up over class-grouped worklists. Afterwards valid[i] is
// kPairContact where raw[i] holds pair i's contact and kPairNone elsewhere, for every pair below
// the PAIRS counter (and the capacity).
// With `side` (a second stream and two events) the wave EPA kernel and the manifolds of its
// polytopes run on the second stream beside the worklist and manifold pass of the other pairs;
// the stream joins back before returning. Without it they run in order on `st`.
int narrowphase(const dyn::W& w, hipStream_t st, bool convex, const SideStream* side = nullptr) {
    using namespace dyn;
    const uint32_t pb = blocks_for(w.pair_cap);
    LAUNCH_ON(st, k_pair_class, pb, w);
    DYN_TRY(worklist(w, st, kWorkAllPairs));
    LAUNCH_ON(st, k_narrow, pb, w);
    if (!convex) return DYN_OK;
    DYN_TRY(worklist(w, st, 0u));
    const uint32_t class_blocks = ceil_div(w.pair_cap ? w.pair_cap : 1u, kClassBlock);
    LAUNCH_ON(st, k_narrow_convex, pb, w, class_blocks);
    DYN_TRY(worklist(w, st, 4u));
    LAUNCH_ON(st, k_narrow_convex_resume, pb, w);
    DYN_TRY(worklist(w, st, 2u));
    {
        constexpr uint32_t kPairsPerBlock = kEpaWaves * kWave / kEpaGroup;
        const uint32_t small_groups = min(ceil_div(w.pair_cap ? w.pair_cap : 1u, kPairsPerBlock), 2048u);
        ProfileScope scope("k_narrow_epa_small", st);
        k_narrow_epa_small<<<small_groups, kEpaWaves * kWave, 0, st>>>(w, g_epa_small_vertices);
    }
    DYN_LAUNCHED("k_narrow_epa_small");
    // the polytopes given a wave are few (the long tail of the expansion count)
    const uint32_t tasks = ceil_div(w.pair_cap ? w.pair_cap : 1u, static_cast<uint32_t>(kWave));
    const uint32_t groups = min(ceil_div(tasks, kEpaWaves), 256u);
    hipStream_t tail = st;
    if (side) {
        DYN_HIP(hipEventRecord(side->fork, st));
        DYN_HIP(hipStreamWaitEvent(side->stream, side->fork, 0));
        tail = side->stream;
    }
    {
        ProfileScope scope("k_narrow_epa", tail);
        k_narrow_epa<<<groups, kEpaWaves * kWave, 0, tail>>>(w);
    }
    DYN_LAUNCHED("k_narrow_epa");
    LAUNCH_ON(tail, k_narrow_finish_tail, min(pb, 64u), w);
    if (side) DYN_HIP(hipEventRecord(side->join, side->stream));
    DYN_TRY(worklist(w, st, 1u));
    LAUNCH_ON(st, k_narrow_finish, pb, w);
    if (side) DYN_HIP(hipStreamWaitEvent(st, side->join, 0));
    return DYN_OK;
}

int world_step(dyn_world* world, float dt);

// Launched flatten rounds that flatten any forest of d rows: a round divides every depth by at
// least kFlattenHops + 1 = 64, and depths are below d, so the smallest r with 64^r >= d. (The
// reference sizes its one-hop rounds as ilog2(max(N_d, 2)) + 1, crates/dynamis-rigid/src/lib.rs:
// 79-81; both end with every row at its root.)
uint32_t flatten_rounds(uint32_t d) {
    uint32_t rounds = 1;
    for (uint64_t reach = dyn::kFlattenHops + 1u; reach < d; reach *= dyn::kFlattenHops + 1u) ++rounds;
    return rounds;
}

// A flatten pass (`pass` picks its round flags): launched round by round, or all rounds in one
// workgroup for a small world; either way it ends once a round's walks all reach their roots.
// With `rounds` 0 it takes flatten_rounds(w.d) rounds and leaves every row at its root.
int island_flatten(const dyn::W& w, hipStream_t st, uint32_t pass, uint32_t rounds = 0) {
    using namespace dyn;
    if (rounds == 0u) rounds = flatten_rounds(w.d);
    if (w.d <= kFlattenGroupRows) {
        {
            ProfileScope scope("k_island_flatten_rounds", st);
            k_island_flatten_rounds<<<1, 1024, 0, st>>>(w, rounds);
        }
        DYN_LAUNCHED("k_island_flatten_rounds");
        return DYN_OK;
    }
    const uint32_t db = blocks_for(w.d);
    uint32_t* flags = w.flat_flags + pass * kFlattenFlags;
    for (uint32_t r = 0; r < rounds; ++r) LAUNCH_ON(st, k_island_flatten, db, w, flags, r);
    return DYN_OK;
}

// The contact part of the island pass, before the resting archive's links and the last flatten:
// the seeding pass (which also lists each contact's linking rows), then kHookRounds hooking
// rounds, a flatten round (walks of up to 64 hops) before each and after the last, then the
// retrying hooks over every contact. The rounds only shorten the hooks' work: neither needs a
// flat forest, and k_island_link alone would end at the same components.
int island_components(const dyn::W& w, hipStream_t st, float slop) {
    using namespace dyn;
    const uint32_t cb = blocks_for(w.contact_cap);
    LAUNCH_ON(st, k_island_seed, cb, w, slop);
    DYN_TRY(island_flatten(w, st, 0u, 1u));
    for (uint32_t h = 0; h < kHookRounds; ++h) {
        LAUNCH_ON(st, k_island_hook, cb, w, slop);
        DYN_TRY(island_flatten(w, st, 1u + h, 1u));
    }
    LAUNCH_ON(st, k_island_link, cb, w, slop);
    return DYN_OK;
}

// SolverReset (solver.rs:89-101) and the block set-up of SolveSubsteps (solver.rs:310): block
// counts, the contact blocks, the rows' block lists.
int solver_prepare(const dyn::W& w, hipStream_t st) {
    using namespace dyn;
    const uint32_t nb = blocks_for(w.n), cb = blocks_for(w.contact_cap);
    LAUNCH_ON(st, k_solver_reset, count_blocks(w.n), w);
    LAUNCH_ON(st, k_blocks, cb, w);
    LAUNCH_ON(st, k_rows_count, count_blocks(w.n), w);
    DYN_TRY(scan(w, st, w.list_counts, nullptr, w.n, w.row_start, &w.scan_total[3], true));
    LAUNCH_ON(st, k_fill_lists, cb, w);
    return DYN_OK;
}

#define LAUNCH_TH(stream, name, items, threads, ...)                                      \
    do {                                                                                  \
        {                                                                                 \
            ::dyn::ProfileScope scope_(#name, (stream));                                  \
            name<<<::dyn::ceil_div((items) ? (items) : 1u, (threads)), (threads), 0, (stream)>>>(__VA_ARGS__); \
        }                                                                                 \
        DYN_LAUNCHED(#name);                                                              \
    } while (0)

// The substeps of SolveSubsteps (solver.rs:311-332), then the accumulated impulses back into the
// contacts. The per-contact and per-row kernels of the substeps index their element by the global
// thread index alone, so they run in workgroups of `threads` (64, 128 or 256; the solver's
// autotuned shape) with the same results.
int solver_substeps(const dyn::W& w, hipStream_t st, uint32_t threads = dyn::kThreads) {
    using namespace dyn;
    const uint32_t nb = blocks_for(w.n);
    // lanes a row in the apply kernels: four for a world of up to 2^18 rows (too few rows to keep
    // the GPU's memory pipes busy otherwise; eight were slower), two above (one was slower)
    const uint32_t lanes_log2 = w.n <= (1u << 18) ? 2u : 1u;
    for (uint32_t sub = 0; sub < w.p.substeps; ++sub) {
        LAUNCH_TH(st, k_integrate, w.d, threads, w);
        LAUNCH_TH(st, k_solver_bodies, w.n, threads, w);
        LAUNCH_TH(st, k_solve_prep, w.contact_cap, threads, w, sub == 0 ? 1u : 0u);
        // the substep's last apply writes the velocities back to the body states
        if (sub == 0) {
            LAUNCH_TH(st, k_warm, w.contact_cap, threads, w);
            LAUNCH_TH(st, k_apply, w.n << lanes_log2, threads, w, w.p.solve_iterations == 0u ? 1u : 0u, lanes_log2);
        }
        for (uint32_t it = 0; it < w.p.solve_iterations; ++it) {
            // solver_block_solve.wgsl:85-87: the residual round is the one where the round
            // counter (reset by integrate, bumped by every apply) equals the iteration count,
            // which only the warm-started first substep reaches.
            LAUNCH_TH(st, k_solve, w.contact_cap, threads, w, (sub == 0 && it + 1u == w.p.solve_iterations) ? 1u : 0u);
            LAUNCH_TH(st, k_apply, w.n << lanes_log2, threads, w, it + 1u == w.p.solve_iterations ? 1u : 0u,
                      lanes_log2);
        }
        LAUNCH_TH(st, k_advance, w.d, threads, w);
        for (uint32_t it = 0; it < w.p.position_iterations; ++it) {
            LAUNCH_TH(st, k_position_block, w.contact_cap, threads, w);
            LAUNCH_TH(st, k_position_apply, w.n << lanes_log2, threads, w, it == 0u ? 1u : 0u,
                      it + 1u == w.p.position_iterations ? 1u : 0u, lanes_log2);
        }
    }
    LAUNCH_TH(st, k_impulses_store, w.contact_cap, threads, w);
    (void)nb;
    return DYN_OK;
}

void drop_graph(dyn_world* world) {
    for (int p = 0; p < 2; ++p) {
        if (world->graph[p]) (void)hipGraphExecDestroy(world->graph[p]);
        world->graph[p] = nullptr;
    }
}

// The step's contacts become the next step's archive (contact_archive.wgsl) and the old archive
// the buffer the next step compacts into.
void swap_contact_buffers(dyn_world* world) {
    std::swap(world->w.contacts, world->w.archive);
    world->parity ^= 1u;
}

// Records one step on the world's stream into the graph of the current parity and instantiates
// it for dt (the recording swaps the buffers, as a launched step does).
int capture_step(dyn_world* world, float dt) {
    const uint32_t p = world->parity;
    if (world->graph[p]) (void)hipGraphExecDestroy(world->graph[p]);
    world->graph[p] = nullptr;
    DYN_HIP(hipStreamBeginCapture(world->stream, hipStreamCaptureModeThreadLocal));
    const int rc = world_step(world, dt);
    hipGraph_t graph = nullptr;
    const hipError_t ended = hipStreamEndCapture(world->stream, &graph);
    if (rc != DYN_OK || ended != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc != DYN_OK ? rc : dyn::hip_fail(ended, "dyn_world_step: graph capture");
    }
    const hipError_t made = hipGraphInstantiate(&world->graph[p], graph, nullptr, nullptr, 0);
    world->graph_dt[p] = dt;
    (void)hipGraphDestroy(graph);
    if (made != hipSuccess) {
        world->graph[p] = nullptr;
        return dyn::hip_fail(made, "dyn_world_step: graph instantiation");
    }
    return DYN_OK;
}

// After a step: its demand record (the counter words from PAIRS through REFUSED_RESTING) copied
// to the ring of pinned records, and the event that marks it complete.
int record_demand(dyn_world* world) {
    const uint32_t slot = static_cast<uint32_t>(world->issued % kPlanRing);
    DYN_HIP(hipMemcpyAsync(world->plan_host + static_cast<size_t>(slot) * kPlanWords,
                           world->w.counters + kPlanFirst * DYN_COUNTER_STRIDE_WORDS, 4u * kPlanWords,
                           hipMemcpyDeviceToHost, world->stream));
    DYN_HIP(hipEventRecord(world->plan_events[slot], world->stream));
    ++world->issued;
    return DYN_OK;
}

// Before a step: the demand of the step kPlanLag steps back, and growth of the streams it
// outgrew (crates/dynamis-domain/src/capacity.rs grown: max(required, 2 x capacity)). The pair
// demand is the uncapped pair total; the contact demand the larger of the uncapped contact and
// resting totals (the resting archive shares the contact capacity, rigid/src/capacity.rs:106).
int plan_capacity(dyn_world* world) {
    if (world->issued < kPlanLag) return DYN_OK;
    const uint32_t slot = static_cast<uint32_t>((world->issued - kPlanLag) % kPlanRing);
    DYN_HIP(hipEventSynchronize(world->plan_events[slot]));
    const uint32_t* rec = world->plan_host + static_cast<size_t>(slot) * kPlanWords;
    auto word = [&](uint32_t counter) { return rec[(counter - kPlanFirst) * DYN_COUNTER_STRIDE_WORDS]; };
    const uint64_t pairs = word(DYN_COUNTER_PAIRS);
    const uint64_t contacts = std::max<uint64_t>(word(DYN_COUNTER_CONTACTS),
                                                 static_cast<uint64_t>(word(DYN_COUNTER_RESTING)) +
                                                     word(DYN_COUNTER_REFUSED_RESTING));
    const dyn::W& w = world->w;
    auto grown = [](uint64_t capacity, uint64_t required) -> uint64_t {
        return required <= capacity ? capacity : std::max(required, 2u * capacity);
    };
    const uint64_t np = std::min<uint64_t>(grown(w.pair_cap, pairs), 1u << 30);
    const uint64_t nc = std::min<uint64_t>(grown(w.contact_cap, contacts), 1u << 30);
    if ((np == w.pair_cap && nc == w.contact_cap) || world->growth_locks) return DYN_OK;
    return grow_streams(world, static_cast<uint32_t>(np), static_cast<uint32_t>(nc));
}

int world_step(dyn_world* world, float dt) {
    using namespace dyn;
    W& w = world->w;
    w.p.dt = dt;
    w.p.substep_dt = dt / static_cast<float>(w.p.substeps);
    const uint32_t nb = blocks_for(w.n), db = blocks_for(w.d), kb = blocks_for(w.k);
    const uint32_t pb = blocks_for(w.pair_cap), cb = blocks_for(w.contact_cap);
    hipStream_t st = world->stream;
    {
        ProfileScope scope("k_reset_counters", st);
        k_reset_counters<<<1, 64, 0, st>>>(w);
    }
    DYN_LAUNCHED("k_reset_counters");
    LAUNCH(k_begin, nb, w);
    LAUNCH(k_aabb, kb, w);
    dyn_broadphase_io io{};
    io.d_aabbs = w.aabbs;
    io.d_groups = w.owner;
    io.d_flags = w.flags;
    io.d_resolution = w.res;
    io.colliders = w.k;
    io.d_entry_keys = world->entry_keys;
    io.d_entry_order = world->entry_order;
    io.d_entries = world->entries;
    io.entry_capacity = world->entry_cap;
    io.d_pair_major = w.pmaj;
    io.d_pair_minor = w.pmin;
    io.pair_capacity = w.pair_cap;
    io.d_counters = w.counters;
    io.sort_pairs = 1;
    io.d_scratch = world->bp_scratch;
    io.scratch_bytes = world->bp_scratch_bytes;
    io.d_worlds = w.worlds;
    DYN_TRY(dyn_broadphase(st, &io));
    DYN_TRY(narrowphase(w, st, world->convex_colliders, &world->side));
    const uint32_t* pair_count = w.counters + DYN_COUNTER_PAIRS * DYN_COUNTER_STRIDE_WORDS;
    DYN_TRY(scan(world, w.valid, pair_count, w.pair_cap, w.dest, &w.scan_total[0]));
    LAUNCH(k_contacts_scatter, pb, w);
    {
        ProfileScope scope("k_contacts_total", st);
        k_contacts_total<<<1, 64, 0, st>>>(w);
    }
    DYN_LAUNCHED("k_contacts_total");
    if (world->tape_states) LAUNCH(k_tape_sources, blocks_for(4u * w.contact_cap), w);
    LAUNCH(k_relay, cb, w);
    LAUNCH(k_contact_begin, cb, w, w.p.slop);
    LAUNCH(k_island_init, db, w);
    DYN_TRY(island_components(w, st, w.p.slop));
    LAUNCH(k_island_link_resting, cb, w);
    DYN_TRY(island_flatten(w, st, kFlattenPasses - 1u));
    LAUNCH(k_island_aggregate, count_blocks(w.d), w);
    LAUNCH(k_island_wake, count_blocks(w.d), w);
    DYN_TRY(solver_prepare(w, st));
    if (world->tape_states) {
        LAUNCH(k_tape_record, blocks_for(std::max(w.n, w.contact_cap)), w, world->tape_states, world->tape_contacts,
               world->tape_count, world->tape_error);
    }
    DYN_TRY(solver_substeps(w, st, world->solver_threads));
    LAUNCH(k_sleep, count_blocks(w.d), w);
    LAUNCH(k_thaw, cb, w);
    LAUNCH(k_archive, cb, w, world->freeze_flags);
    DYN_TRY(scan(world, w.rest_flags, w.nresting, w.contact_cap, world->keep_dest, &w.scan_total[2]));
    DYN_TRY(scan(world, world->freeze_flags, w.ncontacts, w.contact_cap, world->freeze_dest, &w.scan_total[1]));
    LAUNCH(k_archive_count, nb, w);
    LAUNCH(k_rest_gather, cb, w, world->keep_dest, &w.scan_total[2], world->freeze_dest, &world->words[0]);
    // Sort the new resting list by (min id, max id); payload = its slot.
    const uint32_t words = dyn_sort_key_words(w.id_cap ? w.id_cap : 1u);
    DYN_TRY(dyn_sort_pairs(st, w.rest_major, w.rest_minor, w.rest_slot, w.contact_cap, &world->words[0],
                           nullptr, 1u, words, words, world->sort_scratch, world->sort_scratch_bytes));
    LAUNCH(k_rest_place, cb, w, &world->words[0]);
    {
        ProfileScope scope("k_consume", st);
        k_consume<<<1, 64, 0, st>>>(w);
    }
    DYN_LAUNCHED("k_consume");
    swap_contact_buffers(world);
    return DYN_OK;
}

}  // namespace

extern "C" {

int dyn_world_create(dyn_world** out, void* stream, const dyn_step_params* params,
                     const dyn_body_state* states, const dyn_body_desc* descs, uint32_t body_count,
                     uint32_t dynamic_count, const dyn_collider* colliders, const uint32_t* owners,
                     uint32_t collider_count, uint32_t contact_capacity, uint32_t pair_capacity) {
    using namespace dyn;
    if (!out || !params || (body_count && (!states || !descs)) ||
        (collider_count && (!colliders || !owners))) {
        return fail(DYN_E_INVALID, "dyn_world_create: null argument");
    }
    if (dynamic_count > body_count) return fail(DYN_E_INVALID, "dyn_world_create: dynamic_count > body_count");
    if (contact_capacity > (1u << 30) || pair_capacity > (1u << 30)) {
        // list entries hold a contact index shifted by one bit
        return fail(DYN_E_INVALID, "dyn_world_create: capacities above 2^30");
    }
    if (params->substeps == 0 || params->solve_iterations == 0 || params->position_iterations == 0) {
        return fail(DYN_E_INVALID, "dyn_world_create: substeps and iterations must be positive");
    }
    for (uint32_t i = 0; i < collider_count; ++i) {
        if (owners[i] >= body_count) return fail(DYN_E_INVALID, "dyn_world_create: collider owner out of range");
    }
    // the id -> row table beside the rows (pool.rs rows): ids are unique
    uint32_t id_cap = 0;
    for (uint32_t i = 0; i < body_count; ++i) {
        if (states[i].body_id >= DYN_NO_BODY - 1u) return fail(DYN_E_INVALID, "dyn_world_create: body_id out of range");
        id_cap = states[i].body_id + 1u > id_cap ? states[i].body_id + 1u : id_cap;
    }
    std::vector<uint32_t> row_of_body(id_cap, DYN_NO_BODY);
    for (uint32_t i = 0; i < body_count; ++i) {
        if (row_of_body[states[i].body_id] != DYN_NO_BODY) return fail(DYN_E_INVALID, "dyn_world_create: duplicate body_id");
        row_of_body[states[i].body_id] = i;
    }
    *out = nullptr;
    dyn_world* world = new dyn_world();
    auto cleanup = [&](int rc) {
        for (void* p : world->owned) (void)hipFree(p);
        free_streams(world->streams);
        if (world->plan_host) (void)hipHostFree(world->plan_host);
        for (hipEvent_t& e : world->plan_events) {
            if (e) (void)hipEventDestroy(e);
        }
        if (world->side.fork) (void)hipEventDestroy(world->side.fork);
        if (world->side.join) (void)hipEventDestroy(world->side.join);
        if (world->side.stream) (void)hipStreamDestroy(world->side.stream);
        if (world->own_stream && world->stream) (void)hipStreamDestroy(world->stream);
        delete world;
        return rc;
    };
    if (stream) {
        world->stream = static_cast<hipStream_t>(stream);
    } else {
        if (hipStreamCreateWithFlags(&world->stream, hipStreamNonBlocking) != hipSuccess) {
            return cleanup(fail(DYN_E_HIP, "dyn_world_create: stream"));
        }
        world->own_stream = true;
    }
    world->collider_count = collider_count;
    world->host_kinds.resize(collider_count);
    world->host_sources.resize(collider_count);
    for (uint32_t i = 0; i < collider_count; ++i) {
        world->host_kinds[i] = colliders[i].kind;
        world->host_sources[i] = colliders[i].source;
        if (colliders[i].kind == DYN_SHAPE_HULL) ++world->hull_colliders;
        const uint32_t k = colliders[i].kind;
        const bool unit_scale = colliders[i].scale[0] == 1.0f && colliders[i].scale[1] == 1.0f &&
                                colliders[i].scale[2] == 1.0f;
        if (k == DYN_SHAPE_HULL || k == DYN_SHAPE_CYLINDER || k == DYN_SHAPE_CAPSULE || !unit_scale) {
            ++world->convex_colliders;
        }
    }
    W& w = world->w;
    w.id_cap = id_cap;
    w.p = *params;
    w.n = body_count;
    w.d = dynamic_count;
    w.k = collider_count;
    w.pair_cap = pair_capacity;
    w.contact_cap = contact_capacity;
    world->entry_cap = DYN_MAX_CELLS_PER_COLLIDER * (collider_count ? collider_count : 1u);
    int rc = DYN_OK;
#define A(ptr, count) if (rc == DYN_OK) rc = alloc(world, &(ptr), (count))
    A(w.st, body_count);
    A(w.de, body_count);
    A(w.co, collider_count);
    A(w.owner, collider_count);
    A(w.row_of_body, id_cap);
    A(w.wake, body_count);
    A(w.activity, body_count);
    A(w.motion, body_count);
    A(w.counters, DYN_COUNTER_BLOCK_WORDS);
    A(w.aabbs, collider_count);
    A(w.res, collider_count);
    A(w.flags, collider_count);
    A(w.parents, body_count);
    A(w.istate, body_count);
    A(w.flat_flags, kFlattenPasses * kFlattenFlags);
    A(w.block_counts, body_count);
    A(w.sbodies, static_cast<size_t>(kSolverBody) * body_count);
    A(w.sposes, 2u * static_cast<size_t>(body_count));
    A(w.bbodies, 3u * static_cast<size_t>(body_count));
    A(w.ppose, 2u * static_cast<size_t>(body_count));
    A(w.resolution, 4u * static_cast<size_t>(body_count));
    A(w.list_counts, body_count);
    A(w.row_start, body_count);
    A(w.contrib, body_count);
    A(w.scan_total, 4);
    A(w.work_total, 8);
    A(world->snap_header, 4);
    A(world->words, 8);
    A(world->entry_keys, world->entry_cap);
    A(world->entry_order, world->entry_cap);
    A(world->entries, world->entry_cap);
#undef A
    if (rc != DYN_OK) return cleanup(rc);
    w.ncontacts = world->words + 1;
    w.narchive = world->words + 2;
    w.nresting = world->words + 3;
    w.edit_runs_count = world->words + 5;
    world->host_row_of_body = row_of_body;
    world->edit_cap = 64u;
    world->run_cap = 64u;
    if (hipMalloc(&world->edit_buf, sizeof(dyn_body_edit) * world->edit_cap) != hipSuccess ||
        hipMalloc(&world->run_buf, sizeof(uint4) * world->run_cap) != hipSuccess) {
        if (world->edit_buf) (void)hipFree(world->edit_buf);
        return cleanup(fail(DYN_E_HIP, "dyn_world_create: edit buffers"));
    }
    w.edits = world->edit_buf;
    w.edit_runs = world->run_buf;
    if (alloc_streams(world, pair_capacity, contact_capacity) != DYN_OK) {
        return cleanup(fail(DYN_E_HIP, "dyn_world_create: pair and contact streams"));
    }
    if (write_snap_header(world) != DYN_OK) return cleanup(fail(DYN_E_HIP, "dyn_world_create: snapshot header"));
    if (hipHostMalloc(reinterpret_cast<void**>(&world->plan_host), 4u * kPlanWords * kPlanRing) != hipSuccess) {
        world->plan_host = nullptr;
        return cleanup(fail(DYN_E_HIP, "dyn_world_create: demand records"));
    }
    for (uint32_t i = 0; i < kPlanRing; ++i) {
        if (hipEventCreateWithFlags(&world->plan_events[i], hipEventDisableTiming) != hipSuccess) {
            world->plan_events[i] = nullptr;
            return cleanup(fail(DYN_E_HIP, "dyn_world_create: demand events"));
        }
    }
    if (hipStreamCreateWithFlags(&world->side.stream, hipStreamNonBlocking) != hipSuccess) {
        world->side.stream = nullptr;
        return cleanup(fail(DYN_E_HIP, "dyn_world_create: second stream"));
    }
    if (hipEventCreateWithFlags(&world->side.fork, hipEventDisableTiming) != hipSuccess) {
        world->side.fork = nullptr;
        return cleanup(fail(DYN_E_HIP, "dyn_world_create: stream events"));
    }
    if (hipEventCreateWithFlags(&world->side.join, hipEventDisableTiming) != hipSuccess) {
        world->side.join = nullptr;
        return cleanup(fail(DYN_E_HIP, "dyn_world_create: stream events"));
    }
    auto up = [&](void* dst, const void* src, size_t bytes) {
        if (bytes && hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess) rc = DYN_E_HIP;
    };
    up(w.st, states, sizeof(dyn_body_state) * body_count);
    up(w.de, descs, sizeof(dyn_body_desc) * body_count);
    up(w.co, colliders, sizeof(dyn_collider) * collider_count);
    up(w.row_of_body, row_of_body.data(), sizeof(uint32_t) * id_cap);
    up(w.owner, owners, 4u * collider_count);
    if (rc != DYN_OK) return cleanup(fail(DYN_E_HIP, "dyn_world_create: upload"));
    *out = world;
    return DYN_OK;
}

int dyn_batch_create(dyn_world** out, void* stream, const dyn_step_params* params,
                     const dyn_body_state* states, const dyn_body_desc* descs, uint32_t body_count,
                     uint32_t dynamic_count, const dyn_collider* colliders, const uint32_t* owners,
                     const uint32_t* collider_worlds, uint32_t collider_count, uint32_t world_count,
                     uint32_t contact_capacity, uint32_t pair_capacity) {
    using namespace dyn;
    if (!out || (collider_count && !collider_worlds)) return fail(DYN_E_INVALID, "dyn_batch_create: null argument");
    if (collider_count && !owners) return fail(DYN_E_INVALID, "dyn_batch_create: null argument");
    // every collider names a world of the batch, and all colliders of one body the same one
    std::vector<uint32_t> world_of_row(body_count, DYN_NO_BODY);
    for (uint32_t i = 0; i < collider_count; ++i) {
        if (collider_worlds[i] >= world_count) return fail(DYN_E_INVALID, "dyn_batch_create: world id out of range");
        if (owners[i] >= body_count) return fail(DYN_E_INVALID, "dyn_batch_create: collider owner out of range");
        uint32_t& seen = world_of_row[owners[i]];
        if (seen != DYN_NO_BODY && seen != collider_worlds[i]) {
            return fail(DYN_E_INVALID, "dyn_batch_create: a body's colliders lie in different worlds");
        }
        seen = collider_worlds[i];
    }
    dyn_world* world = nullptr;
    DYN_TRY(dyn_world_create(&world, stream, params, states, descs, body_count, dynamic_count, colliders,
                             owners, collider_count, contact_capacity, pair_capacity));
    uint32_t* worlds = nullptr;
    if (alloc(world, &worlds, collider_count) != DYN_OK ||
        (collider_count &&
         hipMemcpy(worlds, collider_worlds, 4u * collider_count, hipMemcpyHostToDevice) != hipSuccess)) {
        dyn_world_destroy(world);
        return fail(DYN_E_HIP, "dyn_batch_create: world ids");
    }
    world->w.worlds = worlds;
    *out = world;
    return DYN_OK;
}

int dyn_world_upload_shapes(dyn_world* world, const dyn_shape_source* sources,
                            uint32_t source_count, const float* vertices, uint32_t vertex_count) {
    using namespace dyn;
    if (!world || (source_count && !sources) || (vertex_count && !vertices)) {
        return fail(DYN_E_INVALID, "dyn_world_upload_shapes: null argument");
    }
    for (uint32_t i = 0; i < source_count; ++i) {
        if (sources[i].vertex_count > DYN_LINEAR_SUPPORT_VERTICES) {
            return fail(DYN_E_INVALID, "dyn_world_upload_shapes: a hull holds more vertices than the linear support scan");
        }
        if (sources[i].vertex_count == 0 ||
            static_cast<uint64_t>(sources[i].vertex_offset) + sources[i].vertex_count > vertex_count) {
            return fail(DYN_E_INVALID, "dyn_world_upload_shapes: a source reaches past the vertex array");
        }
    }
    for (uint32_t i = 0; i < world->collider_count; ++i) {
        if (world->host_kinds[i] == DYN_SHAPE_HULL && world->host_sources[i] >= source_count) {
            return fail(DYN_E_INVALID, "dyn_world_upload_shapes: a hull collider names a missing source");
        }
    }
    DYN_HIP(hipStreamSynchronize(world->stream));
    drop_graph(world);  // the captured step holds the old shape pointers
    if (world->shape_sources) DYN_HIP(hipFree(world->shape_sources));
    if (world->shape_vertices) DYN_HIP(hipFree(world->shape_vertices));
    world->shape_sources = nullptr;
    world->shape_vertices = nullptr;
    const size_t sb = sizeof(dyn_shape_source) * (source_count ? source_count : 1u);
    const size_t vb = 16u * static_cast<size_t>(vertex_count ? vertex_count : 1u);
    DYN_HIP(hipMalloc(&world->shape_sources, sb));
    DYN_HIP(hipMalloc(&world->shape_vertices, vb));
    if (source_count) DYN_HIP(hipMemcpy(world->shape_sources, sources, sizeof(dyn_shape_source) * source_count, hipMemcpyHostToDevice));
    if (vertex_count) DYN_HIP(hipMemcpy(world->shape_vertices, vertices, 16u * static_cast<size_t>(vertex_count), hipMemcpyHostToDevice));
    world->w.sources = static_cast<const dyn_shape_source*>(world->shape_sources);
    world->w.vertices = static_cast<const float4*>(world->shape_vertices);
    world->w.source_count = source_count;
    // the hull sources' support cells (support_cells.h), built on the host and uploaded
    if (world->support_cells) DYN_HIP(hipFree(world->support_cells));
    if (world->support_cands) DYN_HIP(hipFree(world->support_cands));
    world->support_cells = world->support_cands = nullptr;
    world->w.support_cells = nullptr;
    world->w.support_cands = nullptr;
    if (source_count) {
        std::vector<uint32_t> cells(2u * static_cast<size_t>(source_count) * kSupportCells);
        uint32_t total = 0;
        DYN_TRY(dyn_hull_support_cells(sources, source_count, vertices, vertex_count, cells.data(), nullptr, 0u, &total));
        std::vector<uint16_t> cands(total ? total : 1u);
        DYN_TRY(dyn_hull_support_cells(sources, source_count, vertices, vertex_count, cells.data(), cands.data(), total,
                                       &total));
        DYN_HIP(hipMalloc(&world->support_cells, 4u * cells.size()));
        DYN_HIP(hipMalloc(&world->support_cands, 2u * cands.size()));
        DYN_HIP(hipMemcpy(world->support_cells, cells.data(), 4u * cells.size(), hipMemcpyHostToDevice));
        DYN_HIP(hipMemcpy(world->support_cands, cands.data(), 2u * cands.size(), hipMemcpyHostToDevice));
        world->w.support_cells = static_cast<const uint2*>(world->support_cells);
        world->w.support_cands = static_cast<const uint16_t*>(world->support_cands);
    }
    return DYN_OK;
}

int dyn_world_destroy(dyn_world* world) {
    if (!world) return DYN_OK;
    (void)hipStreamSynchronize(world->stream);
    drop_graph(world);
    for (void* p : world->owned) (void)hipFree(p);
    free_streams(world->streams);
    if (world->plan_host) (void)hipHostFree(world->plan_host);
    for (hipEvent_t& e : world->plan_events) {
        if (e) (void)hipEventDestroy(e);
    }
    if (world->shape_sources) (void)hipFree(world->shape_sources);
    if (world->shape_vertices) (void)hipFree(world->shape_vertices);
    if (world->support_cells) (void)hipFree(world->support_cells);
    if (world->support_cands) (void)hipFree(world->support_cands);
    if (world->edit_buf) (void)hipFree(world->edit_buf);
    if (world->run_buf) (void)hipFree(world->run_buf);
    if (world->side.fork) (void)hipEventDestroy(world->side.fork);
    if (world->side.join) (void)hipEventDestroy(world->side.join);
    if (world->side.stream) (void)hipStreamDestroy(world->side.stream);
    if (world->own_stream) (void)hipStreamDestroy(world->stream);
    delete world;
    return DYN_OK;
}

int dyn_world_edit(dyn_world* world, const uint32_t* body_ids, const dyn_body_edit* edits, uint32_t count) {
    using namespace dyn;
    if (!world) return fail(DYN_E_INVALID, "dyn_world_edit: null world");
    if (count == 0u) return DYN_OK;
    if (!body_ids || !edits) return fail(DYN_E_INVALID, "dyn_world_edit: null argument");
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t id = body_ids[i];
        if (id >= world->host_row_of_body.size() || world->host_row_of_body[id] == DYN_NO_BODY) {
            return fail(DYN_E_INVALID, "dyn_world_edit: body id names no body");
        }
        if (edits[i].kind > DYN_EDIT_ANGULAR_IMPULSE) return fail(DYN_E_INVALID, "dyn_world_edit: unknown edit kind");
    }
    for (uint32_t i = 0; i < count; ++i) {
        world->pending_edits.emplace_back(world->host_row_of_body[body_ids[i]], edits[i]);
    }
    // runs in row order, each row's edits in queue order (command.rs compiles one run per body)
    std::vector<std::pair<uint32_t, dyn_body_edit>> ordered = world->pending_edits;
    std::stable_sort(ordered.begin(), ordered.end(),
                     [](const std::pair<uint32_t, dyn_body_edit>& a, const std::pair<uint32_t, dyn_body_edit>& b) {
                         return a.first < b.first;
                     });
    world->edit_upload.clear();
    world->run_upload.clear();
    for (uint32_t i = 0; i < ordered.size(); ++i) {
        if (world->run_upload.empty() || world->run_upload.back().x != ordered[i].first) {
            world->run_upload.push_back(make_uint4(ordered[i].first, i, 0u, 0u));
        }
        world->run_upload.back().z += 1u;
        world->edit_upload.push_back(ordered[i].second);
    }
    const uint32_t edits_n = static_cast<uint32_t>(world->edit_upload.size());
    const uint32_t runs_n = static_cast<uint32_t>(world->run_upload.size());
    if (edits_n > world->edit_cap || runs_n > world->run_cap) {
        // the captured step holds the old buffers' addresses
        DYN_HIP(hipStreamSynchronize(world->stream));
        drop_graph(world);
        while (world->edit_cap < edits_n) world->edit_cap *= 2u;
        while (world->run_cap < runs_n) world->run_cap *= 2u;
        DYN_HIP(hipFree(world->edit_buf));
        DYN_HIP(hipFree(world->run_buf));
        world->edit_buf = nullptr;
        world->run_buf = nullptr;
        DYN_HIP(hipMalloc(&world->edit_buf, sizeof(dyn_body_edit) * world->edit_cap));
        DYN_HIP(hipMalloc(&world->run_buf, sizeof(uint4) * world->run_cap));
        world->w.edits = world->edit_buf;
        world->w.edit_runs = world->run_buf;
    }
    hipStream_t st = world->stream;
    DYN_HIP(hipMemcpyAsync(world->edit_buf, world->edit_upload.data(), sizeof(dyn_body_edit) * edits_n,
                           hipMemcpyHostToDevice, st));
    DYN_HIP(hipMemcpyAsync(world->run_buf, world->run_upload.data(), sizeof(uint4) * runs_n, hipMemcpyHostToDevice, st));
    DYN_HIP(hipMemcpyAsync(world->w.edit_runs_count, &runs_n, 4u, hipMemcpyHostToDevice, st));
    // the host copies above are read when the copies run; keep them until the step is enqueued
    DYN_HIP(hipStreamSynchronize(st));
    return DYN_OK;
}

int dyn_world_step(dyn_world* world, float dt) {
    if (!world) return dyn::fail(DYN_E_INVALID, "dyn_world_step: null world");
    world->pending_edits.clear();  // the step applies (and consumes) every queued edit
    if (!(dt > 0.0f)) return dyn::fail(DYN_E_INVALID, "dyn_world_step: dt must be positive");
    if (world->hull_colliders && world->shape_sources == nullptr) {
        return dyn::fail(DYN_E_INVALID, "dyn_world_step: hull colliders need dyn_world_upload_shapes first");
    }
    if (!world->tuning) DYN_TRY(plan_capacity(world));
    world->last_dt = dt;
    // profiled and recorded steps run eagerly
    if (!world->graphs || dyn::profiling() || world->tape_states) {
        DYN_TRY(world_step(world, dt));
    } else {
        const uint32_t p = world->parity;
        if (!world->graph[p] || world->graph_dt[p] != dt) {
            DYN_TRY(capture_step(world, dt));  // recording the step swapped the buffers
        } else {
            swap_contact_buffers(world);
        }
        DYN_HIP(hipGraphLaunch(world->graph[p], world->stream));
    }
    return world->tuning ? DYN_OK : record_demand(world);
}

int dyn_world_read_work(dyn_world* world, uint32_t* out) {
    if (!world || !out) return dyn::fail(DYN_E_INVALID, "dyn_world_read_work: null argument");
    DYN_HIP(hipStreamSynchronize(world->stream));
    DYN_HIP(hipMemcpy(out, world->w.work_total, 6u * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return DYN_OK;
}

int dyn_world_set_graphs(dyn_world* world, int enabled) {
    if (!world) return dyn::fail(DYN_E_INVALID, "dyn_world_set_graphs: null world");
    DYN_HIP(hipStreamSynchronize(world->stream));
    world->graphs = enabled != 0;
    drop_graph(world);
    return DYN_OK;
}

int dyn_world_capacity(dyn_world* world, uint32_t* pair_capacity, uint32_t* contact_capacity, uint32_t* grows) {
    if (!world) return dyn::fail(DYN_E_INVALID, "dyn_world_capacity: null world");
    if (pair_capacity) *pair_capacity = world->w.pair_cap;
    if (contact_capacity) *contact_capacity = world->w.contact_cap;
    if (grows) *grows = world->grows;
    return DYN_OK;
}

int dyn_world_reserve(dyn_world* world, uint32_t pair_capacity, uint32_t contact_capacity) {
    if (!world) return dyn::fail(DYN_E_INVALID, "dyn_world_reserve: null world");
    if (world->growth_locks && (pair_capacity > world->w.pair_cap || contact_capacity > world->w.contact_cap)) {
        return dyn::fail(DYN_E_INVALID, "dyn_world_reserve: a differentiable rollout holds the world's capacities");
    }
    return grow_streams(world, pair_capacity, contact_capacity);
}

int dyn_world_set_solver_threads(dyn_world* world, uint32_t threads) {
    if (!world) return dyn::fail(DYN_E_INVALID, "dyn_world_set_solver_threads: null world");
    if (threads != 64u && threads != 128u && threads != 256u) {
        return dyn::fail(DYN_E_INVALID, "dyn_world_set_solver_threads: threads must be 64, 128 or 256");
    }
    DYN_HIP(hipStreamSynchronize(world->stream));
    world->solver_threads = threads;
    drop_graph(world);
    return DYN_OK;
}

int dyn_world_tune_solver(dyn_world* world, uint32_t steps, uint32_t* threads_out, float* ms_out) {
    using namespace dyn;
    if (!world) return fail(DYN_E_INVALID, "dyn_world_tune_solver: null world");
    if (!world->pending_edits.empty()) return fail(DYN_E_INVALID, "dyn_world_tune_solver: body edits are queued");
    if (world->tape_states) return fail(DYN_E_INVALID, "dyn_world_tune_solver: a tape is being recorded");
    if (world->hull_colliders && world->shape_sources == nullptr) {
        return fail(DYN_E_INVALID, "dyn_world_tune_solver: hull colliders need dyn_world_upload_shapes first");
    }
    if (steps == 0u) steps = 3u;
    const float dt = world->last_dt > 0.0f ? world->last_dt : 1.0f / 60.0f;
    size_t bytes = 0;
    DYN_TRY(dyn_world_snapshot_bytes(world, &bytes));
    DYN_HIP(hipStreamSynchronize(world->stream));
    void* snap = nullptr;
    DYN_HIP(hipMalloc(&snap, bytes));
    hipEvent_t start = nullptr, stop = nullptr;
    const uint32_t parity = world->parity, kept_threads = world->solver_threads;
    // back to the snapshot: the contact and archive buffers in their order at the snapshot, then
    // the durable state copied in
    auto reset = [&]() {
        if (world->parity != parity) swap_contact_buffers(world);
        return dyn_world_restore(world, snap);
    };
    auto finish = [&](int rc) {
        (void)hipStreamSynchronize(world->stream);
        if (rc == DYN_OK) rc = reset();  // the world as it was
        (void)hipStreamSynchronize(world->stream);
        world->tuning = false;
        drop_graph(world);
        if (start) (void)hipEventDestroy(start);
        if (stop) (void)hipEventDestroy(stop);
        (void)hipFree(snap);
        return rc;
    };
    world->tuning = true;
    int rc = dyn_world_snapshot(world, snap);
    if (rc == DYN_OK && (hipEventCreate(&start) != hipSuccess || hipEventCreate(&stop) != hipSuccess)) {
        rc = fail(DYN_E_HIP, "dyn_world_tune_solver: events");
    }
    const uint32_t options[3] = {64u, 128u, 256u};
    float best_ms = 0.0f;
    uint32_t best = kept_threads;
    for (uint32_t i = 0; i < 3u && rc == DYN_OK; ++i) {
        world->solver_threads = options[i];
        drop_graph(world);
        if ((rc = reset()) != DYN_OK) break;
        if ((rc = dyn_world_step(world, dt)) != DYN_OK) break;  // captures the step
        if (hipEventRecord(start, world->stream) != hipSuccess) rc = fail(DYN_E_HIP, "dyn_world_tune_solver: event");
        for (uint32_t k = 0; k < steps && rc == DYN_OK; ++k) rc = dyn_world_step(world, dt);
        if (rc != DYN_OK) break;
        float ms = 0.0f;
        if (hipEventRecord(stop, world->stream) != hipSuccess || hipEventSynchronize(stop) != hipSuccess ||
            hipEventElapsedTime(&ms, start, stop) != hipSuccess) {
            rc = fail(DYN_E_HIP, "dyn_world_tune_solver: timing");
            break;
        }
        if (ms_out) ms_out[i] = ms / static_cast<float>(steps);
        if (i == 0u || ms < best_ms) {
            best_ms = ms;
            best = options[i];
        }
    }
    rc = finish(rc);
    world->solver_threads = rc == DYN_OK ? best : kept_threads;
    if (threads_out) *threads_out = world->solver_threads;
    return rc;
}

int dyn_world_wait(dyn_world* world) {
    if (!world) return dyn::fail(DYN_E_INVALID, "dyn_world_wait: null world");
    DYN_HIP(hipStreamSynchronize(world->stream));
    return DYN_OK;
}

void* dyn_world_stream(dyn_world* world) { return world ? world->stream : nullptr; }

int dyn_world_read_states(dyn_world* world, dyn_body_state* out) {
    if (!world || !out) return dyn::fail(DYN_E_INVALID, "dyn_world_read_states: null argument");
    DYN_HIP(hipStreamSynchronize(world->stream));
    if (world->w.n) DYN_HIP(hipMemcpy(out, world->w.st, sizeof(dyn_body_state) * world->w.n, hipMemcpyDeviceToHost));
    return DYN_OK;
}

int dyn_world_read_counters(dyn_world* world, uint32_t* out) {
    if (!world || !out) return dyn::fail(DYN_E_INVALID, "dyn_world_read_counters: null argument");
    DYN_HIP(hipStreamSynchronize(world->stream));
    DYN_HIP(hipMemcpy(out, world->w.counters, 4u * DYN_COUNTER_BLOCK_WORDS, hipMemcpyDeviceToHost));
    return DYN_OK;
}

int dyn_world_read_contacts(dyn_world* world, dyn_contact* out, uint32_t capacity, uint32_t* count) {
    if (!world || !count) return dyn::fail(DYN_E_INVALID, "dyn_world_read_contacts: null argument");
    DYN_HIP(hipStreamSynchronize(world->stream));
    uint32_t n = 0;
    DYN_HIP(hipMemcpy(&n, world->w.ncontacts, 4u, hipMemcpyDeviceToHost));
    *count = n;
    const uint32_t copy = n < capacity ? n : capacity;
    // the last step's contacts: since its end they are the archive
    if (copy && out) DYN_HIP(hipMemcpy(out, world->w.archive, sizeof(dyn_contact) * copy, hipMemcpyDeviceToHost));
    return DYN_OK;
}

// The resting archive (frozen manifolds of sleeping pairs), sorted by (min body id, max body id):
// the second half of the reference's contact readback (crates/dynamis-world/src/readback.rs:95-139).
int dyn_world_read_resting(dyn_world* world, dyn_contact* out, uint32_t capacity, uint32_t* count) {
    if (!world || !count) return dyn::fail(DYN_E_INVALID, "dyn_world_read_resting: null argument");
    DYN_HIP(hipStreamSynchronize(world->stream));
    uint32_t n = 0;
    DYN_HIP(hipMemcpy(&n, world->w.nresting, 4u, hipMemcpyDeviceToHost));
    *count = n;
    const uint32_t copy = n < capacity ? n : capacity;
    if (copy && out) DYN_HIP(hipMemcpy(out, world->w.resting, sizeof(dyn_contact) * copy, hipMemcpyDeviceToHost));
    return DYN_OK;
}

namespace {

// dyn_narrowphase's scratch: the kind classes and the worklists.
struct NarrowLayout {
    uint8_t* pkind;
    uint32_t *worklist, *class_hist, *work_total, *work_minor;
    size_t bytes;
};

NarrowLayout narrow_layout(uint32_t pair_capacity, void* base) {
    using namespace dyn;
    NarrowLayout l{};
    size_t at = 0;
    auto take = [&](size_t bytes) {
        const size_t here = at;
        at = (at + bytes + 255u) & ~static_cast<size_t>(255u);
        return static_cast<char*>(base) + here;
    };
    const size_t cap = pair_capacity ? pair_capacity : 1u;
    l.pkind = reinterpret_cast<uint8_t*>(take(cap));
    l.worklist = reinterpret_cast<uint32_t*>(take(4u * cap));
    l.class_hist = reinterpret_cast<uint32_t*>(take(4u * kClasses * (static_cast<size_t>(ceil_div(
                                                                          static_cast<uint32_t>(cap), kClassBlock)) + 1u)));
    l.work_total = reinterpret_cast<uint32_t*>(take(4u * 8u));
    l.work_minor = reinterpret_cast<uint32_t*>(take(4u * cap));
    l.bytes = at;
    return l;
}

}  // namespace

size_t dyn_narrowphase_scratch_bytes(uint32_t pair_capacity) {
    return narrow_layout(pair_capacity, nullptr).bytes;
}

int dyn_set_epa_small_vertices(uint32_t vertices) {
    if (vertices < 4u || vertices > dyn::kEpaSmallVertices) {
        return dyn::fail(DYN_E_INVALID, "dyn_set_epa_small_vertices: outside 4..16");
    }
    dyn::g_epa_small_vertices = vertices;
    return DYN_OK;
}

int dyn_narrowphase(void* stream, const dyn_narrowphase_io* io) {
    using namespace dyn;
    if (!io || !io->params) return fail(DYN_E_INVALID, "dyn_narrowphase: null io or params");
    if (io->pair_capacity > kSortMaxCapacity) return fail(DYN_E_INVALID, "dyn_narrowphase: capacity above 2^30-1");
    if (!io->d_counters || !io->d_scratch || (io->pair_capacity && (!io->d_pair_major || !io->d_pair_minor ||
                                                                    !io->d_contacts || !io->d_valid))) {
        return fail(DYN_E_INVALID, "dyn_narrowphase: null buffer");
    }
    if ((io->bodies && (!io->d_states || !io->d_descs)) || (io->colliders && (!io->d_colliders || !io->d_owners))) {
        return fail(DYN_E_INVALID, "dyn_narrowphase: null records");
    }
    if (io->source_count && (!io->d_sources || !io->d_vertices)) {
        return fail(DYN_E_INVALID, "dyn_narrowphase: null shape sources");
    }
    const NarrowLayout l = narrow_layout(io->pair_capacity, io->d_scratch);
    if (io->scratch_bytes < l.bytes) {
        return fail(DYN_E_CAPACITY, "dyn_narrowphase: scratch smaller than dyn_narrowphase_scratch_bytes");
    }
    if (io->pair_capacity == 0u) return DYN_OK;
    W w{};
    w.p = *io->params;
    w.n = io->bodies;
    w.k = io->colliders;
    // the kernels take the records through the world's non-const view and only read them
    w.st = const_cast<dyn_body_state*>(io->d_states);
    w.de = const_cast<dyn_body_desc*>(io->d_descs);
    w.co = const_cast<dyn_collider*>(io->d_colliders);
    w.owner = const_cast<uint32_t*>(io->d_owners);
    w.sources = io->d_sources;
    w.vertices = reinterpret_cast<const float4*>(io->d_vertices);
    w.support_cells = reinterpret_cast<const uint2*>(io->d_support_cells);
    w.support_cands = io->d_support_candidates;
    w.source_count = io->source_count;
    w.counters = const_cast<uint32_t*>(io->d_counters);
    w.pmaj = const_cast<uint32_t*>(io->d_pair_major);
    w.pmin = const_cast<uint32_t*>(io->d_pair_minor);
    w.pair_cap = io->pair_capacity;
    w.raw = io->d_contacts;
    w.valid = io->d_valid;
    w.pkind = l.pkind;
    w.worklist = l.worklist;
    w.class_hist = l.class_hist;
    w.work_total = l.work_total;
    w.work_minor = l.work_minor;
    return narrowphase(w, static_cast<hipStream_t>(stream), true);
}

namespace {

// dyn_solve's scratch: the solver's per-block and per-row buffers.
struct SolveLayout {
    dyn::W parts;  // only the scratch pointers are set
    size_t bytes;
};

SolveLayout solve_layout(uint32_t bodies, uint32_t contact_capacity, void* base) {
    using namespace dyn;
    SolveLayout l{};
    size_t at = 0;
    auto take = [&](size_t bytes) {
        const size_t here = at;
        at = (at + (bytes ? bytes : 1u) + 255u) & ~static_cast<size_t>(255u);
        return static_cast<char*>(base) + here;
    };
    const size_t n = bodies ? bodies : 1u, c = contact_capacity ? contact_capacity : 1u;
    W& w = l.parts;
    w.block_counts = reinterpret_cast<uint32_t*>(take(4u * n));
    w.blocks = reinterpret_cast<uint32_t*>(take(4u * 2u * c));
    w.targets = reinterpret_cast<float*>(take(4u * 4u * c));
    w.prows = reinterpret_cast<float4*>(take(sizeof(float4) * 3u * c));
    w.srows = reinterpret_cast<float4*>(take(sizeof(float4) * kSolveRow * c));
    w.sbodies = reinterpret_cast<float4*>(take(sizeof(float4) * kSolverBody * n));
    w.sposes = reinterpret_cast<float4*>(take(sizeof(float4) * 2u * n));
    w.bbodies = reinterpret_cast<float4*>(take(sizeof(float4) * 3u * n));
    w.ppose = reinterpret_cast<float4*>(take(sizeof(float4) * 2u * n));
    w.simp = reinterpret_cast<float4*>(take(sizeof(float4) * 3u * c));
    w.sanchor = reinterpret_cast<float4*>(take(sizeof(float4) * 8u * c));
    w.resolution = reinterpret_cast<float*>(take(4u * 4u * n));
    w.slots = reinterpret_cast<uint32_t*>(take(4u * 2u * c));
    w.list_counts = reinterpret_cast<uint32_t*>(take(4u * n));
    w.row_start = reinterpret_cast<uint32_t*>(take(4u * n));
    w.adj = reinterpret_cast<uint32_t*>(take(4u * 2u * c));
    w.vwords = reinterpret_cast<uint32_t*>(take(4u * 12u * c));
    w.pwords = reinterpret_cast<uint4*>(take(sizeof(uint4) * 2u * c));
    w.contrib = reinterpret_cast<uint32_t*>(take(4u * n));
    const size_t limit = n > c ? n : c;
    w.scan_blocks = reinterpret_cast<uint32_t*>(take(4u * (ceil_div(static_cast<uint32_t>(limit), kScanBlock) + 1u)));
    w.scan_total = reinterpret_cast<uint32_t*>(take(4u * 4u));
    w.ncontacts = reinterpret_cast<uint32_t*>(take(4u * 2u));  // the clamped count, the check word
    l.bytes = at;
    return l;
}

}  // namespace

size_t dyn_solve_scratch_bytes(uint32_t bodies, uint32_t contact_capacity) {
    return solve_layout(bodies, contact_capacity, nullptr).bytes;
}

int dyn_solve(void* stream, const dyn_solve_io* io, float dt) {
    using namespace dyn;
    if (!io || !io->params) return fail(DYN_E_INVALID, "dyn_solve: null io or params");
    if (!(dt > 0.0f)) return fail(DYN_E_INVALID, "dyn_solve: dt must be positive");
    if (io->params->substeps == 0u) return fail(DYN_E_INVALID, "dyn_solve: substeps must be at least 1");
    if (io->dynamic_count > io->bodies) return fail(DYN_E_INVALID, "dyn_solve: dynamic_count above bodies");
    if (!io->d_counters || !io->d_scratch || !io->d_contact_count || (io->bodies && (!io->d_states || !io->d_descs)) ||
        (io->contact_capacity && !io->d_contacts) || (io->colliders && !io->d_owners)) {
        return fail(DYN_E_INVALID, "dyn_solve: null buffer");
    }
    SolveLayout l = solve_layout(io->bodies, io->contact_capacity, io->d_scratch);
    if (io->scratch_bytes < l.bytes) {
        return fail(DYN_E_CAPACITY, "dyn_solve: scratch smaller than dyn_solve_scratch_bytes");
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    W w = l.parts;
    w.p = *io->params;
    w.p.dt = dt;
    w.p.substep_dt = dt / static_cast<float>(w.p.substeps);
    w.n = io->bodies;
    w.d = io->dynamic_count;
    w.k = io->colliders;
    w.st = io->d_states;
    w.de = const_cast<dyn_body_desc*>(io->d_descs);
    w.owner = const_cast<uint32_t*>(io->d_owners);
    w.contacts = io->d_contacts;
    w.contact_cap = io->contact_capacity;
    w.counters = io->d_counters;
    uint32_t* words = w.ncontacts;  // scratch: [0] the count the kernels read, [1] the check
    LAUNCH_ON(st, k_solve_stage_reset, blocks_for(w.n), w, io->d_contact_count, words);
    if (w.contact_cap) LAUNCH_ON(st, k_solve_stage_check, blocks_for(w.contact_cap), w, words);
    uint32_t bad = 0;
    DYN_HIP(hipMemcpyAsync(&bad, words + 1, 4u, hipMemcpyDeviceToHost, st));
    DYN_HIP(hipStreamSynchronize(st));
    if (bad) return fail(DYN_E_INVALID, "dyn_solve: a contact names a collider or body outside the records");
    DYN_TRY(solver_prepare(w, st));
    return solver_substeps(w, st);
}

size_t dyn_islands_scratch_bytes(uint32_t bodies) {
    const size_t words = bodies ? bodies : 1u;
    return 3u * ((4u * words + 255u) & ~static_cast<size_t>(255u)) + 4u * dyn::kFlattenPasses * dyn::kFlattenFlags;
}

int dyn_islands(void* stream, const dyn_islands_io* io) {
    using namespace dyn;
    if (!io) return fail(DYN_E_INVALID, "dyn_islands: null io");
    if (io->dynamic_count > io->bodies) return fail(DYN_E_INVALID, "dyn_islands: dynamic_count above bodies");
    if (!io->d_counters || !io->d_scratch || !io->d_contact_count || (io->dynamic_count && !io->d_parents) ||
        (io->contact_capacity && !io->d_contacts) || (io->colliders && !io->d_owners)) {
        return fail(DYN_E_INVALID, "dyn_islands: null buffer");
    }
    if (io->scratch_bytes < dyn_islands_scratch_bytes(io->bodies)) {
        return fail(DYN_E_CAPACITY, "dyn_islands: scratch smaller than dyn_islands_scratch_bytes");
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    const size_t part = (4u * static_cast<size_t>(io->bodies ? io->bodies : 1u) + 255u) & ~static_cast<size_t>(255u);
    char* base = static_cast<char*>(io->d_scratch);
    W w{};
    w.n = io->bodies;
    w.d = io->dynamic_count;
    w.k = io->colliders;
    w.owner = const_cast<uint32_t*>(io->d_owners);
    w.contacts = const_cast<dyn_contact*>(io->d_contacts);
    w.ncontacts = const_cast<uint32_t*>(io->d_contact_count);
    w.contact_cap = io->contact_capacity;
    w.parents = io->d_parents;
    w.counters = io->d_counters;
    w.wake = reinterpret_cast<uint32_t*>(base);
    w.motion = reinterpret_cast<uint32_t*>(base + part);
    w.istate = reinterpret_cast<uint32_t*>(base + 2u * part);
    w.flat_flags = reinterpret_cast<uint32_t*>(base + 3u * part);
    const uint32_t nb = blocks_for(w.n), db = blocks_for(w.d), cb = blocks_for(w.contact_cap);
    LAUNCH_ON(st, k_islands_stage_reset, nb, w);
    if (w.d == 0u) return DYN_OK;
    LAUNCH_ON(st, k_island_init, db, w);
    if (w.contact_cap) {
        DYN_TRY(island_components(w, st, io->slop));
    }
    DYN_TRY(island_flatten(w, st, kFlattenPasses - 1u));
    LAUNCH_ON(st, k_island_aggregate, count_blocks(w.d), w);
    return DYN_OK;
}

int dyn_world_read_pairs(dyn_world* world, uint32_t* major, uint32_t* minor, uint32_t capacity,
                         uint32_t* count) {
    if (!world || !count) return dyn::fail(DYN_E_INVALID, "dyn_world_read_pairs: null argument");
    DYN_HIP(hipStreamSynchronize(world->stream));
    uint32_t pairs = 0;
    DYN_HIP(hipMemcpy(&pairs, world->w.counters + DYN_COUNTER_PAIRS * DYN_COUNTER_STRIDE_WORDS, 4u,
                      hipMemcpyDeviceToHost));
    const uint32_t n = pairs < world->w.pair_cap ? pairs : world->w.pair_cap;
    *count = n;
    const uint32_t copy = n < capacity ? n : capacity;
    if (copy && major) DYN_HIP(hipMemcpy(major, world->w.pmaj, 4u * copy, hipMemcpyDeviceToHost));
    if (copy && minor) DYN_HIP(hipMemcpy(minor, world->w.pmin, 4u * copy, hipMemcpyDeviceToHost));
    return DYN_OK;
}

}  // extern "C"

namespace {

// What a step reads that an earlier step wrote (Retention::Durable streams,
// crates/dynamis-world/src/snapshot.rs:64-74): the body states, the contact archive (warm start),
// the resting archive and its sorted index, the static wake marks, the counters and the count
// words. Everything else is rebuilt by every step.
struct SnapPart {
    void* ptr;
    size_t bytes;
};

std::vector<SnapPart> snapshot_parts(dyn_world* world) {
    dyn::W& w = world->w;
    const size_t cap = w.contact_cap;
    // first the header: the capacities the streams below were sized by
    return {{world->snap_header, 4u * 4u},
            {w.st, sizeof(dyn_body_state) * w.n},
            {w.archive, sizeof(dyn_contact) * cap},
            {w.resting, sizeof(dyn_contact) * cap},
            {w.rest_major, 4u * cap},
            {w.rest_minor, 4u * cap},
            {w.rest_slot, 4u * cap},
            {w.wake, 4u * static_cast<size_t>(w.n)},
            {w.counters, 4u * DYN_COUNTER_BLOCK_WORDS},
            {world->words, 4u * 8u}};
}

size_t snap_align(size_t bytes) { return (bytes + 255u) & ~static_cast<size_t>(255u); }

}  // namespace

extern "C" {

int dyn_world_autotune(dyn_world* world, uint32_t* tilings) {
    using namespace dyn;
    if (!world) return fail(DYN_E_INVALID, "dyn_world_autotune: null world");
    const W& w = world->w;
    // the step's sorts: entry regions (4 key digits), candidate pairs and GJK worklists (pair
    // capacity), the resting index (contact capacity)
    const uint32_t caps[3] = {world->entry_cap, w.pair_cap, w.contact_cap};
    const uint32_t passes[3] = {4u, 2u * dyn_sort_key_words(w.k ? w.k : 1u), 2u * dyn_sort_key_words(w.id_cap ? w.id_cap : 1u)};
    DYN_HIP(hipStreamSynchronize(world->stream));
    for (int i = 0; i < 3; ++i) {
        uint32_t items = 0, groups = 0;
        DYN_TRY(dyn_sort_autotune(world->stream, caps[i] ? caps[i] : 1u, passes[i] < 8u ? passes[i] : 8u, &items, &groups));
        if (tilings) {
            tilings[3 * i] = caps[i];
            tilings[3 * i + 1] = items;
            tilings[3 * i + 2] = groups;
        }
    }
    drop_graph(world);  // the captured step holds the old launch shapes
    return DYN_OK;
}

int dyn_world_snapshot_bytes(dyn_world* world, size_t* bytes) {
    if (!world || !bytes) return dyn::fail(DYN_E_INVALID, "dyn_world_snapshot_bytes: null argument");
    size_t total = 0;
    for (const SnapPart& part : snapshot_parts(world)) total += snap_align(part.bytes);
    *bytes = total;
    return DYN_OK;
}

int dyn_world_snapshot(dyn_world* world, void* d_dst) {
    if (!world || !d_dst) return dyn::fail(DYN_E_INVALID, "dyn_world_snapshot: null argument");
    char* at = static_cast<char*>(d_dst);
    for (const SnapPart& part : snapshot_parts(world)) {
        DYN_HIP(hipMemcpyAsync(at, part.ptr, part.bytes, hipMemcpyDeviceToDevice, world->stream));
        at += snap_align(part.bytes);
    }
    return DYN_OK;
}

int dyn_world_restore(dyn_world* world, const void* d_src) {
    if (!world || !d_src) return dyn::fail(DYN_E_INVALID, "dyn_world_restore: null argument");
    uint32_t header[4] = {0, 0, 0, 0};
    DYN_HIP(hipMemcpyAsync(header, d_src, sizeof header, hipMemcpyDeviceToHost, world->stream));
    DYN_HIP(hipStreamSynchronize(world->stream));
    if (header[0] != kSnapMagicWord || header[1] != world->w.contact_cap || header[2] != world->w.n ||
        header[3] != world->w.k) {
        return dyn::fail(DYN_E_INVALID, "dyn_world_restore: the snapshot was taken at other capacities or of another world");
    }
    const char* at = static_cast<const char*>(d_src);
    for (const SnapPart& part : snapshot_parts(world)) {
        DYN_HIP(hipMemcpyAsync(part.ptr, at, part.bytes, hipMemcpyDeviceToDevice, world->stream));
        at += snap_align(part.bytes);
    }
    return DYN_OK;
}

}  // extern "C"

namespace dyn {

// The differentiable rollout's access to a world (adjoint.hip).
hipStream_t world_stream(dyn_world* world) { return world->stream; }

// A holder of buffers sized by the world's capacities (the differentiable rollout) keeps the
// world's streams from growing while it exists.
void world_lock_growth(dyn_world* world, int delta) {
    world->growth_locks = static_cast<uint32_t>(static_cast<int>(world->growth_locks) + delta);
}

void world_sizes(dyn_world* world, uint32_t* n, uint32_t* d, uint32_t* k, uint32_t* cap) {
    *n = world->w.n;
    *d = world->w.d;
    *k = world->w.k;
    *cap = world->w.contact_cap;
}

const dyn_step_params& world_params(dyn_world* world) { return world->w.p; }
const dyn_body_state* world_states(dyn_world* world) { return world->w.st; }
const dyn_body_desc* world_descs(dyn_world* world) { return world->w.de; }

// Starts (tape non-null) or ends recording: every following step writes its start states and
// contacts into the given device buffers (relay_src: 4 words per contact of capacity).
void world_record(dyn_world* world, float* states, float* contacts, uint32_t* count, uint32_t* error,
                  uint32_t* relay_src) {
    world->tape_states = states;
    world->tape_contacts = contacts;
    world->tape_count = count;
    world->tape_error = error;
    world->w.relay_src = states ? relay_src : nullptr;
}

}  // namespace dyn
