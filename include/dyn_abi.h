/*
 * dyn_abi.h - device record layouts, constants and counter slots of the rigid-body step.
 *
 * These are the byte layouts the reference declares once in WGSL and generates host records
 * from (crates/dynamis-abi/abi/records.wgsl:8-353, crates/dynamis-abi/build.rs:25-90). They are
 * kept byte-identical here so that a caller holding reference-format records can hand them to
 * this library unchanged: every struct below carries a static size check against the WGSL
 * storage-layout size (vec3f aligns to 16 bytes and occupies 12).
 *
 * Plain C11, no HIP or torch types: the header is shared by the HIP library, the CPU oracle and
 * ctypes callers.
 */
#ifndef DYN_ABI_H
#define DYN_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- record layouts (records.wgsl) -------------------------------------------------------- */

/* StepParams, records.wgsl:8-47 (uniform; 37 scalars after a vec4f, rounded to 16 bytes). */
typedef struct dyn_step_params {
    float gravity[4];
    float dt;
    float substep_dt;
    float damping;
    float angular_damping;
    uint32_t body_count;
    uint32_t substeps;
    uint32_t solve_iterations;
    uint32_t constraint_count;
    float relaxation;
    float slop;
    float restitution_threshold;
    float max_velocity;
    float max_angular_velocity;
    uint32_t dynamic_count;
    float sleep_velocity;
    float sleep_angular_velocity;
    float sleep_time;
    uint32_t friction_combine;
    uint32_t restitution_combine;
    uint32_t position_iterations;
    float contact_margin;
    uint32_t collider_count;
    uint32_t soft_iterations;
    uint32_t soft_substeps;
    uint32_t particle_count;
    uint32_t element_count;
    float soft_substep_dt;
    uint32_t soft_body_count;
    uint32_t attachment_count;
    float settle_velocity;
    uint32_t observed_count;
    uint32_t observed_joint_count;
    uint32_t character_count;
    uint32_t vehicle_count;
    uint32_t vehicle_wheel_count;
    uint32_t wake_all;
    uint32_t field_count;
    uint32_t _pad_tail[3];
} dyn_step_params;

/* BodyState, records.wgsl:145-164 (144 bytes). Quaternions are (x, y, z, w). */
typedef struct dyn_body_state {
    float position[3];
    float _pad0;
    float prev_position[3];
    float _pad1;
    float prev_orientation[4];
    float orientation[4];
    float velocity[3];
    float _pad2;
    float angular_velocity[3];
    float _pad3;
    float force[3];
    float _pad4;
    float torque[3];
    float _pad5;
    uint32_t body_id;
    uint32_t generation;
    float sleep_timer;
    uint32_t sleeping;
} dyn_body_state;

/* BodyDescriptor, records.wgsl:166-184 (128 bytes). Symmetric tensors are stored as
 * [xx, xy, xz, yy, yz, zz]. */
typedef struct dyn_body_desc {
    float inverse_mass;
    float linear_damping;
    float angular_damping;
    float gravity_scale;
    float sleep_velocity;
    float sleep_angular_velocity;
    uint32_t flags;
    uint32_t _pad0;
    uint32_t collision_group;
    uint32_t collision_mask;
    uint32_t _pad1;
    uint32_t _pad4;
    float com[3];
    float _pad2;
    float inertia[6];
    float inverse_inertia[6];
    float _pad3[4];
} dyn_body_desc;

/* Collider, records.wgsl:208-228 (112 bytes). */
typedef struct dyn_collider {
    uint32_t slot;
    uint32_t kind;
    uint32_t flags;
    float radius;
    float half_height;
    float relaxation;
    float damping_ratio;
    float impact_force;
    float half_extents[3];
    uint32_t collision_group;
    float local_offset[3];
    uint32_t collision_mask;
    float local_rotation[4];
    float friction;
    float restitution;
    uint32_t source;
    float rolling_friction;
    float scale[3];
    float spin_friction;
} dyn_collider;

/* ShapeSource, records.wgsl:256-273 (80 bytes): where a hull's vertices (float4 rows) live.
 * Hulls of at most DYN_LINEAR_SUPPORT_VERTICES vertices are supported (the reference's linear
 * support scan); triangles, BVH nodes and height-field cells are not read by this path. */
typedef struct dyn_shape_source {
    uint32_t kind;
    uint32_t vertex_offset;
    uint32_t vertex_count;
    uint32_t triangle_offset;
    uint32_t node_offset;
    uint32_t node_count;
    uint32_t cell_offset;
    uint32_t cell_count;
    uint32_t grid_rows;
    uint32_t grid_cols;
    uint32_t _pad0;
    uint32_t _pad1;
    float local_min[3];
    float _pad2;
    float local_max[3];
    float _pad3;
} dyn_shape_source;

/* Aabb, records.wgsl:299-304 (32 bytes). */
typedef struct dyn_aabb {
    float min[3];
    float _pad0;
    float max[3];
    float _pad1;
} dyn_aabb;

/* GridEntry, records.wgsl:306-311 (32 bytes). `info` packs the collider index, cell offset,
 * kind and the MOBILE / AWAKE / PRIMARY flags (constant.rs:73-81). */
typedef struct dyn_grid_entry {
    float min[3];
    uint32_t group;
    float max[3];
    uint32_t info;
} dyn_grid_entry;

/* ManifoldPoint, records.wgsl:313-322 (64 bytes). */
typedef struct dyn_manifold_point {
    float position[3];
    float depth;
    float local_a[3];
    float _pad_a;
    float local_b[3];
    float accumulated_normal;
    float accumulated_tangent_1;
    float accumulated_tangent_2;
    uint32_t feature;
    uint32_t _pad_tail;
} dyn_manifold_point;

#define DYN_CONTACT_MAX_POINTS 4u

/* Contact, records.wgsl:331-353 (352 bytes). */
typedef struct dyn_contact {
    uint32_t a;
    uint32_t b;
    uint32_t point_count;
    uint32_t sensor;
    uint32_t first_body_id;
    uint32_t second_body_id;
    uint32_t first_generation;
    uint32_t second_generation;
    float normal[3];
    uint32_t events;
    uint32_t surface;
    uint32_t triangle;
    float friction;
    float restitution;
    float rolling_friction;
    float spin_friction;
    float carried_normal;
    float carried_tangent;
    float relaxation;
    float damping_ratio;
    uint32_t _pad_head[2];
    dyn_manifold_point points[4];
} dyn_contact;

/* BodyEdit, records.wgsl:193-199 (160 bytes): one host command on one body, applied at the start of
 * the next step (body_edits.wgsl). `kind` is a DYN_EDIT_* value; a DYN_EDIT_PATCH writes the fields
 * of `state` its `mask` (DYN_PATCH_*) names; the force, torque and impulse kinds read
 * state.force / state.torque / state.velocity (linear impulse) / state.angular_velocity (angular
 * impulse) and, for the _AT_POINT kinds, state.position (the world point). */
typedef struct dyn_body_edit {
    uint32_t kind;
    uint32_t mask;
    uint32_t _pad0;
    uint32_t _pad1;
    dyn_body_state state;
} dyn_body_edit;

#if defined(__cplusplus)
static_assert(sizeof(dyn_body_edit) == 160, "BodyEdit is 160 bytes");
static_assert(sizeof(dyn_step_params) == 176, "StepParams is 176 bytes");
static_assert(sizeof(dyn_body_state) == 144, "BodyState is 144 bytes");
static_assert(sizeof(dyn_body_desc) == 128, "BodyDescriptor is 128 bytes");
static_assert(sizeof(dyn_collider) == 112, "Collider is 112 bytes");
static_assert(sizeof(dyn_aabb) == 32, "Aabb is 32 bytes");
static_assert(sizeof(dyn_shape_source) == 80, "ShapeSource is 80 bytes");
static_assert(sizeof(dyn_grid_entry) == 32, "GridEntry is 32 bytes");
static_assert(sizeof(dyn_manifold_point) == 64, "ManifoldPoint is 64 bytes");
static_assert(sizeof(dyn_contact) == 352, "Contact is 352 bytes");
#else
_Static_assert(sizeof(dyn_body_edit) == 160, "BodyEdit is 160 bytes");
_Static_assert(sizeof(dyn_step_params) == 176, "StepParams is 176 bytes");
_Static_assert(sizeof(dyn_body_state) == 144, "BodyState is 144 bytes");
_Static_assert(sizeof(dyn_body_desc) == 128, "BodyDescriptor is 128 bytes");
_Static_assert(sizeof(dyn_collider) == 112, "Collider is 112 bytes");
_Static_assert(sizeof(dyn_aabb) == 32, "Aabb is 32 bytes");
_Static_assert(sizeof(dyn_shape_source) == 80, "ShapeSource is 80 bytes");
_Static_assert(sizeof(dyn_grid_entry) == 32, "GridEntry is 32 bytes");
_Static_assert(sizeof(dyn_manifold_point) == 64, "ManifoldPoint is 64 bytes");
_Static_assert(sizeof(dyn_contact) == 352, "Contact is 352 bytes");
#endif

/* ---- constants (crates/dynamis-abi/src/constant.rs) ---------------------------------------- */

/* body edit kinds and patch fields (constant.rs:6-18) */
#define DYN_EDIT_PATCH 0u
#define DYN_EDIT_FORCE 1u
#define DYN_EDIT_TORQUE 2u
#define DYN_EDIT_IMPULSE 3u
#define DYN_EDIT_SLEEP 4u
#define DYN_EDIT_WAKE 5u
#define DYN_EDIT_FORCE_AT_POINT 6u
#define DYN_EDIT_IMPULSE_AT_POINT 7u
#define DYN_EDIT_ANGULAR_IMPULSE 8u
#define DYN_PATCH_POSITION 1u
#define DYN_PATCH_VELOCITY 2u
#define DYN_PATCH_ORIENTATION 4u
#define DYN_PATCH_ANGULAR_VELOCITY 8u

#define DYN_BODY_KINEMATIC 1u
#define DYN_BODY_CCD 2u
#define DYN_OVERRIDE_SLEEP_LINEAR 4u
#define DYN_OVERRIDE_SLEEP_ANGULAR 8u
#define DYN_NO_COLLISION_FILTER 0xFFFFFFFFu

#define DYN_SHAPE_NONE 0u
#define DYN_SHAPE_SPHERE 1u
#define DYN_SHAPE_CUBOID 2u
#define DYN_SHAPE_CAPSULE 3u
#define DYN_SHAPE_CYLINDER 4u
#define DYN_SHAPE_HULL 5u
#define DYN_SHAPE_PLANE 9u
#define DYN_LINEAR_SUPPORT_VERTICES 4096u
#define DYN_COLLIDER_SENSOR 1u

#define DYN_MAX_CELLS_PER_AXIS 2u
#define DYN_MAX_CELLS_PER_COLLIDER 8u
#define DYN_GRID_REGION_RESTING 0u
#define DYN_GRID_REGION_AWAKE 1u
#define DYN_GRID_REGION_IMMOVABLE 2u
#define DYN_GRID_REGION_COUNT 3u
#define DYN_ENTRY_INDEX_BITS 24u
#define DYN_ENTRY_INDEX_MASK ((1u << 24) - 1u)
#define DYN_ENTRY_CELL_SHIFT 24u
#define DYN_ENTRY_CELL_MASK (7u << 24)
#define DYN_ENTRY_KIND_SHIFT 27u
#define DYN_ENTRY_KIND_MASK (3u << 27)
#define DYN_ENTRY_MOBILE (1u << 29)
#define DYN_ENTRY_AWAKE (1u << 30)
#define DYN_ENTRY_PRIMARY (1u << 31)
#define DYN_ENTRY_KIND_COLLIDER 0u
#define DYN_LEVEL_KEY_SHIFT 27u
#define DYN_CELL_HASH_MASK 0x07FFFFFFu
#define DYN_GRID_SCALE_LIMIT 1e12f
#define DYN_GRID_LEVEL_LIMIT 30u

#define DYN_NO_BODY 0xFFFFFFFFu

/* Per-collider flags of the broadphase stage (the reference derives them from the owner's
 * descriptor and activity, collider_emit.wgsl:31-63). */
#define DYN_BP_MOBILE 1u /* the owning body moves (ENTRY_MOBILE) */
#define DYN_BP_AWAKE 2u  /* the owning body is awake (ENTRY_AWAKE) */

#define DYN_SOLVER_VELOCITY_SCALE 262144.0f
#define DYN_SOLVER_POSITION_SCALE 16777216.0f

/* ---- device counters (crates/dynamis-abi/src/counter.rs:75-128) ----------------------------
 * Each counter owns one 256-byte line of the counter block (COUNTER_STRIDE, counter.rs:241), so
 * word `slot * DYN_COUNTER_STRIDE_WORDS` holds counter `slot`. Slot numbers are the reference's. */
enum dyn_counter_slot {
    DYN_COUNTER_ENTRIES = 0,
    DYN_COUNTER_RESTING_ENTRIES,
    DYN_COUNTER_RESTING_LEVELS,
    DYN_COUNTER_RESTING_REBUILD,
    DYN_COUNTER_RESTING_SORT,
    DYN_COUNTER_AWAKE_BASE,
    DYN_COUNTER_IMMOVABLE_ENTRIES,
    DYN_COUNTER_IMMOVABLE_EMITTED,
    DYN_COUNTER_IMMOVABLE_LEVELS,
    DYN_COUNTER_ENTRY_BASE,
    DYN_COUNTER_PAIRS,
    DYN_COUNTER_GRID_LEVELS,
    DYN_COUNTER_CONTACTS,
    DYN_COUNTER_ARCHIVED,
    DYN_COUNTER_JOINTS,
    DYN_COUNTER_EVENTS,
    DYN_COUNTER_REFUSED_PAIRS,
    DYN_COUNTER_REFUSED_EVENTS,
    DYN_COUNTER_ENTRY_FAULTS,
    DYN_COUNTER_RESTING,
    DYN_COUNTER_REFUSED_RESTING,
    DYN_COUNTER_SLEPT,
    DYN_COUNTER_WOKE,
    DYN_COUNTER_ACTIVE,
    DYN_COUNTER_WOKE_DEFERRED,
    DYN_COUNTER_RESTING_GATHER,
    DYN_COUNTER_RESTING_INDEX,
    DYN_COUNTER_RESTING_PENDING,
    DYN_COUNTER_ISLANDS,
    DYN_COUNTER_ISLAND_FAULTS,
    DYN_COUNTER_BLOCKS,
    DYN_COUNTER_SOLVER_ROWS,
    DYN_COUNTER_SOLVE_LINEAR_RESIDUAL,
    DYN_COUNTER_SOLVE_ANGULAR_RESIDUAL,
    DYN_COUNTER_COARSE_ACTIVE,
    DYN_COUNTER_GRID_SCALE,
    DYN_COUNTER_GRID_EXTENT,
    DYN_COUNTER_BREAKS,
    DYN_COUNTER_PARTICLE_REACH,
    DYN_COUNTER_COARSE_NEIGHBOURS,
    DYN_COUNTER_LIVE,
    DYN_COUNTER_LIVE_FAULTS,
    DYN_COUNTER_ROW_FAULTS,
    DYN_COUNTER_SOFT_ACTIVE,
    DYN_COUNTER_SOFT_SLEPT,
    DYN_COUNTER_SOFT_WOKE,
    DYN_COUNTER_REFUSED_CONTACTS,
    DYN_COUNTER_STEP,
    DYN_COUNTER_IMPACTS,
    DYN_COUNTER_REFUSED_IMPACTS,
    DYN_COUNTER_SOFT_EVENTS,
    DYN_COUNTER_REFUSED_SOFT_EVENTS,
    DYN_COUNTER_DEVICE_COUNT
};

#define DYN_COUNTER_STRIDE_WORDS 64u
#define DYN_COUNTER_BLOCK_WORDS (DYN_COUNTER_DEVICE_COUNT * DYN_COUNTER_STRIDE_WORDS)

#ifdef __cplusplus
}
#endif

#endif /* DYN_ABI_H */
