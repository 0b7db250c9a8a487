/*
 * dynamis.h - C ABI of the MI355X-native rigid-body step library (libdynamis_amd.so).
 *
 * The reference runs every stage of `World::step` as WGSL programs dispatched through wgpu
 * (crates/dynamis-gpu/src/recorder.rs:29-39, `ComputeRecorder::record`). This library replaces
 * that dispatch layer and the WGSL programs with hand-written gfx950 HIP kernels behind the flat
 * entry points below. Conventions:
 *   - every entry point returns an int status: DYN_OK (0) or a negative DYN_E_* code, with a
 *     message from dyn_last_error(); nothing throws or aborts across the ABI;
 *   - pointers named d_* or documented as "device" are HIP device pointers; `stream` is a
 *     hipStream_t passed as void* (NULL = the default stream); work is enqueued asynchronously
 *     unless the comment says otherwise;
 *   - record layouts are the reference's (dyn_abi.h).
 */
#ifndef DYNAMIS_H
#define DYNAMIS_H

#include <stddef.h>
#include <stdint.h>

#include "dyn_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

#define DYN_OK 0
#define DYN_E_INVALID (-1)   /* an argument violates the documented contract */
#define DYN_E_HIP (-2)       /* the HIP runtime reported an error */
#define DYN_E_CAPACITY (-3)  /* a caller-provided buffer is too small */
#define DYN_E_NODEVICE (-4)  /* no usable gfx950 device */

/* ---- plumbing: device, memory, streams (the reference's GpuContext / GpuBuffer,
 *      crates/dynamis-gpu/src/context.rs:185, src/buffer.rs:6) -------------------------------- */

/* Library ABI version; bumps when a signature below changes. */
int dyn_version(void);
/* Text of the last error raised on the calling thread ("" when none). */
const char* dyn_last_error(void);
/* Number of visible HIP devices (0 on a machine without a GPU; never an error then). */
int dyn_device_count(int* count);
int dyn_set_device(int ordinal);
int dyn_malloc(void** d_ptr, size_t bytes);
int dyn_free(void* d_ptr);
/* Synchronous copies between host and device memory. */
int dyn_memcpy_htod(void* d_dst, const void* src, size_t bytes);
int dyn_memcpy_dtoh(void* dst, const void* d_src, size_t bytes);
int dyn_memcpy_dtod(void* d_dst, const void* d_src, size_t bytes);
int dyn_memset(void* d_dst, int value, size_t bytes);
int dyn_stream_create(void** stream);
int dyn_stream_destroy(void* stream);
int dyn_stream_synchronize(void* stream);
/* Events for timing a region of a stream on that stream's own clock. */
int dyn_event_create(void** event);
int dyn_event_destroy(void* event);
int dyn_event_record(void* event, void* stream);
int dyn_event_elapsed_ms(float* ms, void* start, void* stop);

/* ---- kernel timing (the reference's `profile` feature: per-pass timestamps,
 *      crates/dynamis-gpu/src/timing.rs:22-164, crates/dynamis-pass/src/schedule.rs:119-136) ----
 * While enabled, every kernel the library launches is bracketed by events on its own stream.
 * dyn_profile_read synchronizes those events and reports, per kernel name, the launch count and
 * the summed device time; it then forgets them. Disabled by default. */
int dyn_profile_enable(int on);
/* Restricts the bracketing of dyn_profile_enable to the launches of one kernel (NULL or "" for
 * all), so a timed run can measure one kernel's launches at the cost of two events each. */
int dyn_profile_watch(const char* kernel);
/* Number of distinct kernel names recorded since the last read. */
int dyn_profile_collect(int* kernels);
/* Entry i (0 <= i < kernels) of the last collect: name (NUL-terminated, truncated to
 * name_capacity), launch count and total milliseconds. */
int dyn_profile_entry(int i, char* name, size_t name_capacity, uint32_t* launches,
                      double* total_ms);

/* ---- radix sort (crates/dynamis-sort/src/radix.rs) ---------------------------------------- */

/* Replaces `units_for` (radix.rs:63-67): ceil(live/256)+64 clamped to [1, 512]. The sort below
 * produces the same order for every plan; the value is accepted for interface parity. */
uint32_t dyn_sort_units_for(uint32_t live);
/* Replaces `key_words` (radix.rs:71-74): bytes needed to order indices below `elements`. */
uint32_t dyn_sort_key_words(uint32_t elements);
/* Scratch bytes dyn_sort_pairs needs for lanes of `capacity` words and `digit_passes` passes. */
size_t dyn_sort_scratch_bytes(uint32_t capacity, uint32_t digit_passes);
/*
 * Replaces `RadixSort::sort` (radix.rs:397-476) with the `SortChannels` contract of
 * radix.rs:76-85: a stable LSD sort of the records [base, base+len) of three u32 lanes
 * (major, minor, payload), where len = min(*d_count, capacity - base) and base = *d_base (0 when
 * d_base is NULL), both read on the device. The order is by the low `major_words` bytes of
 * major, then the low `minor_words` bytes of minor (minor digits are applied first, then major,
 * radix.rs:312-317); equal keys keep their input order; all three lanes are permuted and the
 * result is left in the primaries. `plan` is the reference's unit budget (1..512) and does not
 * change the result. Errors: major_words + minor_words == 0 (the reference asserts the same,
 * radix.rs:404-407), words > 4, or scratch smaller than dyn_sort_scratch_bytes.
 */
int dyn_sort_pairs(void* stream, uint32_t* d_major, uint32_t* d_minor, uint32_t* d_payload,
                   uint32_t capacity, const uint32_t* d_count, const uint32_t* d_base,
                   uint32_t plan, uint32_t major_words, uint32_t minor_words,
                   void* d_scratch, size_t scratch_bytes);
/* dyn_sort_pairs with no allocation and no caller scratch. No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). The sort's scratch is the library's static device array of
 * dyn_sort_static_scratch_bytes() bytes (17 MiB: at least 2^20 records at any digit count), one
 * per device, held until the process ends; a capacity whose scratch (dyn_sort_scratch_bytes)
 * exceeds it fails with DYN_E_CAPACITY. Every call on a device shares that device's array, so
 * calls that overlap on one device (on different streams, or from several threads) race; calls
 * on one stream are ordered and safe. Everything else is dyn_sort_pairs's contract. */
size_t dyn_sort_static_scratch_bytes(void);
int dyn_sort_pairs_static(void* stream, uint32_t* d_major, uint32_t* d_minor, uint32_t* d_payload,
                          uint32_t capacity, const uint32_t* d_count, const uint32_t* d_base,
                          uint32_t plan, uint32_t major_words, uint32_t minor_words);
/* Sort faults on the current device since the last clear: look-back waits of any sort (this
 * entry point's, the broadphase's, a world step's) that gave up and left that sort's order wrong.
 * Never expected; a caller that must trust an order checks it is 0. Waits for the device first;
 * clears the count when `clear` is non-zero. No counterpart in the reference, whose sort has no
 * cross-workgroup wait (crates/dynamis-sort/src/shaders/sort_scatter.wgsl:45-104). */
int dyn_sort_faults(uint32_t* out, int clear);

/* ---- broadphase (crates/dynamis-broadphase/src/passes.rs, crates/dynamis-rigid/shaders/
 *      collider_entries.wgsl, crates/dynamis-shader/shaders/grid_index.wgsl) -------------------- */

typedef struct dyn_broadphase_io {
    /* inputs (device) */
    const dyn_aabb* d_aabbs;       /* K swept boxes, one per collider */
    const uint32_t* d_groups;      /* K owner ids; colliders of one owner never pair */
    const uint32_t* d_flags;       /* K DYN_BP_* words (dyn_abi.h), or NULL for "all mobile
                                      and awake" */
    const dyn_aabb* d_resolution;  /* K boxes that set the base cell (the reference's upright
                                      tight boxes), or NULL to take the base cell from d_aabbs */
    uint32_t colliders;            /* K */
    /* outputs (device) */
    uint32_t* d_entry_keys;        /* sorted cell keys, capacity entry_capacity */
    uint32_t* d_entry_order;       /* entry slot carried by the sort */
    dyn_grid_entry* d_entries;     /* entry records indexed by slot */
    uint32_t entry_capacity;
    uint32_t* d_pair_major;        /* min collider index of each candidate pair */
    uint32_t* d_pair_minor;        /* max collider index */
    uint32_t pair_capacity;
    uint32_t* d_counters;          /* DYN_COUNTER_BLOCK_WORDS words; ENTRIES, PAIRS,
                                      REFUSED_PAIRS, ENTRY_FAULTS, GRID_LEVELS, COARSE_ACTIVE,
                                      GRID_SCALE and GRID_EXTENT are written */
    uint32_t sort_pairs;           /* non-zero: leave the pairs in canonical (min, max) order, as
                                      the narrowphase's pair sort does (narrowphase.rs:116-125) */
    uint32_t _reserved;
    void* d_scratch;
    size_t scratch_bytes;
    const uint32_t* d_worlds;      /* K world ids of a batch of independent worlds, or NULL for
                                      one world: colliders of different worlds never pair and
                                      their cells hash apart */
} dyn_broadphase_io;

/* Scratch bytes dyn_broadphase needs. */
size_t dyn_broadphase_scratch_bytes(uint32_t colliders, uint32_t entry_capacity,
                                    uint32_t pair_capacity);
/*
 * One broadphase pass over the awake region: grid resolution (broadphase_aabb.wgsl:21-24,
 * core.wgsl:109-126), entry emission (collider_entries.wgsl:9, collider_emit.wgsl:6-63), the
 * region sort (passes.rs:76-89: 4 key bytes, payload = slot), coarser-level links
 * (level_links.wgsl) and same-level cell mates (cell_pairs.wgsl) with the pair filter of
 * pair_emit.wgsl. Entry slots are assigned in collider order, so equal keys leave the sort in
 * collider order. Pairs beyond pair_capacity are counted in REFUSED_PAIRS, entries beyond a
 * collider's 8-cell budget or beyond entry_capacity in ENTRY_FAULTS, as the reference does.
 */
int dyn_broadphase(void* stream, const dyn_broadphase_io* io);

/* ---- narrowphase (crates/dynamis-rigid/src/narrowphase.rs, crates/dynamis-rigid/shaders/
 *      narrowphase.wgsl:518-685, crates/dynamis-shader/shaders/convex.wgsl) ------------------- */

typedef struct dyn_narrowphase_io {
    /* inputs */
    const dyn_step_params* params;       /* host; contact_margin, dt (CCD margins) and the
                                            friction and restitution combine modes are read */
    const dyn_body_state* d_states;      /* bodies by row (device) */
    const dyn_body_desc* d_descs;
    uint32_t bodies;
    uint32_t colliders;
    const dyn_collider* d_colliders;     /* K collider records (device) */
    const uint32_t* d_owners;            /* K body rows, DYN_NO_BODY for none */
    const dyn_shape_source* d_sources;   /* hull shape sources as dyn_world_upload_shapes takes
                                            them, or NULL without hulls */
    const float* d_vertices;             /* their float4 vertex rows */
    uint32_t source_count;
    uint32_t pair_capacity;
    const uint32_t* d_pair_major;        /* candidate pairs (first, second collider), as
                                            dyn_broadphase leaves them */
    const uint32_t* d_pair_minor;
    const uint32_t* d_counters;          /* counter block; PAIRS is read (device) */
    /* outputs (device) */
    dyn_contact* d_contacts;             /* pair_capacity raw contact slots (contacts_raw) */
    uint32_t* d_valid;                   /* pair_capacity words (contact_valid): 1 where the
                                            pair's slot holds its contact, 0 elsewhere */
    void* d_scratch;
    size_t scratch_bytes;
    /* optional: the sources' support cells from dyn_hull_support_cells (device copies of its
       `cells` records and `candidates`), or NULL for the linear support scans; the result is
       the same */
    const uint32_t* d_support_cells;
    const uint16_t* d_support_candidates;
} dyn_narrowphase_io;

/* Scratch bytes dyn_narrowphase needs. */
size_t dyn_narrowphase_scratch_bytes(uint32_t pair_capacity);
/*
 * The narrowphase pass over the first min(PAIRS, pair_capacity) candidate pairs: the reference's
 * per-pair `work` (narrowphase.wgsl:518-685) with its pair filter, the analytic manifolds, GJK,
 * EPA, the direction probe and the clipped convex manifolds, material combination and anchors,
 * writing contact_raw / contact_valid as the reference's pass does (narrowphase.rs:126);
 * compaction is the caller's. Pairs naming colliders, bodies or hull sources outside the records
 * make no contact. The same kernels run inside dyn_world_step.
 */
int dyn_narrowphase(void* stream, const dyn_narrowphase_io* io);
/* Test hook, no counterpart in the reference: the vertex capacity (4 to 16, default 16) of the
 * EPA kernel that gives eight lanes to a polytope; a polytope that would grow past it is expanded
 * again by the kernel that gives it a wave. Results do not depend on it; lowering it routes more
 * polytopes through the second kernel. Process-wide, read when the narrowphase is launched (a
 * world's captured step keeps the value it was captured with). */
int dyn_set_epa_small_vertices(uint32_t vertices);

/* ---- islands (crates/dynamis-rigid/src/islands.rs:237-258, crates/dynamis-rigid/shaders/
 *      island_*.wgsl) ----------------------------------------------------------------------- */

typedef struct dyn_islands_io {
    const dyn_contact* d_contacts;     /* compacted contacts (device) */
    const uint32_t* d_contact_count;   /* device word; min(*count, contact_capacity) are read */
    uint32_t contact_capacity;
    uint32_t colliders;
    const uint32_t* d_owners;          /* K body rows of the colliders (device) */
    uint32_t bodies;                   /* rows; rows [0, dynamic_count) are simulated */
    uint32_t dynamic_count;
    float slop;                        /* a contact links when a point is deeper than -slop */
    uint32_t* d_parents;               /* out: dynamic_count words, each row's island root (the
                                          island's minimum row) */
    uint32_t* d_counters;              /* counter block; ISLANDS (islands among the simulated
                                          rows) and ISLAND_FAULTS are written */
    void* d_scratch;
    size_t scratch_bytes;
} dyn_islands_io;

/* Scratch bytes dyn_islands needs. */
size_t dyn_islands_scratch_bytes(uint32_t bodies);
/*
 * The island pass over contacts alone (islands.rs:237-258 without the resting-archive and
 * constraint links, wake and sleep): island_init, island_link_contacts (a touching, non-sensor
 * contact between two simulated rows joins their islands; the higher root is hooked under the
 * lower), the island_flatten rounds (lib.rs:79-81; here rounds that walk up to 64 hops, as many
 * as flatten any forest of `dynamic_count` rows), island_count. Every row's parent is then its
 * island's minimum row, as in the reference; a seeding pass and hooking rounds before the
 * linking hooks only shorten their work. Contacts naming colliders or bodies outside the records
 * link nothing. The same kernels run inside dyn_world_step.
 */
int dyn_islands(void* stream, const dyn_islands_io* io);

/* ---- solver (crates/dynamis-rigid/src/solver.rs, crates/dynamis-rigid/shaders/solver_*.wgsl,
 *      substep_*.wgsl, position_*.wgsl) --------------------------------------------------------- */

typedef struct dyn_solve_io {
    const dyn_step_params* params;     /* host; substeps, iterations, gravity, damping, clamps,
                                          restitution threshold are read (dt is an argument) */
    dyn_body_state* d_states;          /* bodies by row, read and written (device) */
    const dyn_body_desc* d_descs;
    uint32_t bodies;
    uint32_t dynamic_count;            /* rows [0, dynamic_count) are simulated */
    uint32_t colliders;
    uint32_t contact_capacity;
    const uint32_t* d_owners;          /* K body rows of the colliders (device) */
    dyn_contact* d_contacts;           /* compacted contacts (device): their accumulated impulses
                                          warm-start the first substep and are replaced by the
                                          new ones */
    const uint32_t* d_contact_count;   /* device word; min(*count, contact_capacity) are solved */
    uint32_t* d_counters;              /* counter block; LIVE, SOLVER_ROWS and the residuals are
                                          written */
    void* d_scratch;
    size_t scratch_bytes;
} dyn_solve_io;

/* Scratch bytes dyn_solve needs. */
size_t dyn_solve_scratch_bytes(uint32_t bodies, uint32_t contact_capacity);
/*
 * The solver alone: SolverReset (solver.rs:89-101) and SolveSubsteps (solver.rs:304-334) —
 * the contact blocks and their mass splitting, then per substep integration, the warm start
 * (first substep), the Jacobi velocity iterations, the pose advance and the position
 * iterations — and the new accumulated impulses into the contacts. Rows that are awake move;
 * sleeping rows stay. The contacts are checked first (one synchronization of the stream): a
 * contact naming a collider or body outside the records is refused with DYN_E_INVALID and
 * nothing is solved. The same kernels run inside dyn_world_step.
 */
int dyn_solve(void* stream, const dyn_solve_io* io, float dt);

/* ---- host record encoding (crates/dynamis-model/src/config.rs:12-100 PhysicsConfig,
 *      body.rs:89 BodyDesc; crates/dynamis-abi/src/body.rs:96-139, collider.rs:47-102,
 *      param.rs:110-137; crates/dynamis-model/src/mass.rs) -------------------------------------- */

/* PhysicsConfig (config.rs:12-33). Combine modes: 0 multiply (sqrt(a*b)), 1 min, 2 max, 3 average. */
typedef struct dyn_config {
    float gravity[3];
    float damping;
    float angular_damping;
    uint32_t substeps;
    uint32_t solve_iterations;
    uint32_t position_iterations;
    float relaxation;
    float slop;
    float contact_margin;
    float restitution_threshold;
    float max_velocity;
    float max_angular_velocity;
    float sleep_velocity;
    float sleep_angular_velocity;
    float sleep_time;
    uint32_t friction_combine;
    uint32_t restitution_combine;
} dyn_config;

/* A BodyDesc with one collider at the body origin (model/body.rs:89-108, collider.rs:52-68);
 * shape DYN_SHAPE_SPHERE, _CUBOID, _CAPSULE, _CYLINDER, _HULL or _PLANE.
 * density >= 0 selects BodyDesc::density, otherwise `mass` is the fixed mass (0 = static);
 * damping values < 0 take the configuration's. */
typedef struct dyn_body_spec {
    uint32_t shape; /* DYN_SHAPE_SPHERE, _CUBOID, _CAPSULE or _PLANE */
    float radius;
    float half_height;
    float half_extents[3];
    float position[3];
    float orientation[4];
    float velocity[3];
    float angular_velocity[3];
    float mass;
    float density;
    float friction;
    float restitution;
    float rolling_friction;
    float spin_friction;
    float gravity_scale;
    float linear_damping;
    float angular_damping;
    uint32_t kinematic;
    uint32_t collision_group;
    uint32_t collision_mask;
    /* DYN_SHAPE_HULL only: the shape source index and its solid geometry (dyn_hull_solid) */
    uint32_t source;
    float solid_volume;
    float solid_centroid[3];
    float solid_second[6];
} dyn_body_spec;

/* PhysicsConfig::default (config.rs:76-100). */
void dyn_config_default(dyn_config* config);
/* StepParamsRecord::new without the step length (param.rs:110-137). */
int dyn_step_params_build(const dyn_config* config, uint32_t body_count, uint32_t dynamic_count,
                          uint32_t collider_count, dyn_step_params* out);
/* Quickhull over `count` points (xyz triples), as World::add_hull_from_mesh builds a hull
 * (crates/dynamis-hull/src/hull.rs:51-288): hull vertices (xyz, in order of first use) and
 * outward triangles. Fails on collinear or coplanar points, or when either output exceeds
 * `capacity` entries. */
int dyn_hull_build(const float* points, uint32_t count, float* vertices, uint32_t* vertex_count,
                   uint32_t* triangles, uint32_t* triangle_count, uint32_t capacity);
/* Volume, centroid and unit second moment about the centroid ([xx, xy, xz, yy, yz, zz]) of a
 * closed triangle hull, accumulated in FP64 (crates/dynamis-hull/src/solid.rs:62-105). */
int dyn_hull_solid(const float* vertices, uint32_t vertex_count, const uint32_t* triangles,
                   uint32_t triangle_count, float* volume, float* centroid, float* second);
/* Support cells of hull shape sources, for the device's hull support (no counterpart: the
 * reference scans every vertex, convex.wgsl:1-85, and so does the oracle): per source and cell
 * of a cube map of 6 x 8 x 8 cells over local directions, the ascending list of the vertices
 * that can be the direction's support vertex (the first vertex of largest single-precision
 * dot). A vertex is left out of a cell only when another exceeds it by more than 1e-5 x the
 * hull's largest |x|+|y|+|z| everywhere in the slightly widened cell, far above the rounding of
 * a single-precision dot, so scanning a cell's list gives the linear scan's vertex. Sources of
 * fewer than 12 or more than 4096 vertices, or with coordinates beyond 1e6, keep the linear scan.
 * Inputs are the sources and float4 vertex rows of dyn_world_upload_shapes (host). `cells`
 * receives two words per source and cell: word 0 the candidate count in bits 0-7 (0xFF: the cell
 * keeps the linear scan) and the first two candidates' vertex indices in bits 8-19 and 20-31;
 * word 1 the third and fourth in bits 0-11 and 12-23 for counts up to 4, else the offset of the
 * cell's list in `candidates`. `candidates` (optional) receives every cell's ascending index list
 * (*candidate_count indices); a capacity below that count fails with DYN_E_CAPACITY after setting
 * it. dyn_world_upload_shapes builds them itself. */
int dyn_hull_support_cells(const dyn_shape_source* sources, uint32_t source_count, const float* vertices,
                           uint32_t vertex_count, uint32_t* cells, uint16_t* candidates, uint32_t capacity,
                           uint32_t* candidate_count);
/* inertia_inverse (mass.rs:259-280); fails on a tensor that is not positive definite. */
int dyn_inertia_inverse(const float* inertia, float* inverse);
/* BodyStateRecord::initial + BodyDescriptorRecord::build + ColliderRecord::build for one body
 * (body.rs:34-58, :96-139, collider.rs:47-102). `body_id` is the body's handle id, which the
 * state record carries; `collider_slot` is the collider's index within its body (0 for a body
 * with one collider). */
int dyn_encode_body(const dyn_body_spec* spec, const dyn_config* config, uint32_t body_id,
                    uint32_t collider_slot, dyn_body_state* state, dyn_body_desc* desc,
                    dyn_collider* collider);
/* dyn_encode_body over `count` specs with body ids first_body_id, first_body_id + 1, ... and
 * collider slot 0: one call for a spawn list (the reference builds the records body by body in
 * World::spawn, crates/dynamis-world/src/body.rs:34-58). Stops at the first failing spec and
 * reports its index in the error text. */
int dyn_encode_bodies(const dyn_body_spec* specs, uint32_t count, const dyn_config* config,
                      uint32_t first_body_id, dyn_body_state* states, dyn_body_desc* descs,
                      dyn_collider* colliders);

/* ---- world step (crates/dynamis-world/src/lib.rs:86 World::new, src/step.rs:34 World::step,
 *      src/observation.rs:484 wait, src/body.rs:326 read_state, src/readback.rs:95
 *      inspect_contacts, :215 measured) ---------------------------------------------------------- */

typedef struct dyn_world dyn_world;

/*
 * Creates a world from device-format records (host arrays, uploaded once): bodies in rows
 * [0, body_count) with the simulated ones first in [0, dynamic_count) (the reference's row
 * partition, crates/dynamis-world/src/body.rs:292-321), each carrying a unique body_id (its
 * handle id; the world keeps the id -> row table of body_row.wgsl), colliders in collider-slot
 * order with the rows of their owners (upload.rs:133-175). `params` is the StepParams record (dt and substep_dt are set by each step).
 * `stream` (hipStream_t) orders all of the world's work; NULL creates a private stream.
 * contact_capacity bounds contacts and resting manifolds per step, pair_capacity the candidate
 * pairs (the reference's stream capacities; overflow is counted in REFUSED_CONTACTS,
 * REFUSED_RESTING and REFUSED_PAIRS). The capacities are the starting plan: a step whose demand
 * outgrew them grows them two steps later (see dyn_world_capacity).
 */
int dyn_world_create(dyn_world** world, void* stream, const dyn_step_params* params,
                     const dyn_body_state* states, const dyn_body_desc* descs, uint32_t body_count,
                     uint32_t dynamic_count, const dyn_collider* colliders, const uint32_t* owners,
                     uint32_t collider_count, uint32_t contact_capacity, uint32_t pair_capacity);
int dyn_world_destroy(dyn_world* world);
/* A batch of independent worlds stepped as one: the worlds' rows, colliders and contacts share
 * the buffers of one dyn_world, and collider_worlds[k] names the world of collider k (all
 * colliders of a body in one world). Colliders of different worlds never pair, so every world
 * steps exactly as it would alone. Rows: all worlds' simulated rows first, then the others;
 * body ids unique across the batch. Step, wait, reads and destroy are the dyn_world_* calls.
 * No counterpart in the reference, which steps one world per World. */
int dyn_batch_create(dyn_world** out, void* stream, const dyn_step_params* params,
                     const dyn_body_state* states, const dyn_body_desc* descs, uint32_t body_count,
                     uint32_t dynamic_count, const dyn_collider* colliders, const uint32_t* owners,
                     const uint32_t* collider_worlds, uint32_t collider_count, uint32_t world_count,
                     uint32_t contact_capacity, uint32_t pair_capacity);
/* Uploads the hull shape sources colliders name in their `source` field, with their vertices as
 * float4 rows (x, y, z, 0) (the reference's shape pool flush, crates/dynamis-world/src/
 * shape_pool.rs:243-279). Replaces any earlier upload; call before a step that needs them.
 * Fails when a source holds more than DYN_LINEAR_SUPPORT_VERTICES vertices or reaches past
 * the vertex array, or when a hull collider names a source that is not there. */
int dyn_world_upload_shapes(dyn_world* world, const dyn_shape_source* sources,
                            uint32_t source_count, const float* vertices, uint32_t vertex_count);
/* Queues body edits for the next step (replaces the body-edit half of the reference's command
 * path: World::set_velocity / set_position / apply_force / apply_impulse / sleep / wake compile
 * into BodyEdit runs, crates/dynamis-world/src/command.rs, applied by body_edits.wgsl in
 * ApplyCommands, crates/dynamis-rigid/src/apply_commands.rs:273-365). body_ids[i] names the body
 * edits[i] acts on; the next dyn_world_step applies every queued edit after clearing the bodies'
 * inputs and before their activity flags, a body's edits in the order they were queued (over any
 * number of calls), exactly as body_edits.wgsl does. The edits reach the device on the world's
 * stream after the steps already enqueued; the call waits for those steps. Errors: an id naming
 * no body, an unknown kind. */
int dyn_world_edit(dyn_world* world, const uint32_t* body_ids, const dyn_body_edit* edits, uint32_t count);
/* Enqueues one step of `dt` seconds on the world's stream (non-blocking). The step is captured
 * as a HIP graph on its first call for a given dt and replayed afterwards; steps run eagerly
 * while dyn_profile_enable is on or after dyn_world_set_graphs(world, 0). */
int dyn_world_step(dyn_world* world, float dt);
/* The last step's convex-path work: out[0] pairs given to GJK, out[1] pairs given manifolds
 * by the convex path, out[2] pairs given to EPA, out[3] the GJK pairs with a hull, out[4] the GJK
 * pairs that ran past their first 8 iterations, out[5] the EPA polytopes too large for the
 * eight-lane kernel (6 words). No counterpart in the reference (whose counters do not split the
 * narrowphase); for measurement. */
int dyn_world_read_work(dyn_world* world, uint32_t* out);
/* Turns the graph capture of dyn_world_step on (default) or off. No counterpart in the
 * reference, whose steps are recorded command buffers (crates/dynamis-gpu/src/recorder.rs:29). */
int dyn_world_set_graphs(dyn_world* world, int enabled);
/* Blocks until every enqueued step has finished. */
int dyn_world_wait(dyn_world* world);
/* The workgroup size of the solver substeps' per-contact and per-row kernels (k_integrate,
 * k_solver_bodies, k_solve_prep, k_warm, k_solve, k_apply, k_advance, k_position_block,
 * k_position_apply, k_impulses_store): 64, 128 or 256 (default). Each of them indexes its contact
 * or row by the global thread index alone, so every size gives the same result. No counterpart in
 * the reference, whose programs have a fixed workgroup size. dyn_world_tune_solver times `steps`
 * steps (3 when 0) of the world with each size from a snapshot of its current state (at the last
 * step's dt), keeps the fastest, restores the state, and returns the size and the three
 * milliseconds-per-step (either pointer may be NULL); it records no capacity demand, allocates a
 * snapshot buffer while it runs, and fails while edits are queued or a tape is recorded. */
int dyn_world_set_solver_threads(dyn_world* world, uint32_t threads);
int dyn_world_tune_solver(dyn_world* world, uint32_t steps, uint32_t* threads_out, float* ms_out);
/* Capacity planning (World::apply_plan, crates/dynamis-world/src/backend/mod.rs:194-230, with the
 * stream plans of crates/dynamis-broadphase/src/capacity.rs:27-80 and crates/dynamis-rigid/src/
 * capacity.rs:50-140). Every step's demand (its uncapped pair total, and the larger of its
 * uncapped contact and resting totals) is read back two steps later; before that later step a
 * demand above a stream's capacity grows the stream to max(demand, 2 x capacity), keeping the
 * durable contact and resting archives. The step that overflowed has already refused what did
 * not fit (REFUSED_PAIRS, REFUSED_CONTACTS, REFUSED_RESTING), as in the reference, whose plan also
 * follows a lagged readback. A fixed lag makes growth a function of the steps alone, so a run is
 * deterministic. Streams never shrink (the reference releases capacity once a scene settles).
 * Growth waits for the world's stream and drops its captured graphs. It is held off while a
 * differentiable rollout (dyn_adjoint_create) on the world exists, whose buffers are sized by the
 * contact capacity. Snapshots record the capacities they were taken at; a restore into other
 * capacities fails.
 * dyn_world_capacity reads the current capacities and the number of growths so far (any pointer
 * may be NULL); dyn_world_reserve grows them to at least the given values now. */
int dyn_world_capacity(dyn_world* world, uint32_t* pair_capacity, uint32_t* contact_capacity, uint32_t* grows);
int dyn_world_reserve(dyn_world* world, uint32_t pair_capacity, uint32_t contact_capacity);

/* ---- the literal fixed-cell broadphase (no reference counterpart: the reference's grid is
 * hierarchical, crates/dynamis-broadphase/src/passes.rs:76-160) ----
 * Every box enters each cell of a uniform grid of side cell_size that its AABB touches, with the
 * key (cell hash of grid_index.wgsl:28-33, box index); the keys are radix-sorted (hash major, box
 * minor, cell ordinal payload: entries out in d_entry_*), and every pair of overlapping boxes is
 * emitted once (in the cell holding the overlap's min corner) and canonicalised (first < second,
 * sorted lexicographically). The pair set equals the hierarchical grid's on the same boxes.
 * d_counts receives 4 words: entries, pairs, boxes refused (touching more than 64 cells), pairs
 * refused (past pair_capacity); entries past entry_capacity are dropped. Asynchronous on stream. */
size_t dyn_fixed_grid_scratch_bytes(uint32_t boxes, uint32_t entry_capacity, uint32_t pair_capacity);
int dyn_fixed_grid(void* stream, const dyn_aabb* d_boxes, uint32_t boxes, float cell_size, uint32_t* d_entry_hash,
                   uint32_t* d_entry_box, uint32_t* d_entry_ordinal, uint32_t entry_capacity, uint32_t* d_pair_first,
                   uint32_t* d_pair_second, uint32_t pair_capacity, uint32_t* d_counts, void* d_scratch,
                   size_t scratch_bytes);

/* ---- autotuning (no reference counterpart: the reference's sort has one fixed tiling) ----
 * A sort pass's tiling (keys per lane 8 or 16, cap on persistent workgroups 512/1024/2048) is
 * chosen per (device, ceil(log2(capacity))) from a process-wide table; untuned buckets use the
 * default. dyn_sort_autotune times every tiling on the device for `capacity` keys and
 * `digit_passes` digits (synthetic keys, minimum of three rounds after a warm round) and keeps the
 * fastest; dyn_world_autotune does so for the capacities of a world's sorts (entries, pairs and
 * worklists, the resting index) and drops its captured step; tilings (optional, 9 words) receives
 * capacity, keys per lane, groups for each. The table is exported and imported as 4 words per
 * entry (device, bucket, keys per lane, groups) so a caller can keep it per GPU. Every tiling
 * sorts to the same stable order. */
int dyn_sort_autotune(void* stream, uint32_t capacity, uint32_t digit_passes, uint32_t* items_out,
                      uint32_t* groups_out);
int dyn_sort_tuning_export(uint32_t* out, uint32_t capacity, uint32_t* count);
int dyn_sort_tuning_import(const uint32_t* entries, uint32_t count);
int dyn_world_autotune(dyn_world* world, uint32_t* tilings);
/* The current device's marketing name and gfx target, "name|gfx950", into out (NUL-terminated). */
int dyn_device_name(char* out, uint32_t capacity);

/* Checkpoint / resume on the device (World::snapshot / restore, crates/dynamis-world/src/
 * snapshot.rs:56-123): the durable state a step reads from earlier steps (body states, contact
 * and resting archives with the resting index, static wake marks, counters) copied to or from a
 * caller's device buffer of dyn_world_snapshot_bytes bytes, enqueued on the world's stream.
 * A world restored from a snapshot steps bit for bit as it did after the snapshot was taken. */
int dyn_world_snapshot_bytes(dyn_world* world, size_t* bytes);
int dyn_world_snapshot(dyn_world* world, void* d_dst);
int dyn_world_restore(dyn_world* world, const void* d_src);

/* ---- the differentiable rollout (no reference counterpart: the reference has no adjoint) ----
 *
 * dyn_adjoint_forward steps the world `steps` times with dt (HIP graphs as usual), keeping a
 * snapshot every `checkpoint_every` steps, recording the tape of the last segment as it goes.
 * dyn_adjoint_backward takes the adjoint of the final
 * states (device, 16 floats per row: position xyz _, orientation xyzw, velocity xyz _, spin xyz _)
 * and returns, on the device, the adjoint of the initial states (same layout), of each row's mass
 * (mass and inertia scaled together at fixed shape; zero for rows that do not simulate) and of each
 * collider's friction coefficient. It re-steps each segment from its snapshot while recording the
 * tape (the contacts as the solver received them) and differentiates the step with that contact
 * set frozen: contact pairs, normals, features, body-local anchors, warm-start relay and split
 * counts stay at their forward values; poses, velocities, the contact depths along the frozen
 * normals, the impulses and the mass and friction parameters are differentiated, branches taking
 * the derivative of the branch the forward selected. Soft contacts, rolling or spin friction and
 * sleeping bodies are refused (DYN_E_INVALID). world_ranges: per world of a batch 4 words (first
 * simulated row, simulated rows, first collider, colliders); one world: {0, dynamic, 0, colliders}.
 * The world is left as the forward rollout ended. */
typedef struct dyn_adjoint dyn_adjoint;
int dyn_adjoint_create(dyn_adjoint** out, dyn_world* world, uint32_t steps, uint32_t checkpoint_every,
                       const uint32_t* world_ranges, uint32_t world_count);
int dyn_adjoint_forward(dyn_adjoint* adjoint, float dt);
int dyn_adjoint_backward(dyn_adjoint* adjoint, const float* d_state_grad, float* d_state0_grad,
                         float* d_mass_grad, float* d_friction_grad);
/* Host copies of the tape slots (slot < checkpoint_every): the segment last recorded, by the
 * forward rollout (its last segment) or by the backward pass's recompute; with checkpoint_every >=
 * steps, slot t is step t and the backward pass recomputes nothing. Also the rollout's body
 * records. */
int dyn_adjoint_read_tape(dyn_adjoint* adjoint, uint32_t slot, float* states, float* contacts,
                          uint32_t capacity, uint32_t* count);
int dyn_adjoint_read_bodies(dyn_adjoint* adjoint, float* out);
int dyn_adjoint_destroy(dyn_adjoint* adjoint);
/* The world's hipStream_t. */
void* dyn_world_stream(dyn_world* world);
/* Synchronous reads (they wait for the stream first). */
int dyn_world_read_states(dyn_world* world, dyn_body_state* out);
int dyn_world_read_counters(dyn_world* world, uint32_t* out);
int dyn_world_read_contacts(dyn_world* world, dyn_contact* out, uint32_t capacity, uint32_t* count);
/* The resting archive: manifolds frozen with their sleeping pairs, which the reference's
 * World::inspect_contacts appends to the active ones (crates/dynamis-world/src/readback.rs:95-139). */
int dyn_world_read_resting(dyn_world* world, dyn_contact* out, uint32_t capacity, uint32_t* count);
int dyn_world_read_pairs(dyn_world* world, uint32_t* major, uint32_t* minor, uint32_t capacity,
                         uint32_t* count);

#ifdef __cplusplus
}
#endif

#endif /* DYNAMIS_H */
