/*
 * ostep.h - CPU restatement of one rigid-body World::step. TEST INFRASTRUCTURE ONLY (see
 * oracle.h). Serial FP32, one pass after another in the reference's pass order
 * (crates/dynamis-world/tests/suite/pipeline.rs:6-49) for a world without joints, CCD, soft
 * bodies, characters, vehicles, fields or command edits.
 */
#ifndef DYN_OSTEP_H
#define DYN_OSTEP_H

#include <stdint.h>

#include "../include/dyn_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_world oracle_world;

/* Bodies occupy rows [0, body_count); rows [0, dynamic_count) are the simulated ones
 * (crates/dynamis-world/src/body.rs:292-321). Capacities bound the per-step streams. */
oracle_world* oracle_world_create(const dyn_step_params* params, const dyn_body_state* states,
                                  const dyn_body_desc* descs, uint32_t body_count,
                                  uint32_t dynamic_count, const dyn_collider* colliders,
                                  const uint32_t* owners, uint32_t collider_count,
                                  uint32_t contact_capacity, uint32_t pair_capacity);
void oracle_world_destroy(oracle_world* w);
/* Hull shape sources (records.wgsl ShapeSource) and their vertices as float4 rows; 0 on success,
 * -1 when a source exceeds the linear support bound or its vertex range. */
int oracle_world_shapes(oracle_world* w, const dyn_shape_source* sources, uint32_t source_count,
                        const float* vertices, uint32_t vertex_count);
/* Queues body edits for the next step (body_edits.wgsl semantics); -1 when an id names no body or
 * a kind is unknown. */
int oracle_world_edit(oracle_world* w, const uint32_t* body_ids, const dyn_body_edit* edits, uint32_t count);
/* One step with the given dt (params.dt / substep_dt are set from it). */
void oracle_world_step(oracle_world* w, float dt);
/* Current pair and contact capacities and the growths so far (3 words); see plan_capacity. */
void oracle_world_capacity(const oracle_world* w, uint32_t* out);
void oracle_world_states(const oracle_world* w, dyn_body_state* out);
/* The last step's tape for the differentiable-step checks: the states it started from, its
 * contacts as the solver received them (after the warm-start relay) and, per manifold point,
 * the archive point its impulses were relayed from (4 * archived contact + point, or ~0). */
uint32_t oracle_world_tape(const oracle_world* w, dyn_body_state* start, dyn_contact* pre, uint32_t* relay_src,
                           uint32_t capacity);
/* Contacts of the last step (after the solve), in canonical pair order; returns the count. */
uint32_t oracle_world_contacts(const oracle_world* w, dyn_contact* out, uint32_t capacity);
/* The resting archive (frozen manifolds of sleeping pairs), in list order; returns the count. */
uint32_t oracle_world_resting(const oracle_world* w, dyn_contact* out, uint32_t capacity);
void oracle_world_counters(const oracle_world* w, uint32_t* out);
/* Sorted candidate pairs of the last step; returns the count. */
/* The solver alone over the given contacts (read, then written back with the new impulses) and
 * the current records; -1 when count exceeds the contact capacity or a contact names a collider
 * or body outside the records. */
int oracle_world_solve(oracle_world* w, dyn_contact* contacts, uint32_t count, float dt);
/* The narrowphase alone over the given pairs and the current records (contacts_raw,
 * contact_valid); -1 when count exceeds the pair capacity or a pair names no collider. */
/* Threads the step's parallel loops use: 1 in the serial build, OpenMP's maximum in the
 * OpenMP build (liboracle_omp.so). */
uint32_t oracle_threads(void);
int oracle_world_narrowphase(oracle_world* w, const uint32_t* major, const uint32_t* minor, uint32_t count,
                             dyn_contact* raw, uint32_t* valid);
uint32_t oracle_world_pairs(const oracle_world* w, uint32_t* major, uint32_t* minor,
                            uint32_t capacity);

/* The warm-start relay of identity.wgsl:85-117 alone (test hook): out = current with the
 * impulses of held relayed; picks (4 words, may be null) = the held point relayed into each
 * point, or ~0. */
void oracle_relay_impulses(const dyn_contact* current, const dyn_contact* held, dyn_contact* out, uint32_t* picks);

/* GJK / convex_hit of two colliders at given body states (test hook, no world). */
int oracle_convex_query(const dyn_shape_source* sources, uint32_t source_count, const float* vertices,
                        uint32_t vertex_count, const dyn_collider* ca, const dyn_body_state* sa,
                        const dyn_collider* cb, const dyn_body_state* sb, float* out);

/* Host record encoding (crates/dynamis-model/src/mass.rs, crates/dynamis-abi/src/body.rs). */
void oracle_inertia_inverse(const float* inertia, float* inverse);

#ifdef __cplusplus
}
#endif

#endif
