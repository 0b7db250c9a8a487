// Host record encoding: a body description with one collider becomes the device records the step
// reads, and a physics configuration becomes the StepParams record. Follows
// crates/dynamis-abi/src/body.rs:96-139 (BodyDescriptorRecord::build), :34-58 (BodyStateRecord::
// initial), crates/dynamis-abi/src/collider.rs:47-102 (ColliderRecord::build),
// crates/dynamis-model/src/mass.rs:34-96 (analytic solids), :155-210 (compute_mass_properties),
// :282-342 (second-moment rotation),
// :259-280 (inertia_inverse) and crates/dynamis-abi/src/param.rs:110-137 (StepParamsRecord::new).
// Arithmetic is FP32 in the reference's operation order.
#include <cmath>
#include <cstring>

#include "common.h"

namespace {

constexpr float kPi = 3.14159265358979323846f;

struct Solid {
    float volume;
    float centroid[3];
    float second[6];  // unit second moment about the centroid: xx xy xz yy yz zz
    bool present;
};

// mass.rs:34-96 (analytic shapes) and the hull's solid geometry carried by the spec
Solid solid_of(const dyn_body_spec& b) {
    Solid s{};
    s.present = true;
    if (b.shape == DYN_SHAPE_SPHERE) {
        const float r = b.radius;
        const float moment = r * r / 5.0f;
        s.volume = 4.0f / 3.0f * kPi * r * r * r;
        const float m[6] = {moment, 0.0f, 0.0f, moment, 0.0f, moment};
        std::memcpy(s.second, m, sizeof m);
    } else if (b.shape == DYN_SHAPE_CUBOID) {
        const float ex = b.half_extents[0] * 2.0f;
        const float ey = b.half_extents[1] * 2.0f;
        const float ez = b.half_extents[2] * 2.0f;
        s.volume = ex * ey * ez;
        const float m[6] = {ex * ex / 12.0f, 0.0f, 0.0f, ey * ey / 12.0f, 0.0f, ez * ez / 12.0f};
        std::memcpy(s.second, m, sizeof m);
    } else if (b.shape == DYN_SHAPE_CAPSULE) {
        const float r = b.radius, hh = b.half_height;
        const float cylinder = kPi * r * r * 2.0f * hh;
        const float sphere = 4.0f / 3.0f * kPi * r * r * r;
        const float total = cylinder + sphere;
        const float lateral = (cylinder * r * r / 4.0f + sphere * r * r / 5.0f) / total;
        const float axial = (cylinder * 4.0f * hh * hh / 12.0f +
                             sphere * (r * r / 5.0f + 3.0f / 4.0f * r * hh + hh * hh)) /
                            total;
        s.volume = total;
        const float m[6] = {lateral, 0.0f, 0.0f, axial, 0.0f, lateral};
        std::memcpy(s.second, m, sizeof m);
    } else if (b.shape == DYN_SHAPE_CYLINDER) {
        const float r = b.radius;
        const float height = b.half_height * 2.0f;
        const float lateral = r * r / 4.0f;
        s.volume = kPi * r * r * height;
        const float m[6] = {lateral, 0.0f, 0.0f, height * height / 12.0f, 0.0f, lateral};
        std::memcpy(s.second, m, sizeof m);
    } else if (b.shape == DYN_SHAPE_HULL) {
        s.volume = b.solid_volume;
        std::memcpy(s.centroid, b.solid_centroid, sizeof s.centroid);
        std::memcpy(s.second, b.solid_second, sizeof s.second);
    } else {
        s.present = false;  // planes carry no solid
    }
    return s;
}

// mass.rs:282-342 for the collider's own rotation (the identity here: one collider at the body
// origin), written out so the products and sums round as the reference's do.
using M3 = float[3][3];
void mat_mul(const M3 a, const M3 b, M3 out) {
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) out[r][c] = a[r][0] * b[0][c] + a[r][1] * b[1][c] + a[r][2] * b[2][c];
    }
}
void second_moment_rotate(const float* second, const float* q, float* out) {
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    const M3 r = {{1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y - w * z), 2.0f * (x * z + w * y)},
                  {2.0f * (x * y + w * z), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z - w * x)},
                  {2.0f * (x * z - w * y), 2.0f * (y * z + w * x), 1.0f - 2.0f * (x * x + y * y)}};
    const M3 sm = {{second[0], second[1], second[2]}, {second[1], second[3], second[4]},
                   {second[2], second[4], second[5]}};
    const M3 rt = {{r[0][0], r[1][0], r[2][0]}, {r[0][1], r[1][1], r[2][1]}, {r[0][2], r[1][2], r[2][2]}};
    M3 inner, m;
    mat_mul(sm, rt, inner);
    mat_mul(r, inner, m);
    out[0] = m[0][0];
    out[1] = (m[0][1] + m[1][0]) * 0.5f;
    out[2] = (m[0][2] + m[2][0]) * 0.5f;
    out[3] = m[1][1];
    out[4] = (m[1][2] + m[2][1]) * 0.5f;
    out[5] = m[2][2];
}

// mass.rs:259-280
void inertia_inverse(const float* in, float* out) {
    const float xx = in[0], xy = in[1], xz = in[2], yy = in[3], yz = in[4], zz = in[5];
    const float cxx = yy * zz - yz * yz;
    const float cxy = xz * yz - xy * zz;
    const float cxz = xy * yz - xz * yy;
    const float cyy = xx * zz - xz * xz;
    const float cyz = xy * xz - xx * yz;
    const float czz = xx * yy - xy * xy;
    const float det = xx * cxx + xy * cxy + xz * cxz;
    out[0] = cxx / det;
    out[1] = cxy / det;
    out[2] = cxz / det;
    out[3] = cyy / det;
    out[4] = cyz / det;
    out[5] = czz / det;
}

}  // namespace

extern "C" {

int dyn_inertia_inverse(const float* inertia, float* inverse) {
    if (!inertia || !inverse) return dyn::fail(DYN_E_INVALID, "dyn_inertia_inverse: null argument");
    const float xx = inertia[0], xy = inertia[1], yy = inertia[3];
    const float czz = xx * yy - xy * xy;
    inertia_inverse(inertia, inverse);
    // mass.rs:271-274 asserts positive definiteness
    const float det_sign = inverse[0];
    if (!(xx > 0.0f && czz > 0.0f && std::isfinite(det_sign))) {
        return dyn::fail(DYN_E_INVALID, "inertia tensor must be positive definite");
    }
    return DYN_OK;
}

int dyn_encode_body(const dyn_body_spec* spec, const dyn_config* config, uint32_t body_id,
                    uint32_t collider_slot, dyn_body_state* state, dyn_body_desc* desc,
                    dyn_collider* collider) {
    using dyn::fail;
    if (!spec || !config || !state || !desc || !collider) {
        return fail(DYN_E_INVALID, "dyn_encode_body: null argument");
    }
    const dyn_body_spec& b = *spec;
    if (b.shape != DYN_SHAPE_SPHERE && b.shape != DYN_SHAPE_CUBOID && b.shape != DYN_SHAPE_CAPSULE &&
        b.shape != DYN_SHAPE_CYLINDER && b.shape != DYN_SHAPE_HULL && b.shape != DYN_SHAPE_PLANE) {
        return fail(DYN_E_INVALID, "dyn_encode_body: shape kind not supported by this encoder");
    }
    if (b.shape == DYN_SHAPE_HULL && !(b.solid_volume > 0.0f)) {
        return fail(DYN_E_INVALID, "dyn_encode_body: a hull carries its solid geometry (dyn_hull_solid)");
    }
    const Solid solid = solid_of(b);
    // BodyDesc::effective_mass (model/body.rs:363-368): scale volume (1) x solid volume
    const float volume = solid.present ? 1.0f * 1.0f * 1.0f * solid.volume : 0.0f;
    const float mass = b.density >= 0.0f ? b.density * volume : b.mass;
    const bool kinematic = b.kinematic != 0u;
    const bool simulates = !kinematic && mass > 0.0f;
    // compute_mass_properties (mass.rs:155-210) for one collider at the body origin, unrotated
    // and unscaled
    float com[3] = {0.0f, 0.0f, 0.0f};
    float inertia[6] = {0, 0, 0, 0, 0, 0};
    float inverse[6] = {0, 0, 0, 0, 0, 0};
    if (solid.present && volume > 0.0f && mass > 0.0f) {
        const float identity[4] = {0.0f, 0.0f, 0.0f, 1.0f};
        // centroid_of: offset (0) + rotate(identity, scale (1) x centroid)
        float at[3];
        for (int a = 0; a < 3; ++a) at[a] = 0.0f + 1.0f * solid.centroid[a];
        for (int a = 0; a < 3; ++a) com[a] = (0.0f + at[a] * volume) / volume;
        const float shape_mass = mass * volume / volume;
        float scaled[6];
        for (int i = 0; i < 6; ++i) scaled[i] = solid.second[i] * 1.0f * 1.0f;
        float rotated[6];
        second_moment_rotate(scaled, identity, rotated);
        const float d[3] = {at[0] - com[0], at[1] - com[1], at[2] - com[2]};
        const float translated[6] = {rotated[0] + d[0] * d[0], rotated[1] + d[0] * d[1],
                                     rotated[2] + d[0] * d[2], rotated[3] + d[1] * d[1],
                                     rotated[4] + d[1] * d[2], rotated[5] + d[2] * d[2]};
        float second[6];
        for (int i = 0; i < 6; ++i) second[i] = 0.0f + translated[i] * shape_mass;
        const float xx = second[0], xy = second[1], xz = second[2], yy = second[3], yz = second[4],
                    zz = second[5];
        const float in[6] = {yy + zz, -xy, -xz, xx + zz, -yz, xx + yy};
        std::memcpy(inertia, in, sizeof in);
        if (dyn_inertia_inverse(inertia, inverse) != DYN_OK) return DYN_E_INVALID;
    }
    std::memset(state, 0, sizeof *state);
    std::memcpy(state->position, b.position, 12);
    std::memcpy(state->prev_position, b.position, 12);
    std::memcpy(state->prev_orientation, b.orientation, 16);
    std::memcpy(state->orientation, b.orientation, 16);
    std::memcpy(state->velocity, b.velocity, 12);
    std::memcpy(state->angular_velocity, b.angular_velocity, 12);
    state->body_id = body_id;
    state->generation = 0;
    std::memset(desc, 0, sizeof *desc);
    desc->inverse_mass = simulates ? 1.0f / mass : 0.0f;
    desc->linear_damping = b.linear_damping >= 0.0f ? b.linear_damping : config->damping;
    desc->angular_damping = b.angular_damping >= 0.0f ? b.angular_damping : config->angular_damping;
    desc->gravity_scale = b.gravity_scale;
    desc->flags = kinematic ? DYN_BODY_KINEMATIC : 0u;
    desc->collision_group = b.collision_group;
    desc->collision_mask = b.collision_mask;
    std::memcpy(desc->com, com, sizeof com);
    std::memcpy(desc->inertia, inertia, sizeof inertia);
    std::memcpy(desc->inverse_inertia, inverse, sizeof inverse);
    std::memset(collider, 0, sizeof *collider);
    collider->slot = collider_slot;
    collider->kind = b.shape;
    collider->flags = 2u;  // ContactEventMode::BeginEnd, the ColliderDesc default
    const bool round = b.shape == DYN_SHAPE_SPHERE || b.shape == DYN_SHAPE_CAPSULE || b.shape == DYN_SHAPE_CYLINDER;
    collider->radius = round ? b.radius : 0.0f;
    collider->half_height = (b.shape == DYN_SHAPE_CAPSULE || b.shape == DYN_SHAPE_CYLINDER) ? b.half_height : 0.0f;
    collider->source = b.shape == DYN_SHAPE_HULL ? b.source : 0u;
    collider->relaxation = 0.0f;     // infinite contact frequency
    collider->damping_ratio = 1.0f;  // ColliderDesc default
    collider->impact_force = INFINITY;
    if (b.shape == DYN_SHAPE_CUBOID) std::memcpy(collider->half_extents, b.half_extents, 12);
    collider->collision_group = DYN_NO_COLLISION_FILTER;
    collider->collision_mask = DYN_NO_COLLISION_FILTER;
    collider->local_rotation[3] = 1.0f;
    collider->friction = b.friction;
    collider->restitution = b.restitution;
    collider->rolling_friction = b.rolling_friction;
    collider->spin_friction = b.spin_friction;
    collider->scale[0] = collider->scale[1] = collider->scale[2] = 1.0f;
    return DYN_OK;
}

int dyn_step_params_build(const dyn_config* c, uint32_t body_count, uint32_t dynamic_count,
                          uint32_t collider_count, dyn_step_params* out) {
    if (!c || !out) return dyn::fail(DYN_E_INVALID, "dyn_step_params_build: null argument");
    if (c->substeps == 0 || c->solve_iterations == 0 || c->position_iterations == 0) {
        return dyn::fail(DYN_E_INVALID, "substeps and iterations must be strictly positive");
    }
    if (!(c->relaxation > 0.0f && c->relaxation <= 1.0f)) {
        return dyn::fail(DYN_E_INVALID, "position relaxation must be within (0, 1]");
    }
    std::memset(out, 0, sizeof *out);
    out->gravity[0] = c->gravity[0];
    out->gravity[1] = c->gravity[1];
    out->gravity[2] = c->gravity[2];
    out->damping = c->damping;
    out->angular_damping = c->angular_damping;
    out->substeps = c->substeps;
    out->solve_iterations = c->solve_iterations;
    out->position_iterations = c->position_iterations;
    out->relaxation = c->relaxation;
    out->slop = c->slop;
    out->contact_margin = c->contact_margin;
    out->restitution_threshold = c->restitution_threshold;
    out->max_velocity = c->max_velocity;
    out->max_angular_velocity = c->max_angular_velocity;
    out->sleep_velocity = c->sleep_velocity;
    out->sleep_angular_velocity = c->sleep_angular_velocity;
    out->sleep_time = c->sleep_time;
    out->friction_combine = c->friction_combine;
    out->restitution_combine = c->restitution_combine;
    out->soft_iterations = 1;
    out->soft_substeps = 4;
    out->settle_velocity = 1.0e-4f;
    out->body_count = body_count;
    out->dynamic_count = dynamic_count;
    out->collider_count = collider_count;
    return DYN_OK;
}

void dyn_config_default(dyn_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof *c);
    c->gravity[1] = -9.81f;
    c->damping = 0.05f;
    c->angular_damping = 0.05f;
    c->substeps = 2;
    c->solve_iterations = 8;
    c->position_iterations = 10;
    c->relaxation = 0.8f;
    c->slop = 0.005f;
    c->contact_margin = 0.02f;
    c->restitution_threshold = 1.0f;
    c->max_velocity = 200.0f;
    c->max_angular_velocity = 400.0f;
    c->sleep_velocity = 0.2f;
    c->sleep_angular_velocity = 0.5f;
    c->sleep_time = 0.5f;
    c->friction_combine = 0;     // MaterialCombine::Multiply
    c->restitution_combine = 2;  // MaterialCombine::Max
}

int dyn_encode_bodies(const dyn_body_spec* specs, uint32_t count, const dyn_config* config,
                      uint32_t first_body_id, dyn_body_state* states, dyn_body_desc* descs,
                      dyn_collider* colliders) {
    if (count && (!specs || !states || !descs || !colliders)) {
        return dyn::fail(DYN_E_INVALID, "dyn_encode_bodies: null argument");
    }
    for (uint32_t i = 0; i < count; ++i) {
        const int rc = dyn_encode_body(specs + i, config, first_body_id + i, 0u, states + i, descs + i,
                                       colliders + i);
        if (rc != DYN_OK) {
            return dyn::fail(rc, "dyn_encode_bodies: spec " + std::to_string(i) + ": " + dyn_last_error());
        }
    }
    return DYN_OK;
}

}  // extern "C"
