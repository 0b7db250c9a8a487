"""The device records' layouts against the reference's WGSL structs (crates/dynamis-abi/abi/
records.wgsl): every named field of the reference struct sits at the offset WGSL's layout rules
give it (f32/u32 size 4 align 4; vec3f size 12 align 16; vec4f size 16 align 16; array<T, N> of
the element's stride; a struct aligned to its widest member and sized to a multiple of it), and
the records are the same size. The field lists below restate the structs, with their lines."""

from __future__ import annotations

import numpy as np
import pytest

from dynamis_amd import abi

SCALAR = ("f32", 4, 4)
VEC3 = ("vec3f", 12, 16)
VEC4 = ("vec4f", 16, 16)


def array(element, count):
    kind, size, align = element
    stride = -(-size // align) * align
    return ("array", stride * count, align)


def layout(fields):
    """WGSL offsets of (name, type) fields, the struct size and alignment."""
    offsets, at, widest = {}, 0, 4
    for name, (_, size, align) in fields:
        at = -(-at // align) * align
        offsets[name] = at
        at += size
        widest = max(widest, align)
    return offsets, -(-at // widest) * widest, widest


def struct(fields):
    offsets, size, align = layout(fields)
    return ("struct", size, align), offsets


# records.wgsl:145-164
BODY_STATE = [("position", VEC3), ("_pad0", SCALAR), ("prev_position", VEC3), ("_pad1", SCALAR),
              ("prev_orientation", VEC4), ("orientation", VEC4), ("velocity", VEC3), ("_pad2", SCALAR),
              ("angular_velocity", VEC3), ("_pad3", SCALAR), ("force", VEC3), ("_pad4", SCALAR),
              ("torque", VEC3), ("_pad5", SCALAR), ("body_id", SCALAR), ("generation", SCALAR),
              ("sleep_timer", SCALAR), ("sleeping", SCALAR)]
# records.wgsl:166-183
BODY_DESC = [("inverse_mass", SCALAR), ("linear_damping", SCALAR), ("angular_damping", SCALAR),
             ("gravity_scale", SCALAR), ("sleep_velocity", SCALAR), ("sleep_angular_velocity", SCALAR),
             ("flags", SCALAR), ("_pad0", SCALAR), ("collision_group", SCALAR), ("collision_mask", SCALAR),
             ("_pad1", SCALAR), ("_pad4", SCALAR), ("com", VEC3), ("_pad2", SCALAR),
             ("inertia", array(SCALAR, 6)), ("inverse_inertia", array(SCALAR, 6)), ("_pad3", array(SCALAR, 4))]
# records.wgsl:208-228
COLLIDER = [("slot", SCALAR), ("kind", SCALAR), ("flags", SCALAR), ("radius", SCALAR), ("half_height", SCALAR),
            ("relaxation", SCALAR), ("damping_ratio", SCALAR), ("impact_force", SCALAR), ("half_extents", VEC3),
            ("collision_group", SCALAR), ("local_offset", VEC3), ("collision_mask", SCALAR),
            ("local_rotation", VEC4), ("friction", SCALAR), ("restitution", SCALAR), ("source", SCALAR),
            ("rolling_friction", SCALAR), ("scale", VEC3), ("spin_friction", SCALAR)]
# records.wgsl:299-304, 306-311
AABB = [("min", VEC3), ("_pad0", SCALAR), ("max", VEC3), ("_pad1", SCALAR)]
GRID_ENTRY = [("min", VEC3), ("group", SCALAR), ("max", VEC3), ("info", SCALAR)]
# records.wgsl:313-322
MANIFOLD_POINT = [("position", VEC3), ("depth", SCALAR), ("local_a", VEC3), ("local_b", VEC3),
                  ("accumulated_normal", SCALAR), ("accumulated_tangent_1", SCALAR),
                  ("accumulated_tangent_2", SCALAR), ("feature", SCALAR)]
MANIFOLD_TYPE, _ = struct(MANIFOLD_POINT)
# records.wgsl:331-352 (CONTACT_MAX_POINTS = 4)
CONTACT = [("a", SCALAR), ("b", SCALAR), ("point_count", SCALAR), ("sensor", SCALAR),
           ("first_body_id", SCALAR), ("second_body_id", SCALAR), ("first_generation", SCALAR),
           ("second_generation", SCALAR), ("normal", VEC3), ("events", SCALAR), ("surface", SCALAR),
           ("triangle", SCALAR), ("friction", SCALAR), ("restitution", SCALAR), ("rolling_friction", SCALAR),
           ("spin_friction", SCALAR), ("carried_normal", SCALAR), ("carried_tangent", SCALAR),
           ("relaxation", SCALAR), ("damping_ratio", SCALAR), ("points", array(MANIFOLD_TYPE, 4))]
# records.wgsl:193-199
BODY_STATE_TYPE, _ = struct(BODY_STATE)
BODY_EDIT = [("kind", SCALAR), ("mask", SCALAR), ("_pad0", SCALAR), ("_pad1", SCALAR), ("state", BODY_STATE_TYPE)]

CASES = [("BodyState", BODY_STATE, abi.BODY_STATE, 144), ("BodyDescriptor", BODY_DESC, abi.BODY_DESC, 128),
         ("Collider", COLLIDER, abi.COLLIDER, 112), ("Aabb", AABB, abi.AABB, 32),
         ("GridEntry", GRID_ENTRY, abi.GRID_ENTRY, 32), ("ManifoldPoint", MANIFOLD_POINT, abi.MANIFOLD_POINT, 64),
         ("Contact", CONTACT, abi.CONTACT, 352), ("BodyEdit", BODY_EDIT, abi.BODY_EDIT, 160)]


@pytest.mark.parametrize("name,fields,dtype,size", CASES, ids=[c[0] for c in CASES])
def test_named_fields_sit_at_the_wgsl_offsets(name, fields, dtype, size):
    offsets, wgsl_size, _ = layout(fields)
    assert wgsl_size == size == dtype.itemsize
    for field, _ in fields:
        if field.startswith("_pad"):
            continue
        assert field in dtype.names, (name, field)
        assert dtype.fields[field][1] == offsets[field], (name, field, dtype.fields[field][1], offsets[field])


def test_manifold_points_inside_the_contact():
    offsets, _, _ = layout(CONTACT)
    point = abi.CONTACT.fields["points"][0]
    assert point.base == abi.MANIFOLD_POINT or point.base.itemsize == 64
    assert abi.CONTACT.fields["points"][1] == offsets["points"] == 96
    assert np.prod(point.shape) == 4
