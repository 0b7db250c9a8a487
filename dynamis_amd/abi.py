"""numpy views of the device records and constants of include/dyn_abi.h (the reference's
crates/dynamis-abi/abi/records.wgsl layouts and src/constant.rs / src/counter.rs values)."""

from __future__ import annotations

import numpy as np

AABB = np.dtype([("min", "<f4", 3), ("_pad0", "<f4"), ("max", "<f4", 3), ("_pad1", "<f4")])
GRID_ENTRY = np.dtype([("min", "<f4", 3), ("group", "<u4"), ("max", "<f4", 3), ("info", "<u4")])

assert AABB.itemsize == 32 and GRID_ENTRY.itemsize == 32

# constant.rs:67-87
MAX_CELLS_PER_AXIS = 2
MAX_CELLS_PER_COLLIDER = 8
ENTRY_INDEX_MASK = (1 << 24) - 1
ENTRY_CELL_SHIFT = 24
ENTRY_CELL_MASK = 7 << 24
ENTRY_KIND_SHIFT = 27
ENTRY_KIND_MASK = 3 << 27
ENTRY_MOBILE = 1 << 29
ENTRY_AWAKE = 1 << 30
ENTRY_PRIMARY = 1 << 31
LEVEL_KEY_SHIFT = 27
CELL_HASH_MASK = 0x07FF_FFFF

BP_MOBILE = 1
BP_AWAKE = 2

# counter.rs:75-128 slot order
COUNTER_NAMES = [
    "ENTRIES", "RESTING_ENTRIES", "RESTING_LEVELS", "RESTING_REBUILD", "RESTING_SORT",
    "AWAKE_BASE", "IMMOVABLE_ENTRIES", "IMMOVABLE_EMITTED", "IMMOVABLE_LEVELS", "ENTRY_BASE",
    "PAIRS", "GRID_LEVELS", "CONTACTS", "ARCHIVED", "JOINTS", "EVENTS", "REFUSED_PAIRS",
    "REFUSED_EVENTS", "ENTRY_FAULTS", "RESTING", "REFUSED_RESTING", "SLEPT", "WOKE", "ACTIVE",
    "WOKE_DEFERRED", "RESTING_GATHER", "RESTING_INDEX", "RESTING_PENDING", "ISLANDS",
    "ISLAND_FAULTS", "BLOCKS", "SOLVER_ROWS", "SOLVE_LINEAR_RESIDUAL",
    "SOLVE_ANGULAR_RESIDUAL", "COARSE_ACTIVE", "GRID_SCALE", "GRID_EXTENT", "BREAKS",
    "PARTICLE_REACH", "COARSE_NEIGHBOURS", "LIVE", "LIVE_FAULTS", "ROW_FAULTS", "SOFT_ACTIVE",
    "SOFT_SLEPT", "SOFT_WOKE", "REFUSED_CONTACTS", "STEP", "IMPACTS", "REFUSED_IMPACTS",
    "SOFT_EVENTS", "REFUSED_SOFT_EVENTS",
]
COUNTER = {name: slot for slot, name in enumerate(COUNTER_NAMES)}
# The counters that refuse work, with their labels and shortfall classes (counter.rs:75-128):
# "physics" refusals change the step's result, "report" ones only what it reports, "fatal" ones
# are faults a step cannot lose.
REFUSING_COUNTERS = {
    "REFUSED_PAIRS": ("candidate pairs beyond the pair stream", "physics"),
    "REFUSED_EVENTS": ("contact events beyond the event stream", "report"),
    "ENTRY_FAULTS": ("grid entries beyond their budget", "fatal"),
    "REFUSED_RESTING": ("resting manifolds beyond the resting archive", "physics"),
    "ISLAND_FAULTS": ("island unions beyond their hook budget", "fatal"),
    "LIVE_FAULTS": ("live body rows beyond the row stream", "fatal"),
    "ROW_FAULTS": ("row moves that answer no state their row could hold", "fatal"),
    "REFUSED_CONTACTS": ("contact manifolds beyond the contact stream", "physics"),
    "REFUSED_IMPACTS": ("impacts beyond the impact stream", "report"),
    "REFUSED_SOFT_EVENTS": ("soft contact events beyond the soft event stream", "report"),
}
COUNTER_DEVICE_COUNT = len(COUNTER_NAMES)
COUNTER_STRIDE_WORDS = 64
COUNTER_BLOCK_WORDS = COUNTER_DEVICE_COUNT * COUNTER_STRIDE_WORDS


def counter(block: np.ndarray, name: str) -> int:
    """Value of a named counter in a downloaded counter block."""
    return int(block[COUNTER[name] * COUNTER_STRIDE_WORDS])


def aabbs(mins: np.ndarray, maxs: np.ndarray) -> np.ndarray:
    out = np.zeros(len(mins), dtype=AABB)
    out["min"] = mins
    out["max"] = maxs
    return out


# records.wgsl:102-143 StepParams (176 bytes)
STEP_PARAMS = np.dtype([
    ("gravity", "<f4", 4), ("dt", "<f4"), ("substep_dt", "<f4"), ("damping", "<f4"),
    ("angular_damping", "<f4"), ("body_count", "<u4"), ("substeps", "<u4"),
    ("solve_iterations", "<u4"), ("constraint_count", "<u4"), ("relaxation", "<f4"),
    ("slop", "<f4"), ("restitution_threshold", "<f4"), ("max_velocity", "<f4"),
    ("max_angular_velocity", "<f4"), ("dynamic_count", "<u4"), ("sleep_velocity", "<f4"),
    ("sleep_angular_velocity", "<f4"), ("sleep_time", "<f4"), ("friction_combine", "<u4"),
    ("restitution_combine", "<u4"), ("position_iterations", "<u4"), ("contact_margin", "<f4"),
    ("collider_count", "<u4"), ("soft_iterations", "<u4"), ("soft_substeps", "<u4"),
    ("particle_count", "<u4"), ("element_count", "<u4"), ("soft_substep_dt", "<f4"),
    ("soft_body_count", "<u4"), ("attachment_count", "<u4"), ("settle_velocity", "<f4"),
    ("observed_count", "<u4"), ("observed_joint_count", "<u4"), ("character_count", "<u4"),
    ("vehicle_count", "<u4"), ("vehicle_wheel_count", "<u4"), ("wake_all", "<u4"),
    ("field_count", "<u4"), ("_pad_tail", "<u4", 3),
])

# records.wgsl:145-164 BodyState (144 bytes)
BODY_STATE = np.dtype([
    ("position", "<f4", 3), ("_pad0", "<f4"), ("prev_position", "<f4", 3), ("_pad1", "<f4"),
    ("prev_orientation", "<f4", 4), ("orientation", "<f4", 4), ("velocity", "<f4", 3),
    ("_pad2", "<f4"), ("angular_velocity", "<f4", 3), ("_pad3", "<f4"), ("force", "<f4", 3),
    ("_pad4", "<f4"), ("torque", "<f4", 3), ("_pad5", "<f4"), ("body_id", "<u4"),
    ("generation", "<u4"), ("sleep_timer", "<f4"), ("sleeping", "<u4"),
])

# records.wgsl:166-184 BodyDescriptor (128 bytes)
BODY_DESC = np.dtype([
    ("inverse_mass", "<f4"), ("linear_damping", "<f4"), ("angular_damping", "<f4"),
    ("gravity_scale", "<f4"), ("sleep_velocity", "<f4"), ("sleep_angular_velocity", "<f4"),
    ("flags", "<u4"), ("_pad0", "<u4"), ("collision_group", "<u4"), ("collision_mask", "<u4"),
    ("_pad1", "<u4"), ("_pad4", "<u4"), ("com", "<f4", 3), ("_pad2", "<f4"),
    ("inertia", "<f4", 6), ("inverse_inertia", "<f4", 6), ("_pad3", "<f4", 4),
])

# records.wgsl:208-228 Collider (112 bytes)
COLLIDER = np.dtype([
    ("slot", "<u4"), ("kind", "<u4"), ("flags", "<u4"), ("radius", "<f4"),
    ("half_height", "<f4"), ("relaxation", "<f4"), ("damping_ratio", "<f4"),
    ("impact_force", "<f4"), ("half_extents", "<f4", 3), ("collision_group", "<u4"),
    ("local_offset", "<f4", 3), ("collision_mask", "<u4"), ("local_rotation", "<f4", 4),
    ("friction", "<f4"), ("restitution", "<f4"), ("source", "<u4"),
    ("rolling_friction", "<f4"), ("scale", "<f4", 3), ("spin_friction", "<f4"),
])

# records.wgsl:313-322 ManifoldPoint (64 bytes)
MANIFOLD_POINT = np.dtype([
    ("position", "<f4", 3), ("depth", "<f4"), ("local_a", "<f4", 3), ("_pad_a", "<f4"),
    ("local_b", "<f4", 3), ("accumulated_normal", "<f4"), ("accumulated_tangent_1", "<f4"),
    ("accumulated_tangent_2", "<f4"), ("feature", "<u4"), ("_pad_tail", "<u4"),
])

# records.wgsl:331-353 Contact (352 bytes)
CONTACT = np.dtype([
    ("a", "<u4"), ("b", "<u4"), ("point_count", "<u4"), ("sensor", "<u4"),
    ("first_body_id", "<u4"), ("second_body_id", "<u4"), ("first_generation", "<u4"),
    ("second_generation", "<u4"), ("normal", "<f4", 3), ("events", "<u4"), ("surface", "<u4"),
    ("triangle", "<u4"), ("friction", "<f4"), ("restitution", "<f4"),
    ("rolling_friction", "<f4"), ("spin_friction", "<f4"), ("carried_normal", "<f4"),
    ("carried_tangent", "<f4"), ("relaxation", "<f4"), ("damping_ratio", "<f4"),
    ("_pad_head", "<u4", 2), ("points", MANIFOLD_POINT, 4),
])

# PhysicsConfig mirror (crates/dynamis-world/src/config.rs:76-100)
CONFIG = np.dtype([
    ("gravity", "<f4", 3), ("damping", "<f4"), ("angular_damping", "<f4"), ("substeps", "<u4"),
    ("solve_iterations", "<u4"), ("position_iterations", "<u4"), ("relaxation", "<f4"),
    ("slop", "<f4"), ("contact_margin", "<f4"), ("restitution_threshold", "<f4"),
    ("max_velocity", "<f4"), ("max_angular_velocity", "<f4"), ("sleep_velocity", "<f4"),
    ("sleep_angular_velocity", "<f4"), ("sleep_time", "<f4"), ("friction_combine", "<u4"),
    ("restitution_combine", "<u4"),
])

# A BodyDesc with one collider at the body origin (model/body.rs:89-108)
BODY_SPEC = np.dtype([
    ("shape", "<u4"), ("radius", "<f4"), ("half_height", "<f4"), ("half_extents", "<f4", 3),
    ("position", "<f4", 3), ("orientation", "<f4", 4), ("velocity", "<f4", 3),
    ("angular_velocity", "<f4", 3), ("mass", "<f4"), ("density", "<f4"), ("friction", "<f4"),
    ("restitution", "<f4"), ("rolling_friction", "<f4"), ("spin_friction", "<f4"),
    ("gravity_scale", "<f4"), ("linear_damping", "<f4"), ("angular_damping", "<f4"),
    ("kinematic", "<u4"), ("collision_group", "<u4"), ("collision_mask", "<u4"),
    ("source", "<u4"), ("solid_volume", "<f4"), ("solid_centroid", "<f4", 3),
    ("solid_second", "<f4", 6),
])

# records.wgsl:256-273 ShapeSource (80 bytes)
SHAPE_SOURCE = np.dtype([
    ("kind", "<u4"), ("vertex_offset", "<u4"), ("vertex_count", "<u4"),
    ("triangle_offset", "<u4"), ("node_offset", "<u4"), ("node_count", "<u4"),
    ("cell_offset", "<u4"), ("cell_count", "<u4"), ("grid_rows", "<u4"), ("grid_cols", "<u4"),
    ("_pad0", "<u4"), ("_pad1", "<u4"), ("local_min", "<f4", 3), ("_pad2", "<f4"),
    ("local_max", "<f4", 3), ("_pad3", "<f4"),
])
assert SHAPE_SOURCE.itemsize == 80
LINEAR_SUPPORT_VERTICES = 4096

assert STEP_PARAMS.itemsize == 176 and BODY_STATE.itemsize == 144
assert BODY_DESC.itemsize == 128 and COLLIDER.itemsize == 112
assert MANIFOLD_POINT.itemsize == 64 and CONTACT.itemsize == 352

# records.wgsl:193-199 BodyEdit (160 bytes)
BODY_EDIT = np.dtype([("kind", "<u4"), ("mask", "<u4"), ("_pad", "<u4", 2), ("state", BODY_STATE)])
assert BODY_EDIT.itemsize == 160

# body edit kinds and patch fields (crates/dynamis-abi/src/constant.rs:6-18)
EDIT_PATCH = 0
EDIT_FORCE = 1
EDIT_TORQUE = 2
EDIT_IMPULSE = 3
EDIT_SLEEP = 4
EDIT_WAKE = 5
EDIT_FORCE_AT_POINT = 6
EDIT_IMPULSE_AT_POINT = 7
EDIT_ANGULAR_IMPULSE = 8
PATCH_POSITION = 1
PATCH_VELOCITY = 2
PATCH_ORIENTATION = 4
PATCH_ANGULAR_VELOCITY = 8

# shape kinds and body flags, the values of include/dyn_abi.h
SHAPE_SPHERE = 1
SHAPE_CUBOID = 2
SHAPE_CAPSULE = 3
SHAPE_CYLINDER = 4
SHAPE_HULL = 5
SHAPE_PLANE = 9
BODY_KINEMATIC = 1
NO_COLLISION_FILTER = 0xFFFF_FFFF

# fixed-point scales of the solver's delta words (core.wgsl:6-12)
SOLVER_VELOCITY_SCALE = 262144.0
SOLVER_POSITION_SCALE = 16777216.0
