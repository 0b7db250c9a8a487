"""Host record encoding (dyn_encode_body, dyn_inertia_inverse, dyn_step_params_build) against
closed forms of the mass properties (crates/dynamis-model/src/mass.rs:34-96, :259-280), the
reference's default configuration (crates/dynamis-world/src/config.rs:76-100) and its spawn
bookkeeping (crates/dynamis-world/src/body.rs:108-130, :292-321). Pure host code: no GPU."""

import math

import numpy as np
import pytest

from dynamis_amd import abi
from dynamis_amd._native import DynamisError, call
from dynamis_amd.world import BodyDesc, default_config, encode, spawn_rows


def encode_one(desc, cfg=None):
    scene = encode([desc], cfg)
    return scene.states[0], scene.descs[0], scene.colliders[0]


def test_inertia_inverse_known_answer(oracle):
    tensor = np.float32([1.0, 0.25, 0.0, 2.0, 0.0, 3.0])
    out = np.zeros(6, np.float32)
    call("dyn_inertia_inverse", tensor.ctypes.data, out.ctypes.data)
    expected = np.float32([1.032258, -0.12903225, 0.0, 0.516129, 0.0, 0.33333334])
    assert np.allclose(out, expected, rtol=0, atol=1e-7)
    # the oracle's restatement agrees bit for bit
    assert np.array_equal(out, oracle.inertia_inverse(tensor))
    # and it is the inverse: M * M^-1 = I
    m = np.array([[1, 0.25, 0], [0.25, 2, 0], [0, 0, 3]], np.float64)
    inv = np.array([[out[0], out[1], out[2]], [out[1], out[3], out[4]], [out[2], out[4], out[5]]])
    assert np.allclose(m @ inv, np.eye(3), atol=1e-6)


def test_inertia_inverse_refuses_indefinite_tensor():
    out = np.zeros(6, np.float32)
    with pytest.raises(DynamisError):
        call("dyn_inertia_inverse", np.float32([1, 2, 0, 1, 0, 1]).ctypes.data, out.ctypes.data)


def test_cube_of_density_1000_weighs_a_tonne():
    state, desc, collider = encode_one(BodyDesc.cuboid((0.5, 0.5, 0.5), density=1000.0))
    assert desc["inverse_mass"] == np.float32(1.0 / 1000.0)
    moment = 1000.0 * (1.0 + 1.0) / 12.0
    assert np.allclose(desc["inertia"], [moment, 0, 0, moment, 0, moment], rtol=1e-6)
    assert np.allclose(desc["inverse_inertia"], [1 / moment, 0, 0, 1 / moment, 0, 1 / moment], rtol=1e-6)
    assert collider["kind"] == abi.SHAPE_CUBOID
    assert np.array_equal(collider["half_extents"], np.float32([0.5, 0.5, 0.5]))


def test_box_inertia_closed_form():
    half = (0.5, 1.0, 2.0)
    _, desc, _ = encode_one(BodyDesc.cuboid(half, mass=3.0))
    x, y, z = (2 * h for h in half)
    expected = [3.0 * (y * y + z * z) / 12, 0, 0, 3.0 * (x * x + z * z) / 12, 0, 3.0 * (x * x + y * y) / 12]
    assert np.allclose(desc["inertia"], expected, rtol=1e-6)


def test_sphere_mass_properties():
    _, desc, collider = encode_one(BodyDesc.sphere(0.5, density=2.0))
    mass = 2.0 * 4.0 / 3.0 * math.pi * 0.125
    assert math.isclose(1.0 / desc["inverse_mass"], mass, rel_tol=1e-6)
    assert np.allclose(desc["inertia"][[0, 3, 5]], 0.4 * mass * 0.25, rtol=1e-6)
    assert collider["radius"] == np.float32(0.5)


def test_capsule_mass_properties():
    r, hh, m = 0.3, 0.6, 2.0
    _, desc, collider = encode_one(BodyDesc.capsule(r, hh, mass=m))
    # a cylinder of height 2hh plus two caps that make one sphere, axis along y; the mass
    # splits by volume
    vc = math.pi * r * r * 2 * hh
    vs = 4.0 / 3.0 * math.pi * r ** 3
    mc, ms = m * vc / (vc + vs), m * vs / (vc + vs)
    # about the axis: the cylinder's and the sphere's moments
    axial = mc * r * r / 2 + ms * 2 * r * r / 5
    # across the axis: the cylinder's, and the caps' with the cap offset term the reference
    # uses (mass.rs:57-78 writes 3/4 r hh for the caps' offset)
    across = mc * (r * r / 4 + (2 * hh) ** 2 / 12) + ms * (2 * r * r / 5 + 0.75 * r * hh + hh * hh)
    assert np.allclose(desc["inertia"][[0, 3, 5]], [across, axial, across], rtol=1e-5)
    assert np.array_equal(desc["inertia"][[1, 2, 4]], np.zeros(3, np.float32))
    assert collider["half_height"] == np.float32(hh) and collider["radius"] == np.float32(r)


def test_static_and_kinematic_bodies_carry_no_inverse_mass():
    _, desc, _ = encode_one(BodyDesc.sphere(1.0, mass=0.0))
    assert desc["inverse_mass"] == 0.0
    _, desc, _ = encode_one(BodyDesc.sphere(1.0, kinematic=True))
    assert desc["inverse_mass"] == 0.0 and desc["flags"] & abi.BODY_KINEMATIC


def test_default_configuration_and_params():
    cfg = default_config()[0]
    assert np.allclose(cfg["gravity"], [0, -9.81, 0])
    assert (cfg["substeps"], cfg["solve_iterations"], cfg["position_iterations"]) == (2, 8, 10)
    assert np.isclose(cfg["damping"], 0.05) and np.isclose(cfg["contact_margin"], 0.02)
    scene = encode([BodyDesc.sphere(1.0)], default_config())
    p = scene.params[0]
    assert p["body_count"] == 1 and p["dynamic_count"] == 1 and p["collider_count"] == 1
    assert np.isclose(p["relaxation"], 0.8) and p["friction_combine"] == 0 and p["restitution_combine"] == 2


def test_record_defaults_follow_collider_desc():
    state, desc, collider = encode_one(BodyDesc.sphere(0.5, position=(1, 2, 3)))
    assert np.array_equal(state["position"], np.float32([1, 2, 3]))
    assert np.array_equal(state["prev_position"], np.float32([1, 2, 3]))
    assert np.array_equal(state["orientation"], np.float32([0, 0, 0, 1]))
    assert collider["friction"] == np.float32(0.5) and collider["restitution"] == 0.0
    assert collider["collision_group"] == abi.NO_COLLISION_FILTER
    assert collider["impact_force"] == np.float32(np.inf)
    assert np.isclose(desc["linear_damping"], 0.05)


def test_spawn_rows_swap_simulated_bodies_forward():
    # static S1, S2, then dynamic D1, D2: D1 swaps with row 0 (S1 moves to the tail), then D2
    # swaps with row 1 (S2 moves to the tail)
    assert spawn_rows([False, False, True, True]) == [2, 3, 0, 1]
    assert spawn_rows([True, False, True]) == [0, 2, 1]
    scene = encode([BodyDesc.sphere(1.0, mass=0.0), BodyDesc.sphere(0.5), BodyDesc.sphere(0.5)])
    assert list(scene.rows) == [2, 0, 1]
    assert list(scene.states["body_id"]) == [1, 2, 0]
    assert list(scene.owners) == [2, 0, 1]
