"""Why bodies are thrown: the reference's manifold normal row mass is not positive definite.

solver_contact_block.wgsl:14-28 (manifold_row_mass) solves each normal row of a manifold of m
points against m x (split linear mass) + (I_a^-1 (R_a x n)) . (r_a x n) + (I_b^-1 (R_b x n)) . (r_b x n),
where R is the sum of the manifold's arms and r the point's own arm. When a point's arm points
away from the sum of the arms (a lopsided manifold) the angular terms are negative and can
outweigh the linear term; the Jacobi update (goal - speed) / mass then pushes the wrong way and
every iteration multiplies the error. Four boxes of the differentiation workload (world 142,
start of its 46th step; tests/golden/thrown_boxes.json) show it in one step: the manifold
between the third and fourth box has a point whose row mass is negative (computed here in
FP64 from the formula, independently of either engine), and the step throws both boxes at
about 195 m/s on the oracle and, bit for bit, on the device. The same mechanism throws the
few per cent of bodies the 100k convex pile ejects. It is the reference's algorithm, kept for
parity; nothing in either engine clamps it.
"""

from __future__ import annotations

import json
import pathlib

import numpy as np
import pytest

from dynamis_amd.world import BodyDesc, default_config, encode

FIXTURE = pathlib.Path(__file__).parent / "golden" / "thrown_boxes.json"
DT = np.float32(1.0 / 120.0)


def scene():
    data = json.loads(FIXTURE.read_text())
    cfg = default_config()
    cfg["solve_iterations"] = 8
    cfg["sleep_velocity"] = 0.0
    cfg["sleep_angular_velocity"] = 0.0
    bodies = [BodyDesc.cuboid(tuple(b["half_extents"]), density=1000.0, position=tuple(b["position"]),
                              orientation=tuple(b["orientation"]), velocity=tuple(b["velocity"]),
                              angular_velocity=tuple(b["angular_velocity"]), friction=b["friction"][0])
              for b in data["bodies"]]
    return encode(bodies, cfg)


def rotation(q) -> np.ndarray:
    x, y, z, w = (float(v) for v in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def symmetric(six) -> np.ndarray:
    a = [float(v) for v in six]
    return np.array([[a[0], a[1], a[2]], [a[1], a[3], a[4]], [a[2], a[4], a[5]]])


def row_masses(sc, contact, counts):
    """Per point: (manifold_row_mass with split bodies, point_momentum_mass), in FP64."""
    rows = [int(sc.owners[contact["a"]]), int(sc.owners[contact["b"]])]
    n = contact["normal"].astype(np.float64)
    pc = int(contact["point_count"])
    pose, inv_world, inv_mass = [], [], []
    for r in rows:
        st = sc.states[r]
        R = rotation(st["orientation"])
        pose.append((R, st["position"].astype(np.float64)))
        inv_world.append(R @ symmetric(sc.descs[r]["inverse_inertia"]) @ R.T)
        inv_mass.append(float(sc.descs[r]["inverse_mass"]))
    pts = contact["points"]
    arms = [[pose[0][0] @ pts["local_a"][p].astype(np.float64) for p in range(pc)],
            [pose[1][0] @ pts["local_b"][p].astype(np.float64) for p in range(pc)]]
    reach = [sum(arms[0]), sum(arms[1])]
    out = []
    for p in range(pc):
        split = [counts[rows[0]], counts[rows[1]]]
        row = pc * (split[0] * inv_mass[0] + split[1] * inv_mass[1])
        truth = inv_mass[0] + inv_mass[1]
        for side in (0, 1):
            ra, rr = np.cross(arms[side][p], n), np.cross(reach[side], n)
            row += split[side] * float(np.dot(inv_world[side] @ rr, ra))
            truth += float(np.dot(inv_world[side] @ ra, ra))
        out.append((row, truth))
    return out


def throw(world_factory):
    sc = scene()
    world = world_factory(sc)
    world.step(DT)
    return sc, world


def test_lopsided_manifold_has_a_negative_row_mass(oracle):
    sc, world = throw(lambda s: oracle.world(s))
    contacts = world.inspect_contacts()
    owners = sc.owners
    counts = np.zeros(len(sc.states), np.int64)
    for c in contacts:
        counts[owners[c["a"]]] += 1
        counts[owners[c["b"]]] += 1
    pair = [c for c in contacts if {int(owners[c["a"]]), int(owners[c["b"]])} == {2, 3}]
    assert len(pair) == 1 and int(pair[0]["point_count"]) == 4
    masses = row_masses(sc, pair[0], counts)
    assert all(truth > 0 for _, truth in masses)  # the point's own momentum mass is positive
    assert min(row for row, _ in masses) < 0.0, masses
    # the same manifold with unsplit bodies (no other contacts) stays positive: the split
    # counts of the third box's other two contacts tip it
    assert min(row for row, _ in row_masses(sc, pair[0], np.ones(len(sc.states), np.int64))) > 0.0
    speeds = np.linalg.norm(world.states()["velocity"][:4].astype(np.float64), axis=1)
    assert speeds[2] > 150.0 and speeds[3] > 150.0, speeds
    assert speeds[0] < 5.0 and speeds[1] < 5.0, speeds


@pytest.mark.gpu
def test_device_throws_them_bit_for_bit(device, oracle):
    from dynamis_amd.world import World

    def dev(s):
        w = World()
        w.create(s)
        return w
    _, ref = throw(lambda s: oracle.world(s))
    _, got = throw(dev)
    assert got.states().tobytes() == ref.states().tobytes()
    assert got.inspect_contacts().tobytes() == ref.inspect_contacts().tobytes()
    got.close()
