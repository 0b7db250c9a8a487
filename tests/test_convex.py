"""Hull shape sources and the GJK / EPA / feature-clip narrowphase (convex.wgsl), on both step
engines. The hull builder is pinned against qhull (scipy) and closed-form solids; the convex
contacts against what geometry fixes: overlap depths of boxes that SAT gives exactly,
separations measured by a brute-force support-function search, resting heights on planes
(crates/dynamis-world/tests/suite/plane.rs:112-273) and four-point face manifolds.
"""

from __future__ import annotations

import numpy as np
import pytest
from scipy.spatial import ConvexHull

from dynamis_amd import abi
from dynamis_amd.world import BodyDesc, build_hull, encode
from tests.worlds import DT, config, gravity_config, qrot, static_config

CORNERS = np.array([[x, y, z] for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (-0.5, 0.5)],
                   np.float32)


def ellipsoid_points(rng, n, axes):
    p = rng.normal(size=(n, 3))
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    return (p * np.asarray(axes)).astype(np.float32)


# ---- hull builder and solid (host code of the library, no GPU) ------------------------------

def test_cube_hull_is_the_box():
    rng = np.random.default_rng(0)
    pts = np.concatenate([CORNERS, rng.uniform(-0.4, 0.4, (20, 3)).astype(np.float32)])
    h = build_hull(pts)
    assert len(h.vertices) == 8 and len(h.triangles) == 12
    assert h.volume == pytest.approx(1.0, rel=1e-6)
    assert np.allclose(h.centroid, 0.0, atol=1e-6)
    assert np.allclose(h.second, [1 / 12, 0, 0, 1 / 12, 0, 1 / 12], atol=1e-7)


@pytest.mark.parametrize("n", [16, 32, 64, 300])
def test_hull_matches_qhull(n):
    rng = np.random.default_rng(n)
    pts = ellipsoid_points(rng, n, rng.uniform(0.3, 0.9, 3))
    h = build_hull(pts)
    ch = ConvexHull(pts.astype(np.float64))
    assert set(map(tuple, h.vertices.tolist())) == set(map(tuple, pts[ch.vertices].tolist()))
    assert h.volume == pytest.approx(ch.volume, rel=1e-5)
    # outward triangles: every face normal points away from the interior
    inner = h.vertices.mean(axis=0)
    for t in h.triangles:
        a, b, c = h.vertices[t].astype(np.float64)
        assert np.dot(np.cross(b - a, c - a), a - inner) > 0


def test_hull_solid_of_a_shifted_tetrahedron():
    # a right tetrahedron has volume 1/6 and centroid (1/4,1/4,1/4), shifted by (1,2,3)
    shift = np.float32([1, 2, 3])
    pts = np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]) + shift
    h = build_hull(pts)
    assert h.volume == pytest.approx(1 / 6, rel=1e-6)
    assert np.allclose(h.centroid, shift + 0.25, atol=1e-6)
    # unit second moments about the centroid: xx = 3/80, xy = -1/80 (standard tetrahedron result)
    assert np.allclose(h.second, [3 / 80, -1 / 80, -1 / 80, 3 / 80, -1 / 80, 3 / 80], atol=1e-6)


def test_degenerate_points_are_refused():
    from dynamis_amd._native import DynamisError
    with pytest.raises(DynamisError):
        build_hull(np.float32([[0, 0, 0], [1, 0, 0], [2, 0, 0], [3, 0, 0]]))
    with pytest.raises(DynamisError):
        build_hull(np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [2, 3, 0]]))


def test_hull_body_mass_properties():
    shift = np.float32([0.2, -0.1, 0.05])
    h = build_hull(CORNERS + shift)
    scene = encode([BodyDesc.hull(0, density=1000.0)], config(), [h])
    d = scene.descs[0]
    assert 1.0 / d["inverse_mass"] == pytest.approx(1000.0, rel=1e-6)
    assert np.allclose(d["com"], shift, atol=1e-6)
    assert np.allclose(d["inertia"], [1000 / 6, 0, 0, 1000 / 6, 0, 1000 / 6], rtol=1e-5, atol=1e-3)
    assert scene.colliders[0]["kind"] == abi.SHAPE_HULL and scene.sources[0]["vertex_count"] == 8


# ---- contacts through the GJK path, on both engines -------------------------------------------

@pytest.fixture(params=["oracle", pytest.param("device", marks=pytest.mark.gpu)])
def engine(request):
    if request.param == "oracle":
        oracle = request.getfixturevalue("oracle")
        return lambda scene: oracle.world(scene)
    request.getfixturevalue("device")
    from dynamis_amd.world import World

    def make(scene):
        world = World()
        world.create(scene)
        return world
    return make


def one_step_contacts(engine, bodies, hulls, cfg=None):
    scene = encode(bodies, static_config() if cfg is None else cfg, hulls)
    world = engine(scene)
    world.step(DT)
    return scene, world, world.inspect_contacts()


def test_overlapping_hull_cubes_meet_face_to_face(engine):
    cube = build_hull(CORNERS)
    bodies = [BodyDesc.hull(0, position=(0.0, 0.0, 0.0), mass=0.0),
              BodyDesc.hull(0, position=(0.9, 0.0, 0.0))]
    scene, world, contacts = one_step_contacts(engine, bodies, [cube])
    assert len(contacts) == 1
    c = contacts[0]
    assert np.allclose(np.abs(c["normal"]), [1, 0, 0], atol=1e-5)
    assert c["point_count"] == 4
    # SAT fixes the overlap: 0.1 along x
    assert np.allclose(c["points"]["depth"][:4], 0.1, atol=1e-4)


def brute_separation(va, vb):
    """The largest support-function separation of two vertex sets over directions: a sphere
    grid refined around its best direction (the distance of two convex sets when positive)."""
    dirs = ellipsoid_points(np.random.default_rng(1), 20000, (1, 1, 1)).astype(np.float64)

    def separation(n):
        return (vb @ n).min() - (va @ n).max()
    best = max(dirs, key=separation)
    for scale in (0.05, 0.01, 0.002, 0.0005):
        cand = best + np.random.default_rng(2).normal(scale=scale, size=(4000, 3))
        cand /= np.linalg.norm(cand, axis=1, keepdims=True)
        best = max(list(cand) + [best], key=separation)
    return separation(best)


@pytest.mark.parametrize("gap", [0.005, 0.05, 0.3])
def test_gjk_distance_matches_brute_force(oracle, gap):
    # convex_closest (convex.wgsl:464-521) against a brute-force support-function search
    rng = np.random.default_rng(7)
    a = build_hull(ellipsoid_points(rng, 40, (0.5, 0.4, 0.3)))
    b = build_hull(ellipsoid_points(rng, 40, (0.3, 0.5, 0.4)))
    va = a.vertices.astype(np.float64)
    lo = float(a.vertices[:, 0].max() - b.vertices[:, 0].min()) - 0.3
    hi = lo + 1.3
    for _ in range(30):
        mid = 0.5 * (lo + hi)
        d = brute_separation(va, b.vertices.astype(np.float64) + [mid, 0, 0])
        lo, hi = (mid, hi) if d < gap else (lo, mid)
    offset = float(np.float32(0.5 * (lo + hi)))
    distance = brute_separation(va, b.vertices.astype(np.float64) + [offset, 0, 0])
    scene = encode([BodyDesc.hull(0, mass=0.0), BodyDesc.hull(1, position=(offset, 0.5, 0.0))],
                   static_config(), [a, b])
    # move b back to y = 0 in the state records (the query reads the states as given)
    states = scene.states.copy()
    states[scene.rows[1]]["position"][1] = 0.0
    scene.states = states
    q = oracle.convex_query(scene, scene.rows[0], scene.rows[1])
    assert not q["penetrating"]
    # |v| is a point of the Minkowski difference, so an upper bound of the distance, and the
    # stop rule (support gain <= 1e-4 max(1, |v|^2)) bounds its excess by 1e-4 / |v|
    assert distance - 1e-5 <= q["distance"] <= distance + 1e-4 / distance + 1e-5
    assert np.linalg.norm(q["point_b"] - q["point_a"]) == pytest.approx(q["distance"], abs=1e-5)


@pytest.mark.parametrize("overlap", [0.02, 0.1, 0.3])
def test_epa_depth_of_boxes_is_the_sat_depth(oracle, overlap):
    # two unit cube hulls overlapping along a tilted axis: EPA's depth equals the smallest
    # face-axis overlap, which is what SAT gives for axis-aligned boxes
    cube = build_hull(CORNERS)
    offset = (1.0 - overlap, 0.3, -0.2)
    scene = encode([BodyDesc.hull(0, mass=0.0), BodyDesc.hull(0, position=offset)], static_config(),
                   [cube])
    q = oracle.convex_query(scene, scene.rows[0], scene.rows[1])
    # EPA when GJK ends inside the Minkowski difference, the direction probe otherwise
    # (convex_hit, convex.wgsl:584-618): either way the depth is the face overlap
    assert -q["hit_distance"] == pytest.approx(overlap, abs=1e-4)
    assert np.allclose(q["hit_normal"], [1, 0, 0], atol=1e-4)


def test_sphere_and_capsule_hulls_by_support_maps(oracle):
    # analytic shapes through the GJK path: sphere-sphere distance is |c1 - c2| - r1 - r2
    bodies = [BodyDesc.sphere(0.5, mass=0.0), BodyDesc.sphere(0.25, position=(0.3, 1.0, 0.4))]
    scene = encode(bodies, static_config())
    q = oracle.convex_query(scene, scene.rows[0], scene.rows[1])
    expected = float(np.linalg.norm([0.3, 1.0, 0.4])) - 0.75
    assert q["distance"] == pytest.approx(expected, abs=1e-4)
    # capsule (vertical, hh 0.5, r 0.2) to a sphere beside its side: horizontal gap
    bodies = [BodyDesc.capsule(0.2, 0.5, mass=0.0), BodyDesc.sphere(0.3, position=(1.0, 0.2, 0.0))]
    scene = encode(bodies, static_config())
    q = oracle.convex_query(scene, scene.rows[0], scene.rows[1])
    assert q["distance"] == pytest.approx(1.0 - 0.2 - 0.3, abs=1e-4)


def rest_height(engine, desc, hulls=(), frames=180):
    scene = encode([BodyDesc.plane(), desc], config(), list(hulls))
    world = engine(scene)
    for _ in range(frames):
        world.step(DT)
    return world.states()[scene.rows[1]]


def test_primitives_rest_on_a_plane(engine):
    # plane.rs:112-132: every solid primitive rests on a plane at its own half height
    cube = build_hull(CORNERS)
    cases = [(BodyDesc.hull(0, position=(0, 1.0, 0)), 0.5, [cube]),
             (BodyDesc.cylinder(0.4, 0.3, position=(0, 1.0, 0)), 0.3, []),
             (BodyDesc.capsule(0.3, 0.4, position=(0, 1.5, 0)), 0.7, [])]
    for desc, height, hulls in cases:
        s = rest_height(engine, desc, hulls)
        assert abs(s["position"][1] - height) <= 0.02 + 0.005, (desc.shape, s["position"])


def test_cylinder_meets_a_plane_with_its_cap_ring(engine):
    # plane.rs:230-257: an upright cylinder touches with its cap ring (four kept points)
    scene, world, contacts = one_step_contacts(
        engine, [BodyDesc.plane(), BodyDesc.cylinder(0.5, 0.5, position=(0, 0.49, 0))], [], config())
    assert len(contacts) == 1 and contacts[0]["point_count"] == 4
    assert np.allclose(np.abs(contacts[0]["normal"]), [0, 1, 0])


def test_capsule_lies_on_a_box(engine):
    # capsule-box pairs go through GJK/EPA and the clipped box face
    bodies = [BodyDesc.cuboid((5.0, 0.5, 5.0), mass=0.0, position=(0, -0.5, 0)),
              BodyDesc.capsule(0.25, 0.5, position=(0, 0.3, 0),
                               orientation=(0.0, 0.0, 0.70710677, 0.70710677))]
    scene = encode(bodies, config())
    world = engine(scene)
    for frame in range(120):
        world.step(DT)
        if frame == 20:
            # a lying capsule touches the box face along its segment: two points
            contacts = world.inspect_contacts()
            assert len(contacts) == 1 and contacts[0]["point_count"] == 2
    s = world.states()[scene.rows[1]]
    assert abs(s["position"][1] - 0.25) <= 0.03


def test_hull_pile_settles(engine):
    rng = np.random.default_rng(3)
    hulls = [build_hull(ellipsoid_points(rng, int(rng.integers(16, 65)), rng.uniform(0.3, 0.9, 3)))
             for _ in range(6)]
    bodies = [BodyDesc.plane()]
    for i in range(24):
        bodies.append(BodyDesc.hull(i % 6, density=1000.0,
                                    position=(float((i % 4) * 2.0 - 3), float(1.5 + (i // 4) * 2.0),
                                              float(((i // 2) % 2) * 2.0 - 1))))
    scene = encode(bodies, config(), hulls)
    world = engine(scene)
    for _ in range(240):
        world.step(DT)
    st = world.states()[:scene.dynamic_count]
    # every hull ends above the plane (lowest vertex within the margin and slop) and slow
    for row in range(scene.dynamic_count):
        handle = scene.handles[row]
        h = hulls[bodies[handle].source]
        low = min(qrot(st[row]["orientation"], v)[1] for v in h.vertices) + st[row]["position"][1]
        assert low > -(0.02 + 0.005 + 0.01), (row, low)
    # ellipsoid-like hulls may still roll, but nothing is thrown about
    speeds = np.linalg.norm(st["velocity"], axis=1)
    assert speeds.max() < 3.0 and float(np.median(speeds)) < 0.5
    m = world.measured()
    assert m["REFUSED_CONTACTS"] == 0 and m["ISLAND_FAULTS"] == 0
