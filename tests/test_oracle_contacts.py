"""Known answers for the oracle's analytic contact generators and its warm-start relay.

The device follows the oracle bit for bit, so a mistake the two restatements shared would pass
every parity test; these cases pin the oracle on its own. Each expected value is worked out in
float64 from the geometry (the closest point of a box by clamping in its frame, the closest points
of two segments by Ericson's ClosestPtSegmentSegment, the box corners by their signs, the clipped
face by hand) and compared with the oracle's FP32 pass over a two-body world:

* sphere-sphere, sphere-box (outside a face, an edge, a corner; centre inside), sphere-capsule
  (side and cap), capsule-capsule (crossing and parallel axes): narrowphase.wgsl:33-218;
* box-box SAT: face-face with the incident face inside the reference face, face-face clipped by a
  reference side (points and clip feature ids derived by hand), edge-edge, and which box gives the
  reference face when two faces overlap almost equally: narrowphase.wgsl:325-473;
* plane-convex: box flat and tilted, box before the plane (mirrored features), sphere, capsule
  lying down: narrowphase.wgsl:485-512 with convex.wgsl:639-670,795-818;
* the warm-start relay of identity.wgsl:85-117: impulses follow equal or mirrored features to the
  nearest held point, each held point at most once.

Conventions (narrowphase.wgsl): the normal points from the pair's first collider to its second;
depth > 0 overlaps; an analytic or plane point sits halfway between the two surfaces, while a
box-box face point is the incident point moved a further half depth along the reference face's
inward normal (narrowphase.wgsl:434-436: polygon point - ref_normal * depth / 2); features are
kind << 28 | first field << 14 | second field (crates/dynamis-abi/src/constant.rs:57-66).
"""

from __future__ import annotations

import numpy as np
import pytest

from dynamis_amd import abi
from tests.worlds import BodyDesc, config, encode

FEATURE_POINT = 0
FEATURE_VERTEX = 1 << 28
FEATURE_EDGE = 2 << 28
FEATURE_FACE = 3 << 28
FIELD_BITS = 14
FEATURE_CLIP = 1 << 12
FACE_BIT = 1 << 13
MARGIN = 0.02      # PhysicsConfig::default contact_margin
CLIP_MARGIN = 1e-4  # core.wgsl:615
POS_TOL = 2e-6
DEPTH_TOL = 2e-6


def feature(kind: int, first: int, second: int) -> int:
    return kind | (first << FIELD_BITS) | second


def mirror(f: int) -> int:
    kind = f & 0xF0000000
    if kind == FEATURE_POINT:
        return f
    return kind | ((f & 0x3FFF) << FIELD_BITS) | ((f >> FIELD_BITS) & 0x3FFF)


def quat(axis, degrees) -> tuple:
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    h = np.radians(degrees) / 2.0
    return tuple(float(v) for v in (*(axis * np.sin(h)), np.cos(h)))


def rot(q) -> np.ndarray:
    """Rotation matrix of a unit (x, y, z, w) quaternion, float64."""
    x, y, z, w = (float(v) for v in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def narrow(oracle, bodies):
    """The oracle's narrowphase over the pair (collider 0, collider 1) of a two-body world:
    (contact, valid, scene). Rows are simulated first, colliders in spawn order."""
    scene = encode(bodies, config(gravity=(0.0, 0.0, 0.0)))
    world = oracle.world(scene)
    raw, valid = world.narrowphase(np.array([[0, 1]], np.uint32))
    world.close()
    return raw[0], bool(valid[0]), scene


def body_frame(scene, collider):
    """Centre (float64) and rotation matrix of a collider as the records hold them."""
    row = int(scene.owners[collider])
    s = scene.states[row]
    return np.asarray(s["position"], np.float64), rot(s["orientation"])


def points_of(contact):
    n = int(contact["point_count"])
    p = contact["points"][:n]
    return (np.asarray(p["position"], np.float64), np.asarray(p["depth"], np.float64),
            [int(f) for f in p["feature"]])


def assert_single(contact, normal, point, depth, feat=FEATURE_POINT):
    assert int(contact["point_count"]) == 1
    np.testing.assert_allclose(contact["normal"], normal, atol=POS_TOL)
    pos, dep, feats = points_of(contact)
    np.testing.assert_allclose(pos[0], point, atol=POS_TOL)
    assert abs(dep[0] - depth) <= DEPTH_TOL, (dep[0], depth)
    assert feats == [feat]


# ---- sphere-sphere (narrowphase.wgsl:33-59) -----------------------------------------------------

@pytest.mark.parametrize("gap", [-0.2, -0.01, 0.01, 0.015])
def test_sphere_sphere(oracle, gap):
    ra, rb = 0.5, 0.4
    direction = np.array([0.6, 0.3, 0.2]) / np.linalg.norm([0.6, 0.3, 0.2])
    ca = np.array([0.0, 1.0, 0.0])
    cb = ca + direction * (ra + rb + gap)
    c, valid, scene = narrow(oracle, [BodyDesc.sphere(ra, position=tuple(ca)),
                                      BodyDesc.sphere(rb, position=tuple(cb))])
    pa, _ = body_frame(scene, 0)
    pb, _ = body_frame(scene, 1)
    d = np.linalg.norm(pb - pa)
    n = (pb - pa) / d
    depth = ra + rb - d
    assert valid
    assert_single(c, n, pa + n * (ra - depth / 2), depth)


def test_sphere_sphere_beyond_margin_gives_no_contact(oracle):
    _, valid, _ = narrow(oracle, [BodyDesc.sphere(0.5), BodyDesc.sphere(0.4, position=(0.9 + MARGIN + 0.01, 0, 0))])
    assert not valid


# ---- sphere-box (narrowphase.wgsl:62-117) -------------------------------------------------------

BOX_HALF = np.array([1.0, 0.5, 0.75])
BOX_Q = quat((0.2, 1.0, 0.1), 30.0)
BOX_C = np.array([0.3, 2.0, -0.4])


@pytest.mark.parametrize("where,local", [("face", (0.0, 0.7, 0.0)), ("edge", (1.1, 0.6, 0.0)),
                                         ("corner", (1.1, 0.6, 0.85)), ("inside", (0.2, 0.3, 0.1))])
@pytest.mark.parametrize("sphere_first", [True, False])
def test_sphere_box(oracle, where, local, sphere_first):
    r = 0.3
    R = rot(BOX_Q)
    centre = BOX_C + R @ np.asarray(local)
    sphere = BodyDesc.sphere(r, position=tuple(centre))
    box = BodyDesc.cuboid(tuple(BOX_HALF), position=tuple(BOX_C), orientation=BOX_Q)
    c, valid, scene = narrow(oracle, [sphere, box] if sphere_first else [box, sphere])
    assert valid
    s_col, b_col = (0, 1) if sphere_first else (1, 0)
    sc, _ = body_frame(scene, s_col)
    bc, Rb = body_frame(scene, b_col)
    half = scene.colliders["half_extents"][b_col].astype(np.float64)
    loc = Rb.T @ (sc - bc)
    if where == "inside":
        # the nearest face: the axis of least clearance, facing the side the centre is on
        clearance = half - np.abs(loc)
        axis = int(np.argmin(clearance))
        outward = np.zeros(3)
        outward[axis] = np.sign(loc[axis])
        normal = -(Rb @ outward)          # sphere -> box: into the box through that face
        depth = r + clearance[axis]
        point = sc - normal * depth / 2
    else:
        closest = bc + Rb @ np.clip(loc, -half, half)
        delta = closest - sc
        normal = delta / np.linalg.norm(delta)
        depth = r - np.linalg.norm(delta)
        point = closest - normal * depth / 2
    assert depth > 0
    # a box listed first turns the sphere's normal around (narrowphase.wgsl:627-631)
    assert_single(c, normal if sphere_first else -normal, point, depth)


# ---- sphere-capsule (narrowphase.wgsl:119-139) ---------------------------------------------------

@pytest.mark.parametrize("centre", [(0.2, 0.45, 0.0), (0.8, 0.3, 0.05)])
@pytest.mark.parametrize("sphere_first", [True, False])
def test_sphere_capsule(oracle, centre, sphere_first):
    r, cr, hh = 0.3, 0.2, 0.5
    q = quat((0, 0, 1), 90.0)  # the capsule's axis (local y) along -x
    sphere = BodyDesc.sphere(r, position=centre)
    capsule = BodyDesc.capsule(cr, hh, position=(0.0, 0.0, 0.0), orientation=q)
    c, valid, scene = narrow(oracle, [sphere, capsule] if sphere_first else [capsule, sphere])
    assert valid
    sc, _ = body_frame(scene, 0 if sphere_first else 1)
    cc, Rc = body_frame(scene, 1 if sphere_first else 0)
    axis = Rc @ np.array([0.0, 1.0, 0.0])
    a0, a1 = cc - axis * hh, cc + axis * hh
    t = np.clip(np.dot(sc - a0, a1 - a0) / np.dot(a1 - a0, a1 - a0), 0.0, 1.0)
    closest = a0 + (a1 - a0) * t
    delta = closest - sc
    normal = delta / np.linalg.norm(delta)
    depth = r + cr - np.linalg.norm(delta)
    assert depth > 0
    assert_single(c, normal if sphere_first else -normal, closest - normal * depth / 2, depth)


# ---- capsule-capsule (narrowphase.wgsl:141-218) --------------------------------------------------

def test_capsule_capsule_crossing_axes(oracle):
    r, hh = 0.2, 0.5
    a = BodyDesc.capsule(r, hh, position=(0.0, 0.0, 0.0), orientation=quat((0, 0, 1), 90.0))   # along x
    b = BodyDesc.capsule(r, hh, position=(0.1, 0.3, -0.05), orientation=quat((1, 0, 0), 90.0))  # along z
    c, valid, _ = narrow(oracle, [a, b])
    assert valid
    # closest points of the two axes: (0.1, 0, 0) on the first, (0.1, 0.3, 0) on the second
    assert_single(c, (0.0, 1.0, 0.0), (0.1, 0.15, 0.0), 2 * r - 0.3)


def test_capsule_capsule_parallel_axes(oracle):
    """Parallel segments overlap along a whole interval; Ericson's ClosestPtSegmentSegment
    (Real-Time Collision Detection 5.1.9), which closest_points_segments follows, takes s = 0 on
    the first segment, clamps t on the second to 0, and moves s back to the second segment's start:
    the pair at the overlap's lower end, y = -0.2."""
    r, hh = 0.2, 0.5
    c, valid, _ = narrow(oracle, [BodyDesc.capsule(r, hh), BodyDesc.capsule(r, hh, position=(0.35, 0.3, 0.0))])
    assert valid
    assert_single(c, (1.0, 0.0, 0.0), (0.175, -0.2, 0.0), 2 * r - 0.35)


# ---- box-box SAT (narrowphase.wgsl:325-473) ------------------------------------------------------

def corner_id(local_signs) -> int:
    """box_corner_id: bit k is set when the corner lies on the positive side of the box's axis k."""
    return sum(1 << k for k in range(3) if local_signs[k] > 0)


def test_box_box_face_face_incident_inside(oracle):
    """A small box resting on a larger one, 0.05 into it: the reference face is the lower box's top
    (ties between equal face overlaps keep the first box), the manifold is the upper box's four
    bottom corners (y = 0.45) moved half the depth down (y = 0.425), with features
    face(top face, corner)."""
    lower = BodyDesc.cuboid((1.0, 0.5, 1.0), position=(0.0, 0.0, 0.0))
    upper = BodyDesc.cuboid((0.5, 0.5, 0.4), position=(0.3, 0.95, 0.2))
    c, valid, _ = narrow(oracle, [lower, upper])
    assert valid and int(c["point_count"]) == 4
    np.testing.assert_allclose(c["normal"], (0, 1, 0), atol=POS_TOL)
    top = FACE_BIT | (1 * 2 + 1)  # the +y face of the lower box
    want = {}
    for sx in (-1, 1):
        for sz in (-1, 1):
            p = (0.3 + sx * 0.5, 0.425, 0.2 + sz * 0.4)
            want[feature(FEATURE_FACE, top, corner_id((sx, -1, sz)))] = p
    pos, dep, feats = points_of(c)
    assert sorted(feats) == sorted(want)
    for p, d, f in zip(pos, dep, feats):
        np.testing.assert_allclose(p, want[f], atol=POS_TOL)
        assert abs(d - 0.05) <= DEPTH_TOL
    # manifold_order: equal depths lead with the lowest feature, then the point farthest away
    assert feats[0] == min(feats)
    lead = np.asarray(want[feats[0]])
    assert feats[1] == max(want, key=lambda f: np.linalg.norm(np.asarray(want[f]) - lead))


def test_box_box_face_face_clipped(oracle):
    """The upper box reaches past the lower box's +x side: the side plane x = 0.5 (moved out by
    CLIP_MARGIN) cuts the incident face. Worked by hand (reference corners (0.5, 0.5, 0.5),
    (0.5, 0.5, -0.5), (-0.5, 0.5, -0.5), (-0.5, 0.5, 0.5); side 0 runs along -z with outward normal
    +x; incident corners in box_face order (1.9, z=0.3), (1.9, -0.3), (-0.1, -0.3), (-0.1, 0.3)):
    side 0 keeps the last two corners and adds the crossings of incident edges 1 and 3, whose
    ids are clip(group 2 * 32 + side 0 * 8 + edge), group 2 being the incident -y face."""
    lower = BodyDesc.cuboid((0.5, 0.5, 0.5), position=(0.0, 0.0, 0.0))
    upper = BodyDesc.cuboid((1.0, 0.5, 0.3), position=(0.9, 0.95, 0.0))
    c, valid, _ = narrow(oracle, [lower, upper])
    assert valid and int(c["point_count"]) == 4
    top = FACE_BIT | 3
    x_cut = 0.5 + CLIP_MARGIN
    want = {
        feature(FEATURE_FACE, top, FEATURE_CLIP | (2 * 32 + 0 * 8 + 1)): (x_cut, 0.425, -0.3),
        feature(FEATURE_FACE, top, corner_id((-1, -1, -1))): (-0.1, 0.425, -0.3),
        feature(FEATURE_FACE, top, corner_id((-1, -1, 1))): (-0.1, 0.425, 0.3),
        feature(FEATURE_FACE, top, FEATURE_CLIP | (2 * 32 + 0 * 8 + 3)): (x_cut, 0.425, 0.3),
    }
    pos, dep, feats = points_of(c)
    assert sorted(feats) == sorted(want)
    for p, d, f in zip(pos, dep, feats):
        np.testing.assert_allclose(p, want[f], atol=1e-5)
        assert abs(d - 0.05) <= DEPTH_TOL
    assert feats[0] == feature(FEATURE_FACE, top, corner_id((-1, -1, -1)))  # lowest feature leads
    assert want[feats[1]] == (x_cut, 0.425, 0.3)                            # then the far corner


def test_box_box_edge_edge(oracle):
    """A box turned 45 degrees about z (an edge along z on top) under a box turned 45 degrees about
    x (an edge along x below), the edges crossing 0.05 deep: the edge axis z x x = y separates by
    more than FACE_AXIS_BIAS than any face axis, so the contact is one point on feature
    edge(2, 0), halfway between the two boxes' support points along the normal (a support takes the
    negative half-extent on an axis perpendicular to the direction)."""
    h = 0.5
    s2 = np.sqrt(2.0)
    gap_centre = 2 * h * s2 - 0.05
    lower = BodyDesc.cuboid((h, h, h), orientation=quat((0, 0, 1), 45.0))
    upper = BodyDesc.cuboid((h, h, h), position=(0.0, gap_centre, 0.0), orientation=quat((1, 0, 0), 45.0))
    c, valid, scene = narrow(oracle, [lower, upper])
    assert valid
    la, Ra = body_frame(scene, 0)
    lb, Rb = body_frame(scene, 1)
    n = np.array([0.0, 1.0, 0.0])

    def support(centre, R, d):
        return centre + sum(R[:, k] * (h if np.dot(R[:, k], d) > 1e-9 else -h) for k in range(3))

    point = 0.5 * (support(la, Ra, n) + support(lb, Rb, -n))
    depth = 2 * h * s2 - (lb[1] - la[1])
    assert_single(c, n, point, depth, feature(FEATURE_EDGE, 2, 0))


def stacked_pair(tilt):
    h = (0.5, 0.3, 0.4)
    return [BodyDesc.cuboid(h, position=(0.0, 0.0, 0.0)),
            BodyDesc.cuboid(h, position=(0.05, 0.58, 0.07), orientation=quat((1, 0, 0), tilt))]


@pytest.mark.parametrize("tilt", [0.5, -0.5])
def test_box_box_reference_face_follows_the_smaller_overlap(oracle, tilt):
    """Two equal boxes stacked face to face, the upper one turned half a degree about x. The first
    box's y face overlaps by hA_y + sum_k |yA . kB| hB_k - |delta . yA|, the second's by
    hB_y + sum_k |yB . kA| hA_k - |delta . yB|; narrowphase.wgsl:334-350 keeps the strictly smaller
    one (the first on a tie). The two tilts put the reference face on different boxes, and with
    it every feature id: the features of one manifold are neither equal nor mirrored to the
    other's, so a relay between them carries nothing (identity.wgsl:77-79) - the mechanism by
    which a stacked pair whose faces overlap almost equally loses its warm start."""
    h = np.array([0.5, 0.3, 0.4])
    c, valid, scene = narrow(oracle, stacked_pair(tilt))
    assert valid
    la, Ra = body_frame(scene, 0)
    lb, Rb = body_frame(scene, 1)
    delta = lb - la
    over_a = h[1] + sum(abs(np.dot(Rb[:, k], Ra[:, 1])) * h[k] for k in range(3)) - abs(np.dot(delta, Ra[:, 1]))
    over_b = h[1] + sum(abs(np.dot(Ra[:, k], Rb[:, 1])) * h[k] for k in range(3)) - abs(np.dot(delta, Rb[:, 1]))
    assert abs(over_a - over_b) > 1e-4  # well clear of FP32 rounding
    reference_first = over_a <= over_b
    _, dep, feats = points_of(c)
    assert abs(max(dep) - min(over_a, over_b)) < 0.02
    # the reference face: +y of the lower box (face 3, first field) or -y of the upper box
    # (face 2, second field: feature_face(incident id, reference field))
    fields = {(f >> FIELD_BITS) & 0x3FFF for f in feats} if reference_first else {f & 0x3FFF for f in feats}
    assert fields == {FACE_BIT | (3 if reference_first else 2)}


def test_box_box_reference_flip_relays_nothing(oracle):
    plus, _, _ = narrow(oracle, stacked_pair(0.5))
    minus, _, _ = narrow(oracle, stacked_pair(-0.5))
    _, _, fp = points_of(plus)
    _, _, fm = points_of(minus)
    first_fields = lambda feats: {(f >> FIELD_BITS) & 0x3FFF for f in feats}
    assert first_fields(fp) != first_fields(fm)  # one tilt's reference is the lower box, the other's the upper
    assert not (set(fp) & (set(fm) | {mirror(f) for f in fm}))
    held = minus.copy()
    held["points"]["accumulated_normal"][: int(held["point_count"])] = 1.0
    relayed, picks = oracle.relay(plus, held)
    n = int(plus["point_count"])
    assert (picks[:n] == -1).all()
    assert (relayed["points"]["accumulated_normal"][:n] == 0.0).all()


# ---- plane-convex (narrowphase.wgsl:485-512, 574-586) --------------------------------------------

PLANE_FACE = FACE_BIT | 0  # feature_field_face(0u)


def test_plane_box_flat(oracle):
    h = (0.5, 0.3, 0.4)
    c, valid, _ = narrow(oracle, [BodyDesc.plane(), BodyDesc.cuboid(h, position=(0.2, 0.28, -0.1))])
    assert valid and int(c["point_count"]) == 4
    np.testing.assert_allclose(c["normal"], (0, 1, 0), atol=POS_TOL)
    want = {}
    for sx in (-1, 1):
        for sz in (-1, 1):
            want[feature(FEATURE_VERTEX, PLANE_FACE, corner_id((sx, -1, sz)))] = (0.2 + sx * 0.5, -0.01, -0.1 + sz * 0.4)
    pos, dep, feats = points_of(c)
    assert sorted(feats) == sorted(want)
    for p, d, f in zip(pos, dep, feats):
        np.testing.assert_allclose(p, want[f], atol=POS_TOL)
        assert abs(d - 0.02) <= DEPTH_TOL
    assert feats[0] == min(feats)


def test_plane_box_tilted_keeps_points_within_the_margin(oracle):
    """A box turned 10 degrees about z: its bottom face's corners on the low side reach below the
    plane; those more than the margin above it are left out."""
    h = np.array([0.5, 0.3, 0.4])
    q = quat((0, 0, 1), 10.0)
    centre = np.array([0.0, 0.33, 0.0])
    c, valid, scene = narrow(oracle, [BodyDesc.plane(), BodyDesc.cuboid(tuple(h), position=tuple(centre), orientation=q)])
    assert valid
    bc, R = body_frame(scene, 1)
    want = {}
    for sx in (-1, 1):
        for sz in (-1, 1):
            corner = bc + R @ (h * np.array([sx, -1, sz]))
            depth = -corner[1]
            if depth > -MARGIN:
                want[feature(FEATURE_VERTEX, PLANE_FACE, corner_id((sx, -1, sz)))] = (corner + np.array([0, depth / 2, 0]), depth)
    pos, dep, feats = points_of(c)
    assert len(want) == 2 and sorted(feats) == sorted(want)
    for p, d, f in zip(pos, dep, feats):
        np.testing.assert_allclose(p, want[f][0], atol=POS_TOL)
        assert abs(d - want[f][1]) <= DEPTH_TOL
    assert dep[0] >= dep[1]  # the deeper point leads


def test_plane_second_mirrors_normal_and_features(oracle):
    h = (0.5, 0.3, 0.4)
    c, valid, _ = narrow(oracle, [BodyDesc.cuboid(h, position=(0.2, 0.28, -0.1)), BodyDesc.plane()])
    flat, _, _ = narrow(oracle, [BodyDesc.plane(), BodyDesc.cuboid(h, position=(0.2, 0.28, -0.1))])
    assert valid
    np.testing.assert_allclose(c["normal"], (0, -1, 0), atol=POS_TOL)
    _, _, feats = points_of(c)
    _, _, flat_feats = points_of(flat)
    assert feats == [mirror(f) for f in flat_feats]
    assert all((f >> FIELD_BITS) & 0x3FFF < FACE_BIT and (f & 0x3FFF) == PLANE_FACE for f in feats)


def test_plane_sphere(oracle):
    c, valid, _ = narrow(oracle, [BodyDesc.plane(), BodyDesc.sphere(0.4, position=(1.0, 0.37, 2.0))])
    assert valid
    assert_single(c, (0, 1, 0), (1.0, -0.03 + 0.015, 2.0), 0.03, feature(FEATURE_VERTEX, PLANE_FACE, 0))


def test_plane_capsule_lying(oracle):
    """A capsule lying along x: capsule_feature gives both ends of its lower side line, the upper
    end (local +y, here world -x) first as id 0."""
    r, hh = 0.25, 0.6
    c, valid, _ = narrow(oracle, [BodyDesc.plane(), BodyDesc.capsule(r, hh, position=(0.0, r - 0.01, 0.0),
                                                                       orientation=quat((0, 0, 1), 90.0))])
    assert valid and int(c["point_count"]) == 2
    pos, dep, feats = points_of(c)
    assert feats == [feature(FEATURE_VERTEX, PLANE_FACE, 0), feature(FEATURE_VERTEX, PLANE_FACE, 1)]
    np.testing.assert_allclose(pos[0], (-hh, -0.005, 0.0), atol=POS_TOL)
    np.testing.assert_allclose(pos[1], (hh, -0.005, 0.0), atol=POS_TOL)
    np.testing.assert_allclose(dep, [0.01, 0.01], atol=DEPTH_TOL)


# ---- warm-start relay (identity.wgsl:85-117) ----------------------------------------------------

def manifold(points):
    """A contact whose points are (feature, local_a, local_b, accumulated normal, tangent 1, 2)."""
    c = np.zeros(1, abi.CONTACT)[0]
    c["point_count"] = len(points)
    for i, (f, la, lb, an, t1, t2) in enumerate(points):
        p = c["points"][i]
        p["feature"] = f
        p["local_a"] = la
        p["local_b"] = lb
        p["accumulated_normal"] = an
        p["accumulated_tangent_1"] = t1
        p["accumulated_tangent_2"] = t2
    return c


F1 = feature(FEATURE_FACE, FACE_BIT | 3, 5)
F2 = feature(FEATURE_FACE, FACE_BIT | 3, 1)
F3 = feature(FEATURE_FACE, FACE_BIT | 3, 0)


def test_relay_follows_features_not_order(oracle):
    held = manifold([(F1, (0, 0, 0), (0, 0, 0), 3.0, 0.1, 0.2), (F2, (1, 0, 0), (1, 0, 0), 5.0, 0.3, 0.4)])
    current = manifold([(F2, (1, 0, 0), (1, 0, 0), 0, 0, 0), (F3, (0, 0, 1), (0, 0, 1), 0, 0, 0),
                        (F1, (0, 0, 0), (0, 0, 0), 0, 0, 0)])
    out, picks = oracle.relay(current, held)
    assert list(picks[:3]) == [1, -1, 0]
    pts = out["points"]
    np.testing.assert_array_equal(pts["accumulated_normal"][:3], np.float32([5.0, 0.0, 3.0]))
    np.testing.assert_array_equal(pts["accumulated_tangent_1"][:3], np.float32([0.3, 0.0, 0.1]))
    np.testing.assert_array_equal(pts["accumulated_tangent_2"][:3], np.float32([0.4, 0.0, 0.2]))


def test_relay_accepts_mirrored_features(oracle):
    held = manifold([(F1, (0, 0, 0), (0, 0, 0), 2.5, 0, 0)])
    current = manifold([(mirror(F1), (0, 0, 0), (0, 0, 0), 0, 0, 0)])
    out, picks = oracle.relay(current, held)
    assert picks[0] == 0 and out["points"]["accumulated_normal"][0] == np.float32(2.5)


def test_relay_takes_the_nearest_held_point_once(oracle):
    """Two held points answer the same feature: the first current point takes the one nearest
    its anchors (distance = |dla| + |dlb|), the second current point the one left over; a third
    finds none."""
    held = manifold([(F1, (0, 0, 0), (0, 0, 0), 1.0, 0, 0), (F1, (1, 0, 0), (1, 0, 0), 2.0, 0, 0)])
    current = manifold([(F1, (0.9, 0, 0), (0.9, 0, 0), 0, 0, 0), (F1, (0.1, 0, 0), (0.1, 0, 0), 0, 0, 0),
                        (F1, (0.5, 0, 0), (0.5, 0, 0), 0, 0, 0)])
    out, picks = oracle.relay(current, held)
    assert list(picks[:3]) == [1, 0, -1]
    np.testing.assert_array_equal(out["points"]["accumulated_normal"][:3], np.float32([2.0, 1.0, 0.0]))
