"""Scenes for the narrowphase stage tests: mixed convex piles (spheres, boxes, capsules, hulls,
cylinders) a few steps into their fall, and the fields the narrowphase pass writes."""

from __future__ import annotations

import dataclasses

import numpy as np

from dynamis_amd import abi
from tests.worlds import DT, config, ellipsoid_hulls, encode, mixed_pile

# Fields of a contact the narrowphase pass leaves final: the begin/touch event bits
# (contact_begin.wgsl), the carried impulses (contact_relay.wgsl) and the accumulated impulses
# (the solver) are written by later passes of the step.
LATER_FIELDS = {"events", "carried_normal", "carried_tangent"}
LATER_POINT_FIELDS = {"accumulated_normal", "accumulated_tangent_1", "accumulated_tangent_2"}


def pile_scene(count: int, seed: int = 2, hull_sources: int = 4, spacing: float = 1.6):
    hulls = ellipsoid_hulls(hull_sources, seed=seed + 1)
    return encode(mixed_pile(count, seed=seed, kinds=5, hull_sources=hull_sources, spacing=spacing), config(),
                  hulls)


def with_states(scene, states: np.ndarray):
    return dataclasses.replace(scene, states=np.ascontiguousarray(states, dtype=abi.BODY_STATE))


def narrow_fields_equal(got: np.ndarray, want: np.ndarray) -> list[str]:
    """Names of the narrowphase-written fields in which two contact arrays differ (bitwise)."""
    bad = []
    for f in want.dtype.names:
        if f in LATER_FIELDS:
            continue
        if f == "points":
            for g in want["points"].dtype.names:
                if g not in LATER_POINT_FIELDS and got["points"][g].tobytes() != want["points"][g].tobytes():
                    bad.append("points." + g)
        elif got[f].tobytes() != want[f].tobytes():
            bad.append(f)
    return bad


def convex_path_contacts(scene, contacts: np.ndarray) -> int:
    """Contacts with a hull or a cylinder on either side (pairs GJK/EPA or the convex manifolds
    answer)."""
    kinds = scene.colliders["kind"]
    a, b = kinds[contacts["a"]], kinds[contacts["b"]]
    convex = (abi.SHAPE_HULL, abi.SHAPE_CYLINDER)
    return int(np.isin(a, convex).sum() + np.isin(b, convex).sum())


__all__ = ["DT", "pile_scene", "with_states", "narrow_fields_equal", "convex_path_contacts"]
