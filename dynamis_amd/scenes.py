"""Synthetic scenes for the benchmark and tests (no datasets exist for this path)."""

from __future__ import annotations

import numpy as np

from . import abi


def uniform_boxes(k: int, extent: float, half_lo: float, half_hi: float,
                  seed: int) -> np.ndarray:
    """K boxes with centres uniform in [0, extent)^3 and per-axis half-extents log-uniform in
    [half_lo, half_hi]. At k=2^20, extent=1000, 0.2..2 m this is the isolated sort + broadphase
    workload."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, extent, size=(k, 3)).astype(np.float32)
    halves = np.exp(rng.uniform(np.log(half_lo), np.log(half_hi), size=(k, 3))).astype(np.float32)
    return abi.aabbs(centres - halves, centres + halves)


def ellipsoid_hull_sources(count: int, seed: int):
    """`count` hull shape sources, each the hull of 16-64 points drawn uniformly on an
    ellipsoid with semi-axes ~ U[0.3, 0.9]."""
    from .world import build_hull
    rng = np.random.default_rng(seed)
    hulls = []
    for _ in range(count):
        n = int(rng.integers(16, 65))
        p = rng.normal(size=(n, 3))
        p /= np.linalg.norm(p, axis=1, keepdims=True)
        hulls.append(build_hull((p * rng.uniform(0.3, 0.9, 3)).astype(np.float32)))
    return hulls


def convex_pile(count: int = 100_000, seed: int = 3, hull_sources: int = 512, bin_size: float = 60.0,
                spacing: float = 2.0, layer_height: float = 2.6):
    """The mixed convex pile: `count` bodies, a quarter each of spheres (r ~ U[0.25, 0.75]),
    boxes (half-extents ~ U[0.2, 0.8] per axis), capsules (r ~ U[0.15, 0.4], half-height
    ~ U[0.2, 0.8]) and hulls (drawn from `hull_sources` ellipsoid hulls of 16-64 points), density
    1000, on a jittered lattice inside a bin_size x bin_size bin of four static cuboid walls
    standing on a static plane. Returns (bodies, hulls)."""
    from .world import BodyDesc
    rng = np.random.default_rng(seed)
    hulls = ellipsoid_hull_sources(hull_sources, seed + 1)
    half = bin_size / 2.0
    wall_h = 25.0
    bodies = [BodyDesc.plane()]
    for sx, sz, hx, hz in ((half + 0.5, 0.0, 0.5, half + 1.0), (-half - 0.5, 0.0, 0.5, half + 1.0),
                           (0.0, half + 0.5, half + 1.0, 0.5), (0.0, -half - 0.5, half + 1.0, 0.5)):
        bodies.append(BodyDesc.cuboid((hx, wall_h, hz), mass=0.0, position=(sx, wall_h, sz)))
    per_row = int(bin_size // spacing) - 1
    per_layer = per_row * per_row
    first = -0.5 * (per_row - 1) * spacing
    kinds = rng.permutation(np.arange(count) % 4)
    jitter = rng.uniform(-0.1, 0.1, size=(count, 3))
    for i in range(count):
        layer, cell = divmod(i, per_layer)
        gx, gz = divmod(cell, per_row)
        pos = (float(first + gx * spacing + jitter[i, 0]), float(1.5 + layer * layer_height + jitter[i, 1]),
               float(first + gz * spacing + jitter[i, 2]))
        kind = int(kinds[i])
        if kind == 0:
            bodies.append(BodyDesc.sphere(float(rng.uniform(0.25, 0.75)), position=pos, density=1000.0))
        elif kind == 1:
            bodies.append(BodyDesc.cuboid(tuple(float(v) for v in rng.uniform(0.2, 0.8, 3)), position=pos,
                                          density=1000.0))
        elif kind == 2:
            bodies.append(BodyDesc.capsule(float(rng.uniform(0.15, 0.4)), float(rng.uniform(0.2, 0.8)),
                                           position=pos, density=1000.0))
        else:
            bodies.append(BodyDesc.hull(int(rng.integers(0, hull_sources)), position=pos, density=1000.0))
    return bodies, hulls


def cube_drop(count: int = 1000, seed: int = 1, footprint: float = 20.0, low: float = 1.0,
              high: float = 40.0):
    """The small plain workload: cubes of half-extent 0.5 m, density 1000 kg/m^3, friction 0.5,
    restitution 0, centres uniform in a footprint x footprint square at heights [low, high],
    uniform random orientations (normalised 4-D Gaussians), above a static plane at y = 0."""
    from .world import BodyDesc
    rng = np.random.default_rng(seed)
    half = footprint / 2.0
    xs = rng.uniform(-half, half, count)
    zs = rng.uniform(-half, half, count)
    ys = rng.uniform(low, high, count)
    q = rng.normal(size=(count, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q = q.astype(np.float32)
    bodies = [BodyDesc.plane()]
    for i in range(count):
        bodies.append(BodyDesc.cuboid((0.5, 0.5, 0.5), density=1000.0, friction=0.5, restitution=0.0,
                                      position=(float(xs[i]), float(ys[i]), float(zs[i])),
                                      orientation=tuple(float(v) for v in q[i])))
    return bodies


def batch_worlds(worlds: int = 4096, seed: int = 4, low: int = 8, high: int = 2048,
                 spacing: float = 2.0, per_row: int = 16):
    """The batch of independent worlds: each a static plane and N_w bodies, N_w log-uniform in
    [low, high], half boxes (half-extents ~ U[0.2, 0.8] per axis) and half spheres
    (r ~ U[0.25, 0.75]) at random, density 1000, on a jittered lattice of per_row x per_row
    bodies a layer (spacing apart, layers spacing apart from y = 1). Returns the worlds' scenes
    (each encoded as the reference would spawn it: the plane first)."""
    from .world import BodyDesc, encode_specs
    rng = np.random.default_rng(seed)
    counts = np.floor(np.exp(rng.uniform(np.log(low), np.log(high + 1), worlds))).astype(np.int64)
    counts = np.clip(counts, low, high)
    box = BodyDesc.cuboid((0.5, 0.5, 0.5), density=1000.0).spec()[0]
    ball = BodyDesc.sphere(0.5, density=1000.0).spec()[0]
    plane = BodyDesc.plane().spec()
    scenes = []
    for n in counts:
        kinds = rng.integers(0, 2, n)
        specs = np.zeros(n, dtype=abi.BODY_SPEC)
        specs[kinds == 0] = box
        specs[kinds == 1] = ball
        halves = rng.uniform(0.2, 0.8, (n, 3)).astype(np.float32)
        radii = rng.uniform(0.25, 0.75, n).astype(np.float32)
        specs["half_extents"][kinds == 0] = halves[kinds == 0]
        specs["radius"][kinds == 1] = radii[kinds == 1]
        i = np.arange(n)
        layer, cell = np.divmod(i, per_row * per_row)
        gx, gz = np.divmod(cell, per_row)
        first = -0.5 * (per_row - 1) * spacing
        jitter = rng.uniform(-0.1, 0.1, (n, 3))
        specs["position"][:, 0] = first + gx * spacing + jitter[:, 0]
        specs["position"][:, 1] = 1.0 + layer * spacing + jitter[:, 1]
        specs["position"][:, 2] = first + gz * spacing + jitter[:, 2]
        scenes.append(encode_specs(np.concatenate([plane, specs])))
    return scenes


def adjoint_worlds(worlds: int = 256, boxes: int = 32, seed: int = 5, region: float = 4.0,
                   speed: float = 2.0, friction=(0.2, 0.8)):
    """The differentiation workload: per world a static plane and `boxes` boxes (half-extents
    ~ U[0.2, 0.5] per axis, density 1000) on a jittered 4 x 2 x 4 lattice filling a region^3 box
    whose floor is 0.5 above the plane, initial velocities ~ N(0, speed) per axis, friction
    ~ U[friction]. Returns the worlds' scenes (plane first, as spawned), the configuration
    (8 iterations, sleeping off: the differentiable step has no sleep) and a target point per
    world for the centre-of-mass loss, uniform in the region."""
    from .world import BodyDesc, default_config, encode_specs
    rng = np.random.default_rng(seed)
    config = default_config()
    config["solve_iterations"] = 8
    config["sleep_velocity"] = 0.0
    config["sleep_angular_velocity"] = 0.0
    box = BodyDesc.cuboid((0.5, 0.5, 0.5), density=1000.0).spec()[0]
    plane = BodyDesc.plane().spec()
    nx, ny, nz = 4, 2, 4
    cell = np.array([region / nx, region / ny, region / nz])
    scenes, targets = [], np.zeros((worlds, 3), np.float32)
    for w in range(worlds):
        specs = np.zeros(boxes, dtype=abi.BODY_SPEC)
        specs[:] = box
        specs["half_extents"] = rng.uniform(0.2, 0.5, (boxes, 3)).astype(np.float32)
        i = np.arange(boxes) % (nx * ny * nz)
        gy, rest = np.divmod(i, nx * nz)
        gx, gz = np.divmod(rest, nz)
        jitter = rng.uniform(-0.05, 0.05, (boxes, 3))
        specs["position"][:, 0] = (gx + 0.5) * cell[0] - 0.5 * region + jitter[:, 0]
        specs["position"][:, 1] = 0.5 + (gy + 0.5) * cell[1] + jitter[:, 1]
        specs["position"][:, 2] = (gz + 0.5) * cell[2] - 0.5 * region + jitter[:, 2]
        specs["velocity"] = rng.normal(0.0, speed, (boxes, 3)).astype(np.float32)
        specs["friction"] = rng.uniform(friction[0], friction[1], boxes).astype(np.float32)
        scenes.append(encode_specs(np.concatenate([plane, specs]), config))
        targets[w] = rng.uniform([-0.5 * region, 0.5, -0.5 * region], [0.5 * region, 0.5 + region, 0.5 * region])
    return scenes, config, targets
