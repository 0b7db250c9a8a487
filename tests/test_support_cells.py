"""Support cells of hull shape sources (dyn_hull_support_cells): scanning a direction's cell
candidates in ascending order gives the linear scan's support vertex (convex.wgsl:1-85: the first
vertex of largest single-precision dot, the dot evaluated x*dx + y*dy, then + z*dz, without
contraction), for the pile's 512 ellipsoid hulls and for a symmetric hull whose axis directions
tie several vertices exactly; directions random, on every cell corner and edge, just either side
of them, along the axes, and tiny or huge. The lookup below follows support_cell_of
(support_cells.h) in single precision, as the device evaluates it."""

from __future__ import annotations

import ctypes
import itertools

import numpy as np

from dynamis_amd import abi
from dynamis_amd._native import call

FACE_CELLS = 8
CELLS = 6 * FACE_CELLS * FACE_CELLS
NO_CELL = 0xFF


def support_cells(sources, vertices):
    cells = np.zeros((len(sources), CELLS, 2), np.uint32)
    total = ctypes.c_uint32()
    call("dyn_hull_support_cells", sources.ctypes.data, len(sources), vertices.ctypes.data, len(vertices),
         cells.ctypes.data, None, 0, ctypes.byref(total))
    cands = np.zeros(max(total.value, 1), np.uint16)
    call("dyn_hull_support_cells", sources.ctypes.data, len(sources), vertices.ctypes.data, len(vertices),
         cells.ctypes.data, cands.ctypes.data, total.value, ctypes.byref(total))
    return cells, cands


def cell_of(d):
    """support_cell_of in single precision; None for the linear scan."""
    x, y, z = (np.float32(c) for c in d)
    ax, ay, az = abs(x), abs(y), abs(z)
    if ax >= ay and ax >= az:
        axis, m, c, a, b = 0, ax, x, y, z
    elif ay >= az:
        axis, m, c, a, b = 1, ay, y, x, z
    else:
        axis, m, c, a, b = 2, az, z, x, y
    if not (m >= np.float32(1e-30)) or not (m <= np.float32(1e30)):
        return None
    face = 2 * axis + (1 if c < 0 else 0)
    half = np.float32(FACE_CELLS * 0.5)
    with np.errstate(over="ignore", invalid="ignore"):
        inv = np.float32(1) / m
        fu = (a * inv + np.float32(1)) * half
        fv = (b * inv + np.float32(1)) * half
    iu = min(max(int(fu), 0), FACE_CELLS - 1)
    iv = min(max(int(fv), 0), FACE_CELLS - 1)
    return (face * FACE_CELLS + iu) * FACE_CELLS + iv


def dots(verts, d):
    d = np.asarray(d, np.float32)
    return (verts[:, 0] * d[0] + verts[:, 1] * d[1]) + verts[:, 2] * d[2]


def first_max(values, order):
    best, index = np.float32(-3.402823466e38), None
    for i in order:
        if values[i] > best:
            best, index = values[i], i
    return index


def cell_list(cells_s, cands, cell, offsets):
    """The ascending candidate list of a cell from its two-word record (None: linear scan)."""
    w0, w1 = int(cells_s[cell, 0]), int(cells_s[cell, 1])
    count = w0 & 0xFF
    if count == NO_CELL:
        return None
    if count <= 4:
        return [(w0 >> 8) & 0xFFF, w0 >> 20, w1 & 0xFFF, (w1 >> 12) & 0xFFF][:count]
    return list(cands[w1:w1 + count])


def check_source(verts, cells_s, cands, dirs):
    for d in dirs:
        values = dots(verts, d)
        want = first_max(values, range(len(verts)))
        cell = cell_of(d)
        if cell is None:
            continue
        order = cell_list(cells_s, cands, cell, None)
        if order is None:
            continue
        assert order == sorted(order) and len(set(order)) == len(order)
        got = first_max(values, order)
        assert got == want, (d, cell, got, want)


def boundary_dirs(rng, jitter=True):
    """Directions on every cell corner and edge midpoint of every face, and just either side."""
    grid = np.linspace(-1.0, 1.0, 2 * FACE_CELLS + 1)
    out = []
    for axis, sign in itertools.product(range(3), (1.0, -1.0)):
        others = [a for a in range(3) if a != axis]
        for u, v in itertools.product(grid, grid):
            d = np.zeros(3)
            d[axis] = sign
            d[others[0]], d[others[1]] = u, v
            out.append(d.astype(np.float32))
            if jitter:
                for e in (-1e-6, 1e-6):
                    out.append((d + e * rng.standard_normal(3)).astype(np.float32))
    return out


def test_pile_hulls_match_the_linear_scan():
    from dynamis_amd.scenes import ellipsoid_hull_sources
    from dynamis_amd.world import shape_records
    hulls = ellipsoid_hull_sources(512, 4)
    sources, vertices = shape_records(hulls)
    cells, cands = support_cells(sources, vertices)
    rng = np.random.default_rng(7)
    counts = []
    for s in range(0, len(sources), 8):
        src = sources[s]
        verts = vertices[src["vertex_offset"]:src["vertex_offset"] + src["vertex_count"], :3]
        counts.append(np.mean([len(cell_list(cells[s], cands, c, None)) for c in range(CELLS)]) / len(verts))
        dirs = list(rng.standard_normal((300, 3)).astype(np.float32))
        dirs += list((rng.standard_normal((50, 3)) * 1e-20).astype(np.float32))
        dirs += list((rng.standard_normal((50, 3)) * 1e20).astype(np.float32))
        dirs += boundary_dirs(rng, jitter=(s % 64 == 0))
        check_source(verts, cells[s], cands, dirs)
    print(f"mean candidates per cell: {np.mean(counts):.3f} of the hull's vertices")
    assert np.mean(counts) < 0.3


def test_exact_ties_keep_the_first_vertex():
    # a rhombicuboctahedron-like set: every permutation of (+-1, +-1, +-3), 24 vertices, which tie
    # in groups along the axes and the diagonals
    pts = sorted({tuple(p) for base in itertools.permutations((1.0, 1.0, 3.0))
                  for p in itertools.product(*[(c, -c) for c in base])})
    verts4 = np.zeros((len(pts), 4), np.float32)
    verts4[:, :3] = np.array(pts, np.float32)
    sources = np.zeros(1, abi.SHAPE_SOURCE)
    sources["kind"] = abi.SHAPE_HULL
    sources["vertex_count"] = len(pts)
    cells, cands = support_cells(sources, verts4)
    rng = np.random.default_rng(3)
    dirs = boundary_dirs(rng) + [np.eye(3, dtype=np.float32)[i] * s for i in range(3) for s in (1, -1)]
    dirs += [np.array(d, np.float32) for d in itertools.product((-1, 0, 1), repeat=3) if any(d)]
    check_source(verts4[:, :3], cells[0], cands, dirs)


def test_small_hulls_keep_the_linear_scan():
    verts4 = np.zeros((8, 4), np.float32)
    verts4[:, :3] = np.array(list(itertools.product((-1, 1), repeat=3)), np.float32)
    sources = np.zeros(1, abi.SHAPE_SOURCE)
    sources["kind"] = abi.SHAPE_HULL
    sources["vertex_count"] = 8
    cells, _ = support_cells(sources, verts4)
    assert ((cells[0, :, 0] & 0xFF) == NO_CELL).all()
