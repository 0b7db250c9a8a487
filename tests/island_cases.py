"""Expected islands from contacts by scipy's connected components (an independent library
routine): the components of the graph whose edges are the touching, non-sensor contacts between
two simulated rows, each row mapped to its component's minimum row."""

from __future__ import annotations

import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components


def linking_edges(contacts: np.ndarray, owners: np.ndarray, dynamic_count: int, slop: float) -> np.ndarray:
    if len(contacts) == 0:
        return np.zeros((0, 2), np.int64)
    count = contacts["point_count"].astype(np.int64)
    depth = contacts["points"]["depth"]  # (C, 4)
    live = np.arange(4)[None, :] < count[:, None]
    touching = ((depth >= -np.float32(slop)) & live).any(1)
    keep = (count > 0) & (contacts["sensor"] == 0) & touching
    ra = owners[contacts["a"]].astype(np.int64)
    rb = owners[contacts["b"]].astype(np.int64)
    keep &= (ra < dynamic_count) & (rb < dynamic_count)
    return np.stack([ra[keep], rb[keep]], axis=1)


def expected_roots(contacts: np.ndarray, owners: np.ndarray, dynamic_count: int, slop: float):
    """(root of each simulated row, number of islands)."""
    edges = linking_edges(contacts, owners, dynamic_count, slop)
    d = dynamic_count
    graph = coo_matrix((np.ones(len(edges)), (edges[:, 0], edges[:, 1])), shape=(d, d))
    count, labels = connected_components(graph, directed=False)
    first = np.full(count, d, np.int64)
    np.minimum.at(first, labels, np.arange(d))
    return first[labels].astype(np.uint32), int(count)
