"""The oracle's OpenMP build (liboracle_omp.so, the all-core CPU baseline) against its serial
build: a mixed convex pile (spheres, boxes, capsules, hulls, cylinders) stepped through its drop
and first contacts gives the same body states, contacts and counters byte for byte, whatever the
thread count, since a pair writes only its own slot and the solver's shared words are integer
sums and maxima."""

from __future__ import annotations

import os

import pytest

from tests.narrow_cases import DT, pile_scene
from tests.oracle_binding import load


def run(oracle, scene, steps):
    world = oracle.world(scene)
    for _ in range(steps):
        world.step(DT)
    out = world.states().copy(), world.inspect_contacts().copy(), world.measured()
    world.close()
    return out


@pytest.mark.skipif((os.cpu_count() or 1) < 2, reason="one core: the OpenMP build runs serially")
def test_openmp_build_steps_like_the_serial_one():
    serial, parallel = load(), load(parallel=True)
    assert serial.threads() == 1 and parallel.threads() >= 2
    scene = pile_scene(600, seed=7, hull_sources=8, spacing=1.4)
    s0, c0, m0 = run(serial, scene, 50)
    s1, c1, m1 = run(parallel, scene, 50)
    assert len(c0) > 200
    assert s1.tobytes() == s0.tobytes()
    assert c1.tobytes() == c0.tobytes()
    assert m1 == m0
