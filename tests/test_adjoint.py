"""CPU checks of the differentiable step (dynamis_amd/csrc/adjoint.h run on the host through
oracle/adjoint_host.cpp) on tapes recorded by the CPU oracle's steps.

- The frozen-contact replay of a step reproduces the oracle's step to the resolution of the
  step's fixed-point velocity and position words.
- The reverse, in FP64, matches FP64 central differences of the frozen-contact rollout's loss for
  every kind of parameter (initial positions, orientations, velocities, spins, mass scales,
  collider frictions) through box-box and box-plane contacts, the warm-start relay, the Jacobi
  sweeps, the position phase and the gyroscopic integrator.
- The FP32 reverse (what the device runs) stays within 1e-3 of the FP64 reverse.
"""

from __future__ import annotations

import numpy as np
import pytest

from dynamis_amd import adjoint as adj
from dynamis_amd import scenes

from . import adjoint_host as ah

DT = 1.0 / 120.0


@pytest.fixture(scope="module")
def small():
    """8 boxes dropped onto the plane, 40 steps: box-plane and box-box contacts, relayed impulses."""
    sc, _, _ = scenes.adjoint_worlds(1, 8, seed=1, speed=1.0)
    tape, after = ah.oracle_tape(sc[0], 40, DT)
    assert sum(len(c) for c in tape.contacts) > 40
    return sc[0], tape, after


def test_replay_reproduces_oracle_steps(small):
    scene, tape, after = small
    d = scene.dynamic_count
    for t in range(tape.steps):
        one = ah.Tape(tape.params, tape.ranges, tape.bodies, [tape.contacts[t]], [tape.starts[t]], tape.k,
                      tape.capacity)
        out = one.rollout()[:d]
        ref = adj.state_records(after[t]).astype(np.float64)[:d]
        # velocity words are 2^-18 m/s apart and truncated once per sweep (18 sweeps a step)
        assert np.abs(out[:, 8:15] - ref[:, 8:15]).max() < 2e-4, t
        assert np.abs(out[:, 0:3] - ref[:, 0:3]).max() < 1e-5, t
        assert np.abs(out[:, 4:8] - ref[:, 4:8]).max() < 1e-5, t


def _linear_loss(n, d, seed=0):
    rng = np.random.default_rng(seed)
    w = np.zeros((n, 16))
    w[:d, 0:3] = rng.normal(size=(d, 3))
    w[:d, 4:8] = rng.normal(size=(d, 4)) * 0.1
    w[:d, 8:11] = rng.normal(size=(d, 3)) * 0.1
    w[:d, 12:15] = rng.normal(size=(d, 3)) * 0.05
    return w


def test_reverse_matches_central_differences(small):
    scene, tape, _ = small
    d, n = scene.dynamic_count, tape.n
    w = _linear_loss(n, d)
    gx, gs, gm = tape.gradient(w)
    x0 = tape.x0()
    mu0 = scene.colliders["friction"].astype(np.float64)
    h = 1e-7

    def loss(x):
        return float((w * x).sum())

    def central(make):
        return (loss(make(+h)) - loss(make(-h))) / (2 * h)

    checked = 0
    for r in range(d):
        for c in (0, 1, 2, 4, 5, 6, 7, 8, 9, 10, 12, 13, 14):
            def make(e, r=r, c=c):
                x = x0.copy()
                x[r, c] += e
                return tape.rollout(x)
            fd = central(make)
            assert abs(gx[r, c] - fd) <= 1e-4 * max(abs(fd), 1e-2), (r, c, gx[r, c], fd)
            checked += 1
        def make_s(e, r=r):
            s = np.ones(n)
            s[r] += e
            return tape.rollout(s=s)
        fd = central(make_s)
        assert abs(gs[r] - fd) <= 1e-4 * max(abs(fd), 1e-2), (r, gs[r], fd)
    for c in range(len(mu0)):
        def make_mu(e, c=c):
            m = mu0.copy()
            m[c] += e
            return tape.rollout(mu=m)
        fd = central(make_mu)
        assert abs(gm[c] - fd) <= 1e-4 * max(abs(fd), 1e-2), (c, gm[c], fd)
    assert checked == 13 * d


def test_fp32_reverse_close_to_fp64():
    """A world of the differentiation workload (32 boxes, 64 steps): the FP32 reverse about the
    recorded states against the FP64 one, for the centre-of-mass loss."""
    sc, _, targets = scenes.adjoint_worlds(1, 32, seed=2)
    scene = sc[0]
    tape, _ = ah.oracle_tape(scene, 64, DT)
    d, n = scene.dynamic_count, tape.n
    m = 1.0 / tape.bodies[:d, 0].astype(np.float64)
    end = tape.starts[-1].astype(np.float64)
    com = (m[:, None] * end[:d, 0:3]).sum(0) / m.sum()
    w = np.zeros((n, 16))
    w[:d, 0:3] = 2 * (com - targets[0])[None, :] * (m / m.sum())[:, None]
    g64 = tape.gradient(w, anchored=True)
    g32 = tape.gradient(w, single=True, anchored=True)
    for a, b in ((g64[0][:d, 8:11], g32[0][:d, 8:11]), (g64[1][:d], g32[1][:d]), (g64[2], g32[2])):
        assert np.linalg.norm(a - b) <= 1e-3 * np.linalg.norm(a)


def test_anchored_rollout_is_linearised_about_the_recording(small):
    """The anchored rollout's derivative is the chain of the steps' derivatives at the recorded
    states: central differences of it match the anchored FP64 reverse."""
    scene, tape, _ = small
    d, n = scene.dynamic_count, tape.n
    w = _linear_loss(n, d, seed=1)
    gx, gs, _ = tape.gradient(w, anchored=True)
    x0 = tape.x0()
    h = 1e-7
    for r in range(d):
        for c in (0, 8, 12):
            xp, xm = x0.copy(), x0.copy()
            xp[r, c] += h
            xm[r, c] -= h
            fd = (float((w * tape.rollout(xp, anchored=True)).sum()) -
                  float((w * tape.rollout(xm, anchored=True)).sum())) / (2 * h)
            assert abs(gx[r, c] - fd) <= 1e-4 * max(abs(fd), 1e-2), (r, c, gx[r, c], fd)


def test_friction_partials():
    a = np.array([0.2, 0.5, 0.8], np.float32)
    b = np.array([0.5, 0.5, 0.3], np.float32)
    for mode in range(4):
        c, da, db = adj.friction_partials(a, b, mode)
        e = 1e-3
        cp, _, _ = adj.friction_partials(a + e, b, mode)
        cm, _, _ = adj.friction_partials(a - e, b, mode)
        smooth = np.abs(a - b) > 2 * e
        np.testing.assert_allclose(da[smooth], ((cp - cm) / (2 * e))[smooth], atol=2e-3)


def test_every_parameter_matches_differences_or_sits_at_a_kink():
    """A world of the differentiation workload (32 boxes, 64 steps, centre-of-mass loss), the
    FP64 reverse about the oracle's recording against FP64 differences of the anchored rollout at
    h = 1e-7, for every initial velocity, mass and friction: each agrees with the central difference
    within 1e-3, or sits at a kink (its one-sided differences disagree, and the reverse matches the
    one on the side of the branch the recording took). At the workload's h = 1e-4 a quarter of the
    parameters straddle kinks of the clamps; each of those is one of the above at h = 1e-7."""
    sc, _, targets = scenes.adjoint_worlds(1, 32, seed=5)
    scene = sc[0]
    tape, _ = ah.oracle_tape(scene, 64, DT)
    d, n = scene.dynamic_count, tape.n
    inv = tape.bodies[:, 0].astype(np.float64)
    m0 = 1.0 / inv[:d]
    x0 = tape.x0()
    mu0 = scene.colliders["friction"].astype(np.float64)

    def loss(x, s=None):
        m = m0 * (1.0 if s is None else s[:d])
        com = (m[:, None] * x[:d, 0:3]).sum(0) / m.sum()
        r = com - targets[0]
        return float(r @ r)

    end = tape.rollout(anchored=True)
    com = (m0[:, None] * end[:d, 0:3]).sum(0) / m0.sum()
    r0 = com - targets[0]
    w = np.zeros((n, 16))
    w[:d, 0:3] = 2 * r0[None, :] * (m0 / m0.sum())[:, None]
    direct = 2 * ((end[:d, 0:3] - com) @ r0) / m0.sum()
    gx, gs, gm = tape.gradient(w, anchored=True)
    params = []
    for r in range(d):
        for c in (8, 9, 10):
            def fv(e, r=r, c=c):
                x = x0.copy()
                x[r, c] += e
                return loss(tape.rollout(x, anchored=True))
            params.append((gx[r, c], fv, 1.0))
        def fs(e, r=r):
            s = np.ones(n)
            s[r] += e
            return loss(tape.rollout(s=s, anchored=True), s)
        params.append((gs[r] * inv[r] + direct[r], fs, m0[r]))
    for c in range(len(mu0)):
        def fm(e, c=c):
            mu = mu0.copy()
            mu[c] += e
            return loss(tape.rollout(mu=mu, anchored=True))
        params.append((gm[c], fm, 1.0))
    fine = [ah.differences(f, 1e-7, unit) for _, f, unit in params]
    scale = max(abs(v[0]) for v in fine)
    kinds = [ah.classify(dev, diff, scale) for (dev, _, _), diff in zip(params, fine)]
    assert kinds.count("off") == 0, [i for i, k in enumerate(kinds) if k == "off"]
    coarse = [ah.differences(f, 1e-4, unit) for _, f, unit in params]
    at_config_h = sum(ah.classify(dev, diff, scale) == "smooth" for (dev, _, _), diff in zip(params, coarse))
    print(f"h = 1e-7: {kinds.count('smooth')} smooth, {kinds.count('kink')} at kinks of {len(params)}; "
          f"h = 1e-4: {at_config_h} within 1e-3")
    assert at_config_h >= 0.7 * len(params)
