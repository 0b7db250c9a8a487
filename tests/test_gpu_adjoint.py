"""The differentiable rollout on the device (dyn_adjoint_*, dyn_world_snapshot/restore), checked
against the host reverse and FP64 central differences of the frozen-contact rollout on the
device's own tapes, and through the torch.autograd.Function harness.

Workload: 8 worlds of the differentiation workload (32 boxes on a plane each, initial velocities
~ N(0, 2 m/s), friction ~ U[0.2, 0.8]), dt = 1/120, 8 iterations, 64 steps, loss = sum over worlds
of the squared distance of the final centre of mass to a target.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from dynamis_amd import adjoint as adj
from dynamis_amd import scenes
from dynamis_amd._native import DynamisError, call
from dynamis_amd.world import Batch

from . import adjoint_host as ah

pytestmark = pytest.mark.gpu

DT = 1.0 / 120.0
WORLDS, STEPS = 8, 64


def com_gradient(batch, targets, states):
    """Loss, its adjoint of the final states and its direct derivative in each row's mass."""
    L = batch.layout
    n = len(states)
    g = np.zeros((n, adj.STATE_FLOATS), np.float32)
    direct = np.zeros(n)
    loss = 0.0
    for w in range(batch.worlds):
        rows = L.dynamic_base[w] + np.arange(L.dynamic_count[w])
        m = 1.0 / batch.scene.descs["inverse_mass"][rows].astype(np.float64)
        p = states["position"][rows].astype(np.float64)
        com = (m[:, None] * p).sum(0) / m.sum()
        r = com - targets[w]
        loss += float(r @ r)
        g[rows, 0:3] = 2 * r[None, :] * (m / m.sum())[:, None]
        direct[rows] = 2 * ((p - com) @ r) / m.sum()
    return loss, g, direct


@pytest.fixture(scope="module")
def rollout():
    sc, cfg, targets = scenes.adjoint_worlds(WORLDS, 32, seed=5)
    batch = Batch(sc, cfg)
    a = adj.Adjoint(batch, STEPS, STEPS)  # one segment: every step's tape stays readable
    a.forward(DT)
    end = batch.states().copy()
    loss, g, direct = com_gradient(batch, targets, end)
    g0, gm, gf = a.backward(g)
    tapes = [a.tape(t) for t in range(STEPS)]
    bodies = a.bodies()
    params = adj.adj_params(batch.scene.params, DT)
    ranges = np.zeros(WORLDS, adj.RANGE)
    for w, r in enumerate(adj.world_ranges(batch)):
        ranges[w] = tuple(r)
    tape = ah.Tape(params, ranges, bodies, [c for _, c in tapes], [s for s, _ in tapes], len(batch.scene.colliders),
                   batch.capacity)
    yield dict(batch=batch, adjoint=a, targets=targets, end=end, g=g, direct=direct, g0=g0, gm=gm, gf=gf,
               tape=tape, bodies=bodies, sc=sc, cfg=cfg)
    a.close()


def test_snapshot_restore_is_bit_exact():
    sc, cfg, _ = scenes.adjoint_worlds(4, 32, seed=6)
    batch = Batch(sc, cfg)
    for _ in range(5):
        batch.step(DT)
    size = ctypes.c_size_t()
    call("dyn_world_snapshot_bytes", batch._handle, ctypes.byref(size))
    buf = adj._Device(size.value)
    try:
        call("dyn_world_snapshot", batch._handle, buf.ptr)
        for _ in range(8):
            batch.step(DT)
        first = batch.states().copy()
        first_contacts = batch.inspect_contacts().copy()
        call("dyn_world_restore", batch._handle, buf.ptr)
        for _ in range(8):
            batch.step(DT)
        assert batch.states().tobytes() == first.tobytes()
        assert batch.inspect_contacts().tobytes() == first_contacts.tobytes()
    finally:
        buf.free()


def test_checkpointed_backward_equals_one_segment(rollout):
    sc, cfg = rollout["sc"], rollout["cfg"]
    batch = Batch(sc, cfg)
    a = adj.Adjoint(batch, STEPS, 8)
    try:
        a.forward(DT)
        assert batch.states().tobytes() == rollout["end"].tobytes()
        g0, gm, gf = a.backward(rollout["g"])
        assert np.array_equal(g0, rollout["g0"])
        assert np.array_equal(gm, rollout["gm"])
        assert np.array_equal(gf, rollout["gf"])
        # the world is left where the forward rollout ended
        assert batch.states().tobytes() == rollout["end"].tobytes()
    finally:
        a.close()


def _per_world(rollout):
    L = rollout["batch"].layout
    for w in range(WORLDS):
        rows = L.dynamic_base[w] + np.arange(L.dynamic_count[w])
        cols = L.collider_base[w] + np.arange(L.collider_count[w])
        yield w, rows, cols


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_device_reverse_equals_host_reverse(rollout):
    """The device kernel and the same templates run on the host in FP32 about the same recorded
    states agree (the host reverse is checked against differences in the CPU tests)."""
    tape, bodies = rollout["tape"], rollout["bodies"]
    gx, gs, gm = tape.gradient(rollout["g"].astype(np.float64), single=True, anchored=True)
    mass_host = gs * bodies[:, 0]
    for w, rows, cols in _per_world(rollout):
        assert _rel(rollout["g0"][rows], gx[rows]) <= 1e-6, w
        assert _rel(rollout["gm"][rows], mass_host[rows]) <= 1e-6, w
        assert _rel(rollout["gf"][cols], gm[cols]) <= 1e-6, w


def test_device_gradient_within_1e3_of_fp64(rollout):
    """Initial velocities, masses and frictions: the device's FP32 gradient against the FP64
    reverse about the same recorded rollout, world by world."""
    tape, bodies = rollout["tape"], rollout["bodies"]
    gx, gs, gm = tape.gradient(rollout["g"].astype(np.float64), anchored=True)
    mass64 = gs * bodies[:, 0]
    worst = 0.0
    for w, rows, cols in _per_world(rollout):
        errs = (_rel(rollout["g0"][rows, 8:11], gx[rows, 8:11]), _rel(rollout["gm"][rows], mass64[rows]),
                _rel(rollout["gf"][cols], gm[cols]))
        worst = max(worst, *errs)
        assert max(errs) <= 1e-3, (w, errs)
    print(f"worst per-world relative error against the FP64 reverse: {worst:.2e}")


def test_fp64_central_differences(rollout):
    """FP64 differences of the loss of the frozen-contact rollout about the recorded states, for
    every initial velocity, mass and friction of one world, against the device gradient. Contact
    dynamics is piecewise smooth: at h = 1e-7 every parameter agrees with its central difference
    within 1e-3, or sits at a kink within h - its forward and backward differences disagree by more
    than 1e-3 and the device gradient (the slope of the branch its recording took) agrees with one
    of them within 1e-3. At the workload's h = 1e-4 the differences straddle more of the clamps'
    kinks; at least 70 % of the parameters still agree within 1e-3 there."""
    batch, tape, targets = rollout["batch"], rollout["tape"], rollout["targets"]
    L = batch.layout
    w = 0
    rows = L.dynamic_base[w] + np.arange(L.dynamic_count[w])
    cols = L.collider_base[w] + np.arange(L.collider_count[w])
    one = ah.Tape(tape.params, tape.ranges[w:w + 1], tape.bodies, tape.contacts, tape.starts, tape.k, tape.capacity)
    m0 = 1.0 / tape.bodies[rows, 0].astype(np.float64)
    x0 = one.x0()
    mu0 = batch.scene.colliders["friction"].astype(np.float64)

    def loss(x, s=None):
        m = m0 * (1.0 if s is None else s[rows])
        com = (m[:, None] * x[rows, 0:3]).sum(0) / m.sum()
        r = com - targets[w]
        return float(r @ r)

    params = []
    for i, r in enumerate(rows):
        for c in (8, 9, 10):
            def fv(e, r=r, c=c):
                x = x0.copy()
                x[r, c] += e
                return loss(one.rollout(x, anchored=True))
            params.append((rollout["g0"][r, c], fv, 1.0))
        def fs(e, r=r):
            s = np.ones(tape.n)
            s[r] += e
            return loss(one.rollout(s=s, anchored=True), s)
        params.append((rollout["gm"][r] + rollout["direct"][r], fs, m0[i]))  # d/dmass = d/dscale / mass
    for c in cols:
        def fm(e, c=c):
            mu = mu0.copy()
            mu[c] += e
            return loss(one.rollout(mu=mu, anchored=True))
        params.append((rollout["gf"][c], fm, 1.0))
    fine = [ah.differences(f, 1e-7, unit) for _, f, unit in params]
    scale = max(abs(v[0]) for v in fine)
    kinds = [ah.classify(dev, diff, scale) for (dev, _, _), diff in zip(params, fine)]
    rel = [abs(dev - diff[0]) / max(abs(diff[0]), 1e-2 * scale) for (dev, _, _), diff in zip(params, fine)]
    kinks = [i for i, k in enumerate(kinds) if k == "kink"]
    print(f"h = 1e-7: {kinds.count('smooth')} of {len(params)} parameters within 1e-3 of the central "
          f"difference (median relative error {np.median(rel):.1e}); at kinks: "
          f"{[(i, params[i][0], fine[i]) for i in kinks]}")
    assert kinds.count("off") == 0, [(i, params[i][0], fine[i]) for i, k in enumerate(kinds) if k == "off"]
    coarse = [ah.differences(f, 1e-4, unit) for _, f, unit in params]
    at_config_h = sum(ah.classify(dev, diff, scale) == "smooth" for (dev, _, _), diff in zip(params, coarse))
    print(f"h = 1e-4: {at_config_h} of {len(params)} within 1e-3")
    assert at_config_h >= 0.7 * len(params)


def calm_worlds(L, tapes, worlds: int, speed: float = 15.0) -> np.ndarray:
    """Worlds whose bodies stay below `speed` through the recorded rollout (the initial speeds
    are ~ N(0, 2 m/s) per axis): the others have a body thrown by a non-positive manifold row
    mass (solver_contact_block.wgsl:14-28), whose step amplifies rounding by orders of magnitude,
    so FP32 and FP64 reverses of it part."""
    top = np.zeros(worlds)
    for states, _ in tapes:
        v = np.linalg.norm(states[:, 8:11].astype(np.float64), axis=1)
        for w in range(worlds):
            top[w] = max(top[w], v[L.dynamic_base[w]:L.dynamic_base[w] + L.dynamic_count[w]].max())
    return top <= speed


def test_full_workload_sampled_worlds_against_fp64_reverse():
    """The whole differentiation workload (256 worlds of 32 boxes, 64 steps at dt = 1/120,
    checkpoints every 8): six worlds spread over the batch among those that throw no body, the
    device gradient of each against the FP64 reverse about the device's own recording, within
    1e-3 per world; every world's checkpointed gradient equals its one-segment gradient."""
    worlds = 256
    sc, cfg, targets = scenes.adjoint_worlds(worlds, 32, seed=5)
    batch = Batch(sc, cfg)
    a = adj.Adjoint(batch, STEPS, 8)
    try:
        a.forward(DT)
        end = batch.states().copy()
        _, g, _ = com_gradient(batch, targets, end)
        g0, gm, gf = a.backward(g)
        # the same rollout recorded as one segment (a fresh batch of the same worlds), whose tape
        # the host reverse reads; its gradient equals the checkpointed one bit for bit
        again = Batch(sc, cfg)
        whole = adj.Adjoint(again, STEPS, STEPS)
        try:
            whole.forward(DT)
            assert again.states().tobytes() == end.tobytes()
            w0, wm, wf = whole.backward(g)
            # byte equality: a world whose rollout throws a body (a manifold row whose mass comes
            # out non-positive, solver_contact_block.wgsl:14-28) overflows its FP32 reverse to
            # NaN, identically in both passes
            assert w0.tobytes() == g0.tobytes() and wm.tobytes() == gm.tobytes() and wf.tobytes() == gf.tobytes()
            tapes = [whole.tape(t) for t in range(STEPS)]
            bodies = whole.bodies()
        finally:
            whole.close()
            again.close()
        params = adj.adj_params(batch.scene.params, DT)
        ranges = np.zeros(worlds, adj.RANGE)
        for w, r in enumerate(adj.world_ranges(batch)):
            ranges[w] = tuple(r)
        L = batch.layout
        one = ah.Tape(params, ranges[0:1], bodies, [c for _, c in tapes], [s for s, _ in tapes],
                      len(batch.scene.colliders), batch.capacity)
        calm = calm_worlds(L, tapes, worlds)
        print(f"[adjoint] {int((~calm).sum())} of {worlds} worlds throw a body (speed above 15 m/s)", flush=True)
        sample = [int(np.flatnonzero(calm[start:])[0]) + start for start in (0, 51, 102, 153, 204, 250)]
        for w in sample:
            one.ranges = ranges[w:w + 1].copy()
            gx, gs, gmu = one.gradient(g.astype(np.float64), anchored=True)
            rows = L.dynamic_base[w] + np.arange(L.dynamic_count[w])
            cols = L.collider_base[w] + np.arange(L.collider_count[w])
            errs = (_rel(g0[rows, 8:11], gx[rows, 8:11]), _rel(gm[rows], gs[rows] * bodies[rows, 0]),
                    _rel(gf[cols], gmu[cols]))
            print(f"[adjoint] world {w}: relative errors against the FP64 reverse {errs}", flush=True)
            assert max(errs) <= 1e-3, (w, errs)
    finally:
        a.close()


def test_torch_autograd_function(rollout):
    torch = pytest.importorskip("torch")
    from .adjoint_torch import Rollout, Runner, com_loss
    runner = Runner(worlds=WORLDS, steps=STEPS, every=8, seed=5)
    v0, mass, friction = (torch.tensor(a, dtype=torch.float32, requires_grad=True) for a in runner.initial())
    positions = Rollout.apply(v0, mass, friction, runner)
    loss = com_loss(positions.double(), mass.double(), torch.tensor(runner.targets, dtype=torch.float64))
    loss.backward()
    L = runner.batch.layout
    for w in range(WORLDS):
        rows = L.dynamic_base[w] + np.arange(L.dynamic_count[w])
        cols = L.collider_base[w] + np.arange(L.collider_count[w])
        # The loss adjoint torch hands back rounds differently from the fixture's, and the FP32
        # reverse carries that through 64 steps: the same bar as against FP64, world by world.
        assert _rel(v0.grad[w].numpy(), rollout["g0"][rows, 8:11]) <= 1e-3, w
        assert _rel(mass.grad[w].numpy(), rollout["gm"][rows] + rollout["direct"][rows]) <= 1e-3, w
        assert _rel(friction.grad[w].numpy(), rollout["gf"][cols]) <= 1e-3, w
    runner.adjoint.close()


def test_soft_contacts_are_refused():
    sc, cfg, _ = scenes.adjoint_worlds(2, 8, seed=7)
    for s in sc:
        s.colliders["relaxation"] = 0.1
    batch = Batch(sc, cfg)
    a = adj.Adjoint(batch, 48, 8)
    try:
        a.forward(DT)
        with pytest.raises(DynamisError, match="soft contacts"):
            a.backward(np.zeros((len(batch.scene.states), adj.STATE_FLOATS), np.float32))
    finally:
        a.close()
