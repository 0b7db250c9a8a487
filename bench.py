#!/usr/bin/env python3
"""Throughput benchmark of the rigid-body step on one MI355X.

Workloads (`--workload`):
  pile        (default) the mixed convex pile: 100,000 bodies, a quarter each of spheres, boxes,
              capsules and ellipsoid hulls of 16-64 points, density 1000, on a jittered lattice in
              a 60 x 60 m bin of four static walls on a plane; FP32, dt = 1/60, 8 solver
              iterations (2 substeps, 10 position iterations); the run is 600 steps.
  cubes       1,000 cubes (half-extent 0.5 m, density 1000, friction 0.5) dropped from 1-40 m over
              a 20 x 20 m footprint onto a plane; 4 solver iterations; 300 steps.
  batch       4,096 independent worlds stepped as one batch (dyn_batch_create): each a plane and
              N_w bodies, N_w log-uniform in [8, 2048] (1,550,732 bodies in all), half boxes
              (half-extents ~ U[0.2, 0.8]) and half spheres (r ~ U[0.25, 0.75]), density 1000;
              8 solver iterations, 240 steps. Also times the same worlds stepped one at a time on
              one stream (every 32nd world through the same run, the same window timed, scaled
              to all 4,096).
  broadphase  the isolated sort + broadphase pass over 1,048,576 boxes.
  fixed_grid  the same boxes through the literal fixed-cell mode: a uniform 4 m grid, keys
              (cell hash, box index), LSD radix sort, cell runs, canonical overlapping pairs
              (dyn_fixed_grid; the same pair set as the hierarchical grid).
  copy        calibration only: device-to-device copies of 1 GiB, whose HBM counter readings
              (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE) scale the counters of the other kernels.
  narrowphase the narrowphase pass alone (dyn_narrowphase) over the candidate pairs of the pile's
              600th step, saved by a first run (--stage-file) so that a later run executes the pass
              and nothing else (for rocprofv3 counter passes).
  solve       the solver alone (dyn_solve) over the contacts of the pile's 600th step, from the
              same saved stage inputs (for rocprofv3 counter passes on the solver's kernels).
  adjoint     the differentiable rollout alone: 256 worlds of 32 boxes (half-extents
              ~ U[0.2, 0.5]) on a plane, v0 ~ N(0, 2 m/s), friction ~ U[0.2, 0.8], dt = 1/120,
              8 iterations, 64 steps with a checkpoint every 8; forward and backward times and
              their ratio (also reported as `adjoint` in the default line).
One step = one World::step of every body (broadphase: one pass over every box); body-steps/s =
bodies x timed steps / seconds. Inputs are resident in HBM before the timed region.

The world workloads time a settled window: of the workload's run (pile 600 steps, batch 240,
cubes 300) the first run - steps - warmup steps settle the scene untimed, `--warmup` more steps
are untimed, and the last `--steps` steps are timed by events on the world's own stream as
replays of the captured step graph, with no launch bracketed. The default line (pile) also carries
the batch's settled window with its ratio to the same worlds stepped one at a time (`batch`), the
isolated sort + broadphase pass (`broadphase`) and the adjoint's backward/forward ratio
(`adjoint`).

`roofline` covers the kernel with the largest device time among those with an algorithmic byte
count: the timed steps are re-run eagerly from a snapshot taken at their start with every launch
bracketed by events (the `stages` table), then re-run again with only that kernel's launches
bracketed, which gives its average launch duration (both re-runs end in the timed run's states
bit for bit: `config.reruns_identical`). The second re-run is bracketed by roctxProfilerResume /
roctxProfilerPause marks. Bytes per launch come from the step's counts at the window's start and
end; `traffic` (HBM bytes per launch from rocprofv3 counter passes) comes from
profiles/traffic.json when those passes were taken on the same kernel and workload. `cpu_baseline` times the C
oracle on a bounded sample of the same workload.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "body-steps/s (100k convex pile) + batched vs plain; % HBM peak; bwd/fwd ratio"
DT = np.float32(1.0 / 60.0)
GRAPHS = True  # steps replay the captured graph; --no-graphs launches them eagerly
BOXES = 1 << 20
BOX_SEED = 2


def progress(message: str) -> None:
    """One line on stderr per finished phase (the JSON line alone goes to stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {message}", file=sys.stderr, flush=True)


# ---- algorithmic bytes per launch ---------------------------------------------------------------

def broadphase_bytes(name: str, k: int, entries: int, pairs: int) -> float:
    """Minimal HBM payload of one launch of each broadphase kernel (entry = key 4 + slot 4 +
    record 32 bytes; box = 24 bytes of min/max; pair = 8 bytes)."""
    if name == "entry_sort_pass":        # read key+slot, write key+slot
        return 16.0 * entries
    if name in ("bp_mates_count_kernel", "bp_mates_emit_kernel"):  # every sorted key read once;
        return 4.0 * entries + 8.0 * pairs                             # a pair's record each
    if name in ("bp_links_count_kernel", "bp_links_emit_kernel"):  # box + packed count per collider;
        return 28.0 * k + 8.0 * pairs                                  # a pair's record each
    if name == "bp_emit_kernel":         # box + packed count in; key, slot, record out
        return 28.0 * k + 40.0 * entries
    if name == "bp_count_kernel":        # box in, packed count out
        return 28.0 * k
    if name == "bp_resolution_kernel":   # box in
        return 24.0 * k
    if name == "pair_sort_pass":         # 3 lanes in and out
        return 24.0 * pairs
    if name == "pair_sort_histogram":
        return 8.0 * pairs
    return 0.0


def step_bytes(name: str, n: int, k: int, entries: int, pairs: int, contacts: int, points: float,
               rows: int, gjk: int = 0, epa: int = 0, convex_manifolds: int = 0, gjk_hull: int = 0,
               gjk_resumed: int = 0, epa_wave: int = 0) -> float:
    """Minimal HBM payload of one launch of each step kernel, from the step's own counts:
    N bodies, K colliders, E entries, P pairs, C contacts of m points on average, R solver rows.
    A manifold is M(m) = 48 + 56 m bytes (normal, pair and material words plus per point anchor,
    depth, impulses and feature); pose + velocity 52 B and body constants 44 B per body."""
    manifold = 48.0 + 56.0 * points
    manifold_pairs = convex_manifolds
    if name == "k_solve":               # manifold, two bodies' pose/velocity/constants, targets,
        return contacts * (manifold + 16.0 * points + 240.0)  # delta words
    if name == "k_position_block":      # position row, both sides' resolution, both sides' words
        return contacts * (48.0 + 32.0 + 32.0)
    if name in ("k_apply", "k_position_apply"):
        return rows * 72.0
    if name == "k_blocks":              # manifold, two bodies' 48-byte records, position row,
        return contacts * (manifold + 96.0 + 48.0 + 16.0 + 16.0)  # targets, block and slot words
    if name == "k_solve_prep":          # manifold, two bodies' pose and solver records, the row out
        return contacts * (manifold + 224.0 + 48.0 + 48.0 * points)
    if name == "k_island_seed":         # contact header and depths, owners, the linking rows out
        return contacts * (manifold + 16.0)
    if name in ("k_island_hook", "k_island_link"):  # the linking rows, both rows' parents
        return contacts * 16.0
    if name == "k_narrow":              # worklist entry, two bodies and colliders per pair, manifold out
        return pairs * 320.0 + contacts * manifold
    if name == "k_pair_class":          # the pair, both colliders' kinds, the class out
        return pairs * 17.0
    if name == "k_narrow_convex":       # the GJK worklist: pair record + index, parked results
        return gjk * 320.0 + manifold_pairs * 68.0 + gjk_resumed * 208.0  # and parked GJK states
    if name == "k_narrow_convex_resume":  # pair record + index, GJK state in, results out
        return gjk_resumed * (320.0 + 208.0 + 68.0)
    if name == "k_narrow_epa_small":    # pair record, parked simplex in, result out
        return epa * (316.0 + 48.0 + 20.0)
    if name == "k_narrow_epa":          # the polytopes given a wave: index, pair record, simplex, result
        return epa_wave * (4.0 + 316.0 + 48.0 + 20.0)
    if name == "k_narrow_finish":       # pair record, parked result, contact out
        return manifold_pairs * (316.0 + 68.0 + 352.0)
    if name in ("k_integrate", "k_advance"):
        return n * 126.0
    if name in ("k_relay", "k_contact_begin", "k_contacts_scatter"):
        return 2.0 * contacts * manifold
    if name == "k_archive":             # the archive is the swapped contact buffer: freeze marks only
        return contacts * 40.0
    return broadphase_bytes(name, k, entries, pairs)


# ---- workloads ------------------------------------------------------------------------------------

def pile_scene():
    from dynamis_amd.scenes import convex_pile
    from dynamis_amd.world import default_config, encode
    bodies, hulls = convex_pile(100_000, seed=3)
    return encode(bodies, default_config(), hulls)


def cubes_scene():
    from dynamis_amd.scenes import cube_drop
    from dynamis_amd.world import default_config, encode
    cfg = default_config()
    cfg["solve_iterations"] = 4
    return encode(cube_drop(1000, seed=1), cfg)


BATCH_WORLDS = 4096
BATCH_SEED = 4


def batch_scenes():
    from dynamis_amd.scenes import batch_worlds
    return batch_worlds(BATCH_WORLDS, seed=BATCH_SEED)


WORKLOADS = {
    "pile": {"scene": pile_scene, "steps": 50, "warmup": 10, "run": 600,
             "describe": "mixed convex pile: 100,000 bodies (25% each spheres r~U[0.25,0.75], boxes "
                         "h~U[0.2,0.8]^3, capsules r~U[0.15,0.4] hh~U[0.2,0.8], hulls of 16-64 points "
                         "on ellipsoids with semi-axes ~U[0.3,0.9] from 512 shape sources) on a "
                         "jittered lattice in a 60x60 m walled bin on a plane, density 1000, "
                         "dt=1/60, 8 solver iterations, 2 substeps, 10 position iterations"},
    "cubes": {"scene": cubes_scene, "steps": 50, "warmup": 10, "run": 300,
              "describe": "1,000 cubes (h=0.5 m, density 1000, friction 0.5) dropped from 1-40 m "
                          "over 20x20 m onto a plane, dt=1/60, 4 solver iterations"},
    "batch": {"scene": None, "steps": 40, "warmup": 10, "run": 240,
              "describe": "4,096 independent worlds in one batch, each a plane and N_w bodies with "
                          "N_w log-uniform in [8, 2048] (1,550,732 bodies), 50/50 boxes "
                          "h~U[0.2,0.8]^3 and spheres r~U[0.25,0.75], density 1000, dt=1/60, "
                          "8 solver iterations"},
}


def world_counts(world) -> dict:
    m = world.measured()
    contacts = world.inspect_contacts()
    points = float(np.mean(contacts["point_count"])) if len(contacts) else 0.0
    work = world.convex_work()
    return {"n": len(world.scene.states), "k": len(world.scene.colliders), "entries": m["ENTRIES"],
            "pairs": m["PAIRS"], "contacts": m["CONTACTS"], "points": points, "rows": m["SOLVER_ROWS"],
            "gjk": work["gjk"], "epa": work["epa"], "manifold": work["manifold"], "gjk_hull": work["gjk_hull"],
            "gjk_resumed": work["gjk_resumed"], "epa_wave": work["epa_wave"]}


def kernel_bytes(name: str, c: dict) -> float:
    return step_bytes(name, c["n"], c["k"], c["entries"], c["pairs"], c["contacts"], c["points"], c["rows"],
                      c.get("gjk", 0), c.get("epa", 0), c.get("manifold", 0), c.get("gjk_hull", 0),
                      c.get("gjk_resumed", 0), c.get("epa_wave", 0))


class Roctx:
    """roctxProfilerResume / roctxProfilerPause marks around the eager window whose kernel
    launches the roofline figure times (no-ops when no profiler is attached; absent library:
    nothing is marked)."""

    def __init__(self):
        import ctypes
        self.lib = None
        for name in ("librocprofiler-sdk-roctx.so.1", "librocprofiler-sdk-roctx.so"):
            try:
                self.lib = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if self.lib is not None:
            for fn in (self.lib.roctxProfilerResume, self.lib.roctxProfilerPause):
                fn.argtypes = [ctypes.c_uint64]
                fn.restype = ctypes.c_int

    def resume(self) -> None:
        if self.lib is not None:
            self.lib.roctxProfilerResume(0)

    def pause(self) -> None:
        if self.lib is not None:
            self.lib.roctxProfilerPause(0)


ROCTX = None


def settled_run(world, run: int, steps: int, warmup: int):
    """The last `steps` steps of a `run`-step run, timed as graph replays.

    The first run - steps - warmup steps settle the scene untimed, `warmup` more steps are
    untimed, then the durable state is snapshotted and `steps` steps are timed by events on the
    world's stream (graph replays unless --no-graphs; no launch is bracketed). The same steps are
    then re-run twice from the snapshot, eagerly: once with every launch bracketed by events (the
    stage table; its kernel with the largest device time is the roofline kernel), once with only
    the roofline kernel's launches bracketed (its average launch duration). Each re-run ends in
    the timed run's states bit for bit, so all three windows do the same work."""
    import ctypes

    from dynamis_amd._native import call
    from dynamis_amd.device import profile_collect, profile_enable, profile_watch

    global ROCTX
    if ROCTX is None:
        ROCTX = Roctx()
    stream = ctypes.c_void_p(world.stream())
    settle = max(run - steps - warmup, 0)
    for _ in range(settle):
        world.step(DT)
    world.wait()
    progress(f"{settle} settling steps done")
    first = ctypes.c_void_p()
    call("dyn_event_create", ctypes.byref(first))
    call("dyn_event_record", first, stream)
    for _ in range(warmup):
        world.step(DT)
    world.wait()
    solver = None
    if TUNE:  # the solver's workgroup size, on the settled scene (its state restored afterwards)
        from dynamis_amd import tune
        solver = tune.autotune_solver(world)
        progress(f"solver workgroup size {solver['threads']}")
    snap = world.snapshot()
    before = world_counts(world)
    ms = timed_steps([world], stream, steps)
    # the reference's benchmark.rs figure: its wall time includes the warm-up steps
    last = ctypes.c_void_p()
    call("dyn_event_create", ctypes.byref(last))
    call("dyn_event_record", last, stream)
    all_ms = ctypes.c_float()
    call("dyn_event_elapsed_ms", ctypes.byref(all_ms), first, last)
    call("dyn_event_destroy", first)
    call("dyn_event_destroy", last)
    after = world_counts(world)
    measured = world.measured()
    final_states = world.states()
    progress(f"{steps} timed steps: {ms / steps:.3f} ms per step")
    # every launch of the same steps timed, from the snapshot
    world.restore(snap)
    profile_watch(None)
    profile_enable(True)
    for _ in range(steps):
        world.step(DT)
    world.wait()
    profile_enable(False)
    stages = profile_collect()
    same = world.states().tobytes() == final_states.tobytes()
    mean = {k: 0.5 * (before[k] + after[k]) for k in before}
    dominant = max((n for n in stages if kernel_bytes(n, mean) > 0), key=lambda n: stages[n][1], default=None)
    timed = None
    if dominant:
        world.restore(snap)
        profile_watch(dominant)
        profile_enable(True)
        ROCTX.resume()
        for _ in range(steps):
            world.step(DT)
        world.wait()
        ROCTX.pause()
        profile_enable(False)
        profile_watch(None)
        watched = profile_collect()
        same = same and world.states().tobytes() == final_states.tobytes()
        if dominant in watched:
            launches, total_ms = watched[dominant]
            # bytes per launch: the mean of the counts at the window's start and end
            timed = {"kernel": dominant, "launches": launches, "avg_us": 1000.0 * total_ms / launches,
                     "bytes_per_launch": kernel_bytes(dominant, mean),
                     "window": f"the {steps} timed steps re-run eagerly from a snapshot taken at "
                               f"their start, only this kernel's launches bracketed by events"}
    snap.free()
    progress(f"eager re-runs done (roofline kernel {dominant})")
    return {"ms": ms, "stages": stages, "counts": mean, "measured": measured, "timed": timed,
            "final_states": final_states, "with_warmup_ms": float(all_ms.value), "settle": settle,
            "reruns_identical": same, "solver": solver}


def sleep_off(scene):
    """The same scene with sleeping switched off (both velocity thresholds zero)."""
    import copy
    s = copy.copy(scene)
    s.params = scene.params.copy()
    s.descs = scene.descs.copy()
    for field in ("sleep_velocity", "sleep_angular_velocity"):
        s.params[field] = 0.0
        s.descs[field] = 0.0
    return s


def run_sleep_off(scene, run: int, steps: int, warmup: int) -> float:
    """Milliseconds of the same settled window (the last `steps` of `run`) with sleeping off."""
    import ctypes

    from dynamis_amd.world import World
    world = World()
    world.create(sleep_off(scene))
    world.set_graphs(GRAPHS)
    for _ in range(max(run - steps, 0)):
        world.step(DT)
    world.wait()
    ms = timed_steps([world], ctypes.c_void_p(world.stream()), steps)
    world.close()
    return ms


TUNE = True
TUNING = None


def tune_world(world) -> None:
    """The sorts' tilings from the autotuner (timed on this GPU, or loaded from its cache),
    before any timed step."""
    global TUNING
    if not TUNE:
        return
    from dynamis_amd import tune
    info = tune.autotune_world(world)
    TUNING = {"sort_tilings": {str(c): {"keys_per_lane": t[0], "groups": t[1]} for c, t in info["tilings"].items()},
              "from_cache": info["from_cache"]}


def run_world(workload: str, steps: int, warmup: int):
    import dynamis_amd
    from dynamis_amd._native import call
    from dynamis_amd.world import World

    if dynamis_amd.device_count() < 1:
        raise SystemExit("bench.py needs a HIP device")
    call("dyn_set_device", 0)
    scene = WORKLOADS[workload]["scene"]()
    world = World()
    world.create(scene)
    world.set_graphs(GRAPHS)
    tune_world(world)
    r = settled_run(world, WORKLOADS[workload]["run"], steps, warmup)
    world.close()
    r["scene"] = scene
    r["bodies"] = len(scene.states)
    return r


def timed_steps(worlds, stream, steps: int) -> float:
    """Milliseconds of `steps` rounds of one step of every world, by events on `stream`."""
    import ctypes

    from dynamis_amd._native import call
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    call("dyn_event_create", ctypes.byref(e0))
    call("dyn_event_create", ctypes.byref(e1))
    call("dyn_event_record", e0, stream)
    for _ in range(steps):
        for w in worlds:
            w.step(DT)
    call("dyn_event_record", e1, stream)
    call("dyn_stream_synchronize", stream)
    ms = ctypes.c_float()
    call("dyn_event_elapsed_ms", ctypes.byref(ms), e0, e1)
    call("dyn_event_destroy", e0)
    call("dyn_event_destroy", e1)
    return ms.value


def run_batch(steps: int, warmup: int, serial_every: int):
    """The batch's settled window (the last `steps` of its run) and the same window of every
    `serial_every`-th world stepped one at a time on one stream."""
    import ctypes

    import dynamis_amd
    from dynamis_amd._native import call
    from dynamis_amd.world import Batch, World

    if dynamis_amd.device_count() < 1:
        raise SystemExit("bench.py needs a HIP device")
    call("dyn_set_device", 0)
    run = WORKLOADS["batch"]["run"]
    scenes = batch_scenes()
    batch = Batch(scenes)
    batch.set_graphs(GRAPHS)
    tune_world(batch)
    r = settled_run(batch, run, steps, warmup)
    r["bodies"] = len(batch.scene.states)
    r["ends"] = [batch.world_states(i, r["final_states"]) for i in range(len(scenes))]
    batch.close()
    serial = None
    if serial_every > 0:
        shared = ctypes.c_void_p()
        call("dyn_stream_create", ctypes.byref(shared))
        sampled = scenes[::serial_every]
        worlds = []
        for s in sampled:
            w = World()
            w.create(s, stream=shared.value)
            w.set_graphs(GRAPHS)
            worlds.append(w)
        progress(f"serial: {len(worlds)} worlds created")
        timed_steps(worlds, shared, max(run - steps, 0))  # the batch's untimed steps
        sms = timed_steps(worlds, shared, steps)
        same = all(w.states().tobytes() == end.tobytes()
                   for w, end in zip(worlds, r["ends"][::serial_every]))
        for w in worlds:
            w.close()
        call("dyn_stream_destroy", shared)
        batched = r["ms"] / steps
        per_step = sms / steps * (len(scenes) / len(sampled))
        serial = {"serial_ms_per_step": per_step, "batched_ms_per_step": batched,
                  "batched_vs_serial": per_step / batched, "sampled_worlds_match_batch": same,
                  "sample": f"{len(sampled)} of the {len(scenes)} worlds (every {serial_every}th), each "
                            f"its own world, stepped one at a time on one stream through the same "
                            f"{run}-step run; the same {steps} steps timed, scaled by "
                            f"{len(scenes) / len(sampled):g}"}
    r["serial"] = serial
    r["scenes"] = scenes
    return r


def _cpu_worlds(args):
    """Worker: `sample_steps` oracle steps of each (scene, start states) pair; body-steps and
    seconds of stepping."""
    import copy

    from tests.oracle_binding import load
    pairs, sample_steps = args
    oracle = load()
    bodies, secs = 0, 0.0
    for s, end in pairs:
        start = copy.copy(s)
        start.states = end.copy()
        world = oracle.world(start)
        t0 = time.perf_counter()
        for _ in range(sample_steps):
            world.step(DT)
        secs += time.perf_counter() - t0
        bodies += len(s.states) * sample_steps
        world.close()
    return bodies, secs


def cpu_batch(scenes, ends, every: int, sample_steps: int, workers: int):
    """The oracle on every `every`-th world from the device's end states: the worlds are
    independent, so `workers` processes step disjoint sets of them. Returns body-steps/s over
    the wall time of the pool, the wall time, and the single-process rate (the sum of the
    workers' stepping seconds)."""
    import concurrent.futures
    pairs = list(zip(scenes, ends))[::every]
    chunks = [(pairs[i::workers], sample_steps) for i in range(workers)]
    t0 = time.perf_counter()
    if workers > 1:
        import multiprocessing
        with concurrent.futures.ProcessPoolExecutor(workers, mp_context=multiprocessing.get_context("fork")) as pool:
            results = list(pool.map(_cpu_worlds, chunks))
    else:
        results = [_cpu_worlds(chunks[0])]
    wall = time.perf_counter() - t0
    bodies = sum(r[0] for r in results)
    busy = sum(r[1] for r in results)
    return bodies / wall, wall, bodies / busy


def run_broadphase(steps: int, warmup: int):
    import dynamis_amd
    from dynamis_amd import abi
    from dynamis_amd.broadphase import Broadphase
    from dynamis_amd.device import Event, Stream, profile_collect, profile_enable
    from dynamis_amd.scenes import uniform_boxes

    if dynamis_amd.device_count() < 1:
        raise SystemExit("bench.py needs a HIP device")
    dynamis_amd._native.call("dyn_set_device", 0)
    boxes = uniform_boxes(BOXES, 1000.0, 0.2, 2.0, BOX_SEED)
    stream = Stream()
    bp = Broadphase(BOXES, stream=stream)
    bp.upload(boxes)
    for _ in range(warmup):
        bp.run()
    stream.synchronize()
    res = bp.read()
    entries = abi.counter(res.counters, "ENTRIES")
    pairs = abi.counter(res.counters, "PAIRS")
    start, stop = Event(), Event()
    start.record(stream)
    for _ in range(steps):
        bp.run()
    stop.record(stream)
    ms = start.elapsed_ms(stop)
    stream.synchronize()
    profile_enable(True)
    for _ in range(min(steps, 10)):
        bp.run()
    stream.synchronize()
    profile_enable(False)
    stages = profile_collect()
    counts = {"n": BOXES, "k": BOXES, "entries": entries, "pairs": pairs, "contacts": 0,
              "points": 0.0, "rows": 0}
    return ms, BOXES, stages, counts


def fixed_grid_bytes(name: str, k: int, entries: int, pairs: int) -> float:
    """Minimal HBM payload per launch of the fixed-cell pass: AABB 24 B, entry 12 B (hash, box,
    ordinal), pair 8 B; a sort pass moves its three lanes in and out (24 B per record)."""
    words = max(1, ((k - 1).bit_length() + 7) // 8)
    if name == "k_fixed_count":
        return k * 28.0
    if name == "k_fixed_emit":
        return k * 28.0 + entries * 12.0
    if name == "k_fixed_pairs":
        return entries * (12.0 + 24.0) + pairs * 8.0
    if name == "sort_pass_kernel":  # the entry sort's 4 + words passes and the pair sort's 2 words
        return (entries * 24.0 * (4 + words) + pairs * 24.0 * (2 * words)) / (4 + 3 * words)
    if name == "sort_histogram_kernel":
        return (entries * 8.0 + pairs * 8.0) / 2
    return 0.0


def run_fixed_grid(steps: int, warmup: int):
    import dynamis_amd
    from dynamis_amd.broadphase import FixedGrid
    from dynamis_amd.device import Event, Stream, profile_collect, profile_enable
    from dynamis_amd.scenes import uniform_boxes

    if dynamis_amd.device_count() < 1:
        raise SystemExit("bench.py needs a HIP device")
    dynamis_amd._native.call("dyn_set_device", 0)
    boxes = uniform_boxes(BOXES, 1000.0, 0.2, 2.0, BOX_SEED)
    stream = Stream()
    grid = FixedGrid(BOXES, 4.0, stream=stream)
    grid.upload(boxes)
    for _ in range(warmup):
        grid.run()
    res = grid.read()
    start, stop = Event(), Event()
    start.record(stream)
    for _ in range(steps):
        grid.run()
    stop.record(stream)
    ms = start.elapsed_ms(stop)
    stream.synchronize()
    profile_enable(True)
    for _ in range(min(steps, 10)):
        grid.run()
    stream.synchronize()
    profile_enable(False)
    stages = profile_collect()
    counts = {"k": BOXES, "entries": len(res.entry_hash), "pairs": len(res.pairs),
              "refused": res.refused_boxes + res.refused_pairs}
    return ms, stages, counts


def pile_stage(stage_file: str):
    """The settled pile's stage inputs: the body states step 600 began from, that step's
    candidate pairs and the contacts it solved. Run once and saved in `stage_file` (numpy .npz);
    later runs load them, so a counter pass (rocprofv3 --pmc) over a stage sees no world step."""
    import dynamis_amd
    from dynamis_amd.world import World

    if dynamis_amd.device_count() < 1:
        raise SystemExit("bench.py needs a HIP device")
    dynamis_amd._native.call("dyn_set_device", 0)
    scene = pile_scene()
    if os.path.exists(stage_file):
        data = np.load(stage_file)
        if "contacts" in data.files:
            return scene, data
    world = World()
    world.create(scene)
    world.set_graphs(GRAPHS)
    run = WORKLOADS["pile"]["run"]
    for _ in range(run - 1):
        world.step(DT)
    states = world.states()
    world.step(DT)
    pairs = world.pairs()
    contacts = world.inspect_contacts()
    world.close()
    np.savez(stage_file, states=states, pairs=pairs, contacts=contacts)
    progress(f"saved the pile's step-{run} stage inputs ({len(pairs)} pairs, {len(contacts)} contacts) "
             f"to {stage_file}")
    return scene, np.load(stage_file)


def run_solve(steps: int, warmup: int, stage_file: str) -> dict:
    """The solver alone (dyn_solve: SolverReset and SolveSubsteps, the kernels dyn_world_step
    runs) over the settled pile's step-600 contacts from the states that step began from,
    uploaded once and solved again and again on the device."""
    from dynamis_amd.device import Event, Stream, profile_collect, profile_enable
    from dynamis_amd.solver import Solver

    scene, data = pile_stage(stage_file)
    states, contacts = data["states"], data["contacts"]
    stream = Stream()
    solver = Solver(scene, len(contacts), stream=stream)
    solver.run(states, contacts, DT)
    for _ in range(warmup):
        solver.again(DT)
    stream.synchronize()
    start, stop = Event(), Event()
    start.record(stream)
    for _ in range(steps):
        solver.again(DT)
    stop.record(stream)
    ms = start.elapsed_ms(stop)
    stream.synchronize()
    profile_enable(True)
    for _ in range(min(steps, 10)):
        solver.again(DT)
    stream.synchronize()
    profile_enable(False)
    return {"ms": ms, "contacts": int(len(contacts)), "stages": profile_collect(), "bodies": int(len(scene.states))}


def run_narrowphase(steps: int, warmup: int, stage_file: str) -> dict:
    """The narrowphase pass alone (dyn_narrowphase: the kernels dyn_world_step runs) over the
    settled pile's step-600 candidate pairs from the states that step began from (pile_stage)."""
    from dynamis_amd.device import Event, Stream, profile_collect, profile_enable
    from dynamis_amd.narrowphase import Narrowphase

    scene, data = pile_stage(stage_file)
    states, pairs = data["states"], data["pairs"]
    stream = Stream()
    stage = Narrowphase(scene, len(pairs), stream=stream)
    stage.set_states(states)
    for _ in range(warmup):
        stage.run(pairs)
    stream.synchronize()
    start, stop = Event(), Event()
    start.record(stream)
    for _ in range(steps):
        stage.run(pairs)
    stop.record(stream)
    ms = start.elapsed_ms(stop)
    stream.synchronize()
    profile_enable(True)
    for _ in range(min(steps, 10)):
        stage.run(pairs)
    stream.synchronize()
    profile_enable(False)
    stages = profile_collect()
    _, valid = stage.read()
    return {"ms": ms, "pairs": int(len(pairs)), "contacts": int((valid != 0).sum()), "stages": stages,
            "bodies": int(len(scene.states))}


def run_copy(count: int) -> int:
    """Copies of a known size for the HBM counters' calibration; prints one JSON line."""
    import ctypes

    import dynamis_amd
    from dynamis_amd._native import call
    if dynamis_amd.device_count() < 1:
        raise SystemExit("bench.py needs a HIP device")
    call("dyn_set_device", 0)
    nbytes = 1 << 30
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    call("dyn_malloc", ctypes.byref(src), ctypes.c_size_t(nbytes))
    call("dyn_malloc", ctypes.byref(dst), ctypes.c_size_t(nbytes))
    call("dyn_memset", src, 1, ctypes.c_size_t(nbytes))
    t0 = time.perf_counter()
    for _ in range(count):
        call("dyn_memcpy_dtod", dst, src, ctypes.c_size_t(nbytes))
    secs = time.perf_counter() - t0
    call("dyn_free", src)
    call("dyn_free", dst)
    print(json.dumps({"workload": "copy", "bytes_per_copy": nbytes, "copies": count,
                      "gbs": 2.0 * nbytes * count / secs / 1e9}))
    return 0


# ---- CPU baseline ------------------------------------------------------------------------------

def cpu_world(scene, states, sample_steps: int, parallel: bool = False):
    """The oracle's rate over `sample_steps` steps from the given states: the serial build, or
    with `parallel` its OpenMP build (the same results) on every thread OpenMP is given."""
    import copy

    from tests.oracle_binding import load
    oracle = load(parallel)
    start = copy.copy(scene)
    start.states = states.copy()
    world = oracle.world(start)
    t0 = time.perf_counter()
    for _ in range(sample_steps):
        world.step(DT)
    secs = time.perf_counter() - t0
    world.close()
    return len(scene.states) * sample_steps / secs, secs, oracle.threads()


def cpu_broadphase(passes: int):
    from dynamis_amd.scenes import uniform_boxes
    from tests.oracle_binding import load
    oracle = load()
    boxes = uniform_boxes(BOXES, 1000.0, 0.2, 2.0, BOX_SEED)
    t0 = time.perf_counter()
    for _ in range(passes):
        oracle.broadphase(boxes)
    secs = time.perf_counter() - t0
    return BOXES * passes / secs, secs


# ---- report ------------------------------------------------------------------------------------------

def step_roofline(table: dict, steps: int, ms_per_step: float) -> dict:
    """All kernels' algorithmic bytes of one step (the stage window's launches over its `steps`
    steps) over the timed step time: the whole step against the HBM roof."""
    per_step = sum(v["bytes"] * v["launches"] for v in table.values()) / max(steps, 1)
    gbs = per_step / (ms_per_step * 1e-3) / 1e9
    return {"bytes_per_step": per_step, "achieved": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
            "unit": "GB/s"}


def stage_table(stages, counts, bytes_of, timed=None):
    total = max(sum(t for _, t in stages.values()), 1e-9)
    table = {}
    for name, (launches, total_ms) in stages.items():
        avg_us = 1000.0 * total_ms / max(launches, 1)
        nbytes = bytes_of(name)
        table[name] = {"launches": launches, "avg_us": round(avg_us, 3), "share": round(total_ms / total, 4),
                       "bytes": nbytes,
                       "gbs": round(nbytes / (avg_us * 1e-6) / 1e9, 1) if avg_us > 0 and nbytes else 0.0}
    dominant = max((n for n in table if table[n]["bytes"] > 0), key=lambda n: stages[n][1], default=None)
    roofline = None
    if timed:
        gbs = timed["bytes_per_launch"] / (timed["avg_us"] * 1e-6) / 1e9
        roofline = {"bound": "hbm", "kernel": timed["kernel"], "achieved": round(gbs, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                    "traffic": None, "avg_us": round(timed["avg_us"], 3),
                    "bytes_per_launch": timed["bytes_per_launch"], "launches_timed": timed["launches"],
                    "share_of_device_time": table.get(timed["kernel"], {}).get("share")}
    elif dominant:
        d = table[dominant]
        roofline = {"bound": "hbm", "kernel": dominant, "achieved": d["gbs"], "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(d["gbs"] / HBM_PEAK_GBS, 4), "traffic": None,
                    "avg_us": d["avg_us"], "bytes_per_launch": d["bytes"],
                    "share_of_device_time": d["share"]}
    return table, roofline


def run_adjoint(reps: int = 3) -> dict:
    """Forward (64 steps, graphs) and backward (segment recompute with recording + the per-world
    reverse) of the differentiation workload; wall-clock seconds around each call on an idle
    stream, the minimum over `reps` repetitions after one untimed round."""
    from dynamis_amd import adjoint as adj
    from dynamis_amd import scenes as scene_lib
    from dynamis_amd.world import Batch
    worlds, boxes, steps, every, dt = 256, 32, 64, 8, 1.0 / 120.0
    sc, cfg, targets = scene_lib.adjoint_worlds(worlds, boxes, seed=5)
    batch = Batch(sc, cfg)
    batch.set_graphs(GRAPHS)
    a = adj.Adjoint(batch, steps, every)
    L = batch.layout

    def grad(states):
        g = np.zeros((len(states), adj.STATE_FLOATS), np.float32)
        for w in range(worlds):
            rows = L.dynamic_base[w] + np.arange(L.dynamic_count[w])
            m = 1.0 / batch.scene.descs["inverse_mass"][rows].astype(np.float64)
            p = states["position"][rows].astype(np.float64)
            r = (m[:, None] * p).sum(0) / m.sum() - targets[w]
            g[rows, 0:3] = 2 * r[None, :] * (m / m.sum())[:, None]
        return g

    fwd, bwd = [], []
    for rep in range(reps + 1):
        batch.wait()
        t0 = time.perf_counter()
        a.forward(dt)
        batch.wait()
        t1 = time.perf_counter()
        g = grad(batch.states())
        batch.wait()
        t2 = time.perf_counter()
        a.backward(g)
        t3 = time.perf_counter()
        if rep:
            fwd.append(t1 - t0)
            bwd.append(t3 - t2)
    a.close()
    # the whole rollout as one segment: every step's tape stays in HBM (no recompute)
    a1 = adj.Adjoint(batch, steps, steps)
    f1, b1 = [], []
    for rep in range(reps + 1):
        batch.wait()
        t0 = time.perf_counter()
        a1.forward(dt)
        batch.wait()
        t1 = time.perf_counter()
        g = grad(batch.states())
        batch.wait()
        t2 = time.perf_counter()
        a1.backward(g)
        t3 = time.perf_counter()
        if rep:
            f1.append(t1 - t0)
            b1.append(t3 - t2)
    a1.close()
    one = {"forward_ms": min(f1) * 1e3, "backward_ms": min(b1) * 1e3, "bwd_fwd": min(b1) / min(f1),
           "note": "checkpoint_every = steps: the forward records every step's tape and the backward recomputes none"}
    f, b = min(fwd), min(bwd)
    return {"single_segment": one,
            "workload": f"{worlds} worlds x {boxes} boxes on a plane, dt = 1/120, 8 iterations, {steps} steps, "
                        f"checkpoint every {every}; loss = squared distance of each world's centre of mass "
                        f"to a target",
            "bodies": worlds * boxes, "forward_ms": f * 1e3, "backward_ms": b * 1e3, "bwd_fwd": b / f,
            "forward_body_steps_per_s": worlds * boxes * steps / f}


CPU_WORKERS = 16  # host cores a GPU box grants one GPU's job


TRAFFIC = os.path.join(ROOT, "profiles", "traffic.json")


def measured_traffic(roofline: dict, workload: str) -> None:
    """roofline.traffic from the counter passes recorded in profiles/traffic.json: HBM bytes per
    launch of the roofline kernel on this workload (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE,
    each scaled by what the counter reads for a copy of known size), when they were taken."""
    try:
        with open(TRAFFIC) as f:
            entry = json.load(f).get(workload, {}).get(roofline["kernel"])
    except (OSError, ValueError):
        return
    if entry:
        roofline["traffic"] = entry["bytes_per_launch"]
        roofline["traffic_vs_algorithmic"] = round(entry["bytes_per_launch"] / max(roofline["bytes_per_launch"], 1.0), 2)
        roofline["traffic_source"] = entry["source"]


def world_report(wl: str, r: dict, steps: int, warmup: int, cpu_sample: int | None) -> dict:
    """The line of a single-world workload's settled window (pile, cubes)."""
    spec = WORKLOADS[wl]
    counts = r["counts"]
    table, roofline = stage_table(r["stages"], counts, lambda nm: kernel_bytes(nm, counts), r["timed"])
    ms = r["ms"]
    if roofline is not None:
        roofline["step"] = step_roofline(table, steps, ms / steps)
        measured_traffic(roofline, wl)
    bodies = r["bodies"]
    sleep = None
    if wl == "pile":
        off_ms = run_sleep_off(r["scene"], spec["run"], steps, warmup)
        sleep = {"sleep_off_value": bodies * steps / (off_ms / 1000.0), "sleep_off_ms_per_step": off_ms / steps,
                 "note": "the same window with sleep_velocity = sleep_angular_velocity = 0; the main value "
                         "keeps the reference's default sleeping"}
    sample = (2 if wl == "pile" else 100) if cpu_sample is None else cpu_sample
    cpu = None
    if sample > 0:
        progress(f"CPU baseline: {sample} oracle steps on every core, then on one")
        v, secs, threads = cpu_world(r["scene"], r["final_states"], sample, parallel=True)
        one, one_secs, _ = cpu_world(r["scene"], r["final_states"], sample)
        cpu = {"value": v, "unit": "body-steps/s", "cores": threads, "kind": "port",
               "sample": f"{sample} steps of the same scene through the C oracle's OpenMP build ({threads} "
                         f"threads over the per-pair narrowphase and the per-contact and per-row solver "
                         f"loops; it steps bit-identically to the serial build), started from the device's "
                         f"body states at the end of the timed steps ({secs:.1f} s)",
               "single_core": {"value": one, "cores": 1,
                               "sample": f"the same {sample} steps through the serial oracle ({one_secs:.1f} s)"}}
    run = r["settle"] + warmup + steps
    config = {"workload": spec["describe"], "bodies": bodies, "steps_in_run": run,
              "timed_window": f"steps {run - steps + 1}-{run} of the {run}-step run: {r['settle']} settling "
                              f"steps and {warmup} warm-up steps untimed, then {steps} timed "
                              f"{'graph replays' if GRAPHS else 'eager steps'}",
              "graphs": GRAPHS, "seed": 3 if wl == "pile" else 1, "sleep": sleep,
              "reference_style_value": bodies * (steps + warmup) / ((r["with_warmup_ms"]) / 1000.0),
              "reference_style_note": "benchmark.rs divides by a wall time that includes its warm-up "
                                      "steps: here the warm-up and timed steps together",
              "reruns_identical": r["reruns_identical"],
              "solver_workgroup": r.get("solver"),
              "at_end": {k: r["measured"][k] for k in ("PAIRS", "CONTACTS", "ISLANDS", "SOLVER_ROWS",
                                                      "RESTING", "ACTIVE", "REFUSED_PAIRS",
                                                      "REFUSED_CONTACTS", "ENTRY_FAULTS")}}
    return {"value": bodies * steps / (ms / 1000.0), "ms_per_step": ms / steps, "config": config,
            "roofline": roofline, "cpu_baseline": cpu, "stages": table}


def batch_report(r: dict, steps: int, warmup: int, cpu_sample: int | None) -> dict:
    spec = WORKLOADS["batch"]
    counts = r["counts"]
    table, roofline = stage_table(r["stages"], counts, lambda nm: kernel_bytes(nm, counts), r["timed"])
    ms = r["ms"]
    if roofline is not None:
        roofline["step"] = step_roofline(table, steps, ms / steps)
    sample = 10 if cpu_sample is None else cpu_sample
    cpu = None
    if sample > 0:
        workers = max(1, min(CPU_WORKERS, os.cpu_count() or 1))
        progress(f"CPU baseline: oracle over sampled worlds in {workers} processes")
        v, wall, one = cpu_batch(r["scenes"], r["ends"], 16, sample, workers)
        cpu = {"value": v, "unit": "body-steps/s", "cores": workers, "kind": "port",
               "single_core_value": one,
               "sample": f"{sample} steps of every 16th world through the C oracle in {workers} worker "
                         f"processes (independent worlds), each world started from the device's body "
                         f"states at the end of the timed steps ({wall:.1f} s wall)"}
    run = r["settle"] + warmup + steps
    config = {"workload": spec["describe"], "bodies": r["bodies"], "worlds": BATCH_WORLDS,
              "steps_in_run": run, "seed": BATCH_SEED, "graphs": GRAPHS,
              "timed_window": f"steps {run - steps + 1}-{run} of the {run}-step run",
              "serial": r["serial"], "reruns_identical": r["reruns_identical"],
              "at_end": {k: r["measured"][k] for k in ("PAIRS", "CONTACTS", "ISLANDS", "SOLVER_ROWS",
                                                      "REFUSED_PAIRS", "REFUSED_CONTACTS", "ENTRY_FAULTS")}}
    return {"value": r["bodies"] * steps / (ms / 1000.0), "ms_per_step": ms / steps, "config": config,
            "roofline": roofline, "cpu_baseline": cpu, "stages": table}


def broadphase_report(steps: int, warmup: int, cpu_sample: int | None) -> dict:
    ms, bodies, stages, counts = run_broadphase(steps, warmup)
    table, roofline = stage_table(stages, counts, lambda nm: broadphase_bytes(
        nm, counts["k"], counts["entries"], counts["pairs"]))
    sample = 3 if cpu_sample is None else cpu_sample
    cpu = None
    if sample > 0:
        v, secs = cpu_broadphase(sample)
        cpu = {"value": v, "unit": "boxes/s", "cores": 1, "kind": "port",
               "sample": f"{sample} full passes of the same 2^20-box workload through the "
                         f"single-threaded C oracle ({secs:.1f} s)"}
    config = {"workload": "isolated sort + broadphase pass over 1,048,576 boxes (centres "
                          "U[0,1000)^3 m, half-extents log-U[0.2,2] m)",
              "boxes": BOXES, "entries": counts["entries"], "pairs": counts["pairs"], "seed": BOX_SEED}
    return {"value": bodies * steps / (ms / 1000.0), "unit": "boxes/s", "ms_per_pass": ms / steps,
            "passes": steps, "warmup": warmup, "config": config, "roofline": roofline, "cpu_baseline": cpu,
            "stages": table}


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--workload", choices=["pile", "cubes", "batch", "broadphase", "fixed_grid", "copy",
                                               "adjoint", "narrowphase", "solve"], default="pile")
    parser.add_argument("--steps", type=int, default=None)
    parser.add_argument("--warmup", type=int, default=None)
    parser.add_argument("--no-graphs", action="store_true",
                        help="launch every step eagerly (for external profilers, which trace the "
                             "replayed graph of a captured step badly)")
    parser.add_argument("--no-tune", action="store_true",
                        help="keep the sorts' default tilings instead of autotuning them first")
    parser.add_argument("--cpu-sample", type=int, default=None,
                        help="oracle steps (broadphase: passes) timed for the CPU baseline; 0 skips it")
    parser.add_argument("--only", action="store_true",
                        help="pile: leave out the batch, broadphase and adjoint parts of the line")
    parser.add_argument("--stage-file", default="/tmp/dynamis_pile_stage.npz",
                        help="narrowphase, solve: where the settled pile's stage inputs are saved or found")
    args = parser.parse_args()
    global GRAPHS, TUNE
    GRAPHS = not args.no_graphs
    TUNE = not args.no_tune
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(independent worlds would shard by world range)", file=sys.stderr)
        return 2
    wl = args.workload
    if wl == "copy":
        return run_copy(max(args.steps if args.steps is not None else 5, 1))
    if wl == "adjoint":
        r = run_adjoint()
        print(json.dumps({"metric": "adjoint backward / forward time", "value": r["bwd_fwd"], "unit": "ratio",
                          "n_gpus": 1, "steps": 64, "warmup": 0, "ms_per_step": r["forward_ms"] / 64,
                          "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                          "data": "synthetic", "config": r}))
        return 0
    if wl == "narrowphase":
        steps = max(args.steps if args.steps is not None else 20, 1)
        warmup = max(args.warmup if args.warmup is not None else 3, 0)
        r = run_narrowphase(steps, warmup, args.stage_file)
        print(json.dumps({"metric": "narrowphase pass over the settled pile's candidate pairs", "unit": "pairs/s",
                          "value": r["pairs"] * steps / (r["ms"] / 1000.0), "n_gpus": 1, "steps": steps,
                          "warmup": warmup, "ms_per_step": r["ms"] / steps, "higher_is_better": True,
                          "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                          "config": {"workload": "dyn_narrowphase over the candidate pairs of step 600 of the "
                                                 "100,000-body mixed convex pile", "pairs": r["pairs"],
                                     "contacts": r["contacts"], "bodies": r["bodies"]},
                          "stages": {k: {"launches": v[0], "ms": v[1]} for k, v in r["stages"].items()}}))
        return 0
    if wl == "solve":
        steps = max(args.steps if args.steps is not None else 20, 1)
        warmup = max(args.warmup if args.warmup is not None else 3, 0)
        r = run_solve(steps, warmup, args.stage_file)
        print(json.dumps({"metric": "solver over the settled pile's contacts", "unit": "contact-solves/s",
                          "value": r["contacts"] * steps / (r["ms"] / 1000.0), "n_gpus": 1, "steps": steps,
                          "warmup": warmup, "ms_per_step": r["ms"] / steps, "higher_is_better": True,
                          "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                          "config": {"workload": "dyn_solve (2 substeps, 8 iterations, 10 position iterations) "
                                                 "over the contacts of step 600 of the 100,000-body mixed "
                                                 "convex pile", "contacts": r["contacts"], "bodies": r["bodies"]},
                          "stages": {k: {"launches": v[0], "ms": v[1]} for k, v in r["stages"].items()}}))
        return 0
    if wl == "fixed_grid":
        steps = max(args.steps if args.steps is not None else 50, 1)
        warmup = max(args.warmup if args.warmup is not None else 10, 0)
        ms, stages, counts = run_fixed_grid(steps, warmup)
        table, roofline = stage_table(stages, counts, lambda nm: fixed_grid_bytes(
            nm, counts["k"], counts["entries"], counts["pairs"]))
        config = {"workload": "fixed-cell sort + broadphase pass over 1,048,576 boxes (centres U[0,1000)^3 m, "
                              "half-extents log-U[0.2,2] m), 4 m cells, keys (cell hash, box index)",
                  "boxes": BOXES, "entries": counts["entries"], "pairs": counts["pairs"], "refused": counts["refused"],
                  "seed": BOX_SEED, "cell_size": 4.0}
        part = {"value": BOXES * steps / (ms / 1000.0), "ms_per_step": ms / steps, "config": config,
                "roofline": roofline, "cpu_baseline": None, "stages": table}
        dtype, unit = "u32", "boxes/s"
    elif wl == "broadphase":
        steps = max(args.steps if args.steps is not None else 50, 1)
        warmup = max(args.warmup if args.warmup is not None else 10, 0)
        part = broadphase_report(steps, warmup, args.cpu_sample)
        part["ms_per_step"] = part["ms_per_pass"]
        dtype, unit = "u32", "boxes/s"
    elif wl == "batch":
        spec = WORKLOADS[wl]
        steps = max(args.steps if args.steps is not None else spec["steps"], 1)
        warmup = max(args.warmup if args.warmup is not None else spec["warmup"], 0)
        part = batch_report(run_batch(steps, warmup, 32), steps, warmup, args.cpu_sample)
        dtype, unit = "f32", "body-steps/s"
    else:
        spec = WORKLOADS[wl]
        steps = max(args.steps if args.steps is not None else spec["steps"], 1)
        warmup = max(args.warmup if args.warmup is not None else spec["warmup"], 0)
        part = world_report(wl, run_world(wl, steps, warmup), steps, warmup, args.cpu_sample)
        dtype, unit = "f32", "body-steps/s"
    line = {
        "metric": METRIC,
        "value": part["value"],
        "unit": unit,
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": part["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": part["config"],
        "roofline": part["roofline"],
        "cpu_baseline": part["cpu_baseline"],
    }
    if TUNING is not None:
        line["config"]["autotune"] = TUNING
    if wl == "pile" and not args.only:
        # the rest of the metric: the batch's rate and its ratio to serial stepping over the
        # same window, the isolated sort + broadphase pass, the adjoint's backward/forward ratio
        progress("batch")
        b = batch_report(run_batch(steps, warmup, 32), steps, warmup, args.cpu_sample)
        line["batch"] = {"value": b["value"], "unit": "body-steps/s", "ms_per_step": b["ms_per_step"],
                         "batched_vs_serial": (b["config"]["serial"] or {}).get("batched_vs_serial"),
                         "config": b["config"], "roofline": b["roofline"], "cpu_baseline": b["cpu_baseline"]}
        progress("broadphase")
        bp = broadphase_report(max(steps, 20), max(warmup, 5), args.cpu_sample)
        line["broadphase"] = {k: bp[k] for k in ("value", "unit", "ms_per_pass", "passes", "warmup", "config",
                                                 "roofline", "cpu_baseline")}
        progress("adjoint")
        line["adjoint"] = run_adjoint()
    line["stages"] = part["stages"]
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
