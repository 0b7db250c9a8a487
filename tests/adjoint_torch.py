"""The differentiable rollout as a torch.autograd.Function (test harness only).

Rollout.apply(v0, mass, friction, runner) steps a batch of worlds of the differentiation workload
with the given initial velocities (worlds x boxes x 3), box masses (worlds x boxes, inertia scaled
with them at fixed shape) and collider frictions (worlds x colliders, the plane first) and returns
the boxes' final positions (worlds x boxes x 3). Its backward is dyn_adjoint_backward. A loss built
in torch from the positions and the mass tensor (the centre of mass) gets the masses' direct term
from autograd and their dynamic term from the adjoint.
"""

from __future__ import annotations

import numpy as np
import torch

from dynamis_amd import adjoint as adj
from dynamis_amd import scenes as scene_lib
from dynamis_amd.world import Batch


class Runner:
    def __init__(self, worlds: int = 8, boxes: int = 32, steps: int = 64, every: int = 8, seed: int = 5,
                 dt: float = 1.0 / 120.0):
        self.scenes, self.config, self.targets = scene_lib.adjoint_worlds(worlds, boxes, seed=seed)
        self.worlds, self.boxes, self.steps, self.every, self.dt = worlds, boxes, steps, every, dt
        self.batch = None
        self.adjoint = None

    def initial(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        v0 = np.stack([s.states["velocity"][:self.boxes] for s in self.scenes]).astype(np.float32)
        mass = np.stack([(1.0 / s.descs["inverse_mass"][:self.boxes]).astype(np.float32) for s in self.scenes])
        friction = np.stack([s.colliders["friction"] for s in self.scenes]).astype(np.float32)
        return v0, mass, friction

    def _rows(self, w: int) -> np.ndarray:
        return self.batch.layout.dynamic_base[w] + np.arange(self.boxes)

    def forward(self, v0: np.ndarray, mass: np.ndarray, friction: np.ndarray) -> np.ndarray:
        worlds = []
        for w, s in enumerate(self.scenes):
            s = type(s)(**{f: (getattr(s, f).copy() if isinstance(getattr(s, f), np.ndarray) else getattr(s, f))
                           for f in s.__dataclass_fields__})
            rows = np.arange(self.boxes)  # a world's boxes are its simulated rows, in spawn order
            s.states["velocity"][rows] = v0[w]
            # rows whose mass moved get mass and inertia rescaled; the others keep their records' bits
            old = (1.0 / s.descs["inverse_mass"][rows]).astype(np.float32)
            moved = rows[mass[w] != old]
            if len(moved):
                scale = mass[w][moved].astype(np.float64) / old[moved].astype(np.float64)
                s.descs["inverse_mass"][moved] = (1.0 / mass[w][moved].astype(np.float64)).astype(np.float32)
                s.descs["inverse_inertia"][moved] = (s.descs["inverse_inertia"][moved] / scale[:, None]).astype(np.float32)
                s.descs["inertia"][moved] = (s.descs["inertia"][moved] * scale[:, None]).astype(np.float32)
            s.colliders["friction"] = friction[w]
            worlds.append(s)
        if self.adjoint is not None:
            self.adjoint.close()
        self.batch = Batch(worlds, self.config)
        self.adjoint = adj.Adjoint(self.batch, self.steps, self.every)
        self.adjoint.forward(self.dt)
        states = self.batch.states()
        return np.stack([states["position"][self._rows(w)] for w in range(self.worlds)]).astype(np.float32)

    def backward(self, grad_positions: np.ndarray) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        g = np.zeros((len(self.batch.scene.states), adj.STATE_FLOATS), np.float32)
        for w in range(self.worlds):
            g[self._rows(w), 0:3] = grad_positions[w]
        g0, gm, gf = self.adjoint.backward(g)
        L = self.batch.layout
        gv = np.stack([g0[self._rows(w), 8:11] for w in range(self.worlds)])
        gmass = np.stack([gm[self._rows(w)] for w in range(self.worlds)])
        gfric = np.stack([gf[L.collider_base[w]:L.collider_base[w] + L.collider_count[w]] for w in range(self.worlds)])
        return gv, gmass, gfric


class Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, mass, friction, runner):
        ctx.runner = runner
        positions = runner.forward(v0.detach().cpu().numpy(), mass.detach().cpu().numpy(),
                                   friction.detach().cpu().numpy())
        return torch.from_numpy(positions).to(v0.device)

    @staticmethod
    def backward(ctx, grad_positions):
        gv, gm, gf = ctx.runner.backward(grad_positions.detach().cpu().numpy())
        dev = grad_positions.device
        return torch.from_numpy(gv).to(dev), torch.from_numpy(gm).to(dev), torch.from_numpy(gf).to(dev), None


def com_loss(positions: torch.Tensor, mass: torch.Tensor, targets: torch.Tensor) -> torch.Tensor:
    """Sum over worlds of the squared distance of the final centre of mass to the world's target."""
    com = (mass[..., None] * positions).sum(1) / mass.sum(1, keepdim=True)
    return ((com - targets) ** 2).sum()
