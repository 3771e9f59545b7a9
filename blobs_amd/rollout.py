"""Differentiable rollouts of batches of `Physics` worlds on the GPU.

`Rollout(worlds, targets, checkpoint_every)` copies single-sphere worlds with
springs into device buffers; `forward(delta, n_steps)` runs n_steps calls of
Physics::step(delta) (physics.rs:78-82) keeping a checkpoint every
`checkpoint_every` steps, and `backward()` returns the gradients of the
summed centroid loss |centroid(final positions) - target|^2 with respect to
the initial velocity requests and the spring stiffnesses, by the adjoint
kernels (contact pair sets held fixed, segments recomputed from their
checkpoints).

`RolloutFunction` wraps the same entry points with torch.autograd.Function
semantics: final positions = f(velocity requests, stiffnesses), so any
torch loss on the final positions back-propagates through the GPU adjoint.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Rollout:
    def __init__(self, worlds, targets, checkpoint_every: int = 20):
        self._lib = L.load()
        worlds = list(worlds)
        self._worlds = worlds  # keep the sources alive while the rollout exists
        arr = (C.c_void_p * len(worlds))(*[w._w.value for w in worlds])
        t = np.ascontiguousarray(targets, np.float32).reshape(len(worlds), 2)
        r = C.c_void_p()
        L.check(self._lib.bg_rollout_create(arr, len(worlds), _fp(t), int(checkpoint_every), C.byref(r)), None)
        self._r = r
        nw, nb, ns = C.c_uint32(), C.c_uint32(), C.c_uint32()
        self._chk(self._lib.bg_rollout_sizes(r, C.byref(nw), C.byref(nb), C.byref(ns)))
        self.n_worlds, self.n_bodies, self.n_springs = nw.value, nb.value, ns.value

    def __del__(self):
        r = getattr(self, "_r", None)
        if r:
            self._lib.bg_rollout_destroy(r)
            self._r = None

    def _chk(self, code):
        L.check(code, None)

    def set_params(self, velocity_request=None, stiffness=None) -> None:
        v = None if velocity_request is None else np.ascontiguousarray(velocity_request, np.float32)
        k = None if stiffness is None else np.ascontiguousarray(stiffness, np.float32)
        if v is not None and v.size != 2 * self.n_bodies:
            raise ValueError("velocity_request needs one Vec2 per body")
        if k is not None and k.size != self.n_springs:
            raise ValueError("stiffness needs one value per spring")
        self._chk(self._lib.bg_rollout_set_params(self._r, None if v is None else _fp(v),
                                                  None if k is None else _fp(k)))

    def forward(self, delta: float, n_steps: int) -> np.ndarray:
        """Runs the rollout; returns the loss of every world."""
        loss = np.zeros(self.n_worlds, np.float32)
        self._chk(self._lib.bg_rollout_forward(self._r, float(delta), int(n_steps), _fp(loss)))
        return loss

    def final_positions(self) -> np.ndarray:
        out = np.zeros((self.n_bodies, 2), np.float32)
        self._chk(self._lib.bg_rollout_final_positions(self._r, _fp(out)))
        return out

    def backward(self, dloss_dfinal=None):
        """(d loss / d velocity_request [n_bodies, 2], d loss / d stiffness [n_springs])."""
        gv = np.zeros((self.n_bodies, 2), np.float32)
        gk = np.zeros(self.n_springs, np.float32)
        dl = None if dloss_dfinal is None else np.ascontiguousarray(dloss_dfinal, np.float32)
        self._chk(self._lib.bg_rollout_backward(self._r, None if dl is None else _fp(dl), _fp(gv), _fp(gk)))
        return gv, gk


def rollout_function():
    """torch.autograd.Function: (velocity_request [n,2], stiffness [s]) ->
    final positions [n,2] (float32, CPU tensors; the GPU work happens
    behind the C ABI)."""
    import torch

    class RolloutFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, vreq, k, rollout: Rollout, delta: float, n_steps: int):
            rollout.set_params(vreq.detach().cpu().numpy(), k.detach().cpu().numpy())
            rollout.forward(delta, n_steps)
            ctx.rollout = rollout
            return torch.from_numpy(rollout.final_positions())

        @staticmethod
        def backward(ctx, grad_pos):
            gv, gk = ctx.rollout.backward(grad_pos.detach().cpu().contiguous().numpy())
            return torch.from_numpy(gv), torch.from_numpy(gk), None, None, None

    return RolloutFunction
