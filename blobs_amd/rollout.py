"""Differentiable rollouts of batches of `Physics` worlds on the GPU.

`Rollout(worlds, targets, checkpoint_every)` copies single-sphere worlds with
springs (damped or not) and fixed joints into device buffers; `forward(delta, n_steps)` runs n_steps calls of
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
        self.generation = 0  # forwards run so far (a backward belongs to the latest)
        self.nonfinite_worlds = np.zeros(0, np.int64)

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
        self.generation += 1
        return loss

    def final_positions(self) -> np.ndarray:
        out = np.zeros((self.n_bodies, 2), np.float32)
        self._chk(self._lib.bg_rollout_final_positions(self._r, _fp(out)))
        return out

    def backward(self, dloss_dfinal=None):
        """(d loss / d velocity_request [n_bodies, 2], d loss / d stiffness [n_springs]).
        Worlds whose gradients hold inf or NaN (the fixed-pair derivative of
        the contact projection can outgrow f32 over long horizons) are listed
        in `nonfinite_worlds` after the call."""
        gv = np.zeros((self.n_bodies, 2), np.float32)
        gk = np.zeros(self.n_springs, np.float32)
        dl = None if dloss_dfinal is None else np.ascontiguousarray(dloss_dfinal, np.float32)
        self._chk(self._lib.bg_rollout_backward(self._r, None if dl is None else _fp(dl), _fp(gv), _fp(gk)))
        mask = np.zeros(self.n_worlds, np.uint8)
        self._chk(self._lib.bg_rollout_nonfinite(self._r, mask.ctypes.data_as(C.POINTER(C.c_uint8))))
        self.nonfinite_worlds = np.flatnonzero(mask)
        return gv, gk


def rollout_function():
    """torch.autograd.Function: (velocity_request [n,2], stiffness [s]) ->
    final positions [n,2] (float32, on the device of velocity_request; the
    GPU work happens behind the C ABI, the tensors cross it through host
    memory).  The rollout keeps one forward's checkpoints: a backward after
    another forward on the same Rollout raises instead of returning the later
    call's gradients."""
    import torch

    class RolloutFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, vreq, k, rollout: Rollout, delta: float, n_steps: int):
            rollout.set_params(vreq.detach().cpu().numpy(), k.detach().cpu().numpy())
            rollout.forward(delta, n_steps)
            ctx.rollout = rollout
            ctx.generation = rollout.generation
            ctx.devices = (vreq.device, k.device)
            return torch.from_numpy(rollout.final_positions()).to(vreq.device)

        @staticmethod
        def backward(ctx, grad_pos):
            if ctx.rollout.generation != ctx.generation:
                raise RuntimeError("the Rollout ran another forward after this graph was recorded: its checkpoints "
                                   "now belong to the later call (use one Rollout per pending backward)")
            gv, gk = ctx.rollout.backward(grad_pos.detach().cpu().contiguous().numpy())
            return (torch.from_numpy(gv).to(ctx.devices[0]), torch.from_numpy(gk).to(ctx.devices[1]), None, None,
                    None)

    return RolloutFunction
