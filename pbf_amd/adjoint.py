"""Differentiable rollouts over the C ABI (pbf_adjoint_* of include/pbf.h).

``Rollout`` runs a fixed number of steps from an initial state and returns the
final positions; ``backward`` maps dL/d(final positions) to dL/d(initial
velocities), dL/d(XSPH c) and dL/d(vorticity epsilon) with the hand-written
reverse kernels, recomputing each checkpoint segment. ``RolloutFunction`` puts
the same pair behind ``torch.autograd.Function``. Each step's sort order and
neighbour table are held fixed (not differentiated). The reference has no
differentiable form; its step is src/SPH.cpp:246-334.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check
from .sph import scene_params


class Rollout:
    """steps steps of a scene's solver, with checkpoints every checkpoint_every steps.

    fp64: states, tapes and adjoints in double (used for finite-difference
    checks); otherwise float, the plain step's arithmetic."""

    def __init__(self, n, grid, params, steps, checkpoint_every=8, fp64=False):
        self._lib = _lib.load()
        self.n = int(n)
        self.grid = tuple(int(g) for g in grid)
        self.steps = int(steps)
        self.fp64 = bool(fp64)
        self.dtype = np.float64 if fp64 else np.float32
        self.params = params
        h = C.c_void_p()
        check(self._lib.pbf_adjoint_create(C.byref(h), self.n, *self.grid, C.byref(params), self.steps,
                                           int(checkpoint_every), int(self.fp64)))
        self._h = h
        # Counts the calls that replace what the reverse pass uses (the kept
        # structure and checkpoints of a forward, the coefficients); a replay
        # keeps them. A backward belongs to the forward of one generation only.
        self.generation = 0

    @classmethod
    def for_scene(cls, scene, steps=None, checkpoint_every=8, fp64=False):
        return cls(scene.n, scene.grid, scene_params(scene), scene.steps if steps is None else steps,
                   checkpoint_every, fp64)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pbf_adjoint_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _arr(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.shape == (self.n, 4), a.shape
        return a

    def set_coefficients(self, xsph_c, vorticity_epsilon):
        check(self._lib.pbf_adjoint_set_coefficients(self._h, float(xsph_c), float(vorticity_epsilon)))
        self.generation += 1

    def set_stream(self, stream_handle):
        check(self._lib.pbf_adjoint_set_stream(self._h, stream_handle))

    # host (numpy) interface
    def forward(self, pos4, vel4):
        pos4, vel4 = self._arr(pos4), self._arr(vel4)
        self.generation += 1
        check(self._lib.pbf_adjoint_forward(self._h, pos4.ctypes.data, vel4.ctypes.data, 1))
        out = np.empty((self.n, 4), self.dtype)
        check(self._lib.pbf_adjoint_positions(self._h, out.ctypes.data, 1))
        return out

    def velocities(self):
        out = np.empty((self.n, 4), self.dtype)
        check(self._lib.pbf_adjoint_velocities(self._h, out.ctypes.data, 1))
        return out

    def backward(self, grad_pos4):
        g = self._arr(grad_pos4)
        gv = np.empty((self.n, 4), self.dtype)
        gc, ge = C.c_double(), C.c_double()
        check(self._lib.pbf_adjoint_backward(self._h, g.ctypes.data, 1, gv.ctypes.data, C.byref(gc), C.byref(ge)))
        return gv, gc.value, ge.value

    def replay(self, pos4, vel4):
        pos4, vel4 = self._arr(pos4), self._arr(vel4)
        out = np.empty((self.n, 4), self.dtype)
        check(self._lib.pbf_adjoint_replay(self._h, pos4.ctypes.data, vel4.ctypes.data, 1, out.ctypes.data))
        return out

    def replay_frozen(self, pos4, vel4):
        """replay with the last forward's clamp decisions imposed (the smooth
        piece that backward differentiates); returns (positions, number of
        clamp decisions this replay's path would have taken otherwise)."""
        pos4, vel4 = self._arr(pos4), self._arr(vel4)
        out = np.empty((self.n, 4), self.dtype)
        flips = C.c_uint64()
        check(self._lib.pbf_adjoint_replay_frozen(self._h, pos4.ctypes.data, vel4.ctypes.data, 1, out.ctypes.data,
                                                  C.byref(flips)))
        return out, int(flips.value)

    def jvp(self, pos4, vel4, dvel4, dc=0.0, de=0.0):
        """FP64 handles: (positions, d positions) of replay along (dvel4, dc, de) by
        forward-mode dual numbers."""
        pos4, vel4, dvel4 = self._arr(pos4), self._arr(vel4), self._arr(dvel4)
        out = np.empty((self.n, 4), self.dtype)
        dout = np.empty((self.n, 4), self.dtype)
        check(self._lib.pbf_adjoint_jvp(self._h, pos4.ctypes.data, vel4.ctypes.data, dvel4.ctypes.data, float(dc),
                                        float(de), 1, out.ctypes.data, dout.ctypes.data))
        return out, dout

    # device (torch) interface: tensors on the GPU. The handle works on its own
    # stream, so torch's stream is drained before and the handle's after.
    def _sync(self):
        check(self._lib.pbf_synchronize(self._lib.pbf_adjoint_get_stream(self._h)))

    def forward_device(self, pos4, vel4):
        import torch
        torch.cuda.current_stream().synchronize()
        self.generation += 1
        check(self._lib.pbf_adjoint_forward(self._h, pos4.data_ptr(), vel4.data_ptr(), 0))
        out = torch.empty_like(pos4)
        check(self._lib.pbf_adjoint_positions(self._h, out.data_ptr(), 0))
        self._sync()
        return out

    def backward_device(self, grad_pos4):
        import torch
        torch.cuda.current_stream().synchronize()
        gv = torch.empty_like(grad_pos4)
        gc, ge = C.c_double(), C.c_double()
        check(self._lib.pbf_adjoint_backward(self._h, grad_pos4.data_ptr(), 0, gv.data_ptr(), C.byref(gc),
                                             C.byref(ge)))
        self._sync()
        return gv, gc.value, ge.value


def _autograd_function():
    import torch

    class RolloutFunction(torch.autograd.Function):
        """final positions (n, 3) = rollout(pos0, v0 (n, 3), xsph c, vorticity epsilon)."""

        @staticmethod
        def forward(ctx, v0, xsph_c, vorticity_epsilon, rollout, pos0):
            dt = torch.float64 if rollout.fp64 else torch.float32
            rollout.set_coefficients(float(xsph_c), float(vorticity_epsilon))
            n = rollout.n
            vel4 = torch.zeros((n, 4), dtype=dt, device=v0.device)
            vel4[:, :3] = v0.detach().to(dt)
            pos4 = torch.zeros((n, 4), dtype=dt, device=v0.device)
            pos4[:, :3] = pos0.to(dt)
            out = rollout.forward_device(pos4.contiguous(), vel4.contiguous())
            ctx.rollout = rollout
            ctx.generation = rollout.generation
            ctx.dtypes = (v0.dtype, xsph_c.dtype, vorticity_epsilon.dtype)
            ctx.devices = (xsph_c.device, vorticity_epsilon.device)
            return out[:, :3].to(v0.dtype)

        @staticmethod
        def backward(ctx, grad_pos):
            r = ctx.rollout
            if r.generation != ctx.generation:
                raise RuntimeError("the Rollout handle ran another forward or coefficient change "
                                   "after the forward this gradient belongs to; the reverse pass would use "
                                   "that later rollout. Call backward first, or use one Rollout per rollout")
            dt = torch.float64 if r.fp64 else torch.float32
            g4 = torch.zeros((r.n, 4), dtype=dt, device=grad_pos.device)
            g4[:, :3] = grad_pos.to(dt)
            gv, gc, ge = r.backward_device(g4.contiguous())
            tv, tc, te = ctx.dtypes
            dc, de = ctx.devices
            return (gv[:, :3].to(tv), torch.tensor(gc, dtype=tc, device=dc), torch.tensor(ge, dtype=te, device=de),
                    None, None)

    return RolloutFunction


_fn = None


def rollout_positions(v0, xsph_c, vorticity_epsilon, rollout: Rollout, pos0):
    """Differentiable final positions (n, 3) of ``rollout`` from pos0 (n, 3) and v0 (n, 3).

    The handle keeps one forward's structure and checkpoints: the gradient of a
    result is available until the handle's next forward or coefficient change,
    after which its backward raises."""
    global _fn
    if _fn is None:
        _fn = _autograd_function()
    return _fn.apply(v0, xsph_c, vorticity_epsilon, rollout, pos0)
