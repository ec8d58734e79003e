"""Batched form of the step over the C ABI (pbf_batch_* of include/pbf.h).

Many independent scenes of uneven particle counts advance together, one launch
per stage for all of them; a whole step can be captured in a hipGraph, and an
autotuner picks the launch shapes. Per scene the results equal an ``SPH``
handle's for that scene bit for bit. This form has no counterpart in the
reference, whose ``SPH`` runs one scene (src/SPH.cpp:246-334).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import PbfBatchTuning, PbfSceneDesc, check
from .sph import params_from


def scene_desc(scene) -> PbfSceneDesc:
    d = PbfSceneDesc()
    d.n = scene.n
    d.grid[:] = [int(x) for x in scene.grid]
    p = scene.params
    d.wall_lo[:] = [float(x) for x in p.wall_lo]
    d.wall_hi[:] = [float(g) - float(m) for g, m in zip(scene.grid, p.wall_hi_margin)]
    d.gravity_dir[:] = [float(x) for x in p.gravity_dir]
    d.xsph_viscosity_c = float(p.xsph_viscosity_c)
    return d


class Batch:
    """Scenes (pbf_amd.scene.Scene) advanced together; shared constants from the first scene."""

    def __init__(self, scenes, shared=None):
        self._lib = _lib.load()
        self.scenes = list(scenes)
        descs = (PbfSceneDesc * len(self.scenes))(*[scene_desc(s) for s in self.scenes])
        self.shared = params_from(self.scenes[0].params) if shared is None else shared
        h = C.c_void_p()
        check(self._lib.pbf_batch_create(C.byref(h), len(self.scenes), descs, C.byref(self.shared)))
        self._h = h
        total = C.c_uint32()
        offs = (C.c_uint32 * (len(self.scenes) + 1))()
        check(self._lib.pbf_batch_layout(self._h, C.byref(total), offs))
        self.n = int(total.value)
        self.offsets = np.array(offs[:], dtype=np.int64)
        self.cells = [int(np.prod(s.grid)) for s in self.scenes]

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pbf_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, pos4=None, vel4=None):
        """State of every scene, one after another, each in id order (default: the scenes' own)."""
        pos4 = np.concatenate([s.pos for s in self.scenes]) if pos4 is None else pos4
        vel4 = np.concatenate([s.vel for s in self.scenes]) if vel4 is None else vel4
        pos4 = np.ascontiguousarray(pos4, np.float32)
        vel4 = np.ascontiguousarray(vel4, np.float32)
        assert pos4.shape == (self.n, 4) and vel4.shape == (self.n, 4)
        check(self._lib.pbf_batch_upload_state(self._h, pos4.ctypes.data, vel4.ctypes.data, 1))

    def download(self):
        pos = np.empty((self.n, 4), np.float32)
        vel = np.empty((self.n, 4), np.float32)
        check(self._lib.pbf_batch_download_state(self._h, pos.ctypes.data, vel.ctypes.data, 1))
        return pos, vel

    def split(self, a):
        return [a[self.offsets[k]:self.offsets[k + 1]] for k in range(len(self.scenes))]

    def step(self, nsteps: int = 1):
        check(self._lib.pbf_batch_step(self._h, int(nsteps)))

    def set_graph(self, enable: bool):
        check(self._lib.pbf_batch_set_graph(self._h, int(bool(enable))))

    def tuning(self) -> dict:
        t = PbfBatchTuning()
        check(self._lib.pbf_batch_get_tuning(self._h, C.byref(t)))
        return {"neighbour_block": t.neighbour_block, "stream_block": t.stream_block, "stream_ppt": t.stream_ppt,
                "list_depth": t.list_depth, "persistent": t.persistent}

    def set_tuning(self, neighbour_block=64, stream_block=256, stream_ppt=1, list_depth=56, persistent=0):
        t = PbfBatchTuning(int(neighbour_block), int(stream_block), int(stream_ppt), int(list_depth),
                           int(persistent), 0.0)
        check(self._lib.pbf_batch_set_tuning(self._h, C.byref(t)))

    def autotune(self, steps_per_trial: int = 2):
        """Times the launch shapes from the current state (neighbour shapes, then
        streaming shapes with the best of those) and keeps the fastest.
        Returns (best, trials) as dicts with step_ms."""
        best = PbfBatchTuning()
        res = (PbfBatchTuning * 128)()
        cnt = C.c_int32()
        check(self._lib.pbf_batch_autotune(self._h, int(steps_per_trial), C.byref(best), res, 128, C.byref(cnt)))

        def d(t):
            return {"neighbour_block": t.neighbour_block, "stream_block": t.stream_block, "stream_ppt": t.stream_ppt,
                    "list_depth": t.list_depth, "persistent": t.persistent, "step_ms": t.step_ms}
        return d(best), [d(res[k]) for k in range(min(cnt.value, 128))]

    def end_tables(self) -> np.ndarray:
        out = np.empty(sum(self.cells), np.int32)
        check(self._lib.pbf_batch_get_end_tables(self._h, out.ctypes.data, 1))
        return out

    def set_end_tables(self, end: np.ndarray):
        end = np.ascontiguousarray(end, np.int32)
        assert end.shape == (sum(self.cells),)
        check(self._lib.pbf_batch_set_end_tables(self._h, end.ctypes.data, 1))

    def set_stream(self, stream_handle):
        check(self._lib.pbf_batch_set_stream(self._h, stream_handle))

    def stream(self) -> int:
        return self._lib.pbf_batch_get_stream(self._h) or 0
