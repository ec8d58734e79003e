"""ctypes binding of the C restatement of the reference's step (liboracle.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it as the checker or as the timed CPU baseline; the
product (pbf_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ORDER_JACOBI = 0
ORDER_REFERENCE = 1


class Cfg(C.Structure):
    _fields_ = [("n", C.c_int32), ("grid", C.c_int32 * 3), ("h", C.c_float), ("p", C.c_float * 8),
                ("iterations", C.c_int32), ("vorticity", C.c_int32), ("extforce", C.c_int32),
                ("wall_lo", C.c_float * 3), ("wall_hi", C.c_float * 3), ("gdir", C.c_float * 3),
                ("key_bits", C.c_int32)]


class Stages(C.Structure):
    _fields_ = [("predict", C.c_void_p), ("sorted", C.c_void_p), ("start", C.c_void_p), ("nbr12", C.c_void_p),
                ("lambda_", C.c_void_p), ("solved", C.c_void_p)]


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        cp = C.POINTER(Cfg)
        sig = {
            "pbfo_predict": [cp, vp, vp, vp], "pbfo_keys": [cp, vp, vp], "pbfo_sort": [cp, vp, vp, vp],
            "pbfo_findcells": [cp, vp, vp, vp], "pbfo_neighbourcells": [cp, vp, vp, vp, vp],
            "pbfo_lambda": [cp, vp, vp, vp], "pbfo_updatepos": [cp, vp, vp, vp, vp, C.c_int],
            "pbfo_update": [cp, vp, vp, vp], "pbfo_vorticity": [cp, vp, vp, vp, vp, C.c_int],
            "pbfo_step": [cp, vp, vp, vp, vp, C.c_int, C.POINTER(Stages)],
            "pbfo_step_hl": [cp, vp, vp, vp, vp, C.c_int, C.POINTER(Stages), vp],
            "pbfo_highlight": [cp, vp, vp, vp],
            "pbfo_step_f64": [cp, vp, vp, vp, vp], "pbfo_step_fixed_f64": [cp, vp, vp, vp, vp],
        }
        for name, args in sig.items():
            f = getattr(_lib, name)
            f.argtypes = args
            f.restype = None
        _lib.pbfo_num_threads.restype = C.c_int
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def make_cfg(n, grid, params, key_bits=None, wall_lo=None, wall_hi=None, gdir=None) -> Cfg:
    """Oracle configuration from a pbf_amd.scene.Params (reference wall and gravity by default)."""
    from pbf_amd.scene import key_bits as kb
    c = Cfg()
    c.n = int(n)
    c.grid[:] = [int(g) for g in grid]
    c.h = float(params.h)
    c.p[:] = [float(x) for x in params.ubo()]
    c.iterations = int(params.iterations)
    c.vorticity = int(bool(params.vorticity))
    c.extforce = int(bool(params.extforce))
    lo = params.wall_lo if wall_lo is None else wall_lo
    if wall_hi is None:
        m = params.wall_hi_margin
        wall_hi = tuple(float(g) - float(w) for g, w in zip(grid, m))
    c.wall_lo[:] = [float(x) for x in lo]
    c.wall_hi[:] = [float(x) for x in wall_hi]
    c.gdir[:] = [float(x) for x in (params.gravity_dir if gdir is None else gdir)]
    c.key_bits = int(kb(grid) if key_bits is None else key_bits)
    return c


def cfg_for(scene) -> Cfg:
    return make_cfg(scene.n, scene.grid, scene.params)


def cells(cfg: Cfg) -> int:
    return int(cfg.grid[0]) * int(cfg.grid[1]) * int(cfg.grid[2])


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def predict(cfg, pos, vel):
    out = np.zeros((cfg.n, 4), np.float32)
    pos, vel = f32(pos), f32(vel)
    lib().pbfo_predict(C.byref(cfg), _p(pos), _p(vel), _p(out))
    return out


def keys(cfg, rec):
    rec = f32(rec)
    out = np.zeros(cfg.n, np.uint32)
    lib().pbfo_keys(C.byref(cfg), _p(rec), _p(out))
    return out


def sort(cfg, rec):
    rec = f32(rec)
    out = np.zeros_like(rec)
    perm = np.zeros(cfg.n, np.int32)
    lib().pbfo_sort(C.byref(cfg), _p(rec), _p(out), _p(perm))
    return out, perm


def findcells(cfg, sorted_rec, end):
    """Returns (start, end'); `end` is the persistent table before the call."""
    sorted_rec = f32(sorted_rec)
    start = np.zeros(cells(cfg), np.int32)
    end = np.array(end, dtype=np.int32, copy=True)
    lib().pbfo_findcells(C.byref(cfg), _p(sorted_rec), _p(start), _p(end))
    return start, end


def neighbourcells(cfg, sorted_rec, start, end):
    sorted_rec = f32(sorted_rec)
    start = np.ascontiguousarray(start, np.int32)
    end = np.ascontiguousarray(end, np.int32)
    out = np.zeros((cfg.n, 3, 4), np.int32)
    lib().pbfo_neighbourcells(C.byref(cfg), _p(sorted_rec), _p(start), _p(end), _p(out))
    return out


def lambdas(cfg, rec, nbr):
    rec = f32(rec)
    nbr = np.ascontiguousarray(nbr, np.int32)
    out = np.zeros(cfg.n, np.float32)
    lib().pbfo_lambda(C.byref(cfg), _p(rec), _p(nbr), _p(out))
    return out


def updatepos(cfg, rec, lam, nbr, order=ORDER_JACOBI):
    rec = f32(rec)
    lam = f32(lam)
    nbr = np.ascontiguousarray(nbr, np.int32)
    out = np.zeros_like(rec)
    lib().pbfo_updatepos(C.byref(cfg), _p(rec), _p(lam), _p(nbr), _p(out), int(order))
    return out


def update(cfg, rec, pos, vel):
    rec = f32(rec)
    pos = np.array(pos, np.float32, copy=True)
    vel = np.array(vel, np.float32, copy=True)
    lib().pbfo_update(C.byref(cfg), _p(rec), _p(pos), _p(vel))
    return pos, vel


def vorticity(cfg, rec, nbr, vel, vort, order=ORDER_JACOBI):
    rec = f32(rec)
    nbr = np.ascontiguousarray(nbr, np.int32)
    vel = np.array(vel, np.float32, copy=True)
    vort = np.array(vort, np.float32, copy=True)
    lib().pbfo_vorticity(C.byref(cfg), _p(rec), _p(nbr), _p(vel), _p(vort), int(order))
    return vel, vort


def highlight(cfg, sorted_rec, nbr, flags):
    """The two highlight passes over the flag words (by id); returns the new flags."""
    sorted_rec = f32(sorted_rec)
    nbr = np.ascontiguousarray(nbr, np.int32)
    out = np.array(flags, dtype=np.uint32, copy=True)
    lib().pbfo_highlight(C.byref(cfg), _p(sorted_rec), _p(nbr), _p(out))
    return out


def step(cfg, pos, vel, end=None, vort=None, order=ORDER_JACOBI, stages=False, highlight=None):
    """One step; returns (pos, vel, end, vort[, stage dict]). highlight: flag words
    (uint32, by id) that the step's highlight passes update in place."""
    n = cfg.n
    pos = np.array(pos, np.float32, copy=True)
    vel = np.array(vel, np.float32, copy=True)
    end = np.zeros(cells(cfg), np.int32) if end is None else np.array(end, np.int32, copy=True)
    vort = np.zeros(n, np.float32) if vort is None else np.array(vort, np.float32, copy=True)
    st = None
    sd = None
    if stages:
        it = int(cfg.iterations)
        sd = {"predict": np.zeros((n, 4), np.float32), "sorted": np.zeros((n, 4), np.float32),
              "start": np.zeros(cells(cfg), np.int32), "nbr": np.zeros((n, 3, 4), np.int32),
              "lambda": np.zeros((it, n), np.float32), "solved": np.zeros((it, n, 4), np.float32)}
        st = Stages(_p(sd["predict"]), _p(sd["sorted"]), _p(sd["start"]), _p(sd["nbr"]), _p(sd["lambda"]),
                    _p(sd["solved"]))
    if highlight is not None:
        assert highlight.dtype == np.uint32 and highlight.flags["C_CONTIGUOUS"] and highlight.shape == (n,)
    lib().pbfo_step_hl(C.byref(cfg), _p(pos), _p(vel), _p(end), _p(vort), int(order),
                       C.byref(st) if st is not None else None, _p(highlight))
    if stages:
        return pos, vel, end, vort, sd
    return pos, vel, end, vort


def step_f64(cfg, pos, vel, end=None, vort=None):
    pos = np.array(pos, np.float64, copy=True)
    vel = np.array(vel, np.float64, copy=True)
    end = np.zeros(cells(cfg), np.int32) if end is None else np.array(end, np.int32, copy=True)
    vort = np.zeros(cfg.n, np.float64) if vort is None else np.array(vort, np.float64, copy=True)
    lib().pbfo_step_f64(C.byref(cfg), _p(pos), _p(vel), _p(end), _p(vort))
    return pos, vel, end, vort


def step_fixed_f64(cfg, pos, vel, perm, nbr):
    pos = np.array(pos, np.float64, copy=True)
    vel = np.array(vel, np.float64, copy=True)
    perm = np.ascontiguousarray(perm, np.int32)
    nbr = np.ascontiguousarray(nbr, np.int32)
    lib().pbfo_step_fixed_f64(C.byref(cfg), _p(pos), _p(vel), _p(perm), _p(nbr))
    return pos, vel


def num_threads() -> int:
    return int(lib().pbfo_num_threads())
