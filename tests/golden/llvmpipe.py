"""Driver for oracle/_ref/glsl_runner: runs the reference's own GLSL compute shaders
under Mesa llvmpipe (in the development container only; the shaders are read
from the reference checkout at run time) and loads what they wrote.

Only the fixture generator in this folder and the CPU-baseline notes use it.
"""
from __future__ import annotations

import os
import struct
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
RUNNER = os.path.join(ROOT, "oracle", "_ref", "glsl_runner")
REFDIR = "/root/reference"


def available() -> bool:
    return os.path.exists(RUNNER) and os.path.isdir(os.path.join(REFDIR, "shaders"))


def write_job(path, scene, nsteps, dump_from=0, dump_to=0, pos=None, vel=None, end_init=None,
              vort_init=None, dump_passes=False, warmup=0, highlight_init=None):
    p = scene.params
    gx, gy, gz = scene.grid
    pos = scene.pos if pos is None else pos
    vel = scene.vel if vel is None else vel
    hdr = struct.pack("<8s4if8f11i", b"PBFJOB2\0", pos.shape[0], gx, gy, gz, p.h, *[float(x) for x in p.ubo()],
                      p.iterations, int(p.vorticity), int(p.extforce), nsteps, dump_from, dump_to,
                      int(end_init is not None), int(vort_init is not None), int(dump_passes), warmup,
                      int(highlight_init is not None))
    with open(path, "wb") as f:
        f.write(hdr)
        f.write(np.ascontiguousarray(pos, dtype=np.float32).tobytes())
        f.write(np.ascontiguousarray(vel, dtype=np.float32).tobytes())
        if end_init is not None:
            f.write(np.ascontiguousarray(end_init, dtype=np.int32).tobytes())
        if vort_init is not None:
            f.write(np.ascontiguousarray(vort_init, dtype=np.float32).tobytes())
        if highlight_init is not None:
            f.write(np.ascontiguousarray(highlight_init, dtype=np.uint32).tobytes())


class Run:
    """Outputs of one runner invocation, loaded lazily from its output folder."""

    def __init__(self, outdir, n, cells):
        self.outdir = outdir
        self.n = n
        self.cells = cells

    def _load(self, name, dtype, shape):
        a = np.fromfile(os.path.join(self.outdir, name + ".bin"), dtype=dtype)
        return a.reshape(shape)

    def has(self, name):
        return os.path.exists(os.path.join(self.outdir, name + ".bin"))

    def records(self, name):
        return self._load(name, np.float32, (self.n, 4))

    def floats(self, name):
        return self._load(name, np.float32, (self.n,))

    def words(self, name):
        return self._load(name, np.uint32, (self.n,))

    def table(self, name):
        return self._load(name, np.int32, (self.cells,))

    def nbr(self, name):
        return self._load(name, np.int32, (self.n, 3, 4))

    def timing(self):
        with open(os.path.join(self.outdir, "timing.txt")) as f:
            return f.read()


def run(scene, nsteps, dump_from=0, dump_to=0, threads=1, workdir=None, timeout=3600, **kw) -> Run:
    if not available():
        raise RuntimeError("llvmpipe runner or reference shaders not available")
    workdir = workdir or tempfile.mkdtemp(prefix="pbf_llvmpipe_")
    os.makedirs(workdir, exist_ok=True)
    job = os.path.join(workdir, "job.bin")
    write_job(job, scene, nsteps, dump_from, dump_to, **kw)
    env = dict(os.environ, LP_NUM_THREADS=str(threads))
    subprocess.run([RUNNER, job, workdir, REFDIR], check=True, env=env, timeout=timeout,
                   stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    gx, gy, gz = scene.grid
    return Run(workdir, scene.n if kw.get("pos") is None else kw["pos"].shape[0], gx * gy * gz)
