"""Python mirror of the reference's step classes over the C ABI.

``SPH``, ``RadixSort`` and ``NeighbourCellFinder`` keep the reference's method
names and argument meaning (src/SPH.h:33-241, src/RadixSort.h:33-57,
src/NeighbourCellFinder.h:35-58); getters that return GL buffer names there
return HIP device pointers here. Errors raise ``PbfError`` (``PbfLogicError``
where the reference throws std::logic_error).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import PbfParams, PbfTiming, check


def params_from(p) -> PbfParams:
    """PbfParams from a pbf_amd.scene.Params and a grid-independent wall margin."""
    out = PbfParams()
    out.one_over_rho_0 = p.one_over_rho_0
    out.epsilon = p.epsilon
    out.gravity = p.gravity
    out.timestep = p.timestep
    out.tensile_instability_k = p.tensile_instability_k
    out.tensile_instability_scale = p.tensile_instability_scale
    out.xsph_viscosity_c = p.xsph_viscosity_c
    out.vorticity_epsilon = p.vorticity_epsilon
    out.iterations = int(p.iterations)
    out.vorticity = int(bool(p.vorticity))
    out.extforce = int(bool(p.extforce))
    out.h = p.h
    out.gravity_dir[:] = [float(x) for x in p.gravity_dir]
    return out


def scene_params(scene) -> PbfParams:
    """Full parameter block of a scene: reference wall (16, 0, 16) inside its grid."""
    out = params_from(scene.params)
    lo = scene.params.wall_lo
    m = scene.params.wall_hi_margin
    out.wall_lo[:] = [float(x) for x in lo]
    out.wall_hi[:] = [float(g) - float(w) for g, w in zip(scene.grid, m)]
    return out


def _grid(g):
    return (C.c_int32 * 3)(*[int(x) for x in g])


class SPH:
    """SPH (src/SPH.h:40-241) on one MI355X."""

    def __init__(self, numparticles: int, gridsize=(128, 64, 128)):
        self._lib = _lib.load()
        h = C.c_void_p()
        check(self._lib.pbf_sph_create(C.byref(h), int(numparticles), *[int(x) for x in gridsize]))
        self._h = h
        self.numparticles = int(numparticles)
        self.gridsize = tuple(int(x) for x in gridsize)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pbf_sph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameters (src/SPH.cpp:166-216)
    def params(self) -> PbfParams:
        p = PbfParams()
        check(self._lib.pbf_sph_get_params(self._h, C.byref(p)))
        return p

    def set_params(self, p: PbfParams):
        check(self._lib.pbf_sph_set_params(self._h, C.byref(p)))

    def _set(self, name, value):
        p = self.params()
        setattr(p, name, value)
        self.set_params(p)

    def GetRestDensity(self):
        return 1.0 / self.params().one_over_rho_0

    def SetRestDensity(self, rho):
        self._set("one_over_rho_0", float(np.float32(1.0) / np.float32(rho)))

    def GetCFMEpsilon(self):
        return self.params().epsilon

    def SetCFMEpsilon(self, v):
        self._set("epsilon", v)

    def GetGravity(self):
        return self.params().gravity

    def SetGravity(self, v):
        self._set("gravity", v)

    def GetTimestep(self):
        return self.params().timestep

    def SetTimestep(self, v):
        self._set("timestep", v)

    def GetTensileInstabilityK(self):
        return self.params().tensile_instability_k

    def SetTensileInstabilityK(self, v):
        self._set("tensile_instability_k", v)

    def GetTensileInstabilityScale(self):
        return self.params().tensile_instability_scale

    def SetTensileInstabilityScale(self, v):
        self._set("tensile_instability_scale", v)

    def GetXSPHViscosity(self):
        return self.params().xsph_viscosity_c

    def SetXSPHViscosity(self, v):
        self._set("xsph_viscosity_c", v)

    def GetVorticityEpsilon(self):
        return self.params().vorticity_epsilon

    def SetVorticityEpsilon(self, v):
        self._set("vorticity_epsilon", v)

    def GetNumSolverIterations(self):
        return self.params().iterations

    def SetNumSolverIterations(self, v):
        self._set("iterations", int(v))

    def IsVorticityConfinementEnabled(self):
        return bool(self.params().vorticity)

    def SetVorticityConfinementEnabled(self, flag):
        self._set("vorticity", int(bool(flag)))

    def SetExternalForce(self, state):
        self._set("extforce", int(bool(state)))

    @staticmethod
    def Wpoly6(r, h):
        return _lib.load().pbf_wpoly6(float(r), float(h))

    # ---- buffers
    def GetPositionBuffer(self) -> int:
        ptr = C.c_void_p()
        check(self._lib.pbf_sph_positions_device(self._h, C.byref(ptr), None))
        return ptr.value

    def GetVelocityBuffer(self) -> int:
        ptr = C.c_void_p()
        check(self._lib.pbf_sph_velocities_device(self._h, C.byref(ptr), None))
        return ptr.value

    def export_positions(self, components: int = 3, stride: int | None = None) -> np.ndarray:
        """Render feed (pbf_sph_export_positions): positions in id order, `components` floats
        per particle, `stride` bytes apart (default packed), as a host array."""
        stride = 4 * components if stride is None else int(stride)
        assert stride % 4 == 0
        buf = np.zeros((self.numparticles, stride // 4), np.float32)
        check(self._lib.pbf_sph_export_positions(self._h, buf.ctypes.data, stride, int(components), 1))
        return buf

    def GetHighlightBuffer(self) -> int:
        """SPH::GetHighlightBuffer (src/SPH.h:205-207): uint32 words by particle id."""
        ptr = C.c_void_p()
        check(self._lib.pbf_sph_highlight_device(self._h, C.byref(ptr), None))
        return ptr.value

    def highlight(self) -> np.ndarray:
        out = np.empty(self.numparticles, np.uint32)
        check(self._lib.pbf_sph_get_highlight(self._h, out.ctypes.data, 1))
        return out

    def set_highlight(self, words=None):
        """Writes the highlight words (uint32 by id); None clears them (src/Simulation.cpp:270-272)."""
        if words is None:
            check(self._lib.pbf_sph_set_highlight(self._h, None, 1))
            return
        words = np.ascontiguousarray(words, np.uint32)
        assert words.shape == (self.numparticles,)
        check(self._lib.pbf_sph_set_highlight(self._h, words.ctypes.data, 1))

    def toggle_highlight(self, particle_id: int) -> int:
        """Selection toggle of Simulation::OnMouseDown (src/Simulation.cpp:163-193); the new word."""
        now = C.c_uint32()
        check(self._lib.pbf_sph_toggle_highlight(self._h, int(particle_id), C.byref(now)))
        return now.value

    def upload(self, pos4: np.ndarray, vel4: np.ndarray):
        """State upload in id order (Simulation::ResetParticleBuffer's copies)."""
        pos4 = np.ascontiguousarray(pos4, dtype=np.float32)
        vel4 = np.ascontiguousarray(vel4, dtype=np.float32)
        assert pos4.shape == (self.numparticles, 4) and vel4.shape == (self.numparticles, 4)
        check(self._lib.pbf_sph_upload_state(self._h, pos4.ctypes.data, vel4.ctypes.data, 1))

    def download(self):
        pos = np.empty((self.numparticles, 4), np.float32)
        vel = np.empty((self.numparticles, 4), np.float32)
        check(self._lib.pbf_sph_download_state(self._h, pos.ctypes.data, vel.ctypes.data, 1))
        return pos, vel

    def end_table(self) -> np.ndarray:
        gx, gy, gz = self.gridsize
        out = np.empty(gx * gy * gz, np.int32)
        check(self._lib.pbf_sph_get_end_table(self._h, out.ctypes.data, 1))
        return out

    def set_end_table(self, end: np.ndarray):
        end = np.ascontiguousarray(end, np.int32)
        check(self._lib.pbf_sph_set_end_table(self._h, end.ctypes.data, 1))

    def neighbour_table(self) -> np.ndarray:
        out = np.empty((self.numparticles, 3, 4), np.int32)
        check(self._lib.pbf_sph_neighbour_table(self._h, out.ctypes.data, 1))
        return out

    def sorted_records(self) -> np.ndarray:
        out = np.empty((self.numparticles, 4), np.float32)
        check(self._lib.pbf_sph_sorted_records(self._h, out.ctypes.data, 1))
        return out

    # ---- stepping
    def Run(self):
        check(self._lib.pbf_sph_run(self._h))

    def step(self, nsteps: int = 1):
        check(self._lib.pbf_sph_step(self._h, int(nsteps)))

    def set_stream(self, stream_handle: int | None):
        check(self._lib.pbf_sph_set_stream(self._h, stream_handle))

    def stream(self) -> int:
        return self._lib.pbf_sph_get_stream(self._h) or 0

    def timing(self) -> dict:
        t = PbfTiming()
        check(self._lib.pbf_sph_timing(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in PbfTiming._fields_}

    def set_graph(self, enable: bool):
        """Run each later step as one hipGraph replay (pbf_sph_set_graph)."""
        check(self._lib.pbf_sph_set_graph(self._h, int(bool(enable))))

    def dense_info(self):
        """(threshold, count) of the last step's dense-particle split (pbf_sph_dense_info)."""
        th, cnt = C.c_int32(), C.c_int32()
        check(self._lib.pbf_sph_dense_info(self._h, C.byref(th), C.byref(cnt)))
        return th.value, cnt.value

    def set_kernel_timing(self, enable: bool):
        check(self._lib.pbf_sph_set_kernel_timing(self._h, int(bool(enable))))

    def kernel_times(self) -> dict:
        """{kernel name: (total ms, launches)} since the last call; resets the totals."""
        arr = (_lib.PbfKernelTime * _lib.NUM_KERNEL_IDS)()
        check(self._lib.pbf_sph_kernel_times(self._h, arr))
        return {self._lib.pbf_kernel_name(k).decode(): (arr[k].total_ms, arr[k].launches)
                for k in range(_lib.NUM_KERNEL_IDS)}

    def OutputTiming(self):
        """Prints the five phases like SPH::OutputTiming (src/SPH.cpp:218-240)."""
        t = self.timing()
        print(f"Position prediction: {t['predict_ms']} ms")
        print(f"Sorting: {t['sort_ms']} ms")
        print(f"Neighbour cell search: {t['neighbour_cells_ms']} ms")
        print(f"Solver: {t['solver_ms']} ms")
        print(f"Vorticity confinement: {t['vorticity_ms']} ms")


class RadixSort:
    """RadixSort (src/RadixSort.h:33-57): sorts float4 records by grid-cell key."""

    def __init__(self, blocksize: int, numblocks: int, gridsize):
        self._lib = _lib.load()
        h = C.c_void_p()
        check(self._lib.pbf_radixsort_create(C.byref(h), int(blocksize), int(numblocks),
                                             *[int(x) for x in gridsize]))
        self._h = h
        self.n = int(blocksize) * int(numblocks)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pbf_radixsort_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def GetBuffer(self) -> int:
        ptr = C.c_void_p()
        check(self._lib.pbf_radixsort_buffer(self._h, C.byref(ptr)))
        return ptr.value

    def Run(self):
        check(self._lib.pbf_radixsort_run(self._h))

    def permutation_ptr(self) -> int:
        ptr = C.c_void_p()
        check(self._lib.pbf_radixsort_permutation(self._h, C.byref(ptr)))
        return ptr.value

    def set_stream(self, stream_handle):
        check(self._lib.pbf_radixsort_set_stream(self._h, stream_handle))

    def stream(self) -> int:
        return self._lib.pbf_radixsort_get_stream(self._h) or 0


class NeighbourCellFinder:
    """NeighbourCellFinder (src/NeighbourCellFinder.h:35-58)."""

    def __init__(self, numparticles: int, gridsize):
        self._lib = _lib.load()
        h = C.c_void_p()
        check(self._lib.pbf_ncf_create(C.byref(h), int(numparticles), *[int(x) for x in gridsize]))
        self._h = h
        self.n = int(numparticles)
        self.gridsize = tuple(int(x) for x in gridsize)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pbf_ncf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def FindNeighbourCells(self, particlebuffer: int):
        check(self._lib.pbf_ncf_find(self._h, C.c_void_p(particlebuffer)))

    def GetResult(self) -> int:
        ptr = C.c_void_p()
        check(self._lib.pbf_ncf_result(self._h, C.byref(ptr)))
        return ptr.value

    def tables(self):
        gx, gy, gz = self.gridsize
        start = np.empty(gx * gy * gz, np.int32)
        end = np.empty(gx * gy * gz, np.int32)
        check(self._lib.pbf_ncf_get_tables(self._h, start.ctypes.data, end.ctypes.data, 1))
        return start, end

    def set_end_table(self, end):
        end = np.ascontiguousarray(end, np.int32)
        check(self._lib.pbf_ncf_set_end_table(self._h, end.ctypes.data, 1))

    def set_stream(self, stream_handle):
        check(self._lib.pbf_ncf_set_stream(self._h, stream_handle))

    def stream(self) -> int:
        return self._lib.pbf_ncf_get_stream(self._h) or 0


# ---------------------------------------------------------------- stages
def stage_predict(n, grid, p: PbfParams, pos_dev, vel_dev, rec_dev, stream=None):
    check(_lib.load().pbf_stage_predict(n, _grid(grid), C.byref(p), pos_dev, vel_dev, rec_dev, stream))


def stage_lambda(n, grid, p: PbfParams, rec_dev, nbr12_dev, lambda_dev, stream=None):
    check(_lib.load().pbf_stage_lambda(n, _grid(grid), C.byref(p), rec_dev, nbr12_dev, lambda_dev, stream))


def stage_updatepos(n, grid, p: PbfParams, rec_in, lambda_dev, nbr12_dev, rec_out, stream=None):
    check(_lib.load().pbf_stage_updatepos(n, _grid(grid), C.byref(p), rec_in, lambda_dev, nbr12_dev, rec_out,
                                          stream))


def stage_update(n, p: PbfParams, rec_dev, pos_dev, vel_dev, stream=None):
    check(_lib.load().pbf_stage_update(n, C.byref(p), rec_dev, pos_dev, vel_dev, stream))


def stage_vorticity(n, grid, p: PbfParams, rec_dev, nbr12_dev, vel_dev, vort_dev, stream=None):
    check(_lib.load().pbf_stage_vorticity(n, _grid(grid), C.byref(p), rec_dev, nbr12_dev, vel_dev, vort_dev,
                                          stream))


def stage_vorticity_noalloc(n, grid, p: PbfParams, rec_dev, nbr12_dev, vel_dev, vort_dev, stream=None):
    """stage_vorticity with the library's fixed scratch array instead of device
    allocations (n at most PBF_NOALLOC_MAX_N); asynchronous on `stream`."""
    check(_lib.load().pbf_stage_vorticity_noalloc(n, _grid(grid), C.byref(p), rec_dev, nbr12_dev, vel_dev,
                                                  vort_dev, stream))
