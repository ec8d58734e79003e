"""ctypes binding of libpbf_amd.so (the C ABI of include/pbf.h).

The library is built in-tree (``python -m pbf_amd.build``). Loading fails
loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpbf_amd.so")


class PbfParams(C.Structure):
    """``pbf_params`` of include/pbf.h (sphparams_t, src/SPH.h:252-285, plus host fields)."""
    _fields_ = [("one_over_rho_0", C.c_float), ("epsilon", C.c_float), ("gravity", C.c_float),
                ("timestep", C.c_float), ("tensile_instability_k", C.c_float),
                ("tensile_instability_scale", C.c_float), ("xsph_viscosity_c", C.c_float),
                ("vorticity_epsilon", C.c_float), ("iterations", C.c_uint32), ("vorticity", C.c_int32),
                ("extforce", C.c_int32), ("h", C.c_float), ("wall_lo", C.c_float * 3), ("wall_hi", C.c_float * 3),
                ("gravity_dir", C.c_float * 3)]


NUM_KERNEL_IDS = 11


class PbfKernelTime(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("launches", C.c_uint32)]


class PbfTiming(C.Structure):
    _fields_ = [("predict_ms", C.c_double), ("sort_ms", C.c_double), ("neighbour_cells_ms", C.c_double),
                ("solver_ms", C.c_double), ("vorticity_ms", C.c_double)]


class PbfSceneDesc(C.Structure):
    """``pbf_scene_desc`` of include/pbf.h: one scene of the batched form."""
    _fields_ = [("n", C.c_uint32), ("grid", C.c_int32 * 3), ("wall_lo", C.c_float * 3), ("wall_hi", C.c_float * 3),
                ("gravity_dir", C.c_float * 3), ("xsph_viscosity_c", C.c_float)]


class PbfBatchTuning(C.Structure):
    _fields_ = [("neighbour_block", C.c_int32), ("stream_block", C.c_int32), ("stream_ppt", C.c_int32),
                ("list_depth", C.c_int32), ("persistent", C.c_int32), ("step_ms", C.c_double)]


vp = C.c_void_p
i32 = C.c_int32
u32 = C.c_uint32
pp = C.POINTER(PbfParams)
grid3 = C.POINTER(C.c_int32)

SIGNATURES = {
    "pbf_last_error": ([], C.c_char_p),
    "pbf_version": ([], C.c_char_p),
    "pbf_device_count": ([], C.c_int),
    "pbf_memcpy": ([vp, vp, C.c_size_t, C.c_int, vp], C.c_int),
    "pbf_synchronize": ([vp], C.c_int),
    "pbf_params_default": ([pp, i32, i32, i32], None),
    "pbf_wpoly6": ([C.c_float, C.c_float], C.c_float),
    "pbf_key_bits": ([i32, i32, i32], C.c_int),
    "pbf_selftest_math": ([C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64)], C.c_int),
    "pbf_selftest_xcd_map": ([i32, u32, vp], C.c_int),
    "pbf_sph_create": ([C.POINTER(vp), u32, i32, i32, i32], C.c_int),
    "pbf_sph_destroy": ([vp], C.c_int),
    "pbf_sph_set_params": ([vp, pp], C.c_int),
    "pbf_sph_get_params": ([vp, pp], C.c_int),
    "pbf_sph_upload_state": ([vp, vp, vp, C.c_int], C.c_int),
    "pbf_sph_download_state": ([vp, vp, vp, C.c_int], C.c_int),
    "pbf_sph_run": ([vp], C.c_int),
    "pbf_sph_step": ([vp, i32], C.c_int),
    "pbf_sph_set_graph": ([vp, C.c_int], C.c_int),
    "pbf_sph_dense_info": ([vp, vp, vp], C.c_int),
    "pbf_sph_positions_device": ([vp, C.POINTER(vp), C.POINTER(C.c_size_t)], C.c_int),
    "pbf_sph_velocities_device": ([vp, C.POINTER(vp), C.POINTER(C.c_size_t)], C.c_int),
    "pbf_sph_export_positions": ([vp, vp, C.c_size_t, i32, C.c_int], C.c_int),
    "pbf_sph_highlight_device": ([vp, C.POINTER(vp), C.POINTER(C.c_size_t)], C.c_int),
    "pbf_sph_get_highlight": ([vp, vp, C.c_int], C.c_int),
    "pbf_sph_set_highlight": ([vp, vp, C.c_int], C.c_int),
    "pbf_sph_toggle_highlight": ([vp, u32, C.POINTER(u32)], C.c_int),
    "pbf_sph_timing": ([vp, C.POINTER(PbfTiming)], C.c_int),
    "pbf_sph_set_kernel_timing": ([vp, C.c_int], C.c_int),
    "pbf_sph_kernel_times": ([vp, C.POINTER(PbfKernelTime)], C.c_int),
    "pbf_kernel_name": ([C.c_int], C.c_char_p),
    "pbf_sph_set_stream": ([vp, vp], C.c_int),
    "pbf_sph_get_stream": ([vp], vp),
    "pbf_sph_get_end_table": ([vp, vp, C.c_int], C.c_int),
    "pbf_sph_set_end_table": ([vp, vp, C.c_int], C.c_int),
    "pbf_sph_neighbour_table": ([vp, vp, C.c_int], C.c_int),
    "pbf_sph_sorted_records": ([vp, vp, C.c_int], C.c_int),
    "pbf_radixsort_create": ([C.POINTER(vp), u32, u32, i32, i32, i32], C.c_int),
    "pbf_radixsort_destroy": ([vp], C.c_int),
    "pbf_radixsort_buffer": ([vp, C.POINTER(vp)], C.c_int),
    "pbf_radixsort_run": ([vp], C.c_int),
    "pbf_radixsort_permutation": ([vp, C.POINTER(vp)], C.c_int),
    "pbf_radixsort_set_stream": ([vp, vp], C.c_int),
    "pbf_radixsort_get_stream": ([vp], vp),
    "pbf_ncf_create": ([C.POINTER(vp), u32, i32, i32, i32], C.c_int),
    "pbf_ncf_destroy": ([vp], C.c_int),
    "pbf_ncf_find": ([vp, vp], C.c_int),
    "pbf_ncf_result": ([vp, C.POINTER(vp)], C.c_int),
    "pbf_ncf_get_tables": ([vp, vp, vp, C.c_int], C.c_int),
    "pbf_ncf_set_end_table": ([vp, vp, C.c_int], C.c_int),
    "pbf_ncf_set_stream": ([vp, vp], C.c_int),
    "pbf_ncf_get_stream": ([vp], vp),
    "pbf_batch_create": ([C.POINTER(vp), u32, C.POINTER(PbfSceneDesc), pp], C.c_int),
    "pbf_batch_destroy": ([vp], C.c_int),
    "pbf_batch_layout": ([vp, C.POINTER(u32), C.POINTER(u32)], C.c_int),
    "pbf_batch_upload_state": ([vp, vp, vp, C.c_int], C.c_int),
    "pbf_batch_download_state": ([vp, vp, vp, C.c_int], C.c_int),
    "pbf_batch_step": ([vp, i32], C.c_int),
    "pbf_batch_set_graph": ([vp, C.c_int], C.c_int),
    "pbf_batch_set_tuning": ([vp, C.POINTER(PbfBatchTuning)], C.c_int),
    "pbf_batch_get_tuning": ([vp, C.POINTER(PbfBatchTuning)], C.c_int),
    "pbf_batch_autotune": ([vp, i32, C.POINTER(PbfBatchTuning), C.POINTER(PbfBatchTuning), i32, C.POINTER(i32)],
                           C.c_int),
    "pbf_batch_get_end_tables": ([vp, vp, C.c_int], C.c_int),
    "pbf_batch_set_end_tables": ([vp, vp, C.c_int], C.c_int),
    "pbf_batch_set_stream": ([vp, vp], C.c_int),
    "pbf_batch_get_stream": ([vp], vp),
    "pbf_adjoint_create": ([C.POINTER(vp), u32, i32, i32, i32, pp, i32, i32, i32], C.c_int),
    "pbf_adjoint_destroy": ([vp], C.c_int),
    "pbf_adjoint_set_coefficients": ([vp, C.c_double, C.c_double], C.c_int),
    "pbf_adjoint_forward": ([vp, vp, vp, C.c_int], C.c_int),
    "pbf_adjoint_positions": ([vp, vp, C.c_int], C.c_int),
    "pbf_adjoint_velocities": ([vp, vp, C.c_int], C.c_int),
    "pbf_adjoint_backward": ([vp, vp, C.c_int, vp, C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int),
    "pbf_adjoint_replay": ([vp, vp, vp, C.c_int, vp], C.c_int),
    "pbf_adjoint_replay_frozen": ([vp, vp, vp, C.c_int, vp, C.POINTER(C.c_uint64)], C.c_int),
    "pbf_adjoint_jvp": ([vp, vp, vp, vp, C.c_double, C.c_double, C.c_int, vp, vp], C.c_int),
    "pbf_adjoint_set_stream": ([vp, vp], C.c_int),
    "pbf_adjoint_get_stream": ([vp], vp),
    "pbf_stage_predict": ([u32, grid3, pp, vp, vp, vp, vp], C.c_int),
    "pbf_stage_lambda": ([u32, grid3, pp, vp, vp, vp, vp], C.c_int),
    "pbf_stage_updatepos": ([u32, grid3, pp, vp, vp, vp, vp, vp], C.c_int),
    "pbf_stage_update": ([u32, pp, vp, vp, vp, vp], C.c_int),
    "pbf_stage_vorticity": ([u32, grid3, pp, vp, vp, vp, vp, vp], C.c_int),
    "pbf_stage_vorticity_noalloc": ([u32, grid3, pp, vp, vp, vp, vp, vp], C.c_int),
}

# Entry points added by later units are bound when present (see OPTIONAL).
OPTIONAL = {}

_lib = None


class PbfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pbf error {code}: {msg}")
        self.code = code


class PbfLogicError(PbfError):
    """PBF_ERR_LOGIC, the reference's std::logic_error (src/RadixSort.cpp:41-42)."""


def load() -> C.CDLL:
    """Loads libpbf_amd.so; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m pbf_amd.build` "
                           "(the HIP extension is required; there is no CPU path)")
    # PyTorch-ROCm ships its own HIP runtime under the same soname
    # (libamdhip64.so.7). Loaded first, it is the one this library binds to,
    # so torch tensors and these kernels share one runtime in the process;
    # loaded after this library it would start a second runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    for name, (args, res) in SIGNATURES.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    for name, (args, res) in OPTIONAL.items():
        if hasattr(lib, name):
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
    _lib = lib
    return lib


def check(rc: int):
    if rc != 0:
        msg = load().pbf_last_error().decode(errors="replace")
        if rc == -4:
            raise PbfLogicError(rc, msg)
        raise PbfError(rc, msg)
    return rc


def exported_symbols() -> list:
    return list(SIGNATURES)
