"""Device-buffer helpers for the GPU tests (torch is only the allocator here)."""
import ctypes as C

import numpy as np


def torch():
    import torch as T
    return T


def dev(a: np.ndarray):
    """Device copy of a host array (torch tensor on cuda:0), complete on return."""
    T = torch()
    t = T.from_numpy(np.ascontiguousarray(a)).to("cuda:0")
    T.cuda.synchronize()
    return t


def empty_like_host(shape, dtype):
    T = torch()
    tdt = {np.float32: T.float32, np.int32: T.int32, np.uint32: T.int32}[np.dtype(dtype).type]
    return T.empty(shape, dtype=tdt, device="cuda:0")


def host(t, dtype=None) -> np.ndarray:
    torch().cuda.synchronize()
    a = t.cpu().numpy()
    return a.view(dtype) if dtype is not None else a


def ptr(t) -> int:
    return t.data_ptr()


def stream() -> int:
    return torch().cuda.current_stream().cuda_stream


def copy_from_device(dev_ptr: int, shape, dtype, stream: int) -> np.ndarray:
    """Copy out of a library-owned buffer, ordered after the work on `stream`."""
    from pbf_amd import _lib
    out = np.empty(shape, dtype)
    _lib.check(_lib.load().pbf_memcpy(out.ctypes.data, C.c_void_p(dev_ptr), out.nbytes, 1, C.c_void_p(stream)))
    return out


def copy_to_device(dev_ptr: int, a: np.ndarray, stream: int):
    from pbf_amd import _lib
    a = np.ascontiguousarray(a)
    _lib.check(_lib.load().pbf_memcpy(C.c_void_p(dev_ptr), a.ctypes.data, a.nbytes, 0, C.c_void_p(stream)))
