"""The neighbour passes' workgroup renumbering (xcd_block, pbf_amd/csrc/
pbf_internal.h) as the device computes it, through pbf_selftest_xcd_map: for
every workgroup width the library launches and the workgroup counts of the
workloads, the renumbering is a permutation (every particle is processed
exactly once), and each chunk of 2048 consecutive sorted particles ran on one
XCD, consecutive chunks on different ones."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CHUNK_PARTICLES = 2048


def device_map(lib, width, nb):
    from pbf_amd import _lib
    out = np.zeros(2 * nb, np.uint32)
    _lib.check(lib.pbf_selftest_xcd_map(width, nb, out.ctypes.data))
    return out[0::2].astype(np.int64), out[1::2].astype(np.int64)


@pytest.mark.parametrize("width", [64, 128, 256, 512])
@pytest.mark.parametrize("n", [512, 4096, 32768, 1 << 20, 3 * (1 << 20) + 512, 4 << 20, 4_277_760])
def test_renumbering_is_a_permutation(gpu_lib, width, n):
    nb = (n + width - 1) // width
    mapped, _ = device_map(gpu_lib, width, nb)
    assert np.array_equal(np.sort(mapped), np.arange(nb))


@pytest.mark.parametrize("width", [64, 256])
def test_chunks_stay_on_one_xcd(gpu_lib, width):
    nb = (1 << 20) // width
    mapped, xcc = device_map(gpu_lib, width, nb)
    chunk = CHUNK_PARTICLES // width
    assert len(np.unique(xcc)) == 8, f"XCDs seen: {np.unique(xcc)}"
    # the XCD that ran each chunk's workgroups
    per_chunk = {}
    for m, x in zip(mapped, xcc):
        per_chunk.setdefault(int(m) // chunk, []).append(int(x))
    one = sum(len(set(v)) == 1 for v in per_chunk.values())
    print(f"[xcd map width {width}] {one} of {len(per_chunk)} chunks ran on one XCD", flush=True)
    assert one == len(per_chunk)
    first = [per_chunk[c][0] for c in range(16)]
    assert len(set(first[:8])) == 8, f"XCDs of the first 8 chunks: {first[:8]}"
