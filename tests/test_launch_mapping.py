"""Host-side checks of the neighbour passes' workgroup renumbering
(xcd_block in pbf_amd/csrc/pbf_kernels.hip): for every workgroup width the
library launches and for the grid sizes of the workloads, the renumbering is
a permutation of the workgroups, so every particle is processed exactly once."""
import pytest

CHUNK_PARTICLES = 2048


def renumber(b, nb, width):
    chunk = CHUNK_PARTICLES // width
    group = 8 * chunk
    if b >= nb // group * group:
        return b
    x, k = b % 8, b // 8
    return (k // chunk) * group + x * chunk + (k % chunk)


@pytest.mark.parametrize("width", [64, 128, 256, 512])
@pytest.mark.parametrize("n", [512, 4096, 32768, 1 << 20, 3 * (1 << 20) + 512, 4 << 20, 4_280_320])
def test_renumbering_is_a_permutation(width, n):
    nb = (n + width - 1) // width
    seen = sorted(renumber(b, nb, width) for b in range(nb))
    assert seen == list(range(nb))


def test_chunks_alternate_over_xcds():
    # workgroup b runs on XCD b % 8; consecutive chunks of sorted particles go
    # to consecutive XCDs
    width, nb = 64, 1 << 14
    chunk = CHUNK_PARTICLES // width
    xcd_of_chunk = {}
    for b in range(nb):
        xcd_of_chunk.setdefault(renumber(b, nb, width) // chunk, set()).add(b % 8)
    assert all(len(s) == 1 for s in xcd_of_chunk.values())
    assert [next(iter(xcd_of_chunk[c])) for c in range(16)] == [c % 8 for c in range(16)]
