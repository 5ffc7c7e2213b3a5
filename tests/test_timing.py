"""Warm-up chunking of the benchmark drivers (no GPU needed)."""
from tcstencil_amd.timing import warm_chunk


def test_warm_chunk_keeps_full_even_launches():
    # unfused: an even sweep count of at least the request
    assert [warm_chunk(s) for s in (0, 1, 2, 5, 8)] == [2, 2, 2, 6, 8]
    # fused: a multiple of 2 * fuse, so the run splits into full launches of
    # `fuse` sweeps with an even launch count
    assert [warm_chunk(s, 8) for s in (1, 8, 16, 17)] == [16, 16, 16, 32]
    assert warm_chunk(5, 3) == 6 and warm_chunk(7, 3) == 12
    assert warm_chunk(4, 0) == warm_chunk(4, 1) == 4
