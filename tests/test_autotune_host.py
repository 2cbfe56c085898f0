"""Host logic of the autotuner: bucket construction from per-size winners."""
from bem2d_amd import autotune


def test_tiles():
    assert autotune.tiles(8) == 1 and autotune.tiles(64) == 36 and autotune.tiles(65) == 45


def test_buckets_merge_and_bounds():
    best = [(36, 4), (136, 4), (528, 2), (1176, 2), (2080, 1), (4656, 1)]
    b = autotune.buckets_from(best)
    assert b == [(int((136 * 528) ** 0.5), 4), (int((1176 * 2080) ** 0.5), 2)]
    # a one-wavefront size between others stays a bucket of its own
    b = autotune.buckets_from([(36, 4), (136, 1), (528, 2), (1176, 1)])
    assert [s for _, s in b] == [4, 1, 2]
    assert all(x[0] < y[0] for x, y in zip(b, b[1:]))
    # four wavefronts at every measured size: the last bucket reaches beyond them
    assert autotune.buckets_from([(36, 4), (136, 4)]) == [(1 << 62, 4)]
