"""Host model of the FP32 LU's launch schedule (panels grouped in twos, fours or eights): every trailing
tile receives every earlier panel's contribution exactly once, before that tile is itself solved
as part of a panel, and the algorithmic FLOP total does not depend on the grouping."""
import itertools

import pytest

import bench


def schedule(nblk, pair_tiles, quad_tiles, oct_tiles=1 << 30):
    """(step, panels, tiles) per update launch, in the grouping rule of bem2d_internal_lu_advance."""
    out, pend, group = [], 0, 1
    for b in range(nblk - 1):
        tiles = nblk - b - 1
        if pend == 0:
            group = 1
            if b + 8 <= nblk - 1 and tiles >= oct_tiles:
                group = 8
            elif b + 4 <= nblk - 1 and tiles >= quad_tiles:
                group = 4
            elif b + 2 <= nblk - 1 and tiles >= pair_tiles:
                group = 2
        panels = range(b - pend, b + 1)
        trail = range(b + 1, nblk)
        if pend + 1 < group:
            cells = [(b + 1, j) for j in trail] + [(i, b + 1) for i in trail if i != b + 1]
            pend += 1
        else:
            cells = list(itertools.product(trail, trail))
            pend = 0
        out.append((b, panels, cells))
    return out


@pytest.mark.parametrize("nblk,pair,quad,oct_", [(9, 1, 1, 1 << 30), (9, 1, 1 << 30, 1 << 30),
                                                 (12, 3, 6, 1 << 30), (13, 5, 1, 1 << 30),
                                                 (7, 1 << 30, 1 << 30, 1 << 30),
                                                 (30, 10, 20, 1 << 30), (30, 3, 10, 20),
                                                 (17, 1, 1, 1), (40, 1 << 30, 1 << 30, 1)])
def test_each_contribution_once_and_in_time(nblk, pair, quad, oct_):
    seen = {}
    for step, panels, cells in schedule(nblk, pair, quad, oct_):
        for p in panels:
            for (i, j) in cells:
                assert (p, i, j) not in seen, (p, i, j)
                seen[(p, i, j)] = step
    for i in range(nblk):
        for j in range(nblk):
            for p in range(min(i, j)):
                # block (i, j) is final once min(i, j) - 1 has been stepped; step s factorises
                # panel s + 1's diagonal block, so the contribution must come by step min(i, j) - 1
                assert seen.get((p, i, j), 1 << 30) <= min(i, j) - 1, (p, i, j)
    assert len(seen) == sum(min(i, j) for i in range(nblk) for j in range(nblk))


def test_flop_total_independent_of_grouping():
    n = 32000
    nblk = (n + 63) // 64
    closed = sum(8 * 64 * (64 * (nblk - b - 1)) ** 2 for b in range(nblk - 1))
    for pair, quad, oct_ in [(1 << 30, 1 << 30, 1 << 30), (100, 1 << 30, 1 << 30),
                             (100, 200, 1 << 30), (100, 200, 250), (1, 1, 1)]:
        total, launches = bench.lu_update_flops(n, pair, quad, oct_)
        assert total == closed and launches == nblk - 1
