"""The strip layout of a one-mesh launch (which wave computes which rows),
queried through fvm2d_strip_layout on the host: every window's rows are
covered exactly once whatever the geometry, the round-filling cut and the
older/younger strip weights apply where they should.  No device needed."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def lay(fvm):
    """The library loaded on the host (the layout query uses no device)."""
    fvm.lib()
    return fvm


def _check_cover(L, nx, ny):
    nw = (nx - 4 + 59) // 60
    assert set(L[:, 0].tolist()) == set(range(nw))
    for w in range(nw):
        s = L[L[:, 0] == w]
        s = s[np.argsort(s[:, 1])]
        assert s[0, 1] == 2 and s[-1, 2] == ny - 2, (nx, ny, w)
        assert np.all(s[:, 2] > s[:, 1]), (nx, ny, w)           # no empty strip
        assert np.all(s[1:, 1] == s[:-1, 2]), (nx, ny, w)        # no gap, no overlap


CASES = [(nx, ny, R) for nx in (64, 133, 517, 1000, 2000, 2900) for ny in (37, 256, 999, 2000, 3100)
         for R in (1, 4, 16, 33, 34, 64)]
# extreme aspect ratios: one window of two interior columns, a few interior
# rows over many windows, strips taller than the mesh
CASES += [(nx, ny, R) for nx in (6, 7, 64, 2999) for ny in (6, 7, 3001) for R in (1, 3, 34, 96)]


@pytest.mark.parametrize("wpb,pers", [(4, 0), (2, 0), (4, 1), (1, 0)])
def test_rows_covered_once(lay, wpb, pers):
    blocks = {1: 8, 2: 4, 4: 2}[wpb]
    for nx, ny, R in CASES:
        L = lay.strip_layout(nx, ny, R, wpb, pers, 256, blocks)
        _check_cover(L, nx, ny)


def test_round_filling_cut_and_weights(lay):
    # 2000 x 2000, 34-row strips: 34 windows x 59 strips = 2006 tasks for
    # 2048 wave slots -> cut into exactly 2048 strips; the first 1024 (the
    # older wave of each SIMD) about twice as high as the other 1024
    L = lay.strip_layout(2000, 2000, 34, 4, 0, 256, 2)
    assert L.shape[0] == 2048
    h = L[:, 2] - L[:, 1]
    ratio = h[:1024].mean() / h[1024:].mean()
    assert 1.9 < ratio < 2.1, ratio
    # second round in mirrored order: level-major order puts 30 full levels
    # (1020 strips) and windows 0-3 of level 30 in the first round, so the
    # reversed second round starts at the last window and ends at window 4
    assert L[1024, 0] == 33 and L[-1, 0] == 1024 - 30 * 34
    # two waves per block: same cut, equal heights within each window
    L2 = lay.strip_layout(2000, 2000, 34, 2, 0, 256, 4)
    assert L2.shape[0] == 2048
    h2 = L2[:, 2] - L2[:, 1]
    for w in range(34):
        hw = h2[L2[:, 0] == w]
        assert hw.max() - hw.min() <= 1
    # persistent tickets: plain strips of R rows, no cut
    L3 = lay.strip_layout(2000, 2000, 34, 4, 1, 256, 2)
    assert L3.shape[0] == 34 * 59


def test_no_cut_outside_the_band(lay):
    # 2000 x 2000 at 16-row strips: 34 x 125 = 4250 tasks > 2 rounds -> plain
    L = lay.strip_layout(2000, 2000, 16, 4, 0, 256, 2)
    assert L.shape[0] == 34 * 125
    h = L[:, 2] - L[:, 1]
    assert h.max() == 16
    # a small mesh far below one round: plain strips
    L = lay.strip_layout(133, 97, 8, 4, 0, 256, 2)
    assert L.shape[0] == 3 * 12


def test_bad_query_rejected(lay):
    with pytest.raises(lay.FVM2DError):
        lay.strip_layout(2000, 2000, 0)
    with pytest.raises(lay.FVM2DError):
        lay.strip_layout(2000, 2000, 34, 5)
    with pytest.raises(lay.FVM2DError):
        lay.strip_layout(5, 64, 8)          # below the 6 x 6 minimum of a mesh
    assert lay.strip_layout(6, 6, 1).tolist() == [[0, 2, 3], [0, 3, 4]]
    with pytest.raises(lay.FVM2DError, match="too low"):
        lay.strip_layout(200000, 200000, 1)   # more tasks than int positions hold
