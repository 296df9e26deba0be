"""Host-side logic of the learned coarse-to-fine path (no GPU): the cell
pooling ranges and the mode windows of the per-mode ridge agreement."""
import numpy as np

from ccfj import learn


def test_pooling_ranges_by_hand():
    # cells at 0, 10, 20 (half-width 5); samples at 1, 4, 6, 14, 16, 17, 30
    c_sub = np.array([1.0, 4.0, 6.0, 14.0, 16.0, 17.0, 30.0])
    cells = np.array([0.0, 10.0, 20.0])
    rg = learn.pooling_ranges(c_sub, cells)
    # [-5, 5): 1, 4 -> [0, 2);  [5, 15): 6, 14 -> [2, 4);  [15, 25): 16, 17 -> [4, 6)
    assert rg.tolist() == [[0, 2], [2, 4], [4, 6]]
    # a cell with no sample inside takes the two either side of its centre
    rg = learn.pooling_ranges(np.array([0.0, 1.0, 9.0, 30.0]), np.array([0.0, 20.0, 40.0]))
    assert rg[1].tolist() == [2, 4]          # 9 and 30 around the centre 20
    assert rg[0].tolist() == [0, 3]          # 0, 1, 9 inside [-10, 10)


def test_pooling_ranges_cover_the_slowness_grid():
    m_c = np.linspace(learn.C_MIN, learn.C_MAX, 64)
    dc = m_c[1] - m_c[0]
    s = np.linspace(1 / (learn.C_MIN - dc / 2), 1 / (learn.C_MAX + dc / 2), 512)
    c_sub = 1 / s
    rg = learn.pooling_ranges(c_sub, m_c)
    assert np.all(rg[:, 1] > rg[:, 0])
    # the ranges tile the samples: each sub-velocity in exactly one cell,
    # but for the last, which lies on the upper edge C_MAX + dc/2 of the
    # half-open cell intervals
    inside = rg[:, 1] - rg[:, 0]
    assert rg[0, 0] == 0 and rg[-1, 1] == len(c_sub) - 1
    assert inside.sum() == len(c_sub) - 1
    assert np.all(rg[1:, 0] == rg[:-1, 1])


def test_mode_windows_by_hand():
    f = np.array([0.1, 0.5, 1.0])
    c_dense = np.linspace(1000.0, 5000.0, 4001)          # 1 m/s steps
    # c(f) = 2000 + 1000 exp(-f / 0.5): 2818.73, 2367.88, 2135.34 m/s
    modes = np.array([[2000.0, 1000.0, 0.5, 1.0],
                      [6000.0, 100.0, 0.5, 0.5],           # off the grid: no window
                      [3000.0, 500.0, 0.2, 0.0]])          # amplitude 0: unused
    fi, lo, hi = learn.mode_windows(modes, f, c_dense, 10.0)
    assert fi.tolist() == [0, 1, 2]
    cm = 2000.0 + 1000.0 * np.exp(-f / 0.5)
    for k in range(3):
        inside = np.nonzero(np.abs(c_dense - cm[k]) <= 10.0)[0]
        assert (lo[k], hi[k]) == (inside[0], inside[-1] + 1)
    spec = np.zeros((4001, 3))
    spec[lo[1] + 3, 1] = -2.0
    spec[lo[1] + 5, 1] = 1.5
    picks = learn.Coarse2Fine.mode_picks(spec, (fi, lo, hi))
    assert picks[1] == lo[1] + 3                           # |F-J|: the sign does not matter
