"""The die-map oracle (oracle/viewer.py) pinned by hand-worked values.

The palette is five RGB stops, linear between them, truncated to bytes
(viewer.cu:53-66); the hottest cell is the largest (float bits << 32 | cell)
key (viewer.cu:68-87). Values worked by hand below: s = 0.125 is halfway
between stops 0 and 1, s = 0.3 is a fifth of the way from stop 1 to stop 2.
"""
import numpy as np

import viewer as V


def test_palette_hand_values():
    got = V.palette(np.array([0.0, 0.125, 0.3, 0.5, 0.75, 1.0, 2.0, -1.0], np.float32))
    want = [(0, 0, 4), (43, 8, 57), (107, 23, 104), (188, 55, 84), (249, 142, 9), (252, 255, 164),
            (252, 255, 164), (0, 0, 4)]
    assert [tuple(int(c) for c in px[:3]) for px in got] == want
    assert (got[:, 3] == 255).all()


def test_hottest_cell_is_the_argmax():
    rng = np.random.default_rng(3)
    nx, ny, nz = 23, 17, 6
    T = rng.random(nz * ny * nx) * 40.0
    img, hot, cell = V.colorize(T, nx, ny, nz, 1, 3, 30.0)
    v = T.reshape(nz, ny, nx)[1:4].astype(np.float32).max(axis=0)
    assert cell == int(np.argmax(v)) and hot == float(v.max())
    assert img.shape == (ny, nx, 4)
    # colours saturate above rise_hi
    assert (img[v >= 30.0][:, :3] == (252, 255, 164)).all()


def test_ties_take_the_last_cell_and_negatives_read_zero():
    nx, ny, nz = 4, 3, 2
    T = np.full(nz * ny * nx, -5.0)
    T3 = T.reshape(nz, ny, nx)
    T3[0, 1, 2] = 7.0
    T3[1, 2, 0] = 7.0  # same value, later cell
    img, hot, cell = V.colorize(T, nx, ny, nz, 0, 1, 10.0)
    assert hot == 7.0 and cell == 2 * nx + 0
    assert tuple(img[0, 0, :3]) == (0, 0, 4)  # a column below 0 reads as 0
