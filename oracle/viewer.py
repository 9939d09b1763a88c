"""CPU oracle of the die map (test infrastructure only).

Restates in numpy the reference viewer's k_colorize (src/viewer.cu:68-87) and
its colormap (src/viewer.cu:53-66): for each column the hottest temperature
rise over a range of planes (starting from 0, as fmaxf from 0.0f), coloured
on [0, rise_hi] with five RGB stops, linear between them and truncated to
bytes, and the hottest cell as the largest 64-bit key (float bits << 32 |
cell index). Every step is done in float32 with the reference's operation
order, so the bytes are the reference's. Pinned in
tests/test_viewer_oracle.py by hand-worked palette values and a brute-force
argmax.
"""
import numpy as np

STOPS = np.array([[0, 0, 4], [87, 16, 110], [188, 55, 84], [249, 142, 9], [252, 255, 164]], np.float32)


def palette(s):
    """RGBA bytes (..., 4) of scale values s (viewer.cu:53-66)."""
    s = np.asarray(s, np.float32)
    s = np.where(s < 0, np.float32(0), np.where(s > 1, np.float32(1), s)).astype(np.float32)
    f = (s * np.float32(4)).astype(np.float32)
    k = np.minimum(f.astype(np.int32), 3)
    f = (f - k.astype(np.float32)).astype(np.float32)
    lo, hi = STOPS[k], STOPS[k + 1]
    rgb = (lo + f[..., None] * (hi - lo)).astype(np.float32)
    out = np.empty(s.shape + (4,), np.uint8)
    out[..., :3] = rgb.astype(np.uint8)  # truncation, as the float-to-byte cast
    out[..., 3] = 255
    return out


def colorize(T, nx, ny, nz, l0, l1, rise_hi):
    """(rgba (ny, nx, 4), hottest rise, hottest cell j*nx + i) of the rise T
    (natural layout, nz x ny x nx) over planes l0..l1 (viewer.cu:68-87)."""
    T3 = np.asarray(T, np.float64).reshape(nz, ny, nx)
    v = np.zeros((ny, nx), np.float32)
    for l in range(l0, l1 + 1):
        v = np.fmax(v, T3[l].astype(np.float32))
    rgba = palette((v / np.float32(rise_hi)).astype(np.float32))
    key = (v.view(np.uint32).astype(np.uint64) << np.uint64(32)) | np.arange(nx * ny, dtype=np.uint64).reshape(ny, nx)
    best = int(key.max())
    return rgba, float(np.uint32(best >> 32).view(np.float32)), best & 0xFFFFFFFF
