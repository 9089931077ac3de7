"""Deterministic sets of 4x4 test blocks, including the edge cases of the
reference encoders: single colours, two colours, duplicate texels, black
texels (three-colour mode), zero weights (partial edge blocks and
transparency), colours straddling the 1/256 merge threshold, out-of-range
floats, collinear ramps and non-unit colour weights."""
import numpy as np

KINDS = ("uniform", "u8", "cluster", "levels", "blacks", "single", "two", "partial",
         "alpha", "allblack", "threshold", "hdr", "ramp", "grey")


def bc1_block(rng, kind):
    """Returns (colors[16,4] float32, weights[16] float32)."""
    w = np.ones(16, np.float32)
    if kind == "uniform":
        c = rng.random((16, 4), dtype=np.float32)
    elif kind == "u8":
        c = rng.integers(0, 256, (16, 4)).astype(np.float32) / np.float32(255)
    elif kind == "cluster":
        c = np.clip(rng.random(4) + rng.normal(0, 0.02, (16, 4)), 0, 1).astype(np.float32)
    elif kind == "levels":
        c = (rng.integers(0, 4, (16, 4)) * 60).astype(np.float32) / np.float32(255)
    elif kind == "blacks":
        c = rng.integers(0, 256, (16, 4)).astype(np.float32) / np.float32(255)
        sel = rng.random(16) < 0.4
        c[sel, :3] = (rng.random((int(sel.sum()), 3)) * 0.12).astype(np.float32)
    elif kind == "single":
        c = np.repeat(rng.random((1, 4), dtype=np.float32), 16, axis=0)
    elif kind == "two":
        pal = rng.random((2, 4), dtype=np.float32)
        c = pal[rng.integers(0, 2, 16)]
    elif kind == "partial":
        c = rng.integers(0, 256, (16, 4)).astype(np.float32) / np.float32(255)
        bw, bh = int(rng.integers(1, 4)), int(rng.integers(1, 5))
        for y in range(4):
            for x in range(4):
                if x >= bw or y >= bh:
                    c[4 * y + x] = 0
                    w[4 * y + x] = 0
    elif kind == "alpha":
        c = rng.random((16, 4), dtype=np.float32)
        w = np.where(rng.random(16) < 0.3, 0.0, rng.random(16)).astype(np.float32)
    elif kind == "allblack":
        c = (rng.random((16, 4)) * 0.12).astype(np.float32)
    elif kind == "threshold":
        base = rng.random(4).astype(np.float32)
        d = np.float32(1.0 / 256)
        c = base + rng.choice(np.array([0, d * 0.999, d, d * 1.001, -d * 0.5], np.float32), (16, 4))
        c = c.astype(np.float32)
    elif kind == "hdr":
        c = (rng.random((16, 4)) * 1.6 - 0.3).astype(np.float32)
    elif kind == "ramp":
        a, b = rng.random(4), rng.random(4)
        t = np.sort(rng.random(16))[:, None]
        c = (a + (b - a) * t).astype(np.float32)
    elif kind == "grey":
        g = rng.integers(0, 256, 16).astype(np.float32) / np.float32(255)
        c = np.stack([g, g, g, np.ones(16, np.float32)], axis=1)
    else:
        raise ValueError(kind)
    return np.ascontiguousarray(c, np.float32), np.ascontiguousarray(w, np.float32)


def bc1_set(seed, count):
    rng = np.random.default_rng(seed)
    cols = np.zeros((count, 16, 4), np.float32)
    wts = np.zeros((count, 16), np.float32)
    kinds = np.zeros(count, np.int32)
    for i in range(count):
        k = i % len(KINDS)
        cols[i], wts[i] = bc1_block(rng, KINDS[k])
        kinds[i] = k
    return cols, wts, kinds


def bgra_set(seed, count):
    """8-bit BGRA blocks (nv::Color32 order) for the BC3/BC4 paths."""
    rng = np.random.default_rng(seed)
    out = np.zeros((count, 16, 4), np.uint8)
    for i in range(count):
        k = i % 6
        px = rng.integers(0, 256, (16, 4)).astype(np.uint8)
        if k == 1:
            px[:] = px[0]
            sel = rng.random(16) < 0.2
            px[sel, :3] = rng.integers(0, 256, 3)
        elif k == 2:
            px = (px // 64 * 64).astype(np.uint8)
        elif k == 3:
            px[:] = px[0]  # single colour
        elif k == 4:
            base = rng.integers(0, 256, 4)
            px = np.clip(base + rng.integers(-3, 4, (16, 4)), 0, 255).astype(np.uint8)
        elif k == 5:
            px[:, 3] = np.where(rng.random(16) < 0.5, 0, 255)
        out[i] = px
    return out


def channel_set(seed, count):
    """16-value channel blocks for the BC4 / BC5 / BC3-alpha paths: uniform
    noise, narrow clusters, two and three levels, 0/255 extremes with a middle
    value, ramps, near-constant blocks and exact constants (kind = index % 8)."""
    rng = np.random.default_rng(seed)
    n = count
    v = np.empty((n, 16), np.int32)
    kind = np.arange(n) % 8
    base = rng.integers(0, 256, (n, 1))
    v[kind == 0] = rng.integers(0, 256, (np.sum(kind == 0), 16))
    k = kind == 1
    v[k] = base[k] + rng.integers(-12, 13, (np.sum(k), 16))
    k = kind == 2
    lv = rng.integers(0, 256, (np.sum(k), 2))
    v[k] = np.take_along_axis(lv, rng.integers(0, 2, (np.sum(k), 16)), 1)
    k = kind == 3
    lv = rng.integers(0, 256, (np.sum(k), 3))
    v[k] = np.take_along_axis(lv, rng.integers(0, 3, (np.sum(k), 16)), 1)
    k = kind == 4
    lv = np.concatenate([np.zeros((np.sum(k), 1), int), np.full((np.sum(k), 1), 255),
                         rng.integers(0, 256, (np.sum(k), 1))], 1)
    v[k] = np.take_along_axis(lv, rng.integers(0, 3, (np.sum(k), 16)), 1)
    k = kind == 5
    a, b = rng.integers(0, 256, (np.sum(k), 1)), rng.integers(0, 256, (np.sum(k), 1))
    v[k] = a + ((b - a) * np.arange(16)[None] // 15)
    k = kind == 6
    v[k] = base[k] + rng.integers(0, 3, (np.sum(k), 16))
    k = kind == 7
    v[k] = base[k]
    return np.clip(v, 0, 255).astype(np.uint8)
