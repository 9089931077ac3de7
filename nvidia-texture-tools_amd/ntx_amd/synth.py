"""Seeded synthetic textures for the parity tests and the benchmark.

The generator defines every fixture and benchmark input, so its steps are
fixed here:
  value noise  -- a lattice of uniform values every `period` texels (drawn
                  from numpy's PCG64 with the given seed), smoothstep-weighted
                  bilinear interpolation, sampled at texel centres;
  colour       -- per channel the sum of 4 octaves (periods 128, 64, 32, 16,
                  amplitudes 1, 1/2, 1/4, 1/8) divided by 1.875, mapped to
                  [0,255] by round-half-up;
  rectangles   -- 32 axis-aligned rectangles, top-left uniform over the image,
                  extents uniform in [max(1, n/32), max(1, n/4)], uniform RGB
                  colour, drawn in order with hard edges;
  alpha        -- 255, or (with transparent=True) one noise octave of period
                  64 thresholded so that 30% of texels are 0 and the rest 255;
  normal maps  -- a 4-octave fBm height field, central differences with wrap,
                  n = normalize(-dh/dx * 8, -dh/dy * 8, 1), R = 0.5x+0.5,
                  G = 0.5y+0.5 (as 8-bit), B = 0, A = 255;
  single channel -- one 4-octave noise channel in R, G = B = 0, A = 255.
Images are (h, w, 4) uint8 RGBA; to_planar() turns them into the planar FP32
layout of nvtt::Surface (Surface::setImage with InputFormat_BGRA_8UB divides
every byte by 255).
"""
from __future__ import annotations

import numpy as np

PERIODS = (128, 64, 32, 16)
AMPS = (1.0, 0.5, 0.25, 0.125)


def value_noise(rng: np.random.Generator, h: int, w: int, period: int) -> np.ndarray:
    gh, gw = h // period + 2, w // period + 2
    grid = rng.random((gh, gw), dtype=np.float32)
    ys = (np.arange(h, dtype=np.float32) + 0.5) / period
    xs = (np.arange(w, dtype=np.float32) + 0.5) / period
    iy = np.floor(ys).astype(np.int64)
    ix = np.floor(xs).astype(np.int64)
    fy = ys - iy
    fx = xs - ix
    sy = (fy * fy * (3 - 2 * fy))[:, None]
    sx = (fx * fx * (3 - 2 * fx))[None, :]
    # x interpolation once per lattice row, then rows gathered and blended in
    # y: the same operations (float64, as the weights are) as interpolating
    # every texel directly, rounded to float32 at the end.
    rowx = grid[:, ix] * (1 - sx) + grid[:, ix + 1] * sx
    top = rowx[iy]
    bot = rowx[iy + 1]
    top *= 1 - sy
    bot *= sy
    top += bot
    return top.astype(np.float32)


def fbm(rng: np.random.Generator, h: int, w: int) -> np.ndarray:
    acc = np.zeros((h, w), np.float32)
    for p, a in zip(PERIODS, AMPS):
        acc += np.float32(a) * value_noise(rng, h, w, p)
    return acc / np.float32(sum(AMPS))


def _to_u8(v: np.ndarray) -> np.ndarray:
    return np.floor(np.clip(v, 0.0, 1.0) * 255.0 + 0.5).astype(np.uint8)


def noise_rect_rgba(w: int, h: int, seed: int = 1, transparent: bool = False, rects: int = 32) -> np.ndarray:
    """Colour texture: 4-octave noise per channel plus hard-edged rectangles."""
    rng = np.random.default_rng(seed)
    img = np.empty((h, w, 4), np.uint8)
    for c in range(3):
        img[:, :, c] = _to_u8(fbm(rng, h, w))
    img[:, :, 3] = 255
    for _ in range(rects):
        x0 = int(rng.integers(0, w))
        y0 = int(rng.integers(0, h))
        rw = int(rng.integers(max(1, w // 32), max(1, w // 4) + 1))
        rh = int(rng.integers(max(1, h // 32), max(1, h // 4) + 1))
        col = rng.integers(0, 256, 3).astype(np.uint8)
        img[y0:y0 + rh, x0:x0 + rw, :3] = col
    if transparent:
        a = value_noise(rng, h, w, 64)
        thr = np.quantile(a, 0.30)
        img[:, :, 3] = np.where(a < thr, 0, 255).astype(np.uint8)
    return img


def normal_map_rgba(w: int, h: int, seed: int = 1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    hf = fbm(rng, h, w)
    dx = (np.roll(hf, -1, axis=1) - np.roll(hf, 1, axis=1)) * np.float32(8.0)
    dy = (np.roll(hf, -1, axis=0) - np.roll(hf, 1, axis=0)) * np.float32(8.0)
    n = np.stack([-dx, -dy, np.ones_like(hf)], axis=-1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    img = np.zeros((h, w, 4), np.uint8)
    img[:, :, 0] = _to_u8(0.5 * n[:, :, 0] + 0.5)
    img[:, :, 1] = _to_u8(0.5 * n[:, :, 1] + 0.5)
    img[:, :, 3] = 255
    return img


def single_channel_rgba(w: int, h: int, seed: int = 1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    img = np.zeros((h, w, 4), np.uint8)
    img[:, :, 0] = _to_u8(fbm(rng, h, w))
    img[:, :, 3] = 255
    return img


def to_bgra(img: np.ndarray) -> np.ndarray:
    """(h, w, 4) RGBA -> (h, w, 4) B,G,R,A bytes (nvtt's InputFormat_BGRA_8UB)."""
    return np.ascontiguousarray(img[:, :, [2, 1, 0, 3]])


def to_planar(img: np.ndarray) -> np.ndarray:
    """(h, w, 4) RGBA uint8 -> (4, h, w) float32 planes, value / 255 (Surface.cpp:745-761)."""
    return np.ascontiguousarray(np.moveaxis(img, -1, 0).astype(np.float32) / np.float32(255.0))
