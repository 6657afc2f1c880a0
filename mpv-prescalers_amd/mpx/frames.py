"""Seeded synthetic test frames.

One plane is the clipped, quantised sum of
  * a linear gradient in a random direction, amplitude U(0, 0.3), around a mid-grey base;
  * 16 anti-aliased step edges, amplitude U(0.1, 0.6) with a random sign: a straight step
    at a random angle through a random point (a one-pixel linear ramp across the line),
    confined to a disc of radius U(0.15, 0.4)*min(W, H) around that point (also with a
    one-pixel ramp), so that the edges overlap locally rather than saturate the frame;
  * a zone plate 0.2*sin(pi*k*r^2), k = U(0.5, 2)/width, r = distance in pixels from a
    random centre inside the frame;
  * Gaussian noise, sigma = 0.01;
clipped to [0, 1] and quantised to ``bits`` (value = round(v*(2^bits-1))/(2^bits-1)).
Every draw comes from ``numpy.random.Generator(PCG64(seed))`` in the order listed, so a
(seed, size, bits) triple names one frame exactly.
"""
import numpy as np


def make_plane(width, height, seed, bits=8):
    rng = np.random.Generator(np.random.PCG64(seed))
    y, x = np.mgrid[0:height, 0:width].astype(np.float64)
    x += 0.5
    y += 0.5
    diag = np.hypot(width, height)
    # linear gradient
    th = rng.uniform(0.0, 2.0 * np.pi)
    amp = rng.uniform(0.0, 0.3)
    t = ((x - width / 2) * np.cos(th) + (y - height / 2) * np.sin(th)) / diag
    v = 0.5 + amp * t * 2.0
    # 16 anti-aliased step edges
    for _ in range(16):
        phi = rng.uniform(0.0, np.pi)
        px = rng.uniform(0.0, width)
        py = rng.uniform(0.0, height)
        a = rng.uniform(0.1, 0.6) * (1.0 if rng.uniform() < 0.5 else -1.0)
        rad = rng.uniform(0.15, 0.4) * min(width, height)
        dist = (x - px) * np.cos(phi) + (y - py) * np.sin(phi)
        disc = np.clip(rad - np.hypot(x - px, y - py) + 0.5, 0.0, 1.0)
        v += a * (np.clip(dist + 0.5, 0.0, 1.0) - 0.5) * disc
    # zone plate
    k = rng.uniform(0.5, 2.0) / width
    cx = rng.uniform(0.0, width)
    cy = rng.uniform(0.0, height)
    r2 = (x - cx) ** 2 + (y - cy) ** 2
    v += 0.2 * np.sin(np.pi * k * r2)
    # noise, clip, quantise
    v += rng.normal(0.0, 0.01, size=v.shape)
    v = np.clip(v, 0.0, 1.0)
    q = float((1 << bits) - 1)
    return np.round(v * q) / q


def make_frame(width, height, seed, bits=8, channels=1):
    """(H, W) for one channel, (3, H, W) with independent seeds seed, seed+1, seed+2."""
    if channels == 1:
        return make_plane(width, height, seed, bits)
    return np.stack([make_plane(width, height, seed + c, bits) for c in range(channels)])
