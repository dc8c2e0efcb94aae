"""Seeded synthetic inputs for the benchmark configurations.

Generator: numpy PCG64 seeded with 1000 + config number.  The "smooth
random field" is 0.5 + (1/32) * sum of 16 sinusoids with random
frequencies u, v ~ U[1, 64] cycles per image and phases ~ U[0, 2pi), drawn
per channel, plus N(0, 0.02) noise, clipped to [0, 1]."""

import numpy as np


def rng(config):
    return np.random.Generator(np.random.PCG64(1000 + config))


def smooth_field(g, h, w, channels=1, dtype=np.float32):
    i = np.arange(h, dtype=np.float64)[:, None] / h
    j = np.arange(w, dtype=np.float64)[None, :] / w
    out = np.empty((channels, h, w), dtype=np.float64)
    for ch in range(channels):
        u = g.uniform(1, 64, 16)
        v = g.uniform(1, 64, 16)
        phi = g.uniform(0, 2 * np.pi, 16)
        acc = np.zeros((h, w))
        for k in range(16):
            acc += np.sin(2 * np.pi * (u[k] * i + v[k] * j) + phi[k])
        out[ch] = 0.5 + acc / 32.0
    out += g.normal(0, 0.02, out.shape)
    np.clip(out, 0.0, 1.0, out=out)
    return out.astype(dtype)


def harris_input(rows=2048, cols=2048, config=1):
    g = rng(config)
    return g.random((rows, cols), dtype=np.float64).astype(np.float32)


def unsharp_input(rows=2848, cols=4256, config=2):
    g = rng(config)
    return smooth_field(g, rows, cols, 3)


def bilateral_input(rows=2560, cols=1536, config=3):
    g = rng(config)
    return smooth_field(g, rows, cols, 1)[0]


def laplacian_input(rows=2560, cols=1536, config=4):
    g = rng(config)
    f = smooth_field(g, rows, cols, 3)
    return np.ascontiguousarray(np.round(f * 65535.0).astype(np.uint16).transpose(1, 2, 0))


def campipe_frames(n=48, config=5):
    g = rng(config)
    frames = []
    for _ in range(n):
        w = 32 * int(g.integers(16, 82))
        h = 32 * int(g.integers(12, 62))
        frames.append(g.integers(0, 1024, size=(h, w)).astype(np.uint16))
    return frames
