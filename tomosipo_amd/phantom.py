"""Test phantoms.

``hollow_box`` follows tomosipo ``phantom.py:4-25`` (box from 20 % to 80 %,
hollow from 40 % to 60 % of each axis). ``box_noise`` and ``ellipsoids`` are
the seeded phantoms of the benchmark configurations.
"""
import numpy as np


def hollow_box(vd):
    shape = np.array(vd.data.shape)
    s20, s40 = shape * 20 // 100, shape * 40 // 100
    vd.data[...] = 0.0
    vd.data[tuple(slice(a, n - a) for a, n in zip(s20, shape))] = 1.0
    vd.data[tuple(slice(a, n - a) for a, n in zip(s40, shape))] = 0.0
    return vd


def box_noise(shape, lo, hi, sigma=0.01, seed=0):
    """Ones on ``[lo:hi]`` in every axis plus seeded N(0, sigma) noise, float32."""
    rng = np.random.default_rng(seed)
    x = np.zeros(shape, dtype=np.float32)
    x[lo:hi, lo:hi, lo:hi] = 1.0
    x += rng.normal(0.0, sigma, size=shape).astype(np.float32)
    return x


def ellipsoids(n, count=20, seed=0):
    """Sum of ``count`` random ellipsoids in an ``n^3`` unit-cube volume, float32.

    Centres U(-0.4, 0.4)^3, semi-axes U(0.05, 0.3), random rotation, density
    U(0.1, 1).
    """
    rng = np.random.default_rng(seed)
    c = (np.arange(n) + 0.5) / n - 0.5
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    P = np.stack([Z, Y, X], axis=-1)
    out = np.zeros((n, n, n), dtype=np.float32)
    for _ in range(count):
        centre = rng.uniform(-0.4, 0.4, 3)
        axes = rng.uniform(0.05, 0.3, 3)
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        dens = rng.uniform(0.1, 1.0)
        local = (P - centre) @ q
        inside = np.sum((local / axes) ** 2, axis=-1) <= 1.0
        out[inside] += dens
    return out
