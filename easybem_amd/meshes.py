"""Deterministic surface-mesh generators for the benchmark and test meshes.

Every mesh is a closed, consistently oriented triangle surface with the
outward normal (P2-P1)x(P3-P1), the orientation the reference examples
assume (python/examples/laplace3dDirichlet.py:40-42).  Elements are returned
as int32 and 1-based, like the reference's ``elements`` arrays
(python/bindings.cpp:46-56).
"""
from __future__ import annotations

import math

import numpy as np


def icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    p = np.array(
        [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
         [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array(
        [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
         [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
         [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
         [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]],
        dtype=np.int64)
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    return p, f


def _subdivide(p, f):
    """Split every triangle into four through edge midpoints (1->4)."""
    pts = [q for q in p]
    cache = {}

    def mid(a, b):
        key = (a, b) if a < b else (b, a)
        k = cache.get(key)
        if k is None:
            k = len(pts)
            pts.append(0.5 * (p[a] + p[b]))
            cache[key] = k
        return k

    out = np.empty((4 * len(f), 3), dtype=np.int64)
    for n, (a, b, c) in enumerate(f):
        ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
        out[4 * n + 0] = (a, ab, ca)
        out[4 * n + 1] = (ab, b, bc)
        out[4 * n + 2] = (ca, bc, c)
        out[4 * n + 3] = (ab, bc, ca)
    return np.array(pts), out


def icosphere(level: int, radius: float = 1.0):
    """Icosahedron subdivided ``level`` times, vertices projected to the
    sphere after every step.  Level 3: 1280 triangles / 642 vertices,
    level 5: 20480 / 10242, level 6: 81920 / 40962."""
    p, f = icosahedron()
    for _ in range(level):
        p, f = _subdivide(p, f)
        p /= np.linalg.norm(p, axis=1, keepdims=True)
    return np.ascontiguousarray(radius * p), np.ascontiguousarray(
        (f + 1).astype(np.int32))


def torus(R: float = 1.0, r: float = 0.4, nu: int = 256, nv: int = 128):
    """Torus on an nu x nv parametric grid, each quad split into two
    triangles (2*nu*nv triangles, nu*nv vertices, all of valence 6)."""
    i = np.arange(nu)
    j = np.arange(nv)
    th = 2.0 * np.pi * i / nu
    ph = 2.0 * np.pi * j / nv
    T, P = np.meshgrid(th, ph, indexing="ij")
    x = (R + r * np.cos(P)) * np.cos(T)
    y = (R + r * np.cos(P)) * np.sin(T)
    z = r * np.sin(P)
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    idx = lambda a, b: (a % nu) * nv + (b % nv)  # noqa: E731
    I, J = np.meshgrid(i, j, indexing="ij")
    I = I.ravel()
    J = J.ravel()
    a = idx(I, J)
    b = idx(I + 1, J)
    c = idx(I + 1, J + 1)
    d = idx(I, J + 1)
    tri = np.empty((2 * len(a), 3), dtype=np.int64)
    tri[0::2] = np.stack([a, b, c], axis=1)
    tri[1::2] = np.stack([a, c, d], axis=1)
    # orient outward: compare the normal with the radial direction from the
    # tube centre line
    P1, P2, P3 = pts[tri[:, 0]], pts[tri[:, 1]], pts[tri[:, 2]]
    n = np.cross(P2 - P1, P3 - P1)
    cen = (P1 + P2 + P3) / 3.0
    ring = cen.copy()
    ring[:, 2] = 0.0
    ring *= (R / np.linalg.norm(ring, axis=1))[:, None]
    flip = np.einsum("ij,ij->i", n, cen - ring) < 0
    tri[flip] = tri[flip][:, [0, 2, 1]]
    return np.ascontiguousarray(pts), np.ascontiguousarray(
        (tri + 1).astype(np.int32))


def real_sph_harm(l: int, m: int, xyz):
    """Real spherical harmonic Y_lm (orthonormal) at unit vectors xyz."""
    from scipy.special import sph_harm_y
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    theta = np.arccos(np.clip(z, -1.0, 1.0))
    phi = np.arctan2(y, x)
    Y = sph_harm_y(l, abs(m), theta, phi)
    if m > 0:
        return math.sqrt(2.0) * (-1) ** m * Y.real
    if m < 0:
        return math.sqrt(2.0) * (-1) ** m * Y.imag
    return Y.real


def perturbed_icosphere(level: int, seed: int):
    """Serving-stream mesh: icosphere whose radius is 1 + 0.1 * sum over
    l <= 4 of real Y_lm with N(0,1) coefficients, then each axis scaled by a
    U(0.5, 1.5) factor; returns (nodes, elements, source_point) where the
    source point is an exterior point at twice the largest radius."""
    rng = np.random.default_rng(seed)
    p, f = icosphere(level)
    rad = np.ones(len(p))
    for l in range(5):
        for m in range(-l, l + 1):
            rad += 0.1 * rng.standard_normal() * real_sph_harm(l, m, p)
    rad = np.maximum(rad, 0.2)
    q = p * rad[:, None]
    q *= rng.uniform(0.5, 1.5, size=3)[None, :]
    d = rng.standard_normal(3)
    d /= np.linalg.norm(d)
    x0 = 2.0 * np.linalg.norm(q, axis=1).max() * d
    return np.ascontiguousarray(q), f, x0


def centroids(nodes, elems):
    T = elems - 1
    return nodes[T].mean(axis=1)


def areas(nodes, elems):
    T = elems - 1
    P = nodes[T]
    return 0.5 * np.linalg.norm(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]),
                                axis=1)
