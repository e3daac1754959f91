"""Surface mesh files and refinement, restating the reference's example
helpers (python/examples/surface3dMesh.py:47-221) with the same names,
arguments and results:

* ``load3dSurface(filename)`` -- the text format: dimension (3), vertices
  per element (3), the vertex count and coordinates, the triangle count and
  1-based triangles, then the count again and one ``index label`` line per
  triangle (``surface3dMesh.py:52-83``);
* ``load3dSurfaceFromNetgen(filename)`` -- the ``points`` and
  ``surfaceelementsgi`` sections of a Netgen ``.vol`` / ``.mesh3d`` file,
  whose element lines read ``surfnr bcnr domin domout np p1 p2 p3``
  (``surface3dMesh.py:86-139``);
* ``refineMesh(P, T, Tidx)`` -- uniform 1-to-4 refinement through edge
  midpoints (``surface3dMesh.py:142-221``).

All return ``(P, T, Tidx)`` with ``T`` 1-based int32, as the operators
expect.  One difference from the reference: ``refineMesh`` keeps each
child's label with its parent's (the reference stacks the children by kind
but repeats the labels per parent, so its labels are out of step with its
triangles).
"""
from __future__ import annotations

import numpy as np


def _nonempty(lines):
    for line in lines:
        s = line.strip()
        if s:
            yield s


def load3dSurface(filename):
    """(P, T, Tidx) from the reference's text format."""
    with open(filename, "r") as f:
        it = _nonempty(f)
        dim, per = int(next(it)), int(next(it))
        if dim != 3 or per != 3:
            raise ValueError("expected a 3-D triangular surface "
                             f"(got dimension {dim}, {per} vertices per element)")
        n = int(next(it))
        P = np.array([[float(v) for v in next(it).split()[:3]] for _ in range(n)])
        m = int(next(it))
        T = np.array([[int(v) for v in next(it).split()[:3]] for _ in range(m)],
                     dtype=np.int32)
        if int(next(it)) != m:
            raise ValueError("triangle count mismatch between the element and "
                             "label blocks")
        Tidx = np.array([int(next(it).split()[1]) for _ in range(m)], dtype=np.int32)
    return P, T, Tidx


def load3dSurfaceFromNetgen(filename):
    """(P, T, Tidx) from the ``points`` and ``surfaceelementsgi`` sections of
    a Netgen mesh file; Tidx is the surface number."""
    with open(filename, "r") as f:
        lines = [l.strip() for l in f]
    low = [l.lower() for l in lines]

    def section(name):
        try:
            return low.index(name)
        except ValueError:
            raise ValueError(f"{filename}: no '{name}' section") from None

    i = section("points")
    n = int(lines[i + 1])
    P = np.array([[float(v) for v in lines[i + 2 + k].split()[:3]] for k in range(n)])
    i = section("surfaceelementsgi")
    m = int(lines[i + 1])
    T = np.zeros((m, 3), dtype=np.int32)
    Tidx = np.zeros(m, dtype=np.int32)
    for k in range(m):
        f = lines[i + 2 + k].split()
        if int(f[4]) != 3:
            raise ValueError(f"{filename}: surface element {k + 1} has {f[4]} "
                             "vertices; only triangles are supported")
        T[k] = [int(v) for v in f[5:8]]
        Tidx[k] = int(f[0])
    return P, T, Tidx


def refineMesh(P, T, Tidx):
    """Uniform 1-to-4 refinement: every edge gets its midpoint, every
    triangle (a, b, c) becomes (a, ab, ca), (ab, b, bc), (ca, bc, c) and
    (ab, bc, ca), stacked by kind as in the reference.  Returns
    (P2, T2 1-based, Tidx2)."""
    P = np.asarray(P, dtype=np.float64)
    T0 = np.asarray(T, dtype=np.int64) - 1
    Tidx = np.asarray(Tidx)
    m, n = T0.shape[0], P.shape[0]
    a, b, c = T0[:, 0], T0[:, 1], T0[:, 2]
    e = np.concatenate([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)])
    e.sort(axis=1)
    uniq, inv = np.unique(e, axis=0, return_inverse=True)
    inv = inv.ravel() + n
    ab, bc, ca = inv[:m], inv[m:2 * m], inv[2 * m:]
    P2 = np.vstack([P, 0.5 * (P[uniq[:, 0]] + P[uniq[:, 1]])])
    T2 = np.vstack([np.stack([a, ab, ca], 1), np.stack([ab, b, bc], 1),
                    np.stack([ca, bc, c], 1), np.stack([ab, bc, ca], 1)])
    return P2, (T2 + 1).astype(np.int32), np.tile(Tidx, 4)
