"""Mesh files and refinement (easybem_amd.mesh_formats) on the reference's
shipped example meshes: counts, labels and geometric invariants (CPU)."""
import os

import numpy as np
import pytest

from easybem_amd import mesh_formats as mf

MESHES = os.path.join(os.path.dirname(__file__), "golden", "meshes")


def areas_volume(P, T):
    A, B, C = (P[T[:, k] - 1] for k in range(3))
    cr = np.cross(B - A, C - A)
    # divergence theorem: volume = sum over faces of (centroid . n) area / 3
    vol = np.sum(((A + B + C) / 3.0) * cr * 0.5) / 3.0
    return 0.5 * np.linalg.norm(cr, axis=1), vol


def euler(P, T):
    e = np.sort(np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]]), axis=1)
    return P.shape[0] - len(np.unique(e, axis=0)) + T.shape[0]


def test_text_format_cube():
    P, T, Tidx = mf.load3dSurface(os.path.join(MESHES, "cube01-surface.txt"))
    assert P.shape == (8, 3) and T.shape == (12, 3) and T.min() == 1
    assert Tidx.tolist() == [1, 1, 3, 3, 2, 2, 0, 0, 4, 4, 5, 5]
    a, vol = areas_volume(P, T)
    assert abs(a.sum() - 6.0) < 1e-14 and abs(vol - 1.0) < 1e-14
    assert euler(P, T) == 2


def test_netgen_format_shaft():
    P, T, Tidx = mf.load3dSurfaceFromNetgen(
        os.path.join(MESHES, "shaft-surface_verycoarse.vol"))
    assert P.shape == (273, 3) and T.shape == (542, 3)
    assert T.min() == 1 and T.max() == 273
    assert set(Tidx.tolist()) >= {1, 2, 3}
    a, vol = areas_volume(P, T)
    assert np.all(a > 0) and vol > 0  # outward orientation
    assert euler(P, T) == 2


@pytest.mark.parametrize("name", ["cube01-surface.txt", "shaft-surface_verycoarse.vol"])
def test_refinement_invariants(name):
    path = os.path.join(MESHES, name)
    load = mf.load3dSurface if name.endswith(".txt") else mf.load3dSurfaceFromNetgen
    P, T, Tidx = load(path)
    P2, T2, I2 = mf.refineMesh(P, T, Tidx)
    ne = len(np.unique(np.sort(np.concatenate([T[:, [0, 1]], T[:, [1, 2]],
                                               T[:, [2, 0]]]), axis=1), axis=0))
    assert T2.shape == (4 * T.shape[0], 3) and P2.shape == (P.shape[0] + ne, 3)
    assert np.array_equal(P2[:P.shape[0]], P)
    a, vol = areas_volume(P, T)
    a2, vol2 = areas_volume(P2, T2)
    assert abs(a2.sum() - a.sum()) < 1e-12 * a.sum()
    assert abs(vol2 - vol) < 1e-12 * abs(vol)
    assert euler(P2, T2) == euler(P, T)
    m = T.shape[0]
    for k in range(4):  # children stacked by kind, labels follow parents
        assert np.array_equal(I2[k * m:(k + 1) * m], Tidx)
        assert np.allclose(a2[k * m:(k + 1) * m], a / 4)
    # the first three children keep the parent's vertices a, b, c
    assert np.array_equal(T2[:m, 0], T[:, 0])
    assert np.array_equal(T2[m:2 * m, 1], T[:, 1])
    assert np.array_equal(T2[2 * m:3 * m, 2], T[:, 2])


def test_meshes_pass_the_operator_validation():
    from easybem_amd import _lib as L
    P, T, Tidx = mf.load3dSurfaceFromNetgen(
        os.path.join(MESHES, "shaft-surface_verycoarse.vol"))
    for _ in range(2):
        info = L.mesh_check(P, T)
        assert info["coincident"] == T.shape[0] and info["edge"] == 3 * T.shape[0]
        P, T, Tidx = mf.refineMesh(P, T, Tidx)
