"""Elasticity double layer and hypersingular operators (EBH_KELVIN_DL /
EBH_KELVIN_HS) and the dense get_elasticity3d against the reference's
getElasticity3d (compiled from its sources, golden fixtures)."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
CASES = [("kelvin_ico1_o4", 4), ("kelvin_torus8x4_o3", 3)]


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def L():
    from easybem_amd import _lib
    return _lib


@pytest.mark.parametrize("name,order", CASES)
@pytest.mark.parametrize("kernel,gold,test,trial", [("kelvin_K", "K", 0, 1),
                                                    ("kelvin_D", "D", 1, 1),
                                                    ("kelvin", "V", 0, 0)])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
def test_apply_matches_reference(L, name, order, kernel, gold, test, trial, prec, tol):
    z = np.load(os.path.join(GOLD, f"{name}.npz"))
    P, T, nu = z["nodes"], z["elements"], float(z["nu"])
    print(f"case {name} {kernel} {prec}", flush=True)
    A = z[gold]
    x = np.random.default_rng(4).uniform(-1, 1, A.shape[1])
    m = L.Mesh(P, T)
    op = L.Operator(m, kernel, test, trial, prec, line_rule=order, tri_rule=order, nu=nu)
    err = rel(op(x), A @ x)
    print("rel err", err)
    assert err <= tol


@pytest.mark.parametrize("kernel,test,trial", [("kelvin_K", 0, 1), ("kelvin_D", 1, 1)])
def test_batched_equals_single(L, kernel, test, trial):
    from easybem_amd import meshes
    P, T = meshes.icosphere(2)
    m = L.Mesh(P, T)
    one = L.Operator(m, kernel, test, trial, "fp64", nu=0.3)
    two = L.Operator(m, kernel, test, trial, "fp64", nu=0.3, max_rhs=2)
    X = np.random.default_rng(5).uniform(-1, 1, (2, 3 * P.shape[0]))
    dx = L.DeviceArray.from_host(X.ravel())
    nrow = 3 * (P.shape[0] if test else T.shape[0])
    dy = L.DeviceArray(2 * nrow)
    two.apply(dx, dy, nrhs=2)
    Y = dy.to_host().reshape(2, nrow)
    for r in range(2):
        assert rel(Y[r], one(X[r])) <= 1e-13


@pytest.mark.parametrize("name,order", CASES)
def test_get_elasticity3d_matches_reference(name, order):
    sys.path.insert(0, os.path.join(ROOT, "easybem_amd"))
    import easybem
    z = np.load(os.path.join(GOLD, f"{name}.npz"))
    V, K, D, M = easybem.get_elasticity3d(z["nodes"], z["elements"].astype(np.int32),
                                          float(z["nu"]), order)
    for A, key in ((V, "V"), (K, "K"), (D, "D"), (M, "M")):
        err = np.abs(A - z[key]).max() / np.abs(z[key]).max()
        print(key, err)
        assert A.shape == z[key].shape and err <= 1e-12
