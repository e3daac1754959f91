"""Off-surface evaluation (ebh_evaluate_laplace) against the reference's
evaluateLaplace3d, the oracle and a harmonic function."""
import os

import numpy as np
import pytest

from easybem_amd import meshes

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def L():
    from easybem_amd import _lib
    return _lib


@pytest.mark.parametrize("order", [3, 4])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
def test_matches_reference(L, order, prec, tol):
    z = np.load(os.path.join(GOLD, "evaluate_ico2.npz"))
    m = L.Mesh(z["nodes"], z["elements"])
    y = L.evaluate(m, z["points"], z["t"], z["u"], tri_rule=order, precision=prec)
    want = z[f"result_o{order}"]
    err = np.abs(y - want).max() / np.abs(want).max()
    print(prec, "err", err)
    assert err <= tol


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
def test_many_points_match_oracle(L, oracle, prec, tol):
    """4000 points inside and outside a perturbed sphere (a chunked launch
    with many target blocks)."""
    P, T, _ = meshes.perturbed_icosphere(3, 21)
    rng = np.random.default_rng(3)
    d = rng.normal(size=(4000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pts = d * rng.choice([0.3, 0.6, 2.0, 3.0], size=(4000, 1))
    t = rng.uniform(-1, 1, T.shape[0])
    u = rng.uniform(-1, 1, P.shape[0])
    m = L.Mesh(P, T)
    y = L.evaluate(m, pts, t, u, tri_rule=4, precision=prec)
    yr = oracle.evaluate(P, T, pts, t, u, tri_idx=4)
    err = np.linalg.norm(y - yr) / np.linalg.norm(yr)
    print(prec, "err", err)
    assert err <= tol


def test_recovers_harmonic_function_from_solved_trace(L):
    """Solve the Dirichlet problem for u = 1/|x - x0| on the GPU, then
    evaluate the representation formula inside: the field matches u, with
    an error that shrinks under refinement."""
    x0 = np.array([1.2, -1.4, 0.9])
    pts = np.array([[0.0, 0.0, 0.0], [0.3, 0.2, -0.1], [-0.4, 0.1, 0.2]])
    exact = 1.0 / np.linalg.norm(pts - x0, axis=1)
    errs = []
    for lev in (3, 4):
        P, T = meshes.icosphere(lev)
        m = L.Mesh(P, T)
        u = 1.0 / np.linalg.norm(P - x0, axis=1)
        V64 = L.Operator(m, "V", 0, 0, "fp64")
        K = L.Operator(m, "K", 0, 1, "fp64")
        b = 0.5 * m.mass(u, 0) + K(u)
        t, st = L.Solver(V64).solve(b, tol=1e-10, mixed=False)
        y = L.evaluate(m, pts, t, u)
        errs.append(np.abs(y - exact).max() / exact.max())
    print("errors", errs)
    assert errs[1] < 0.5 * errs[0] and errs[1] < 5e-3
