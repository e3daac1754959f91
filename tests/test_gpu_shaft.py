"""The reference's shipped shaft geometry (Netgen surface mesh, refined
once and twice) as a real-world case: operator parity against the oracle
and Dirichlet / Neumann solves through the serving loop."""
import os

import numpy as np
import pytest

from easybem_amd import mesh_formats as mf

pytestmark = pytest.mark.gpu

MESHES = os.path.join(os.path.dirname(__file__), "golden", "meshes")


def shaft(nref):
    P, T, I = mf.load3dSurfaceFromNetgen(
        os.path.join(MESHES, "shaft-surface_verycoarse.vol"))
    for _ in range(nref):
        P, T, I = mf.refineMesh(P, T, I)
    return P, T


@pytest.fixture(scope="module")
def L():
    from easybem_amd import _lib
    return _lib


@pytest.mark.parametrize("kernel,test,trial", [("V", 0, 0), ("V", 1, 1), ("K", 0, 1),
                                               ("K", 1, 1)])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
def test_apply_matches_oracle_on_shaft(L, oracle, kernel, test, trial, prec, tol):
    P, T = shaft(1)
    print(f"case shaft1 {kernel}{test}{trial} {prec}", flush=True)
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, P.shape[0] if trial else T.shape[0])
    nrow = P.shape[0] if test else T.shape[0]
    rows = np.sort(rng.choice(nrow, 200, replace=False))
    m = L.Mesh(P, T)
    y = L.Operator(m, kernel, test, trial, prec, line_rule=4, tri_rule=4)(x)
    yr = oracle.apply(kernel, test, trial, P, T, x, rows=rows, line_idx=4, tri_idx=4)
    err = np.linalg.norm(y[rows] - yr) / np.linalg.norm(yr)
    print("rel err", err)
    assert err <= tol


@pytest.mark.parametrize("trial", [0, 1])
def test_fp32_adjoint_double_layer_on_shaft(L, trial):
    """The FP32 adjoint double layer (pair-symmetric, hi/lo centroid
    differences) on a surface with sharp edges: against the FP64 adjoint,
    itself checked against the reference's K^T elsewhere."""
    P, T = shaft(1)
    x = np.random.default_rng(3).uniform(-1, 1, P.shape[0] if trial else T.shape[0])
    m = L.Mesh(P, T)
    y64 = L.Operator(m, "Kt", 1, trial, "fp64", line_rule=4, tri_rule=4)(x)
    y32 = L.Operator(m, "Kt", 1, trial, "fp32", line_rule=4, tri_rule=4)(x)
    err = np.linalg.norm(y32 - y64) / np.linalg.norm(y64)
    print("Kt fp32 vs fp64 on shaft1", err)
    assert err <= 1e-6


def test_dirichlet_and_neumann_solves_on_shaft(L):
    """u = 1/|x - x0| with x0 just outside the shaft (6 units above its top
    face; far sources make u nearly constant, whose zero flux the P0
    discretisation resolves only to its quadrature error): both solves
    reach 1e-8, the Dirichlet flux approaches the exact one under
    refinement."""
    P0, _ = shaft(0)
    x0 = P0.mean(0) + np.array([0.0, 0.0, 0.6]) * np.ptp(P0, axis=0)
    srv = L.Server(max_slots=2)
    errs = []
    for nref in (1, 2):
        P, T = shaft(nref)
        rd = srv.submit(P, T, x0, problem="dirichlet")
        rn = srv.submit(P, T, x0, problem="neumann")
        res, t = srv.wait(rd)
        assert res["status"] == 0 and res["relres"] <= 1e-8, res
        res, u = srv.wait(rn)
        assert res["status"] == 0 and res["relres"] <= 1e-8, res
        A, B, C = (P[T[:, k] - 1] for k in range(3))
        n = np.cross(B - A, C - A)
        n /= np.linalg.norm(n, axis=1, keepdims=True)
        d = (A + B + C) / 3 - x0
        t_exact = -(d * n).sum(1) / np.linalg.norm(d, axis=1) ** 3
        errs.append(np.linalg.norm(t - t_exact) / np.linalg.norm(t_exact))
    srv.close()
    print("flux errors", errs)
    assert errs[1] < 0.8 * errs[0] and errs[1] < 0.1
