"""Device GMRES and mixed-precision refinement (through the C-ABI)."""
import numpy as np
import pytest

from easybem_amd import meshes

pytestmark = pytest.mark.gpu


def dirichlet_rhs(L, m, P, T, x0, test_basis, prec="fp64"):
    """b = (1/2 M + K) u for u = 1/|x - x0| at the nodes."""
    u = 1.0 / np.linalg.norm(P - x0, axis=1)
    K = L.Operator(m, "K", test_basis, 1, prec)
    return 0.5 * m.mass(u, test_basis) + K(u), u


@pytest.mark.parametrize("basis", [0, 1])
def test_gmres_fp64_matches_dense_solve(oracle, basis):
    from easybem_amd import _lib as L
    P, T = meshes.icosphere(2)
    m = L.Mesh(P, T)
    V = L.Operator(m, "V", basis, basis, "fp64")
    b, _ = dirichlet_rhs(L, m, P, T, np.array([0.0, 0.0, 2.0]), basis)
    s = L.Solver(V, None, restart=50)
    t, st = s.solve(b, tol=1e-10)
    print(st)
    assert st["converged"] and st["relres"] <= 1e-10
    # dense matrix of the same operator (columns through the oracle)
    n = V.ncols
    A = np.column_stack([oracle.apply("V", basis, basis, P, T, e)
                         for e in np.eye(n)])
    t_ref = np.linalg.solve(A, b)
    assert np.linalg.norm(t - t_ref) / np.linalg.norm(t_ref) < 1e-8


@pytest.mark.parametrize("basis", [0, 1])
def test_mixed_refinement_reaches_fp64_residual(basis):
    from easybem_amd import _lib as L
    P, T = meshes.icosphere(4)
    m = L.Mesh(P, T)
    V64 = L.Operator(m, "V", basis, basis, "fp64")
    V32 = L.Operator(m, "V", basis, basis, "fp32")
    b, _ = dirichlet_rhs(L, m, P, T, np.array([0.0, 0.0, 2.0]), basis)
    t64, st64 = L.Solver(V64, None).solve(b, tol=1e-8)
    t32, st32 = L.Solver(V64, V32).solve(b, tol=1e-8, mixed=True)
    print(st64, st32)
    assert st64["converged"] and st32["converged"]
    assert st32["relres"] <= 1e-8
    # independent check of the residual of the mixed solution
    r = b - V64(t32)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-8
    assert np.linalg.norm(t32 - t64) / np.linalg.norm(t64) < 1e-6


def test_dirichlet_solution_converges_to_exact_flux():
    """Interior Dirichlet problem, u = 1/|x - x*| with x* outside: the P0
    Neumann trace converges to the exact normal derivative."""
    from easybem_amd import _lib as L
    xs = np.array([2.0, 2.0, 2.0])
    errs = []
    for lev in (2, 3, 4):
        P, T = meshes.icosphere(lev)
        m = L.Mesh(P, T)
        V = L.Operator(m, "V", 0, 0, "fp64")
        b, _ = dirichlet_rhs(L, m, P, T, xs, 0)
        t, st = L.Solver(V).solve(b, tol=1e-10)
        c = meshes.centroids(P, T)
        A = meshes.areas(P, T)
        Tr = T - 1
        Pt = P[Tr]
        N = np.cross(Pt[:, 1] - Pt[:, 0], Pt[:, 2] - Pt[:, 0])
        N /= np.linalg.norm(N, axis=1, keepdims=True)
        tex = -np.sum((c - xs) * N, axis=1) / np.linalg.norm(c - xs, axis=1) ** 3
        errs.append(np.sqrt(np.sum((t - tex) ** 2 * A) / np.sum(tex ** 2 * A)))
    print(errs)
    assert errs[1] < 0.6 * errs[0] and errs[2] < 0.6 * errs[1]
