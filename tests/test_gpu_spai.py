"""Near-field sparse approximate inverse of the single layer
(ebh_precond_spai): the device-built M against the same construction from
the dense FP64 matrix, and GMRES with M as right preconditioner."""
import numpy as np
import pytest

from easybem_amd import meshes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    from easybem_amd import _lib
    return _lib


def _dense(op, n):
    from easybem_amd import _lib as L
    cols = []
    for c in range(n):
        e = np.zeros(n)
        e[c] = 1.0
        cols.append(op(e))
    return np.array(cols).T


def _patches(P, T, p1, rings):
    """DOFs within `rings` adjacency steps: P1 nodes sharing a triangle,
    P0 triangles sharing a vertex."""
    n = P.shape[0] if p1 else T.shape[0]
    tris_of = [[] for _ in range(P.shape[0])]
    for i, tri in enumerate(T - 1):
        for a in tri:
            tris_of[a].append(i)
    if p1:
        adj = [{b for i in tris_of[a] for b in T[i] - 1} for a in range(n)]
    else:
        adj = [{j for a in T[i] - 1 for j in tris_of[a]} for i in range(n)]
    out = []
    for c in range(n):
        ball = {c}
        for _ in range(rings):
            ball |= {b for a in ball for b in adj[a]}
        out.append(sorted(ball))
    return out


@pytest.mark.parametrize("p1,rings", [(0, 1), (0, 2), (1, 2), (1, 3)])
def test_spai_matches_dense_construction(L, p1, rings):
    P, T = meshes.icosphere(2)
    m = L.Mesh(P, T)
    V = L.Operator(m, "V", p1, p1, "fp64")
    n = V.ncols
    A = _dense(V, n)
    M = np.zeros((n, n))
    for c, p in enumerate(_patches(P, T, p1, rings)):
        e = np.zeros(len(p))
        e[p.index(c)] = 1.0
        M[c, p] = np.linalg.solve(A[np.ix_(p, p)].T, e)
    S = L.precond_spai(V, rings=rings)
    x = np.random.default_rng(3).uniform(-1, 1, n)
    err = np.linalg.norm(S(x) - M @ x) / np.linalg.norm(M @ x)
    print("spai vs dense construction", err)
    assert err <= 1e-9


@pytest.mark.parametrize("p1", [0, 1])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_spai_preconditioned_solve(L, p1, prec):
    """Mixed-precision (or FP64) GMRES on V with M: residual <= 1e-8, the
    unpreconditioned solution, and far fewer iterations."""
    P, T = meshes.icosphere(4)
    m = L.Mesh(P, T)
    V64 = L.Operator(m, "V", p1, p1, "fp64")
    V32 = L.Operator(m, "V", p1, p1, "fp32")
    if p1:
        u = 1.0 / np.linalg.norm(P - np.array([0.0, 0.0, 2.0]), axis=1)
        K = L.Operator(m, "K", 1, 1, "fp64")
        b = 0.5 * m.mass(u, 1) + K(u)
    else:
        b = np.random.default_rng(5).uniform(-1, 1, T.shape[0])
    mixed = prec == "fp32"
    t0, st0 = L.Solver(V64, V32).solve(b, tol=1e-8, mixed=mixed)
    S = L.precond_spai(V64)
    s = L.Solver(V64, V32)
    s.set_precond(S, S)
    t1, st1 = s.solve(b, tol=1e-8, mixed=mixed)
    print("iterations", st0["iters"], "->", st1["iters"], st1)
    assert st1["converged"] and st1["relres"] <= 1e-8
    r = b - V64(t1)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-8
    assert np.linalg.norm(t1 - t0) / np.linalg.norm(t0) <= 1e-6
    assert st1["iters"] < 0.5 * st0["iters"]


def test_spai_batched_solve(L):
    P, T = meshes.icosphere(3)
    m = L.Mesh(P, T)
    V64 = L.Operator(m, "V", 0, 0, "fp64", max_rhs=4)
    V32 = L.Operator(m, "V", 0, 0, "fp32", max_rhs=4)
    S = L.precond_spai(V64, max_rhs=4)
    s = L.Solver(V64, V32, restart=50, batch=4)
    s.set_precond(S, S)
    B = np.random.default_rng(9).uniform(-1, 1, 4 * T.shape[0])
    db = L.DeviceArray.from_host(B)
    dx = L.DeviceArray.from_host(np.zeros_like(B))
    st = s.solve_device(db, dx, tol=1e-8, mixed=True, nb=4)
    X = dx.to_host().reshape(4, -1)
    for k in range(4):
        bk = B[k * T.shape[0]:(k + 1) * T.shape[0]]
        assert np.linalg.norm(bk - V64(X[k])) / np.linalg.norm(bk) <= 1e-8
    print([v["iters"] for v in st])


@pytest.mark.parametrize("kernel,test,trial", [("K", 1, 1), ("V", 0, 1), ("kelvin", 0, 0)])
def test_spai_rejects_other_operators(L, kernel, test, trial):
    P, T = meshes.icosphere(1)
    op = L.Operator(L.Mesh(P, T), kernel, test, trial, "fp64")
    with pytest.raises(L.EbhError) as e:
        L.precond_spai(op, rings=2)
    assert e.value.code == -4


def test_spai_rejects_bad_radius_and_far_field_calls(L):
    P, T = meshes.icosphere(1)
    V = L.Operator(L.Mesh(P, T), "V", 1, 1, "fp64")
    for bad in [0, 5]:
        with pytest.raises(L.EbhError):
            L.precond_spai(V, rings=bad)
    S = L.precond_spai(V, rings=2)
    with pytest.raises(L.EbhError):
        S.set_launch(256, 64, 4)
    with pytest.raises(L.EbhError):
        S.set_schedule("symmetric")


def test_spai_hypersingular_matches_dense_construction(L):
    """The sparse inverse of D + m m^T: entries from the curls and V0."""
    P, T = meshes.icosphere(2)
    m = L.Mesh(P, T)
    D = L.Operator(m, "D", 1, 1, "fp64", stab=1.0)
    n = D.ncols
    A = _dense(D, n)
    M = np.zeros((n, n))
    for c, p in enumerate(_patches(P, T, 1, 3)):
        e = np.zeros(len(p))
        e[p.index(c)] = 1.0
        M[c, p] = np.linalg.solve(A[np.ix_(p, p)].T, e)
    S = L.precond_spai(D, rings=3)
    x = np.random.default_rng(4).uniform(-1, 1, n)
    err = np.linalg.norm(S(x) - M @ x) / np.linalg.norm(M @ x)
    print("hypersingular spai vs dense construction", err)
    assert err <= 1e-9


def test_spai_hypersingular_solve(L):
    """Mixed-precision GMRES on D + m m^T with its sparse inverse: residual
    <= 1e-8, fewer iterations than without a preconditioner, more than with
    the dual Calderon preconditioner (D's small eigenvalues belong to smooth
    modes, which a local inverse cannot reach)."""
    P, T, _ = meshes.perturbed_icosphere(4, 11)
    m = L.Mesh(P, T)
    D64 = L.Operator(m, "D", 1, 1, "fp64", stab=1.0)
    D32 = L.Operator(m, "D", 1, 1, "fp32", stab=1.0)
    u = np.random.default_rng(2).uniform(-1, 1, P.shape[0])
    b = D64(u)
    s0 = L.Solver(D64, D32)
    s0.set_precond(L.precond_calderon_hs(m, "fp64"), L.precond_calderon_hs(m, "fp32"))
    _, st0 = s0.solve(b, tol=1e-8, mixed=True)
    S = L.precond_spai(D64)
    s1 = L.Solver(D64, D32)
    s1.set_precond(S, S)
    x, st1 = s1.solve(b, tol=1e-8, mixed=True)
    _, st2 = L.Solver(D64, D32).solve(b, tol=1e-8, mixed=True)
    print("iterations: dual Calderon", st0["iters"], "sparse inverse", st1["iters"],
          "none", st2["iters"])
    assert st1["converged"] and st1["relres"] <= 1e-8
    assert np.linalg.norm(b - D64(x)) / np.linalg.norm(b) <= 1e-8
    assert st0["iters"] < st1["iters"] < st2["iters"]
