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


@pytest.mark.parametrize("basis", [0, 1])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 1e-6)])
def test_multi_rhs_apply_matches_single(basis, prec, tol):
    from easybem_amd import _lib as L
    P, T = meshes.icosphere(3)
    m = L.Mesh(P, T)
    one = L.Operator(m, "V", basis, basis, prec)
    many = L.Operator(m, "V", basis, basis, prec, max_rhs=64)
    n = one.ncols
    rng = np.random.default_rng(3)
    for nrhs in (2, 3, 8, 20, 64):
        print(f"case basis={basis} {prec} nrhs={nrhs}", flush=True)
        X = rng.uniform(-1, 1, (nrhs, n))
        Y = many(X.ravel()).reshape(nrhs, n)
        for r in range(nrhs):
            ref = one(X[r])
            assert np.linalg.norm(Y[r] - ref) / np.linalg.norm(ref) <= tol


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 1e-6)])
@pytest.mark.parametrize("rules", [(4, 4), (4, 3), (2, 1)])
def test_multi_rhs_p0_rules(prec, tol, rules):
    """The P0 multi-RHS kernel sums each triangle pair's point pairs before
    the right-hand sides (FP32 with 6- and 7-point median rules: r^2 from
    triangle-pair terms); rule sizes 7, 6 and 3 against the single-RHS
    apply."""
    from easybem_amd import _lib as L
    P, T = meshes.icosphere(3)
    m = L.Mesh(P, T)
    kw = dict(line_rule=rules[0], tri_rule=rules[1])
    one = L.Operator(m, "V", 0, 0, prec, **kw)
    many = L.Operator(m, "V", 0, 0, prec, max_rhs=40, **kw)
    X = np.random.default_rng(4).uniform(-1, 1, (37, one.ncols))
    Y = many(X.ravel()).reshape(X.shape)
    for r in range(X.shape[0]):
        ref = one(X[r])
        assert np.linalg.norm(Y[r] - ref) / np.linalg.norm(ref) <= tol


def test_batched_solve_matches_individual_solves():
    from easybem_amd import _lib as L
    P, T = meshes.icosphere(3)
    m = L.Mesh(P, T)
    V = L.Operator(m, "V", 0, 0, "fp64", max_rhs=8)
    rng = np.random.default_rng(9)
    B = rng.uniform(-1, 1, (5, T.shape[0]))
    X, st = L.Solver(V, None, batch=8).solve(B.ravel(), tol=1e-9)
    X = X.reshape(5, -1)
    single = L.Solver(V, None, batch=1)
    for r in range(5):
        x, s1 = single.solve(B[r], tol=1e-9)
        assert st[r]["converged"] and s1["converged"]
        assert np.linalg.norm(X[r] - x) / np.linalg.norm(x) < 1e-7
        assert np.linalg.norm(B[r] - V(X[r])) / np.linalg.norm(B[r]) <= 1e-9


def test_graph_replays_match_eager_bitwise():
    """Each Arnoldi step is a captured hipGraph; replays (interleaved with
    eager work) must reproduce the eager solve exactly."""
    from easybem_amd import _lib as L
    P, T = meshes.icosphere(3)
    m = L.Mesh(P, T)
    V64 = L.Operator(m, "V", 0, 0, "fp64")
    V32 = L.Operator(m, "V", 0, 0, "fp32")
    b = np.random.default_rng(4).uniform(-1, 1, T.shape[0])
    eager = L.Solver(V64, V32, graphs=False)
    x0, s0 = eager.solve(b, tol=1e-8, mixed=True)
    graph = L.Solver(V64, V32, graphs=True)
    for rep in range(3):
        x1, s1 = graph.solve(b, tol=1e-8, mixed=True)
        assert s1["iters"] == s0["iters"] and s1["outer"] == s0["outer"]
        assert np.array_equal(x1, x0), rep
        V64(b)  # eager work between replays


def test_calderon_preconditioner_cuts_iterations_keeps_solution():
    """Right preconditioning by M_L^-1 (D + m m^T/(4|Gamma|)) M_L^-1 solves
    the P1 Dirichlet problem to the same FP64 residual in far fewer
    iterations, with the same solution (to the solve tolerance)."""
    from easybem_amd import _lib as L
    from easybem_amd import meshes
    P, T = meshes.icosphere(4)
    x0 = np.array([0.0, 0.0, 2.0])
    u = 1.0 / np.linalg.norm(P - x0, axis=1)
    m = L.Mesh(P, T)
    V64 = L.Operator(m, "V", 1, 1, "fp64")
    V32 = L.Operator(m, "V", 1, 1, "fp32")
    K = L.Operator(m, "K", 1, 1, "fp64")
    b = 0.5 * m.mass(u, 1) + K(u)
    t0, st0 = L.Solver(V64, V32).solve(b, tol=1e-8, mixed=True)
    s = L.Solver(V64, V32)
    s.set_precond(L.precond_calderon(m, "fp64", line_rule=2, tri_rule=0),
                  L.precond_calderon(m, "fp32", line_rule=2, tri_rule=0))
    for mixed in (True, False):
        t1, st1 = s.solve(b, tol=1e-8, mixed=mixed)
        print(mixed, st0["iters"], st1["iters"], st1["precond"], st1["relres"])
        assert st1["converged"] and st1["relres"] <= 1e-8
        assert st1["iters"] < 0.6 * st0["iters"] and st1["precond"] > 0
        assert np.linalg.norm(t1 - t0) <= 1e-6 * np.linalg.norm(t0)
    # the true residual of the returned solution, recomputed independently
    assert np.linalg.norm(V64(t1) - b) <= 1e-8 * np.linalg.norm(b)


def test_preconditioned_graph_replays_match_eager():
    from easybem_amd import _lib as L
    from easybem_amd import meshes
    P, T = meshes.icosphere(3)
    m = L.Mesh(P, T)
    V64 = L.Operator(m, "V", 1, 1, "fp64")
    b = np.random.default_rng(2).uniform(-1, 1, P.shape[0])
    pc = L.precond_calderon(m, "fp64", line_rule=2, tri_rule=0)
    out = []
    for graphs in (True, False):
        s = L.Solver(V64, graphs=graphs)
        s.set_precond(pc, None)
        for _ in range(2):
            x, st = s.solve(b, tol=1e-10, mixed=False)
            out.append(x)
    for x in out[1:]:
        assert np.array_equal(x, out[0])


def test_graphs_dropped_when_preconditioners_change():
    """Captured Arnoldi steps hold the addresses of the preconditioner's
    buffers and of the kept z_j (released when the preconditioners are
    removed): after removing the preconditioner and setting it again, and
    after switching to another one and back, the solves replay fresh
    captures and equal the first solve bit for bit."""
    from easybem_amd import _lib as L
    from easybem_amd import meshes
    P, T = meshes.icosphere(3)
    m = L.Mesh(P, T)
    V64 = L.Operator(m, "V", 1, 1, "fp64")
    b = np.random.default_rng(4).uniform(-1, 1, P.shape[0])
    pc = L.precond_calderon(m, "fp64", line_rule=2, tri_rule=0)
    pc2 = L.precond_calderon(m, "fp64", line_rule=3, tri_rule=1)
    s = L.Solver(V64, graphs=True)
    s.set_precond(pc, None)
    x0, st0 = s.solve(b, tol=1e-10, mixed=False)
    s.set_precond(None, None)
    x1, st1 = s.solve(b, tol=1e-10, mixed=False)
    assert st1["iters"] > st0["iters"]
    s.set_precond(pc, None)
    x2, st2 = s.solve(b, tol=1e-10, mixed=False)
    s.set_precond(pc2, None)
    s.solve(b, tol=1e-10, mixed=False)
    s.set_precond(pc, None)
    x3, st3 = s.solve(b, tol=1e-10, mixed=False)
    x4, st4 = s.solve(b, tol=1e-10, mixed=False)
    print(st0, st2, st3, st4)
    assert st4["graph_steps"] > 0
    for x in (x2, x3, x4):
        assert np.array_equal(x, x0)


def test_batched_calderon_solve_matches_individual_solves():
    """Four right-hand sides in one batched mixed solve with the Calderon
    preconditioner built for four (its Chebyshev mass inverse and the
    hypersingular apply take the batch at once) against one-at-a-time
    solves; graphs replayed twice."""
    from easybem_amd import _lib as L
    from easybem_amd import meshes
    P, T = meshes.icosphere(3)
    m = L.Mesh(P, T)
    n = P.shape[0]
    V64 = L.Operator(m, "V", 1, 1, "fp64", max_rhs=4)
    V32 = L.Operator(m, "V", 1, 1, "fp32", max_rhs=4)
    B = np.random.default_rng(21).uniform(-1, 1, (4, n))
    s = L.Solver(V64, V32, batch=4)
    s.set_precond(L.precond_calderon(m, "fp64", max_rhs=4),
                  L.precond_calderon(m, "fp32", max_rhs=4))
    for _ in range(2):
        X, st = s.solve(B.ravel(), tol=1e-8, mixed=True)
    X = X.reshape(4, n)
    one = L.Solver(V64, V32)
    one.set_precond(L.precond_calderon(m, "fp64"), L.precond_calderon(m, "fp32"))
    for r in range(4):
        x, st1 = one.solve(B[r], tol=1e-8, mixed=True)
        assert st[r]["converged"] and st[r]["relres"] <= 1e-8
        err = np.linalg.norm(X[r] - x) / np.linalg.norm(x)
        print(r, st[r]["iters"], st1["iters"], err)
        assert err <= 1e-6
        assert np.linalg.norm(V64(X[r]) - B[r]) <= 1e-8 * np.linalg.norm(B[r])
