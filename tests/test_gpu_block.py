"""Block GMRES (ebh_solver_set_block): batched solves in one shared block
Krylov space.  Checked against the lockstep batched solver (independent
Krylov spaces) and against residuals recomputed with a separate FP64
operator; graph replays equal eager runs bit for bit; dependent right-hand
sides (a zero pivot in the block QR) still converge; config 4 (icosphere L6,
64 right-hand sides, mixed precision) at full size."""
import numpy as np
import pytest

from easybem_amd import meshes

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _recomputed(op, X, B):
    return max(rel(op(X[r]), B[r]) for r in range(B.shape[0]))


def test_block_fp64_matches_lockstep_solves():
    from easybem_amd import _lib as L
    P, T = meshes.icosphere(3)
    m = L.Mesh(P, T)
    V = L.Operator(m, "V", 0, 0, "fp64", max_rhs=8)
    one = L.Operator(m, "V", 0, 0, "fp64")
    B = np.random.default_rng(31).uniform(-1, 1, (8, T.shape[0]))
    lock = L.Solver(V, None, batch=8)
    X0, st0 = lock.solve(B.ravel(), tol=1e-10)
    blk = L.Solver(V, None, batch=8)
    blk.set_block(True)
    X1, st1 = blk.solve(B.ravel(), tol=1e-10)
    X0, X1 = X0.reshape(8, -1), X1.reshape(8, -1)
    print("lockstep iterations", max(s["iters"] for s in st0),
          "block steps", max(s["iters"] for s in st1))
    assert all(s["converged"] for s in st1)
    assert _recomputed(one, X1, B) <= 1e-10
    assert max(rel(X1[r], X0[r]) for r in range(8)) <= 1e-8
    assert max(s["iters"] for s in st1) <= max(s["iters"] for s in st0)


@pytest.mark.parametrize("nb", [5, 70])
def test_block_odd_and_wide_blocks(nb):
    """Block widths that are not a multiple of 16 (5) and wider than the
    projection kernel's 64 columns (70): the split projection V^T W covers
    partial and several column blocks."""
    from easybem_amd import _lib as L
    P, T = meshes.icosphere(2)
    m = L.Mesh(P, T)
    V = L.Operator(m, "V", 0, 0, "fp64", max_rhs=nb)
    one = L.Operator(m, "V", 0, 0, "fp64")
    B = np.random.default_rng(nb).uniform(-1, 1, (nb, T.shape[0]))
    blk = L.Solver(V, None, batch=nb)
    blk.set_block(True)
    print(f"case block width {nb}", flush=True)
    X1, st1 = blk.solve(B.ravel(), tol=1e-10)
    X1 = X1.reshape(nb, -1)
    print("block steps", max(s["iters"] for s in st1))
    assert all(s["converged"] for s in st1)
    assert _recomputed(one, X1, B) <= 1e-10


def test_block_graph_replays_equal_eager():
    from easybem_amd import _lib as L
    P, T = meshes.icosphere(3)
    m = L.Mesh(P, T)
    V = L.Operator(m, "V", 0, 0, "fp64", max_rhs=4)
    B = np.random.default_rng(32).uniform(-1, 1, (4, T.shape[0]))
    out = []
    for graphs in (False, True, True):
        s = L.Solver(V, None, batch=4, graphs=graphs)
        s.set_block(True)
        X, st = s.solve(B.ravel(), tol=1e-10)
        out.append(X)
        # a second solve on the same solver replays the captured steps
        X2, _ = s.solve(B.ravel(), tol=1e-10)
        assert np.array_equal(X, X2)
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[1], out[2])


def test_block_dependent_right_hand_sides():
    """Columns 1 and 3 equal columns 0 and 2: the block QR meets exactly
    dependent columns, whose pivots it sets to zero; every system still
    converges, the duplicates to the same solutions."""
    from easybem_amd import _lib as L
    P, T = meshes.icosphere(3)
    m = L.Mesh(P, T)
    V = L.Operator(m, "V", 0, 0, "fp64", max_rhs=4)
    one = L.Operator(m, "V", 0, 0, "fp64")
    rng = np.random.default_rng(33)
    a, b = rng.uniform(-1, 1, (2, T.shape[0]))
    B = np.stack([a, a, b, b])
    s = L.Solver(V, None, batch=4)
    s.set_block(True)
    X, st = s.solve(B.ravel(), tol=1e-10)
    X = X.reshape(4, -1)
    assert all(v["converged"] for v in st)
    assert _recomputed(one, X, B) <= 1e-10
    assert rel(X[1], X[0]) <= 1e-9 and rel(X[3], X[2]) <= 1e-9


def test_block_mixed_p1_with_calderon_preconditioner():
    from easybem_amd import _lib as L
    P, T = meshes.icosphere(3)
    m = L.Mesh(P, T)
    V64 = L.Operator(m, "V", 1, 1, "fp64", max_rhs=4)
    V32 = L.Operator(m, "V", 1, 1, "fp32", max_rhs=4)
    one = L.Operator(m, "V", 1, 1, "fp64")
    P64 = L.precond_calderon(m, "fp64", line_rule=2, tri_rule=0, max_rhs=4)
    B = np.random.default_rng(34).uniform(-1, 1, (4, P.shape[0]))
    s = L.Solver(V64, V32, batch=4)
    s.set_precond(P64, P64)
    s.set_block(True)
    X, st = s.solve(B.ravel(), tol=1e-8, mixed=True)
    X = X.reshape(4, -1)
    print("block steps", st[0]["iters"], "outer", st[0]["outer"])
    assert all(v["converged"] for v in st)
    assert _recomputed(one, X, B) <= 1.01e-8


def test_config4_block_mixed_solve():
    """Config 4: icosphere L6 (81920 triangles), P0 V, 64 right-hand sides,
    near-field sparse inverse on 2 triangle rings, mixed precision (64-RHS
    FP32 block applies inside, FP64 residuals) by block GMRES, against the
    lockstep batched solve; residuals of sampled systems recomputed with a
    separate single-RHS FP64 operator."""
    import time
    from easybem_amd import _lib as L
    P, T = meshes.icosphere(6)
    n = T.shape[0]
    m = L.Mesh(P, T)
    V64 = L.Operator(m, "V", 0, 0, "fp64", line_rule=4, tri_rule=4, max_rhs=64)
    V32 = L.Operator(m, "V", 0, 0, "fp32", line_rule=4, tri_rule=4, max_rhs=64)
    S = L.precond_spai(V64, rings=2, max_rhs=64)
    B = np.random.default_rng(20261203).uniform(-1, 1, (64, n))
    res = {}
    for mode in ("lockstep", "block"):
        s = L.Solver(V64, V32, restart=50, batch=64)
        s.set_precond(S, S)
        if mode == "block":
            s.set_block(True)
        print(f"case config 4 {mode} solve", flush=True)
        t0 = time.perf_counter()
        X, st = s.solve(B.ravel(), tol=1e-8, mixed=True, max_iters=3000)
        res[mode] = (X.reshape(64, n), st, time.perf_counter() - t0)
        s.close()
    X, st, dt = res["block"]
    _, st0, dt0 = res["lockstep"]
    print(f"lockstep: {max(v['iters'] for v in st0)} steps, {dt0:.2f} s; block: "
          f"{max(v['iters'] for v in st)} block steps, {dt:.2f} s")
    assert all(v["converged"] and v["relres"] <= 1e-8 for v in st)
    one = L.Operator(m, "V", 0, 0, "fp64", line_rule=4, tri_rule=4)
    for r in (0, 31, 63):
        assert rel(one(X[r]), B[r]) <= 1.01e-8
    assert max(v["iters"] for v in st) <= max(v["iters"] for v in st0)
