"""The benchmark configurations' own workloads, at full size, on the HIP path.

Config 1 (icosphere L3, P0 V, 7-point far / 4-point Gauss near): the HIP
apply against the golden vectors the compiled reference wrote
(tests/golden/hybrid_ico3.npz: y = V_hyb x and V_hyb t for t = Y_2^0 at the
centroids), and the Rayleigh quotient t'Vt / t'Mt of Y_2^0 approaching the
sphere eigenvalue 1/5 from levels 2 to 4.

Config 2 (icosphere L5 P1 Dirichlet, the headline): the mixed solve with
the benchmark's settings (Calderon preconditioner, first FP32 pass to the
target), its residual recomputed with a separate FP64 operator, its
solution against the plain FP64 solve's, and a second refinement pass when
the first pass is cut short.

Config 4 (icosphere L6, 81920 triangles, P0 V, 64 right-hand sides in one
launch): the block apply against the single-RHS apply column by column, the
FP64 block apply against oracle rows, and the batched 64-RHS mixed solve to
1e-8 with residuals recomputed by a separate FP64 operator.

Config 5 (serving): one L5 Dirichlet, one L5 Neumann and one L6 Dirichlet
request through the server, each residual recomputed here with FP64
operators built outside the server.
"""
import os

import numpy as np
import pytest

from easybem_amd import meshes

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


# ---------------------------------------------------------------- config 2
@pytest.fixture(scope="module")
def headline():
    from easybem_amd import _lib as L
    P, T = meshes.icosphere(5)
    u = 1.0 / np.linalg.norm(P - np.array([0.0, 0.0, 2.0]), axis=1)
    m = L.Mesh(P, T)
    V64 = L.Operator(m, "V", 1, 1, "fp64")
    V32 = L.Operator(m, "V", 1, 1, "fp32")
    b = 0.5 * m.mass(u, 1) + L.Operator(m, "K", 1, 1, "fp64")(u)
    s = L.Solver(V64, V32, restart=50)
    s.set_precond(L.precond_calderon(m, "fp64", line_rule=2, tri_rule=0),
                  L.precond_calderon(m, "fp32", line_rule=2, tri_rule=0))
    return L, m, b, s


def test_config2_mixed_solve_certified_by_fp64_residual(headline):
    """First FP32 pass to the target (the benchmark's setting): one FP64
    residual; the residual recomputed with an FP64 operator built here is
    at most 1e-8, and the solution equals the FP64 solve's to the accuracy
    both residuals allow."""
    L, m, b, s = headline
    x, st = s.solve(b, tol=1e-8, mixed=True, inner_tol=1e-8)
    print("mixed", st)
    assert st["converged"] and st["relres"] <= 1e-8
    V = L.Operator(m, "V", 1, 1, "fp64", line_rule=4, tri_rule=4)
    r = rel(V(x), b)
    print("independent FP64 residual", r)
    assert r <= 1e-8
    x64, st64 = s.solve(b, tol=1e-8, mixed=False)
    print("fp64", st64, "solutions differ by", rel(x, x64))
    assert st64["converged"] and rel(x, x64) <= 1e-5


def test_config2_second_pass_when_the_first_is_cut_short(headline):
    """A first pass stopped at 1e-4 leaves an FP64 residual above the
    target; the refinement then runs more passes until the FP64 residual
    certifies the solve."""
    L, m, b, s = headline
    x, st = s.solve(b, tol=1e-8, mixed=True, inner_tol=1e-4)
    print(st)
    assert st["converged"] and st["relres"] <= 1e-8
    assert st["outer"] >= 2 and st["matvec64"] >= 2
    V = L.Operator(m, "V", 1, 1, "fp64", line_rule=4, tri_rule=4)
    assert rel(V(x), b) <= 1e-8


# ---------------------------------------------------------------- config 1
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
def test_config1_hip_apply_matches_reference_golden(prec, tol):
    from easybem_amd import _lib as L
    g = np.load(os.path.join(GOLD, "hybrid_ico3.npz"))
    P, T = g["nodes"], g["elements"]
    m = L.Mesh(P, T)
    op = L.Operator(m, "V", 0, 0, prec, line_rule=3, tri_rule=4)
    ey, et = rel(op(g["x"]), g["y"]), rel(op(g["t"]), g["Vt"])
    print(f"config 1 {prec}: V x {ey:.2e}, V Y20 {et:.2e}")
    assert ey <= tol and et <= tol


def test_config1_rayleigh_quotient_of_y20_converges():
    """t'Vt / t'Mt -> 1/5 for t = Y_2^0 (3 z^2 - 1 at the centroid
    directions); the error is the discretisation error and shrinks with h."""
    from easybem_amd import _lib as L
    errs = []
    for lev in (2, 3, 4):
        P, T = meshes.icosphere(lev)
        m = L.Mesh(P, T)
        c = meshes.centroids(P, T)
        c /= np.linalg.norm(c, axis=1, keepdims=True)
        t = 3 * c[:, 2] ** 2 - 1
        Vt = L.Operator(m, "V", 0, 0, "fp64", line_rule=3, tri_rule=4)(t)
        q = t @ Vt / np.sum(meshes.areas(P, T) * t * t)
        errs.append(abs(q - 0.2))
    print("Rayleigh quotient errors L2..L4", errs)
    assert errs[1] < 0.5 * errs[0] and errs[2] < 0.5 * errs[1]
    assert errs[1] < 5e-3


# ---------------------------------------------------------------- config 4
@pytest.fixture(scope="module")
def ico6():
    from easybem_amd import _lib as L
    P, T = meshes.icosphere(6)
    m = L.Mesh(P, T)
    return P, T, m


def test_config4_block_apply_matches_single_rhs(ico6):
    from easybem_amd import _lib as L
    P, T, m = ico6
    rng = np.random.default_rng(20261203)
    X = rng.uniform(-1, 1, (64, T.shape[0]))
    cols = [0, 17, 40, 63]
    for prec, tol in (("fp64", 1e-13), ("fp32", 1e-6)):
        many = L.Operator(m, "V", 0, 0, prec, line_rule=4, tri_rule=4, max_rhs=64)
        print(f"case config 4 {prec} 64-RHS block apply", flush=True)
        Y = many(X.ravel()).reshape(64, -1)
        del many
        one = L.Operator(m, "V", 0, 0, prec, line_rule=4, tri_rule=4)
        for c in cols:
            e = rel(Y[c], one(X[c]))
            print(f"  column {c}: {e:.2e}")
            assert e <= tol


def test_config4_fp64_block_apply_matches_oracle_rows(ico6, oracle):
    from easybem_amd import _lib as L
    P, T, m = ico6
    rng = np.random.default_rng(41)
    X = rng.uniform(-1, 1, (64, T.shape[0]))
    op = L.Operator(m, "V", 0, 0, "fp64", line_rule=4, tri_rule=4, max_rhs=64)
    Y = op(X.ravel()).reshape(64, -1)
    rows = np.sort(rng.choice(T.shape[0], size=24, replace=False)).astype(np.int32)
    for c in (5, 50):
        yo = np.asarray(oracle.apply("V", 0, 0, P, T, X[c], rows=rows, line_idx=4,
                                     tri_idx=4)).ravel()
        e = rel(Y[c][rows], yo)
        print(f"column {c}: FP64 block vs oracle rows {e:.2e}")
        assert e <= 1e-12


def test_config4_batched_mixed_solve(ico6):
    """64 systems V x_r = b_r in one batched mixed-precision solve (the
    64-RHS FP32 block apply inside, FP64 refinement residuals), near-field
    sparse inverse on 2 triangle rings; each sampled solution's residual is
    recomputed with a separate single-RHS FP64 operator."""
    from easybem_amd import _lib as L
    P, T, m = ico6
    n = T.shape[0]
    V64 = L.Operator(m, "V", 0, 0, "fp64", line_rule=4, tri_rule=4, max_rhs=64)
    V32 = L.Operator(m, "V", 0, 0, "fp32", line_rule=4, tri_rule=4, max_rhs=64)
    s = L.Solver(V64, V32, restart=50, batch=64)
    S = L.precond_spai(V64, rings=2, max_rhs=64)
    s.set_precond(S, S)
    B = np.random.default_rng(20261203).uniform(-1, 1, (64, n))
    print("case config 4 batched solve, 64 systems", flush=True)
    X, st = s.solve(B.ravel(), tol=1e-8, mixed=True, max_iters=3000)
    X = X.reshape(64, n)
    assert all(v["converged"] and v["relres"] <= 1e-8 for v in st)
    one = L.Operator(m, "V", 0, 0, "fp64", line_rule=4, tri_rule=4)
    for r in (0, 21, 42, 63):
        res = rel(one(X[r]), B[r])
        print(f"system {r}: reported {st[r]['relres']:.2e}, recomputed {res:.2e}")
        assert res <= 1.01e-8


# ---------------------------------------------------------------- config 5
def _dirichlet_residual(L, P, T, x0, t):
    m = L.Mesh(P, T)
    u = 1.0 / np.linalg.norm(P - x0, axis=1)
    b = 0.5 * m.mass(u, 0) + L.Operator(m, "K", 0, 1, "fp64")(u)
    return rel(L.Operator(m, "V", 0, 0, "fp64")(t), b)


def _neumann_residual(L, P, T, x0, u):
    from test_gpu_server import _neumann_data
    m = L.Mesh(P, T)
    t, Mt = _neumann_data(P, T, x0)
    b = 0.5 * Mt - L.Operator(m, "Kt", 1, 0, "fp64")(t)
    return rel(L.Operator(m, "D", 1, 1, "fp64", stab=1.0)(u), b)


def test_config5_served_requests_at_levels_5_and_6():
    from easybem_amd import _lib as L
    rng = np.random.default_rng(20261205)
    reqs = []
    for lev, problem in ((5, "dirichlet"), (5, "neumann"), (6, "dirichlet")):
        P, T, x0 = meshes.perturbed_icosphere(lev, int(rng.integers(1 << 30)))
        reqs.append((lev, problem, P, T, x0))
    srv = L.Server(max_slots=3)
    ids = [srv.submit(P, T, x0, problem=pb) for _, pb, P, T, x0 in reqs]
    for rid, (lev, pb, P, T, x0) in zip(ids, reqs):
        print(f"case config 5 L{lev} {pb}", flush=True)
        res, v = srv.wait(rid)
        assert res["status"] == 0, res
        r = (_dirichlet_residual if pb == "dirichlet" else _neumann_residual)(
            L, P, T, x0, v)
        print(f"L{lev} {pb}: reported {res['relres']:.2e}, recomputed {r:.2e}, "
              f"iters {res['iters']}, {res['total_ms']:.0f} ms, measured "
              f"{res['device_peak'] / 2**20:.0f} MiB of {res['footprint'] / 2**20:.0f} "
              "estimated")
        assert res["relres"] <= 1e-8 and r <= 1.01e-8
        assert 0 < res["device_peak"] <= res["footprint"]
    st = srv.stats()
    srv.close()
    assert st["completed"] == 3 and st["device_peak"] <= st["peak_bytes"]
