"""Hypersingular operator D = sum_c C_c^T V C_c (+ stab m m^T) on the GPU
(EBH_LAPLACE_HS through the C-ABI) against the reference's dense D and the
oracle.  Bars: FP64 <= 1e-12 relative (2-norm), FP32 far field <= 1e-6."""
import os

import numpy as np
import pytest

from easybem_amd import meshes

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def L():
    from easybem_amd import _lib
    return _lib


@pytest.mark.parametrize("mesh", ["ico1", "torus8x4"])
@pytest.mark.parametrize("order", [3, 4])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_matches_reference_dense_d(L, mesh, order, prec):
    z = np.load(os.path.join(GOLD, f"laplace_{mesh}_o{order}.npz"))
    P, T, D = z["nodes"], z["elements"], z["D"]
    print(f"case {mesh} order={order} {prec}", flush=True)
    x = np.random.default_rng(12).uniform(-1, 1, P.shape[0])
    m = L.Mesh(P, T)
    op = L.Operator(m, "D", 1, 1, prec, line_rule=order, tri_rule=order)
    err = rel(op(x), D @ x)
    print("rel err", err)
    assert err <= (1e-12 if prec == "fp64" else 1e-6)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_matches_oracle_on_perturbed_sphere(L, oracle, prec):
    P, T = meshes.perturbed_icosphere(2, 7)[:2]
    x = np.random.default_rng(2).uniform(-1, 1, P.shape[0])
    m = L.Mesh(P, T)
    op = L.Operator(m, "D", 1, 1, prec, line_rule=4, tri_rule=4)
    err = rel(op(x), oracle.hypersingular_apply(P, T, x, line_idx=4, tri_idx=4))
    print("rel err", err)
    assert err <= (1e-12 if prec == "fp64" else 1e-6)


def test_stabilisation_adds_rank_one_term(L):
    from oracle import lumped_areas
    P, T = meshes.icosphere(2)
    x = np.random.default_rng(3).uniform(-1, 1, P.shape[0])
    m = L.Mesh(P, T)
    d0 = L.Operator(m, "D", 1, 1, "fp64")
    d1 = L.Operator(m, "D", 1, 1, "fp64", stab=0.7)
    mv = lumped_areas(P, T)
    want = d0(x) + 0.7 * mv * (mv @ x)
    assert rel(d1(x), want) <= 1e-14
    # constants lie in the kernel of D and are carried by the stabilisation
    one = np.ones(P.shape[0])
    assert np.linalg.norm(d0(one)) <= 1e-12 * np.linalg.norm(d0(x))
    assert rel(d1(one), 0.7 * mv * mv.sum()) <= 1e-12


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batched_matches_single(L, prec):
    """Four right-hand sides at once run the wrapped single layer on twelve
    curl densities through the multi-RHS kernel."""
    P, T = meshes.icosphere(3)
    X = np.random.default_rng(5).uniform(-1, 1, (4, P.shape[0]))
    m = L.Mesh(P, T)
    one = L.Operator(m, "D", 1, 1, prec, stab=1.0)
    many = L.Operator(m, "D", 1, 1, prec, max_rhs=4, stab=1.0)
    dx = L.DeviceArray.from_host(X.ravel())
    dy = L.DeviceArray(X.size)
    many.apply(dx, dy, nrhs=4)
    Y = dy.to_host().reshape(4, -1)
    for r in range(4):
        assert rel(Y[r], one(X[r])) <= (1e-13 if prec == "fp64" else 1e-6)


def test_rejects_p0_and_nonzero_stab_elsewhere(L):
    P, T = meshes.icosphere(1)
    m = L.Mesh(P, T)
    with pytest.raises(L.EbhError):
        L.Operator(m, "D", 0, 1, "fp64")
    with pytest.raises(L.EbhError):
        L.Operator(m, "V", 0, 0, "fp64", stab=1.0)


# ------------------------------------------------ adjoint double layer
@pytest.mark.parametrize("mesh", ["ico1", "torus8x4"])
@pytest.mark.parametrize("order", [3, 4])
def test_adjoint_double_layer_is_reference_k_transposed(L, mesh, order):
    """P1-test, P0-trial adjoint double layer against K^T of the
    reference's dense K (getLaplace3d)."""
    z = np.load(os.path.join(GOLD, f"laplace_{mesh}_o{order}.npz"))
    P, T, K = z["nodes"], z["elements"], z["K"]
    t = np.random.default_rng(13).uniform(-1, 1, T.shape[0])
    m = L.Mesh(P, T)
    op = L.Operator(m, "Kt", 1, 0, "fp64", line_rule=order, tri_rule=order)
    err = rel(op(t), K.T @ t)
    print("rel err", err)
    assert err <= 1e-12


def test_adjoint_double_layer_p1_trial(L, oracle):
    """P1 x P1: the transpose of the oracle's P1 x P1 double layer."""
    P, T = meshes.perturbed_icosphere(1, 2)[:2]
    n = P.shape[0]
    Kd = np.stack([oracle.apply("K", 1, 1, P, T, e, line_idx=4, tri_idx=4)
                   for e in np.eye(n)], axis=1)
    u = np.random.default_rng(6).uniform(-1, 1, n)
    m = L.Mesh(P, T)
    op = L.Operator(m, "Kt", 1, 1, "fp64", line_rule=4, tri_rule=4)
    err = rel(op(u), Kd.T @ u)
    print("rel err", err)
    assert err <= 1e-12


def test_adjoint_double_layer_limits(L):
    """A P0 test space is rejected; the FP32 adjoint runs on the
    pair-symmetric schedule only (rules of at most 7 points, one right-hand
    side, no chunked schedule), where its (y - x).n comes from hi/lo
    centroid differences."""
    P, T = meshes.icosphere(1)
    m = L.Mesh(P, T)
    with pytest.raises(L.EbhError):
        L.Operator(m, "Kt", 0, 0, "fp64")
    with pytest.raises(L.EbhError, match="pair-symmetric"):
        L.Operator(m, "Kt", 1, 0, "fp32", line_rule=4, tri_rule=5)
    with pytest.raises(L.EbhError, match="pair-symmetric"):
        L.Operator(m, "Kt", 1, 0, "fp32", max_rhs=2)
    op = L.Operator(m, "Kt", 1, 0, "fp32")
    assert op.schedule == "symmetric"
    with pytest.raises(L.EbhError, match="pair-symmetric"):
        op.set_schedule("chunked")
    with pytest.raises(L.EbhError, match="pair-symmetric"):
        op.set_launch(256, 64, 64)
    assert op.schedule == "symmetric"


# ------------------------------------- dual Calderon preconditioner of D
def _p1_mass(P, T):
    n = P.shape[0]
    M = np.zeros((n, n))
    for tri in T - 1:
        p = P[tri]
        area = 0.5 * np.linalg.norm(np.cross(p[1] - p[0], p[2] - p[0]))
        for a in range(3):
            for c in range(3):
                M[tri[a], tri[c]] += area / 12 * (2 if a == c else 1)
    return M, M.sum(axis=1)


def _cheb_mass_inverse(M, ML, v, steps=3, lo=0.25, hi=1.0):
    """Chebyshev iteration for M z = v preconditioned by diag(ML) (Saad,
    alg. 12.1); the eigenvalues of diag(ML)^-1 M lie in [1/4, 1]."""
    th, de = (hi + lo) / 2, (hi - lo) / 2
    sig = th / de
    rho = 1 / sig
    r = v.copy()
    d = v / ML / th
    z = d.copy()
    for _ in range(steps - 1):
        r = r - M @ d
        rn = 1 / (2 * sig - rho)
        d = rn * rho * d + 2 * rn / de * (r / ML)
        z = z + d
        rho = rn
    return z


def test_calderon_hs_is_chebyshev_mass_sandwich(L):
    """P = Q V Q with Q five Chebyshev steps for the consistent P1 mass,
    against the same polynomial in numpy with the dense mass matrix."""
    P, T, _ = meshes.perturbed_icosphere(2, 5)
    m = L.Mesh(P, T)
    Pc = L.precond_calderon_hs(m, "fp64")
    V = L.Operator(m, "V", 1, 1, "fp64", line_rule=3, tri_rule=1)
    M, ML = _p1_mass(P, T)
    ev = np.linalg.eigvals(M / ML[:, None]).real
    assert ev.min() >= 0.25 - 1e-12 and ev.max() <= 1 + 1e-12
    x = np.random.default_rng(15).uniform(-1, 1, P.shape[0])
    y = Pc(x)
    yr = _cheb_mass_inverse(M, ML, V(_cheb_mass_inverse(M, ML, x, steps=5)), steps=5)
    err = rel(y, yr)
    print("Q V Q vs numpy", err)
    assert err <= 1e-13
    # Q is close to M^-1: five steps leave at most 1 / T_5(5/3) < 0.01 of
    # the error in the M norm (T_5 the Chebyshev polynomial, 5/3 the
    # interval's centre over its half-width)
    q = _cheb_mass_inverse(M, ML, x, steps=5)
    z = np.linalg.solve(M, x)
    assert rel(q, z) <= 0.05


def test_calderon_hs_neumann_iterations(L):
    """FP64 GMRES(50) on D + m m^T of an L4 serving mesh: the preconditioned
    solve converges in few iterations."""
    P, T, x0 = meshes.perturbed_icosphere(4, 3)
    m = L.Mesh(P, T)
    D64 = L.Operator(m, "D", 1, 1, "fp64", stab=1.0)
    u = 1.0 / np.linalg.norm(P - x0, axis=1)
    b = D64(u)
    s = L.Solver(D64)
    s.set_precond(L.precond_calderon_hs(m, "fp64"))
    x, st = s.solve(b, tol=1e-8)
    print("iterations", st["iters"])
    assert st["converged"] and st["iters"] <= 10
    assert rel(D64(x), b) <= 1e-8


def test_calderon_v_is_chebyshev_mass_sandwich(L):
    """The single layer's Calderon preconditioner P = Q (D + m m^T / (4
    |Gamma|)) Q, Q five Chebyshev steps for the consistent P1 mass, against
    numpy with the dense mass matrix; two right-hand sides at once."""
    P, T, _ = meshes.perturbed_icosphere(2, 6)
    m = L.Mesh(P, T)
    Pc = L.precond_calderon(m, "fp64", line_rule=2, tri_rule=0, max_rhs=2)
    M, ML = _p1_mass(P, T)
    area = sum(0.5 * np.linalg.norm(np.cross(P[t[1]] - P[t[0]], P[t[2]] - P[t[0]]))
               for t in T - 1)
    D = L.Operator(m, "D", 1, 1, "fp64", line_rule=2, tri_rule=0)
    rng = np.random.default_rng(16)
    X = rng.uniform(-1, 1, (2, P.shape[0]))
    Y = Pc(X.ravel()).reshape(2, -1)
    for x, y in zip(X, Y):
        q = _cheb_mass_inverse(M, ML, x, steps=5)
        yr = _cheb_mass_inverse(M, ML, D(q) + 0.25 / area * ML * (ML @ q), steps=5)
        err = rel(y, yr)
        print("Q (D + stab m m^T) Q vs numpy", err)
        assert err <= 1e-13
