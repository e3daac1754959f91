"""The fused reduction and near-field pass of the pair-symmetric schedule.

When requested (set_reduction(True); the separate passes are the default),
the symmetric far kernel stores each P1 test DOF's group partial sums at
the DOF's node-incidence positions, and one kernel (k_reduce_finish) sums
them and adds the near CSR row.  Its summation order
is that of the separate reduction (k_reduce_wide) and gather (k_finish), so
the two paths must agree bitwise; checked for both bases and precisions,
several step splits, the FP64 double layer and its adjoint, and the
three-density launch inside the hypersingular operator.  A mesh with a node
of valence above the fused pass's LDS stage (1024 incidences) keeps the
separate passes, checked against the oracle.
"""
import numpy as np
import pytest

from easybem_amd import meshes

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def L():
    from easybem_amd import _lib
    return _lib


MESHES = {
    "ico0": lambda: meshes.icosphere(0),
    "torus12x6": lambda: meshes.torus(1.0, 0.4, 12, 6),
    "blob3": lambda: meshes.perturbed_icosphere(3, 11)[:2],
    "ico4": lambda: meshes.icosphere(4),
}


def bipyramid(n):
    """Two apexes over an n-gon: the apexes have valence n."""
    a = 2 * np.pi * np.arange(n) / n
    P = np.vstack([[0.0, 0.0, 1.0], [0.0, 0.0, -1.0],
                   np.c_[np.cos(a), np.sin(a), np.zeros(n)]])
    i = np.arange(n)
    e, f = 2 + i, 2 + (i + 1) % n
    T = np.vstack([np.c_[np.zeros(n, int), e, f], np.c_[np.ones(n, int), f, e]])
    return P, (T + 1).astype(np.int32)


def _both(op, x):
    assert not op.fused  # the default: separate passes
    op.set_reduction(True)
    assert op.fused
    y1 = op(x)
    op.set_reduction(False)
    assert not op.fused
    y0 = op(x)
    op.set_reduction(True)
    return y1, y0


@pytest.mark.parametrize("mesh", list(MESHES))
@pytest.mark.parametrize("basis", [0, 1])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("split", [1, 4])
def test_fused_pass_bitwise_equals_separate_passes(L, mesh, basis, prec, split):
    P, T = MESHES[mesh]()
    print(f"case {mesh} basis {basis} {prec} split {split}", flush=True)
    m = L.Mesh(P, T)
    op = L.Operator(m, "V", basis, basis, prec, line_rule=4, tri_rule=4)
    op.set_schedule("symmetric", split)
    x = np.random.default_rng(3).uniform(-1, 1, P.shape[0] if basis else T.shape[0])
    y1, y0 = _both(op, x)
    assert np.array_equal(y1, y0)
    op.set_schedule("chunked")
    assert not op.fused
    assert rel(y1, op(x)) <= (1e-14 if prec == "fp64" else 1e-6)


@pytest.mark.parametrize("kernel,test,trial", [("K", 0, 1), ("K", 1, 1), ("Kt", 1, 0)])
def test_fused_pass_double_layers(L, oracle, kernel, test, trial):
    P, T = meshes.perturbed_icosphere(3, 5)[:2]
    print(f"case {kernel} {test}{trial}", flush=True)
    m = L.Mesh(P, T)
    op = L.Operator(m, kernel, test, trial, "fp64", line_rule=4, tri_rule=4)
    op.set_schedule("symmetric")
    x = np.random.default_rng(4).uniform(-1, 1, P.shape[0] if trial else T.shape[0])
    y1, y0 = _both(op, x)
    assert np.array_equal(y1, y0)
    op.set_schedule("chunked")
    assert rel(y1, op(x)) <= 1e-13


def test_fused_pass_three_densities_in_hypersingular(L):
    """D = sum_c C_c^T V C_c applies V (P0, symmetric) to three curl
    densities per launch; its fused pass covers three right-hand-side planes."""
    P, T = meshes.icosphere(3)
    m = L.Mesh(P, T)
    D = L.Operator(m, "D", 1, 1, "fp64", line_rule=4, tri_rule=4, max_rhs=3)
    D.set_schedule("symmetric")
    u = np.random.default_rng(6).uniform(-1, 1, P.shape[0])
    y1, y0 = _both(D, u)
    assert np.array_equal(y1, y0)
    U = np.random.default_rng(7).uniform(-1, 1, (3, P.shape[0]))
    D.set_reduction(True)
    Y1 = D(U.ravel())
    D.set_reduction(False)
    Y0 = D(U.ravel())
    assert np.array_equal(Y1, Y0)


def test_high_valence_mesh_keeps_separate_passes(L, oracle):
    P, T = bipyramid(1030)
    m = L.Mesh(P, T)
    rows = np.array([0, 1, 2, 500, 1031], dtype=np.int32)
    x = np.random.default_rng(8).uniform(-1, 1, P.shape[0])
    for prec, tol in (("fp64", 1e-12), ("fp32", 1e-6)):
        op = L.Operator(m, "V", 1, 1, prec, line_rule=4, tri_rule=4)
        op.set_schedule("symmetric")
        op.set_reduction(True)
        assert not op.fused
        y = op(x)
        op.set_schedule("chunked")
        yc = op(x)
        if prec == "fp64":
            assert rel(y, yc) <= 1e-14
        yo = np.asarray(oracle.apply("V", 1, 1, P, T, x, rows=rows, line_idx=4,
                                     tri_idx=4)).ravel()
        e = rel(y[rows], yo)
        print(f"bipyramid {prec}: symmetric vs oracle rows {e:.2e}")
        assert e <= tol
