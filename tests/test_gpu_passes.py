"""The pair-symmetric schedule in passes (bounded partial sums).

In one pass the symmetric schedule keeps ceil(ne/64) x split x outputs x ne
FP64 partial sums, quadratic in the mesh.  Above the operator's limit it
runs in passes of block rows, each followed by a reduction into 8 slice
sums that sees every partial row in the order k_reduce_wide sums them, so
y must equal the one-pass y bitwise.  Checked for both bases and
precisions, step splits 1 and 2, limits giving from 2 passes to one block
row per pass, the FP64 double layers, the three-density launch inside the
hypersingular operator, and the icosphere L6 P1 single layer (2.5 GB in
one pass) under a 512 MiB limit.
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
    "torus40x20": lambda: meshes.torus(1.0, 0.4, 40, 20),   # 1600: 25 groups
    "blob3": lambda: meshes.perturbed_icosphere(3, 11)[:2],  # 1280: 20 groups
    "ico4": lambda: meshes.icosphere(4),                     # 5120: 80 groups
}


def _passes(op, x, limit):
    op.set_partial_limit(0)
    one = op.partial
    assert one["passes"] == 1
    y1 = op(x)
    op.set_partial_limit(limit)
    many = op.partial
    yp = op(x)
    op.set_partial_limit(0)
    return y1, yp, one, many


@pytest.mark.parametrize("mesh", list(MESHES))
@pytest.mark.parametrize("basis", [0, 1])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("split", [1, 2])
@pytest.mark.parametrize("frac", [0.6, 0.1, 1e-6])
def test_passes_bitwise_equal_one_pass(L, mesh, basis, prec, split, frac):
    P, T = MESHES[mesh]()
    m = L.Mesh(P, T)
    op = L.Operator(m, "V", basis, basis, prec, line_rule=4, tri_rule=4)
    op.set_schedule("symmetric", split)
    x = np.random.default_rng(9).uniform(-1, 1, P.shape[0] if basis else T.shape[0])
    ne = T.shape[0]
    full = ((ne + 63) // 64) * split * (3 if basis else 1) * ne * 8
    limit = max(1, int(frac * full))
    print(f"case {mesh} basis {basis} {prec} split {split} limit {limit}", flush=True)
    y1, yp, one, many = _passes(op, x, limit)
    print(f"  passes {many['passes']}, partial bytes {one['bytes']} -> {many['bytes']}")
    assert many["passes"] > 1
    assert np.array_equal(y1, yp)
    if frac < 0.2:
        assert many["bytes"] < one["bytes"]


@pytest.mark.parametrize("kernel,test,trial", [("K", 0, 1), ("K", 1, 1), ("Kt", 1, 0)])
def test_passes_double_layers(L, kernel, test, trial):
    P, T = meshes.perturbed_icosphere(3, 5)[:2]
    m = L.Mesh(P, T)
    op = L.Operator(m, kernel, test, trial, "fp64", line_rule=4, tri_rule=4)
    op.set_schedule("symmetric")
    x = np.random.default_rng(4).uniform(-1, 1, P.shape[0] if trial else T.shape[0])
    y1, yp, one, many = _passes(op, x, 200_000)
    assert many["passes"] > 1 and np.array_equal(y1, yp)


def test_passes_three_densities_in_hypersingular(L):
    P, T = meshes.icosphere(3)
    m = L.Mesh(P, T)
    D = L.Operator(m, "D", 1, 1, "fp64", line_rule=4, tri_rule=4, max_rhs=3)
    D.set_schedule("symmetric")
    U = np.random.default_rng(7).uniform(-1, 1, (3, P.shape[0])).ravel()
    y1, yp, one, many = _passes(D, U, 300_000)
    assert many["passes"] > 1 and np.array_equal(y1, yp)


def test_passes_at_icosphere_l6_p1(L, oracle):
    """2.5 GB of partial sums in one pass; under a 512 MiB limit the passes'
    buffers stay below it, y is the same, and it matches oracle rows."""
    P, T = meshes.icosphere(6)
    m = L.Mesh(P, T)
    x = np.random.default_rng(6).uniform(-1, 1, P.shape[0])
    rows = np.arange(0, P.shape[0], 2731, dtype=np.int32)
    for prec, tol in (("fp64", 1e-12), ("fp32", 1e-6)):
        op = L.Operator(m, "V", 1, 1, prec, line_rule=4, tri_rule=4)
        op.set_schedule("symmetric")
        print(f"case L6 {prec}", flush=True)
        y1, yp, one, many = _passes(op, x, 512 << 20)
        print(f"L6 {prec}: {many['passes']} passes, partial bytes {one['bytes']} -> "
              f"{many['bytes']}")
        assert many["passes"] > 1 and many["bytes"] < (512 + 64) << 20
        assert np.array_equal(y1, yp)
        if prec == "fp64":
            yo = np.asarray(oracle.apply("V", 1, 1, P, T, x, rows=rows, line_idx=4,
                                         tri_idx=4)).ravel()
            e = rel(yp[rows], yo)
            print(f"L6 passes vs oracle rows {e:.2e}")
            assert e <= tol
