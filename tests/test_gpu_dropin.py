"""The reference's pybind11 surface (module `easybem`, python/bindings.cpp)
served by this build's module, against the oracle and the reference's own
dense matrices."""
import os
import sys

import numpy as np
import pytest

from easybem_amd import meshes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def eb():
    sys.path.insert(0, os.path.join(ROOT, "easybem_amd"))
    import easybem
    yield easybem
    easybem.close_gpu()


@pytest.mark.parametrize("mesh", ["ico1", "torus8x4"])
@pytest.mark.parametrize("order", [3, 4])
def test_applies_match_reference_dense(eb, mesh, order):
    """init_gpu + apply_gpu_laplace3d_{1layer,2layer,hypersing} against the
    reference's V, K and D (getLaplace3d, compiled from its sources)."""
    z = np.load(os.path.join(GOLD, f"laplace_{mesh}_o{order}.npz"))
    P, T = z["nodes"], z["elements"].astype(np.int32)
    eb.init_gpu(P, T, order, 240, 32)
    rng = np.random.default_rng(order)
    t = rng.uniform(-1, 1, T.shape[0])
    u = rng.uniform(-1, 1, P.shape[0])
    assert rel(eb.apply_gpu_laplace3d_1layer(t), z["V"] @ t) <= 1e-12
    assert rel(eb.apply_gpu_laplace3d_2layer(u), z["K"] @ u) <= 1e-12
    assert rel(eb.apply_gpu_laplace3d_hypersing(u), z["D"] @ u) <= 1e-12
    eb.close_gpu()


def test_linears_and_fp32(eb, oracle):
    P, T = meshes.icosphere(2)
    eb.init_gpu(P, T.astype(np.int32), 4, 240, 32)
    u = np.random.default_rng(1).uniform(-1, 1, P.shape[0])
    want = oracle.apply("V", 1, 1, P, T, u, line_idx=4, tri_idx=4)
    assert rel(eb.apply_gpu_laplace3d_1layer_linears(u), want) <= 1e-12
    eb.set_far_precision(1)
    assert rel(eb.apply_gpu_laplace3d_1layer_linears(u), want) <= 1e-6
    eb.set_far_precision(0)
    eb.close_gpu()


def test_mass_and_version(eb):
    z = np.load(os.path.join(GOLD, "laplace_torus8x4_o4.npz"))
    M = eb.get_laplace3d_mass(z["nodes"], z["elements"].astype(np.int32))
    assert np.abs(M - z["M"]).max() <= 1e-15 * np.abs(z["M"]).max()
    assert eb.version() == "1.0"


@pytest.mark.parametrize("order", [3, 4])
def test_eval_laplace3d_matches_reference(eb, order):
    """eval_laplace3d against evaluateLaplace3d's values (golden)."""
    z = np.load(os.path.join(GOLD, "evaluate_ico2.npz"))
    out = eb.eval_laplace3d(z["nodes"], z["elements"].astype(np.int32),
                            z["points"], z["t"], z["u"], order)
    assert isinstance(out, tuple) and len(out) == 1
    assert out[0].shape == (z["points"].shape[0], 1)
    want = z[f"result_o{order}"]
    assert np.abs(out[0][:, 0] - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("mesh", ["ico1", "torus8x4"])
@pytest.mark.parametrize("order", [3, 4])
def test_get_laplace3d_matches_reference(eb, mesh, order):
    """Dense V, K, D, M assembled on the device against getLaplace3d (V as
    [test, trial])."""
    z = np.load(os.path.join(GOLD, f"laplace_{mesh}_o{order}.npz"))
    V, K, D, M = eb.get_laplace3d(z["nodes"], z["elements"].astype(np.int32), order)
    for A, name in ((V, "V"), (K, "K"), (D, "D"), (M, "M")):
        ref = z[name]
        err = np.abs(A - ref).max() / np.abs(ref).max()
        print(name, err)
        assert A.shape == ref.shape and err <= 1e-13


def test_dense_assembly_agrees_with_applies():
    """Each dense matrix times a vector equals the matrix-free apply."""
    from easybem_amd import _lib as L
    P, T, _ = meshes.perturbed_icosphere(2, 4)
    m = L.Mesh(P, T)
    V, K, D, M = L.assemble(m, 4)
    rng = np.random.default_rng(0)
    t = rng.uniform(-1, 1, T.shape[0])
    u = rng.uniform(-1, 1, P.shape[0])
    assert rel(V @ t, L.Operator(m, "V", 0, 0, "fp64")(t)) <= 1e-13
    assert rel(K @ u, L.Operator(m, "K", 0, 1, "fp64")(u)) <= 1e-13
    assert rel(D @ u, L.Operator(m, "D", 1, 1, "fp64")(u)) <= 1e-13
    assert rel(M @ u, m.mass(u, 0)) <= 1e-15
