"""HIP operator apply against the CPU oracle (through the C-ABI).

Bars: FP64 path <= 1e-12 relative (2-norm), FP32 far field <= 1e-6.
"""
import numpy as np
import pytest

from easybem_amd import meshes

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-6


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def L():
    from easybem_amd import _lib
    return _lib


MESHES = {
    "ico2": lambda: meshes.icosphere(2),
    "torus12x6": lambda: meshes.torus(1.0, 0.4, 12, 6),
    "blob1": lambda: meshes.perturbed_icosphere(1, 3)[:2],
}

CASES = [("V", 0, 0), ("V", 1, 1), ("V", 0, 1), ("V", 1, 0), ("K", 0, 1),
         ("K", 1, 1), ("kelvin", 0, 0), ("kelvin", 1, 1)]


def _x(kernel, trial, P, T, seed):
    rng = np.random.default_rng(seed)
    n = P.shape[0] if trial else T.shape[0]
    nc = 3 if kernel == "kelvin" else 1
    return rng.uniform(-1, 1, nc * n)


@pytest.mark.parametrize("mesh", list(MESHES))
@pytest.mark.parametrize("kernel,test,trial", CASES)
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_apply_matches_oracle(L, oracle, mesh, kernel, test, trial, prec):
    P, T = MESHES[mesh]()
    print(f"case mesh={mesh} kernel={kernel} test={test} trial={trial} {prec}",
          flush=True)
    x = _x(kernel, trial, P, T, 11)
    m = L.Mesh(P, T)
    op = L.Operator(m, kernel, test, trial, prec, line_rule=4, tri_rule=4,
                    nu=0.3)
    y = op(x)
    yr = oracle.apply(kernel, test, trial, P, T, x, line_idx=4, tri_idx=4,
                      nu=0.3).ravel()
    err = rel(y, yr)
    print("rel err", err)
    assert err <= (TOL64 if prec == "fp64" else TOL32)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_hybrid_rule_icosphere3(L, oracle, prec):
    """Config 1: P0 V on icosphere level 3, 7-point far rule, 4-point Gauss
    near rule (line index 3, triangle index 4)."""
    P, T = meshes.icosphere(3)
    rng = np.random.default_rng(20261203)
    x = rng.uniform(-1, 1, T.shape[0])
    m = L.Mesh(P, T)
    op = L.Operator(m, "V", 0, 0, prec, line_rule=3, tri_rule=4)
    y = op(x)
    yr = oracle.apply("V", 0, 0, P, T, x, line_idx=3, tri_idx=4)
    err = rel(y, yr)
    print("rel err", err)
    assert err <= (TOL64 if prec == "fp64" else TOL32)


def test_launch_configurations_agree(L):
    """Every launch shape gives the same FP64 result to rounding."""
    P, T = meshes.icosphere(3)
    x = np.random.default_rng(5).uniform(-1, 1, T.shape[0])
    m = L.Mesh(P, T)
    op = L.Operator(m, "V", 0, 0, "fp64")
    y0 = op(x)
    for blk, tile, ch in [(64, 64, 1), (128, 16, 7), (256, 64, 3), (64, 1, 40),
                          (256, 37, 0)]:
        op.set_launch(blk, tile, ch)
        assert rel(op(x), y0) < 1e-14


def test_mesh_rejects_open_surface(L):
    P, T = meshes.icosphere(1)
    with pytest.raises(L.EbhError) as e:
        L.Mesh(P, T[:-1])
    assert e.value.code == -2
