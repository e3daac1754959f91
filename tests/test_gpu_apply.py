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


@pytest.mark.parametrize("tri_rule", range(7))
@pytest.mark.parametrize("kernel,test,trial", [("V", 0, 0), ("V", 1, 1),
                                               ("K", 0, 1)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_every_triangle_rule(L, oracle, tri_rule, kernel, test, trial, prec):
    """Far-field rules 0..6 (1, 3, 4, 6, 7, 12, 13 points).  Rules of up to
    7 points run the single layer in the centroid-median form after the
    operator reorders their points; 12 and 13 points use the general form."""
    P, T = meshes.perturbed_icosphere(1, 5)[:2]
    print(f"case rule={tri_rule} {kernel}{test}{trial} {prec}", flush=True)
    x = _x(kernel, trial, P, T, 3)
    m = L.Mesh(P, T)
    op = L.Operator(m, kernel, test, trial, prec, line_rule=3,
                    tri_rule=tri_rule)
    yr = oracle.apply(kernel, test, trial, P, T, x, line_idx=3,
                      tri_idx=tri_rule).ravel()
    err = rel(op(x), yr)
    print("rel err", err)
    assert err <= (TOL64 if prec == "fp64" else TOL32)


@pytest.fixture
def tune_table(L, tmp_path):
    """Saves the process-wide tuning table and restores it afterwards."""
    saved = tmp_path / "table.txt"
    L.tune_save(str(saved))
    yield
    L.tune_clear()
    L.tune_load(str(saved))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_autotune_keeps_results_and_records_shape(L, prec, tune_table):
    """The tuner's chosen shape gives the same apply (FP64: to rounding of
    the chunk summation order) and is picked up by a new operator of the
    same key."""
    L.tune_clear()
    P, T = meshes.icosphere(3)
    x = np.random.default_rng(8).uniform(-1, 1, P.shape[0])
    m = L.Mesh(P, T)
    op = L.Operator(m, "V", 1, 1, prec)
    y0 = op(x)
    r = L.Operator.autotune(op, budget_ms=500.0)
    print(r)
    assert r["block"] in (64, 128, 256) and 1 <= r["tile"] <= 64
    assert r["tried"] >= 1 and r["ms"] <= r["ms_default"]
    assert op.launch() == (r["block"], r["tile"], r["chunks"])
    assert op.schedule == ("symmetric" if r["schedule"] else "chunked")
    assert op.split == r["split"]
    assert rel(op(x), y0) < (1e-14 if prec == "fp64" else 1e-6)
    op2 = L.Operator(m, "V", 1, 1, prec)
    assert op2.launch() == op.launch() and op2.schedule == op.schedule
    L.tune_clear()
    op3 = L.Operator(m, "V", 1, 1, prec)
    assert op3.launch()[1] == 64


def test_autotune_multi_rhs(L, tune_table):
    L.tune_clear()
    P, T = meshes.icosphere(3)
    m = L.Mesh(P, T)
    op = L.Operator(m, "V", 0, 0, "fp32", max_rhs=16)
    X = np.random.default_rng(9).uniform(-1, 1, (16, T.shape[0]))
    dx = L.DeviceArray.from_host(X.ravel())
    dy = L.DeviceArray(X.size)
    op.apply(dx, dy, nrhs=16)
    y0 = dy.to_host()
    r = op.autotune(budget_ms=500.0)
    print(r)
    assert r["mgrid"] > 0 and r["mms"] <= r["mms_default"]
    op.apply(dx, dy, nrhs=16)
    assert rel(dy.to_host(), y0) < 1e-6
    L.tune_clear()


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
