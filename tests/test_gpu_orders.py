"""The reference's full `order` range on the HIP path.

Triangle rules 7..19 (16 to 79 points, the far-field rules of the
reference's `order` 7 and above, src/quadrature.cpp:1047-1048) run the
point-lane far kernel; their near field uses the Gauss rules of the same
index.  Checked here:
  - every operator kind at several high rules against the oracle (rows),
    FP64 <= 1e-12 and FP32 <= 1e-6 relative;
  - the drop-in module's dense matrices and applies at orders 7, 8 and 9
    against the compiled reference's (tests/golden/laplace_*_o{7,8,9}.npz);
  - the C++ dense surface (include/BEM3dLaplace.h, BEM3dElasticity.h)
    through a linked C++ caller, layouts included;
  - the calls of python/examples/laplace3dDirichletGPU.py:44-81 (shaft
    mesh, order 9 and 10), against the reference's dense solution.
"""
import os
import subprocess

import numpy as np
import pytest

from easybem_amd import meshes

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b))


def relmax(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


KINDS = [("V", 0, 0), ("V", 1, 1), ("K", 0, 1), ("K", 1, 1), ("kelvin", 0, 0),
         ("kelvin", 1, 1)]


@pytest.mark.parametrize("rule", [7, 9, 12, 19])
@pytest.mark.parametrize("kernel,test,trial", KINDS)
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
def test_high_rule_apply_matches_oracle(oracle, rule, kernel, test, trial, prec, tol):
    from easybem_amd import _lib as L
    P, T = meshes.icosphere(2)
    print(f"case rule {rule} {kernel}{test}{trial} {prec}", flush=True)
    m = L.Mesh(P, T)
    op = L.Operator(m, kernel, test, trial, prec, line_rule=min(rule, 9), tri_rule=rule,
                    nu=0.3)
    rng = np.random.default_rng(rule)
    nc = 3 if kernel == "kelvin" else 1
    x = rng.uniform(-1, 1, nc * (P.shape[0] if trial else T.shape[0]))
    y = op(x)
    nrow = P.shape[0] if test else T.shape[0]
    rows = np.sort(rng.choice(nrow, 12, replace=False)).astype(np.int32)
    yo = np.asarray(oracle.apply(kernel, test, trial, P, T, x, rows=rows,
                                 line_idx=min(rule, 9), tri_idx=rule, nu=0.3))
    yr = np.stack([y[c * nrow + rows] for c in range(3)]) if nc == 3 else y[rows]
    e = rel(yr, yo)
    print(f"rule {rule} {kernel}{test}{trial} {prec}: {e:.2e}")
    assert e <= tol


@pytest.mark.parametrize("rule", range(7, 20))
def test_every_high_rule_p0_single_layer(oracle, rule):
    from easybem_amd import _lib as L
    P, T = meshes.torus(1.0, 0.4, 16, 8)
    m = L.Mesh(P, T)
    x = np.random.default_rng(rule).uniform(-1, 1, T.shape[0])
    rows = np.arange(0, T.shape[0], 7, dtype=np.int32)
    yo = oracle.apply("V", 0, 0, P, T, x, rows=rows, line_idx=4, tri_idx=rule)
    for prec, tol in (("fp64", 1e-12), ("fp32", 1e-6)):
        y = L.Operator(m, "V", 0, 0, prec, line_rule=4, tri_rule=rule)(x)
        assert rel(y[rows], yo) <= tol


def test_hypersingular_and_adjoint_at_rule_9(oracle):
    """The composite hypersingular operator and the adjoint double layer at
    order 9 against the reference's dense D and K^T."""
    from easybem_amd import _lib as L
    g = np.load(os.path.join(GOLD, "laplace_torus8x4_o9.npz"))
    P, T = g["nodes"], g["elements"]
    m = L.Mesh(P, T)
    u = np.random.default_rng(5).uniform(-1, 1, P.shape[0])
    t = np.random.default_rng(6).uniform(-1, 1, T.shape[0])
    D = L.Operator(m, "D", 1, 1, "fp64", line_rule=9, tri_rule=9)
    assert rel(D(u), g["D"] @ u) <= 1e-12
    Kt = L.Operator(m, "Kt", 1, 0, "fp64", line_rule=9, tri_rule=9)
    assert rel(Kt(t), g["K"].T @ t) <= 1e-12


@pytest.mark.parametrize("mesh", ["ico1", "torus8x4"])
@pytest.mark.parametrize("order", [7, 8, 9])
def test_dropin_module_at_high_orders(mesh, order):
    import easybem_amd.easybem as eb
    g = np.load(os.path.join(GOLD, f"laplace_{mesh}_o{order}.npz"))
    P, T = g["nodes"], g["elements"]
    V, K, D, M = eb.get_laplace3d(P, T, order)
    print(f"case drop-in {mesh} order {order}", flush=True)
    assert relmax(V, g["V"]) < 1e-12 and relmax(K, g["K"]) < 1e-12
    assert relmax(D, g["D"]) < 1e-12 and relmax(M, g["M"]) < 1e-15
    eb.init_gpu(P, T, order, 240, 32)
    rng = np.random.default_rng(order)
    x, u = rng.uniform(-1, 1, T.shape[0]), rng.uniform(-1, 1, P.shape[0])
    assert rel(eb.apply_gpu_laplace3d_1layer(x), g["V"] @ x) <= 1e-12
    assert rel(eb.apply_gpu_laplace3d_2layer(u), g["K"] @ u) <= 1e-12
    assert rel(eb.apply_gpu_laplace3d_hypersing(u), g["D"] @ u) <= 1e-12
    eb.close_gpu()


def _run_cpp(tmp_path, P, T, order, points, tn, ud, nu=0.0, elastic=False):
    """Runs tests/cpp/dropin_check; returns its outputs by name."""
    exe = os.path.join(HERE, "cpp", "dropin_check")
    assert os.path.exists(exe), "build first: python -m easybem_amd.build"
    np_, ne, npts = P.shape[0], T.shape[0], points.shape[0]
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(src, "wb") as f:
        np.array([np_, ne, order, npts, int(elastic)], np.int32).tofile(f)
        np.array([nu], np.float64).tofile(f)
        np.ascontiguousarray(P, np.float64).tofile(f)
        np.ascontiguousarray(T, np.int32).tofile(f)
        np.ascontiguousarray(points, np.float64).tofile(f)
        np.ascontiguousarray(tn, np.float64).tofile(f)
        np.ascontiguousarray(ud, np.float64).tofile(f)
    subprocess.run([exe, str(src), str(dst)], check=True, timeout=120)
    v = np.fromfile(dst, np.float64)
    sizes = [("V", ne * ne), ("K", ne * np_), ("D", np_ * np_), ("M", ne * np_),
             ("Mmass", ne * np_), ("C1", ne * np_), ("C2", ne * np_), ("C3", ne * np_),
             ("result", npts)]
    if elastic:
        sizes += [("Ve", 9 * ne * ne), ("Ke", 9 * ne * np_), ("De", 9 * np_ * np_),
                  ("Me", 9 * ne * np_)]
    assert v.size == sum(n for _, n in sizes)
    out, o = {}, 0
    for name, n in sizes:
        out[name] = v[o:o + n]
        o += n
    return out


def _col(a, rows, cols):
    """A column-major flat array as a (rows, cols) matrix."""
    return a.reshape(cols, rows).T


def test_cpp_dense_surface_matches_reference(oracle, tmp_path):
    g = np.load(os.path.join(GOLD, "laplace_ico1_o9.npz"))
    P, T = g["nodes"], g["elements"]
    ne, np_ = T.shape[0], P.shape[0]
    pts = np.array([[0.0, 0.0, 0.0], [0.2, -0.1, 0.3], [2.0, 1.0, 0.5]])
    rng = np.random.default_rng(8)
    tn, ud = rng.uniform(-1, 1, ne), rng.uniform(-1, 1, np_)
    o = _run_cpp(tmp_path, P, T, 9, pts, tn, ud)
    # V[i*ne + j] is (test i, trial j); K, D, M, C column-major
    assert relmax(o["V"].reshape(ne, ne), g["V"]) < 1e-12
    assert relmax(_col(o["K"], ne, np_), g["K"]) < 1e-12
    assert relmax(_col(o["D"], np_, np_), g["D"]) < 1e-12
    assert relmax(_col(o["M"], ne, np_), g["M"]) < 1e-15
    assert relmax(_col(o["Mmass"], ne, np_), g["M"]) < 1e-15
    for c in range(3):
        assert relmax(_col(o[f"C{c + 1}"], ne, np_), g[f"C{c + 1}"]) < 1e-14
    ev = oracle.evaluate(P, T, pts, tn, ud, tri_idx=9)
    assert rel(o["result"], ev) < 1e-12
    # elasticity (order 4) and the reference's evaluation golden values
    k = np.load(os.path.join(GOLD, "kelvin_ico1_o4.npz"))
    o = _run_cpp(tmp_path, k["nodes"], k["elements"], 4, pts, tn, ud, nu=float(k["nu"]),
                 elastic=True)
    assert relmax(_col(o["Ve"], 3 * ne, 3 * ne), k["V"]) < 1e-12
    assert relmax(_col(o["Ke"], 3 * ne, 3 * np_), k["K"]) < 1e-12
    assert relmax(_col(o["De"], 3 * np_, 3 * np_), k["D"]) < 1e-12
    assert relmax(_col(o["Me"], 3 * ne, 3 * np_), k["M"]) < 1e-15
    e = np.load(os.path.join(GOLD, "evaluate_ico2.npz"))
    o = _run_cpp(tmp_path, e["nodes"], e["elements"], 4, e["points"], e["t"], e["u"])
    assert rel(o["result"], e["result_o4"]) < 1e-12


@pytest.mark.parametrize("npass", [0, 1])
def test_reference_gpu_example_calls(npass):
    """python/examples/laplace3dDirichletGPU.py:44-81 on the shaft mesh:
    init_gpu(P, T, order + i, 240, 32), b = M u / 2 + apply_gpu_laplace3d_2layer(u),
    then CG on apply_gpu_laplace3d_1layer (here to rtol 1e-11 instead of the
    example's 1e-4), against the reference's b and its dense solution of
    V t = b at the same order."""
    import easybem_amd.easybem as eb
    from scipy.sparse.linalg import LinearOperator, cg
    g = np.load(os.path.join(GOLD, "example_shaft.npz"))
    P, T = g[f"nodes{npass}"], g[f"elements{npass}"]
    order, u = int(g[f"order{npass}"]), g[f"u{npass}"]
    ne = T.shape[0]
    print(f"case example pass {npass}: {ne} triangles, order {order}", flush=True)
    MM = eb.get_laplace3d_mass(P, T)
    eb.init_gpu(P, T, order, 240, 32)
    b = 0.5 * (MM @ u) + eb.apply_gpu_laplace3d_2layer(u)
    assert rel(b, g[f"b{npass}"]) <= 1e-12
    assert rel(eb.apply_gpu_laplace3d_1layer(g[f"x{npass}"]), g[f"Vx{npass}"]) <= 1e-12
    VV = LinearOperator(shape=(ne, ne), matvec=eb.apply_gpu_laplace3d_1layer,
                        dtype=np.float64)
    t, info = cg(VV, b, x0=np.zeros(ne), rtol=1e-11, maxiter=5000)
    eb.close_gpu()
    e = rel(t, g[f"t{npass}"])
    print(f"pass {npass}: CG info {info}, t vs reference dense solve {e:.2e}")
    assert info == 0 and e <= 1e-7
