"""Pair-symmetric far-field schedule (k_far_sym) of the single layer and of
the FP64 double layer and its adjoint.

The symmetric schedule sums every unordered pair of 64-triangle groups once
and uses each rsqrt for both directions, so it computes the same matrix-free
apply in another order.  Bars: against the oracle as the chunked schedule
(FP64 <= 1e-12, FP32 <= 1e-6) and against the chunked FP64 apply <= 1e-14
(single layer) or 1e-13 (double layer: r^-3 and the kernel's cancelling
dot product).
Meshes cover one partial group (20 triangles), group counts that are not a
multiple of 4 (the workgroup's 4 target groups) and a partial last group.
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
    "ico0": lambda: meshes.icosphere(0),           # 20 triangles, 1 group
    "blob1": lambda: meshes.perturbed_icosphere(1, 3)[:2],  # 80: 2 groups
    "torus12x6": lambda: meshes.torus(1.0, 0.4, 12, 6),     # 144: 3 groups
    "ico2": lambda: meshes.icosphere(2),           # 320: 5 groups
    "torus40x20": lambda: meshes.torus(1.0, 0.4, 40, 20),   # 1600: 25 groups
}


def _x(basis, P, T, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, P.shape[0] if basis else T.shape[0])


@pytest.mark.parametrize("mesh", list(MESHES))
@pytest.mark.parametrize("basis", [0, 1])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_symmetric_matches_oracle_and_chunked(L, oracle, mesh, basis, prec):
    P, T = MESHES[mesh]()
    print(f"case mesh={mesh} ne={T.shape[0]} basis={basis} {prec}", flush=True)
    x = _x(basis, P, T, 5)
    m = L.Mesh(P, T)
    op = L.Operator(m, "V", basis, basis, prec, line_rule=4, tri_rule=4)
    op.set_schedule("chunked")
    y_chunk = op(x)
    op.set_schedule("symmetric")
    assert op.schedule == "symmetric"
    y = op(x)
    yr = oracle.apply("V", basis, basis, P, T, x, line_idx=4, tri_idx=4).ravel()
    err = rel(y, yr)
    print("rel err vs oracle", err, "vs chunked", rel(y, y_chunk))
    assert err <= (1e-12 if prec == "fp64" else 1e-6)
    if prec == "fp64":
        assert rel(y, y_chunk) <= 1e-14
    # a second apply gives the same bits (fixed summation order)
    assert np.array_equal(op(x), y)


@pytest.mark.parametrize("tri_rule", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_symmetric_every_median_rule(L, tri_rule, prec):
    P, T = meshes.icosphere(2)
    print(f"case ico2 P1 tri_rule={tri_rule} {prec}", flush=True)
    x = _x(1, P, T, 7)
    m = L.Mesh(P, T)
    ref = L.Operator(m, "V", 1, 1, "fp64", line_rule=3, tri_rule=tri_rule)
    ref.set_schedule("chunked")
    y_ref = ref(x)
    op = L.Operator(m, "V", 1, 1, prec, line_rule=3, tri_rule=tri_rule)
    op.set_schedule("symmetric")
    err = rel(op(x), y_ref)
    print("rel err", err)
    assert err <= (1e-14 if prec == "fp64" else 1e-6)


@pytest.mark.parametrize("split", [1, 2, 8, 64])
def test_symmetric_step_split(L, split):
    """Every split of the 64 steps into ranges gives the same apply."""
    P, T = meshes.torus(1.0, 0.4, 40, 20)
    m = L.Mesh(P, T)
    x = _x(1, P, T, 4)
    op = L.Operator(m, "V", 1, 1, "fp64")
    op.set_schedule("chunked")
    y0 = op(x)
    op.set_schedule("symmetric", split)
    assert op.split == split
    assert rel(op(x), y0) <= 1e-14


def test_symmetric_split_rejected(L):
    P, T = meshes.icosphere(1)
    op = L.Operator(L.Mesh(P, T), "V", 0, 0, "fp64")
    for bad in [3, 128, -2]:
        with pytest.raises(L.EbhError):
            op.set_schedule("symmetric", bad)


@pytest.mark.parametrize("kernel,test,trial,tri_rule,prec", [
    ("K", 0, 1, 5, "fp32"), ("K", 1, 1, 5, "fp64"), ("V", 0, 1, 4, "fp64"),
    ("V", 1, 0, 4, "fp64"), ("kelvin", 0, 0, 5, "fp64"), ("V", 0, 0, 5, "fp64")])
def test_symmetric_rejected_where_it_does_not_apply(L, kernel, test, trial,
                                                    tri_rule, prec):
    P, T = meshes.icosphere(1)
    m = L.Mesh(P, T)
    op = L.Operator(m, kernel, test, trial, prec, tri_rule=tri_rule)
    with pytest.raises(L.EbhError) as e:
        op.set_schedule("symmetric")
    assert e.value.code == -4
    assert op.schedule == "chunked"


def test_set_launch_returns_to_chunked(L):
    P, T = meshes.icosphere(2)
    m = L.Mesh(P, T)
    op = L.Operator(m, "V", 1, 1, "fp64")
    x = _x(1, P, T, 3)
    op.set_schedule("symmetric")
    ks = op.split
    assert ks >= 1 and op.launch() == (256, 64, 5 * ks)
    y = op(x)
    op.set_launch(128, 32, 7)
    assert op.schedule == "chunked"
    assert rel(op(x), y) <= 1e-14


def test_symmetric_hypersingular_inner(L):
    """The hypersingular operator applies its P0 single layer through the
    schedule of that inner operator."""
    P, T = meshes.icosphere(2)
    m = L.Mesh(P, T)
    D = L.Operator(m, "D", 1, 1, "fp64", line_rule=2, tri_rule=0)
    x = _x(1, P, T, 9)
    D.set_schedule("chunked")
    y0 = D(x)
    D.set_schedule("symmetric")
    assert D.schedule == "symmetric"
    assert rel(D(x), y0) <= 1e-14


def test_symmetric_mixed_solve(L):
    """Mixed-precision GMRES with both operators on the symmetric schedule
    (captured graphs included) reaches the FP64 residual and the chunked
    solution."""
    P, T = meshes.icosphere(3)
    m = L.Mesh(P, T)
    u = 1.0 / np.linalg.norm(P - np.array([0.0, 0.0, 2.0]), axis=1)
    K = L.Operator(m, "K", 1, 1, "fp64")
    b = 0.5 * m.mass(u, 1) + K(u)
    sols = {}
    for sched in ["chunked", "symmetric"]:
        V64 = L.Operator(m, "V", 1, 1, "fp64")
        V32 = L.Operator(m, "V", 1, 1, "fp32")
        V64.set_schedule(sched)
        V32.set_schedule(sched)
        t, st = L.Solver(V64, V32).solve(b, tol=1e-8, mixed=True)
        print(sched, st)
        assert st["converged"] and st["relres"] <= 1e-8
        r = b - V64(t)
        assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-8
        sols[sched] = t
    assert rel(sols["symmetric"], sols["chunked"]) < 1e-6


@pytest.mark.parametrize("kernel,test", [("D", 1), ("kelvin_K", 0),
                                         ("kelvin_D", 1)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_symmetric_three_rhs_in_composites(L, kernel, test, prec):
    """Composite operators apply their single layer to 3 or 9 densities per
    right-hand side, three per symmetric launch."""
    P, T = meshes.torus(1.0, 0.4, 20, 10)
    m = L.Mesh(P, T)
    nc = 3 if kernel.startswith("kelvin") else 1
    x = np.random.default_rng(12).uniform(-1, 1, nc * P.shape[0])
    ref = L.Operator(m, kernel, test, 1, "fp64", line_rule=3, tri_rule=4,
                     nu=0.3)
    ref.set_schedule("chunked")
    y0 = ref(x)
    op = L.Operator(m, kernel, test, 1, prec, line_rule=3, tri_rule=4, nu=0.3)
    op.set_schedule("symmetric")
    assert op.schedule == "symmetric"
    err = rel(op(x), y0)
    print(kernel, prec, "rel err", err)
    assert err <= (1e-13 if prec == "fp64" else 2e-6)


def test_symmetric_full_size_l5(L):
    """The bench mesh (20480 triangles, 320 groups): symmetric against
    chunked in FP64, and the FP32 symmetric apply against FP64."""
    P, T = meshes.icosphere(5)
    m = L.Mesh(P, T)
    x = _x(1, P, T, 21)
    op = L.Operator(m, "V", 1, 1, "fp64")
    op.set_schedule("chunked")
    y0 = op(x)
    op.set_schedule("symmetric", 1)
    y1 = op(x)
    op32 = L.Operator(m, "V", 1, 1, "fp32")
    op32.set_schedule("symmetric", 1)
    e64, e32 = rel(y1, y0), rel(op32(x), y0)
    print("fp64 symmetric vs chunked", e64, "fp32 symmetric vs fp64", e32)
    assert e64 <= 1e-14 and e32 <= 1e-6


@pytest.mark.parametrize("mesh", list(MESHES))
@pytest.mark.parametrize("test", [0, 1])
def test_symmetric_double_layer(L, oracle, mesh, test):
    """FP64 double layer (P1 trial, P0 or P1 test) on the symmetric
    schedule against the chunked schedule and the oracle."""
    P, T = MESHES[mesh]()
    print(f"case K mesh={mesh} ne={T.shape[0]} test={test}", flush=True)
    x = _x(1, P, T, 8)
    m = L.Mesh(P, T)
    op = L.Operator(m, "K", test, 1, "fp64", line_rule=4, tri_rule=4)
    op.set_schedule("chunked")
    y0 = op(x)
    op.set_schedule("symmetric")
    assert op.schedule == "symmetric"
    y = op(x)
    yr = oracle.apply("K", test, 1, P, T, x, line_idx=4, tri_idx=4).ravel()
    print("rel err vs chunked", rel(y, y0), "vs oracle", rel(y, yr))
    assert rel(y, y0) <= 1e-13
    assert rel(y, yr) <= 1e-12
    assert np.array_equal(op(x), y)


@pytest.mark.parametrize("tri_rule", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("split", [1, 4, 64])
def test_symmetric_double_layer_rules_and_splits(L, tri_rule, split):
    P, T = meshes.torus(1.0, 0.4, 40, 20)
    m = L.Mesh(P, T)
    x = _x(1, P, T, 10)
    op = L.Operator(m, "K", 1, 1, "fp64", line_rule=3, tri_rule=tri_rule)
    op.set_schedule("chunked")
    y0 = op(x)
    op.set_schedule("symmetric", split)
    assert op.split == split
    err = rel(op(x), y0)
    print(f"K rule {tri_rule} split {split}: rel err {err}")
    assert err <= 1e-13


def test_symmetric_double_layer_full_size_l5(L):
    """The bench's right-hand side operator (P1 x P1 K on 20480 triangles)
    and its default schedule."""
    P, T = meshes.icosphere(5)
    m = L.Mesh(P, T)
    u = 1.0 / np.linalg.norm(P - np.array([0.0, 0.0, 2.0]), axis=1)
    op = L.Operator(m, "K", 1, 1, "fp64")
    op.set_schedule("chunked")
    y0 = op(u)
    op.set_schedule("symmetric", 1)
    err = rel(op(u), y0)
    print("K L5 symmetric vs chunked", err)
    assert err <= 1e-13


@pytest.mark.parametrize("mesh", list(MESHES))
@pytest.mark.parametrize("trial", [0, 1])
def test_symmetric_adjoint_double_layer(L, mesh, trial):
    """FP64 adjoint double layer (P1 test, P0 or P1 trial) on the symmetric
    schedule against the chunked schedule (the chunked adjoint is checked
    against the reference's K^T in the hypersingular tests)."""
    P, T = MESHES[mesh]()
    print(f"case Kt mesh={mesh} ne={T.shape[0]} trial={trial}", flush=True)
    x = _x(trial, P, T, 9)
    m = L.Mesh(P, T)
    op = L.Operator(m, "Kt", 1, trial, "fp64", line_rule=4, tri_rule=4)
    op.set_schedule("chunked")
    y0 = op(x)
    op.set_schedule("symmetric")
    assert op.schedule == "symmetric"
    y = op(x)
    print("rel err vs chunked", rel(y, y0))
    assert rel(y, y0) <= 1e-13


@pytest.mark.parametrize("split", [1, 8])
def test_symmetric_adjoint_double_layer_l5(L, split):
    """The Neumann right-hand side operator at 20480 triangles."""
    P, T = meshes.icosphere(5)
    m = L.Mesh(P, T)
    t = np.random.default_rng(14).uniform(-1, 1, T.shape[0])
    op = L.Operator(m, "Kt", 1, 0, "fp64")
    op.set_schedule("chunked")
    y0 = op(t)
    op.set_schedule("symmetric", split)
    err = rel(op(t), y0)
    print("Kt L5 symmetric vs chunked", err)
    assert err <= 1e-13

# ------------------------------------------- FP32 double layers, symmetric
@pytest.mark.parametrize("mesh", list(MESHES))
@pytest.mark.parametrize("kernel,test,trial", [("K", 0, 1), ("K", 1, 1),
                                               ("Kt", 1, 0), ("Kt", 1, 1)])
def test_fp32_symmetric_double_layers(L, oracle, mesh, kernel, test, trial):
    """FP32 double layer and adjoint on the pair-symmetric schedule (the
    default for them): against the oracle's double layer (K), its
    transpose (Kt, P0 trial: the columns of the oracle's P1 x P0 matrix are
    not available, so against the FP64 adjoint, itself checked against the
    reference's K^T), to the FP32 bar 1e-6; bitwise repeatable."""
    P, T = MESHES[mesh]()
    print(f"case {kernel}{test}{trial} fp32 mesh={mesh} ne={T.shape[0]}", flush=True)
    x = _x(trial, P, T, 31)
    m = L.Mesh(P, T)
    op = L.Operator(m, kernel, test, trial, "fp32", line_rule=4, tri_rule=4)
    assert op.schedule == "symmetric"
    y = op(x)
    if kernel == "K":
        yr = oracle.apply("K", test, 1, P, T, x, line_idx=4, tri_idx=4).ravel()
    else:
        yr = L.Operator(m, "Kt", test, trial, "fp64", line_rule=4, tri_rule=4)(x)
    err = rel(y, yr)
    print("rel err", err)
    assert err <= 1e-6
    assert np.array_equal(op(x), y)
    if kernel == "K":
        # the chunked FP32 double layer (hi/lo points) agrees to the same bar
        op.set_schedule("chunked")
        assert rel(op(x), yr) <= 1e-6


@pytest.mark.parametrize("kernel,test,trial", [("K", 1, 1), ("Kt", 1, 0)])
@pytest.mark.parametrize("tri_rule", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("split", [1, 4, 64])
def test_fp32_symmetric_double_layer_rules_and_splits(L, kernel, test, trial,
                                                      tri_rule, split):
    P, T = meshes.torus(1.0, 0.4, 40, 20)
    m = L.Mesh(P, T)
    x = _x(trial, P, T, 32)
    y64 = L.Operator(m, kernel, test, trial, "fp64", line_rule=3, tri_rule=tri_rule)(x)
    op = L.Operator(m, kernel, test, trial, "fp32", line_rule=3, tri_rule=tri_rule)
    op.set_schedule("symmetric", split)
    assert op.split == split
    err = rel(op(x), y64)
    print(f"{kernel} fp32 rule {tri_rule} split {split}: rel err vs fp64 {err}")
    assert err <= 1e-6


@pytest.mark.parametrize("kernel,test,trial", [("K", 1, 1), ("Kt", 1, 0)])
def test_fp32_symmetric_double_layer_l5(L, kernel, test, trial):
    """Full size (20480 triangles): the FP32 double layers against FP64 on
    the bench's Dirichlet data and on a random Neumann density."""
    P, T = meshes.icosphere(5)
    m = L.Mesh(P, T)
    if kernel == "K":
        x = 1.0 / np.linalg.norm(P - np.array([0.0, 0.0, 2.0]), axis=1)
    else:
        x = np.random.default_rng(33).uniform(-1, 1, T.shape[0])
    y64 = L.Operator(m, kernel, test, trial, "fp64")(x)
    op = L.Operator(m, kernel, test, trial, "fp32")
    assert op.schedule == "symmetric"
    err = rel(op(x), y64)
    print(f"{kernel} L5 fp32 symmetric vs fp64: {err}")
    assert err <= 1e-6


# --------------------------------- P0-factored single layer (symmetric)
@pytest.mark.parametrize("tri_rule", [1, 2, 3, 4])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("split", [1, 4])
def test_p0_factored_every_rule(L, oracle, tri_rule, prec, split):
    """P0 test and trial functions on the symmetric schedule sum each
    triangle pair's kernel values once, G = sum_b w'_b J' sum_q w_q J / r,
    and apply G to the densities in both directions (rules of 3 to 7
    points; the one-point rule keeps its own form): against the oracle and
    the chunked FP64 schedule."""
    P, T = meshes.torus(1.0, 0.4, 40, 20)
    print(f"case P0 torus tri_rule={tri_rule} {prec} split={split}", flush=True)
    x = _x(0, P, T, 41)
    m = L.Mesh(P, T)
    ref = L.Operator(m, "V", 0, 0, "fp64", line_rule=3, tri_rule=tri_rule)
    ref.set_schedule("chunked")
    y_ref = ref(x)
    op = L.Operator(m, "V", 0, 0, prec, line_rule=3, tri_rule=tri_rule)
    op.set_schedule("symmetric", split)
    y = op(x)
    yo = oracle.apply("V", 0, 0, P, T, x, line_idx=3, tri_idx=tri_rule).ravel()
    print("rel err vs chunked", rel(y, y_ref), "vs oracle", rel(y, yo))
    assert rel(y, y_ref) <= (1e-14 if prec == "fp64" else 1e-6)
    assert rel(y, yo) <= (1e-12 if prec == "fp64" else 1e-6)
    assert np.array_equal(op(x), y)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_p0_factored_three_densities_through_d(L, prec):
    """The hypersingular operator applies the P0 single layer to three curl
    densities in one symmetric launch (the factored form with three
    right-hand sides): equal to the same operator on the chunked
    schedule."""
    P, T, _ = meshes.perturbed_icosphere(3, 5)
    m = L.Mesh(P, T)
    u = np.random.default_rng(42).uniform(-1, 1, P.shape[0])
    D64 = L.Operator(m, "D", 1, 1, "fp64", stab=1.0)
    y0 = D64(u)
    D = L.Operator(m, "D", 1, 1, prec, stab=1.0)
    assert D.schedule == "symmetric"
    y = D(u)
    D.set_schedule("chunked")
    yc = D(u)
    print("symmetric vs fp64", rel(y, y0), "chunked vs fp64", rel(yc, y0))
    assert rel(y, y0) <= (1e-14 if prec == "fp64" else 1e-6)
    assert rel(yc, y0) <= (1e-14 if prec == "fp64" else 1e-6)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_p0_factored_in_passes(L, prec):
    """Bounded partial sums (passes of block rows) give the one-pass
    result bitwise in the factored form too."""
    P, T = meshes.icosphere(4)
    m = L.Mesh(P, T)
    x = _x(0, P, T, 43)
    op = L.Operator(m, "V", 0, 0, prec)
    op.set_schedule("symmetric", 1)
    y1 = op(x)
    op.set_partial_limit(1 << 20)
    assert op.partial["passes"] > 1
    assert np.array_equal(op(x), y1)
