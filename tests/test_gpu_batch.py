"""Persistent multi-mesh far-field launch (ebh_batch_*): several meshes of
uneven size in one launch give the same results as their own applies."""
import numpy as np
import pytest

from easybem_amd import meshes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    from easybem_amd import _lib
    return _lib


def uneven_meshes(n, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        lev = [1, 2, 3, 4][k % 4]
        P, T, _ = meshes.perturbed_icosphere(lev, int(rng.integers(1 << 30)))
        out.append((P, T))
    out.append(meshes.torus(1.0, 0.4, 24, 12))
    return out


@pytest.mark.parametrize("kernel,test,trial", [("V", 0, 0), ("V", 1, 1),
                                               ("K", 0, 1), ("kelvin", 1, 1),
                                               ("Kt", 1, 0)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_equals_individual_applies(L, kernel, test, trial, prec):
    print(f"case {kernel}{test}{trial} {prec}", flush=True)
    ms = uneven_meshes(6, seed=3)
    rng = np.random.default_rng(1)
    ops, xs, hx = [], [], []
    for k, (P, T) in enumerate(ms):
        m = L.Mesh(P, T)
        op = L.Operator(m, kernel, test, trial, prec, nu=0.3)
        nc = 3 if kernel == "kelvin" else 1
        x = rng.uniform(-1, 1, nc * (P.shape[0] if trial else T.shape[0]))
        ops.append(op)
        hx.append(x)
        xs.append(L.DeviceArray.from_host(x))
    if kernel == "Kt" and prec == "fp32":
        # the FP32 adjoint double layer runs on the pair-symmetric schedule
        # only; batches run double layers chunked
        with pytest.raises(RuntimeError, match="pair-symmetric"):
            L.Batch(ops)
        return
    b = L.Batch(ops)
    print(b.info())
    # the members' own applies use the batch's launch shapes
    want = [op(x) for op, x in zip(ops, hx)]
    ys = [L.DeviceArray(w.size) for w in want]
    for rep in range(2):
        b.apply(xs, ys)
        for y, w in zip(ys, want):
            assert np.array_equal(y.to_host(), w)
    b.close()


def test_batch_follows_launch_shape_changes(L):
    ms = uneven_meshes(3, seed=4)
    ops, xs, ys = [], [], []
    for P, T in ms:
        op = L.Operator(L.Mesh(P, T), "V", 0, 0, "fp64")
        ops.append(op)
        xs.append(L.DeviceArray.from_host(np.ones(T.shape[0])))
        ys.append(L.DeviceArray(T.shape[0]))
    b = L.Batch(ops)
    ops[1].set_launch(128, 16, 5)
    b.apply(xs, ys)
    for op, x, y in zip(ops, xs, ys):
        assert np.array_equal(y.to_host(), op(x.to_host()))


def test_batch_results_match_oracle(L, oracle):
    """The batch shapes keep parity: FP64 members against the oracle."""
    ms = uneven_meshes(3, seed=5)
    ops, xs, ys, hx = [], [], [], []
    for P, T in ms:
        op = L.Operator(L.Mesh(P, T), "V", 1, 1, "fp64", line_rule=4, tri_rule=4)
        x = np.random.default_rng(2).uniform(-1, 1, P.shape[0])
        ops.append(op)
        hx.append(x)
        xs.append(L.DeviceArray.from_host(x))
        ys.append(L.DeviceArray(P.shape[0]))
    b = L.Batch(ops)
    b.apply(xs, ys)
    for (P, T), x, y in zip(ms, hx, ys):
        yr = oracle.apply("V", 1, 1, P, T, x, line_idx=4, tri_idx=4)
        assert np.linalg.norm(y.to_host() - yr) <= 1e-12 * np.linalg.norm(yr)


def test_batch_rejects_mixed_kinds(L):
    P, T = meshes.icosphere(1)
    m = L.Mesh(P, T)
    with pytest.raises(L.EbhError):
        L.Batch([L.Operator(m, "V", 0, 0, "fp64"), L.Operator(m, "V", 0, 0, "fp32")])
    with pytest.raises(L.EbhError):
        L.Batch([L.Operator(m, "D", 1, 1, "fp64")])
    op = L.Operator(m, "V", 0, 0, "fp64")
    with pytest.raises(L.EbhError):
        L.Batch([op, op])
