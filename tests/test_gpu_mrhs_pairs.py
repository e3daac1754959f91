"""Pair-symmetric multi-RHS schedule of the P0 single layer (EBH_MRHS_PAIRS):
each unordered triangle pair's point sums formed once into 64 x 64 tiles,
then applied to the right-hand sides in both directions.  Checked against
the per-target schedule (EBH_MRHS_TARGET), the single-RHS apply and the
oracle, on meshes whose triangle count is and is not a multiple of 64, with
one and several right-hand-side groups, every rule size the schedule takes,
and in one pass and in many (a small partial limit)."""
import numpy as np
import pytest

from easybem_amd import meshes

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def mesh_case(name):
    if name == "ico0":
        return meshes.icosphere(0)          # 20 triangles: one partial group
    if name == "ico1":
        return meshes.icosphere(1)          # 80 triangles: a group and 16
    if name == "ico3":
        return meshes.icosphere(3)          # 1280 triangles: 20 whole groups
    if name == "torus":
        return meshes.torus(1.0, 0.4, 20, 13)  # 520 triangles: last group of 8
    P, T, _ = meshes.perturbed_icosphere(2, 7)  # 320 triangles: 5 groups
    return P, T


def pair_ops(L, m, prec, max_rhs, **kw):
    tgt = L.Operator(m, "V", 0, 0, prec, max_rhs=max_rhs, **kw)
    tgt.set_mrhs_schedule("target")
    prs = L.Operator(m, "V", 0, 0, prec, max_rhs=max_rhs, **kw)
    prs.set_mrhs_schedule("pairs")
    assert tgt.mrhs_schedule == ("target", 0)
    assert prs.mrhs_schedule[0] == "pairs" and prs.mrhs_schedule[1] >= 1
    return tgt, prs


@pytest.mark.parametrize("name", ["ico0", "ico1", "ico3", "torus", "blob"])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-14), ("fp32", 5e-7)])
def test_pairs_match_target_schedule(name, prec, tol):
    from easybem_amd import _lib as L
    P, T = mesh_case(name)
    m = L.Mesh(P, T)
    tgt, prs = pair_ops(L, m, prec, 64)
    rng = np.random.default_rng(11)
    for nrhs in (2, 7, 64):
        print(f"case {name} {prec} nrhs={nrhs}", flush=True)
        X = rng.uniform(-1, 1, (nrhs, T.shape[0]))
        Yt = tgt(X.ravel()).reshape(nrhs, -1)
        Yp = prs(X.ravel()).reshape(nrhs, -1)
        e = max(rel(Yp[r], Yt[r]) for r in range(nrhs))
        print(f"  pairs vs target: {e:.2e}")
        assert e <= tol


@pytest.mark.parametrize("max_rhs,nrhs", [(8, 5), (12, 12), (20, 19), (100, 100), (100, 70)])
def test_pairs_rhs_groups(max_rhs, nrhs):
    """8, 16, 32 and 64 RHS per group; two groups of 64 (100 RHS, and 70
    RHS in a 100-RHS operator: a partial second group)."""
    from easybem_amd import _lib as L
    P, T = mesh_case("torus")
    m = L.Mesh(P, T)
    for prec, tol in (("fp64", 1e-14), ("fp32", 5e-7)):
        tgt, prs = pair_ops(L, m, prec, max_rhs)
        X = np.random.default_rng(nrhs).uniform(-1, 1, (nrhs, T.shape[0]))
        print(f"case max_rhs={max_rhs} nrhs={nrhs} {prec}", flush=True)
        Yt = tgt(X.ravel()).reshape(nrhs, -1)
        Yp = prs(X.ravel()).reshape(nrhs, -1)
        e = max(rel(Yp[r], Yt[r]) for r in range(nrhs))
        print(f"  pairs vs target: {e:.2e}")
        assert e <= tol


@pytest.mark.parametrize("rules", [(4, 4), (4, 3), (2, 1), (2, 2), (1, 0)])
def test_pairs_rule_sizes(rules):
    """Far rules of 7, 6, 3, 4 and 1 points (FP32 with 6 and 7 points: the
    median form of the pair sums) against the single-RHS apply."""
    from easybem_amd import _lib as L
    P, T = mesh_case("blob")
    m = L.Mesh(P, T)
    kw = dict(line_rule=rules[0], tri_rule=rules[1])
    X = np.random.default_rng(5).uniform(-1, 1, (9, T.shape[0]))
    for prec, tol in (("fp64", 1e-13), ("fp32", 1e-6)):
        one = L.Operator(m, "V", 0, 0, prec, **kw)
        many = L.Operator(m, "V", 0, 0, prec, max_rhs=16, **kw)
        many.set_mrhs_schedule("pairs")
        print(f"case rules={rules} {prec}", flush=True)
        Y = many(X.ravel()).reshape(X.shape)
        e = max(rel(Y[r], one(X[r])) for r in range(X.shape[0]))
        print(f"  pairs vs single-RHS: {e:.2e}")
        assert e <= tol


def test_pairs_fp64_match_oracle(oracle):
    from easybem_amd import _lib as L
    P, T = mesh_case("torus")
    m = L.Mesh(P, T)
    op = L.Operator(m, "V", 0, 0, "fp64", max_rhs=4)
    op.set_mrhs_schedule("pairs")
    X = np.random.default_rng(3).uniform(-1, 1, (4, T.shape[0]))
    Y = op(X.ravel()).reshape(4, -1)
    for r in range(4):
        yo = np.asarray(oracle.apply("V", 0, 0, P, T, X[r], line_idx=4, tri_idx=4)).ravel()
        e = rel(Y[r], yo)
        print(f"column {r}: pairs FP64 vs oracle {e:.2e}")
        assert e <= 1e-12


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_pairs_in_passes(prec):
    """A small partial limit splits the tiles into passes (blocks of 32 tile
    rows with their items; icosphere L5: 320 groups, 10 blocks); every
    output sums its items in the same order whatever the passes, so y is
    bitwise that of one pass, and repeated applies are bitwise equal."""
    from easybem_amd import _lib as L
    P, T = meshes.icosphere(5)
    m = L.Mesh(P, T)
    one = L.Operator(m, "V", 0, 0, prec, max_rhs=16)
    one.set_mrhs_schedule("pairs")
    many = L.Operator(m, "V", 0, 0, prec, max_rhs=16)
    many.set_mrhs_schedule("pairs")
    many.set_partial_limit(1 << 20)
    npass = many.mrhs_schedule[1]
    print(f"passes: one {one.mrhs_schedule[1]}, limited {npass}", flush=True)
    assert one.mrhs_schedule[1] == 1 and npass == 10
    X = np.random.default_rng(9).uniform(-1, 1, (16, T.shape[0]))
    Y1 = one(X.ravel())
    Yp = many(X.ravel())
    assert np.array_equal(Yp, many(X.ravel()))
    assert np.array_equal(Yp, Y1)
    # back to the default limit: one pass again
    many.set_partial_limit(0)
    assert many.mrhs_schedule[1] == 1
    assert np.array_equal(many(X.ravel()), Y1)


def test_pairs_schedule_rejected_where_it_does_not_apply():
    from easybem_amd import _lib as L
    P, T = mesh_case("blob")
    m = L.Mesh(P, T)
    p1 = L.Operator(m, "V", 1, 1, "fp64", max_rhs=8)
    with pytest.raises(RuntimeError):
        p1.set_mrhs_schedule("pairs")
    assert p1.mrhs_schedule == ("target", 0)
    big_rule = L.Operator(m, "V", 0, 0, "fp64", max_rhs=8, line_rule=4, tri_rule=6)
    with pytest.raises(RuntimeError):
        big_rule.set_mrhs_schedule("pairs")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_resident_tiles_equal_per_apply_tiles(prec):
    """Resident tiles (formed once, when enabled) against tiles formed in
    every apply: y bitwise equal in one pass and against a 10-pass apply,
    for repeated applies and a different number of right-hand sides; a
    limit the tiles do not fit in leaves the passes; 0 turns it off."""
    from easybem_amd import _lib as L
    P, T = meshes.icosphere(5)
    m = L.Mesh(P, T)
    ref = L.Operator(m, "V", 0, 0, prec, max_rhs=16)
    ref.set_mrhs_schedule("pairs")
    ref.set_partial_limit(1 << 20)
    op = L.Operator(m, "V", 0, 0, prec, max_rhs=16)
    assert op.resident_tiles == (False, 0)
    op.set_resident_tiles(1 << 40)
    res, nbytes = op.resident_tiles
    ng = (T.shape[0] + 63) // 64
    print(f"case {prec}: resident {res}, {nbytes} bytes, passes {op.mrhs_schedule[1]}",
          flush=True)
    assert res and op.mrhs_schedule == ("pairs", 1)
    assert nbytes == ng * (ng + 1) // 2 * 4096 * (4 if prec == "fp32" else 8)
    rng = np.random.default_rng(17)
    for nrhs in (16, 16, 5):
        X = rng.uniform(-1, 1, (nrhs, T.shape[0]))
        assert np.array_equal(op(X.ravel()), ref(X.ravel()))
    # a limit the tiles do not fit in: per-apply tiles in passes again
    op.set_partial_limit(1 << 20)
    op.set_resident_tiles(1 << 24)
    assert op.resident_tiles == (False, 0) and op.mrhs_schedule == ("pairs", 10)
    X = rng.uniform(-1, 1, (16, T.shape[0]))
    Y = ref(X.ravel())
    assert np.array_equal(op(X.ravel()), Y)
    op.set_resident_tiles(1 << 40)
    assert op.resident_tiles[0]
    assert np.array_equal(op(X.ravel()), Y)
    op.set_resident_tiles(0)
    assert op.resident_tiles == (False, 0) and op.mrhs_schedule[1] == 10
    assert np.array_equal(op(X.ravel()), Y)


def test_resident_tiles_in_block_solves():
    """A block GMRES solve (captured block steps, replayed) with resident
    tiles returns bitwise the iterates of the solve that forms its tiles in
    every apply, in FP64 and in mixed precision."""
    from easybem_amd import _lib as L
    P, T = meshes.icosphere(3)
    m = L.Mesh(P, T)
    B = np.random.default_rng(23).uniform(-1, 1, (8, T.shape[0]))
    out = {}
    for res in (False, True):
        V64 = L.Operator(m, "V", 0, 0, "fp64", max_rhs=8)
        V32 = L.Operator(m, "V", 0, 0, "fp32", max_rhs=8)
        if res:
            V64.set_resident_tiles(1 << 34)
            V32.set_resident_tiles(1 << 34)
            assert V64.resident_tiles[0] and V32.resident_tiles[0]
        for mixed in (False, True):
            s = L.Solver(V64, V32 if mixed else None, batch=8)
            s.set_block(True)
            for k in range(2):  # the second solve replays the captured steps
                X, st = s.solve(B.ravel(), tol=1e-10, mixed=mixed)
                assert all(v["converged"] for v in st)
                out[res, mixed, k] = X
            s.close()
    for mixed in (False, True):
        for k in range(2):
            print(f"case mixed={mixed} solve {k}", flush=True)
            assert np.array_equal(out[True, mixed, k], out[False, mixed, k])
