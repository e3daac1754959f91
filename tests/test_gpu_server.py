"""Solve server: continuous batching, admission by device footprint,
per-request failure isolation, results identical alone and in a mix."""
import numpy as np
import pytest

from easybem_amd import meshes

pytestmark = pytest.mark.gpu


def stream(n, seed=0, levels=(2, 3, 4)):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        lev = levels[k % len(levels)]
        P, T, x0 = meshes.perturbed_icosphere(lev, int(rng.integers(1 << 30)))
        out.append((P, T, x0))
    return out


@pytest.mark.parametrize("inner_tol", [0.0, 1e-5])
def test_server_result_matches_direct_solve(inner_tol):
    """A served P0 Dirichlet solve equals the same mixed solve run
    directly, with the server's default first-pass tolerance for Dirichlet
    requests (5e-9: one FP64 residual) and with one given."""
    from easybem_amd import _lib as L
    P, T, x0 = stream(1, seed=5, levels=(3,))[0]
    srv = L.Server(max_slots=1, inner_tol=inner_tol)
    rid = srv.submit(P, T, x0)
    res, t = srv.wait(rid)
    srv.close()
    assert res["status"] == 0 and res["relres"] <= 1e-8
    m = L.Mesh(P, T)
    V64 = L.Operator(m, "V", 0, 0, "fp64")
    V32 = L.Operator(m, "V", 0, 0, "fp32")
    K = L.Operator(m, "K", 0, 1, "fp64")
    u = 1.0 / np.linalg.norm(P - x0, axis=1)
    b = 0.5 * m.mass(u, 0) + K(u)
    # the server's preconditioner: the sparse inverse on 2 triangle rings
    S = L.precond_spai(V64, rings=2)
    s = L.Solver(V64, V32)
    s.set_precond(S, S)
    t2, st = s.solve(b, tol=1e-8, mixed=True, inner_tol=inner_tol or 5e-9)
    assert np.array_equal(t, t2)


def test_results_identical_alone_and_in_mix():
    from easybem_amd import _lib as L
    reqs = stream(9, seed=1)
    alone = []
    srv = L.Server(max_slots=1)
    for P, T, x0 in reqs:
        res, t = srv.wait(srv.submit(P, T, x0))
        assert res["status"] == 0, res
        alone.append(t)
    srv.close()
    srv = L.Server(max_slots=4)
    ids = [srv.submit(P, T, x0) for P, T, x0 in reqs]
    for rid, t0 in zip(ids, alone):
        res, t = srv.wait(rid)
        print(res)
        assert res["status"] == 0 and res["relres"] <= 1e-8
        assert np.array_equal(t, t0)
    st = srv.stats()
    srv.close()
    assert st["completed"] == 9 and st["failed"] == 0


def test_admission_respects_budget_and_bad_mesh_is_isolated():
    from easybem_amd import _lib as L
    reqs = stream(6, seed=2, levels=(3, 4))
    P, T, x0 = reqs[0]
    srv = L.Server(max_slots=4, hbm_budget=260 << 20)
    ids = [srv.submit(P, T, x0) for P, T, x0 in reqs]
    Pb, Tb, xb = reqs[1]
    bad = srv.submit(Pb, Tb[:-1], xb)          # open surface
    ids2 = [srv.submit(P, T, x0) for P, T, x0 in reqs[:2]]
    res_bad, t_bad = srv.wait(bad)
    assert res_bad["status"] == -2 and "closed 2-manifold" in res_bad["error"]
    for rid in ids + ids2:
        res, _ = srv.wait(rid)
        assert res["status"] == 0 and res["relres"] <= 1e-8
    st = srv.stats()
    srv.close()
    assert st["peak_bytes"] <= 260 << 20
    assert st["failed"] == 1


def test_resident_mesh_reused():
    from easybem_amd import _lib as L
    P, T, x0 = stream(1, seed=7, levels=(3,))[0]
    srv = L.Server(max_slots=2)
    r1, t1 = srv.wait(srv.submit(P, T, x0, key=42))
    r2, t2 = srv.wait(srv.submit(P, T, x0 + 0.5, key=42))
    r3, t3 = srv.wait(srv.submit(P, T, x0, key=42))
    srv.close()
    assert r1["cache_hit"] == 0 and r2["cache_hit"] == 1 and r3["cache_hit"] == 1
    assert np.array_equal(t1, t3) and not np.array_equal(t1, t2)
    assert r3["setup_ms"] < r1["setup_ms"]


def _neumann_data(P, T, x0, tri_rule=4):
    """Triangle means of du/dn for u = 1/|x - x0| by the 7-point rule."""
    from oracle import Oracle
    pts, w = Oracle().tri_rule(tri_rule)
    A, B, Cc = (P[T[:, k] - 1] for k in range(3))
    n = np.cross(B - A, Cc - A)
    J = np.linalg.norm(n, axis=1)
    n /= J[:, None]
    t = np.zeros(T.shape[0])
    for (s1, s2), wq in zip(pts, w):
        x = A + s1 * (B - A) + s2 * (Cc - A)
        d = x - x0
        r = np.linalg.norm(d, axis=1)
        t += 2 * wq * (-(d * n).sum(1) / r ** 3)
    Mt = np.zeros(P.shape[0])
    np.add.at(Mt, T - 1, (J / 6 * t)[:, None])
    return t, Mt


def _weighted_zero_mean(P, T, v):
    from oracle import lumped_areas
    m = lumped_areas(P, T)
    return v - (m @ v) / m.sum()


def test_neumann_server_matches_direct_solve():
    from easybem_amd import _lib as L
    P, T, x0 = stream(1, seed=8, levels=(3,))[0]
    srv = L.Server(max_slots=1)
    res, u = srv.wait(srv.submit(P, T, x0, problem="neumann"))
    srv.close()
    assert res["status"] == 0 and res["relres"] <= 1e-8 and u.size == P.shape[0]
    m = L.Mesh(P, T)
    D64 = L.Operator(m, "D", 1, 1, "fp64", stab=1.0)
    D32 = L.Operator(m, "D", 1, 1, "fp32", stab=1.0)
    Kt = L.Operator(m, "Kt", 1, 0, "fp64")
    t, Mt = _neumann_data(P, T, x0)
    b = 0.5 * Mt - Kt(t)
    s = L.Solver(D64, D32)
    s.set_precond(L.precond_calderon_hs(m, "fp64"), L.precond_calderon_hs(m, "fp32"))
    u2, st = s.solve(b, tol=1e-8, mixed=True)
    # the server forms b in its own host loop, so the two right-hand sides
    # differ in the last bits; the FP32 inner solves can round those
    # differences apart, far below the 1e-8 solve tolerance
    err = np.linalg.norm(u - u2) / np.linalg.norm(u2)
    print("server vs direct", err)
    assert err <= 1e-9
    # without the preconditioner: the same solution to the solve tolerance
    u3, st3 = L.Solver(D64, D32).solve(b, tol=1e-8, mixed=True)
    assert np.linalg.norm(u - u3) <= 1e-6 * np.linalg.norm(u3)


def test_dual_calderon_cuts_neumann_iterations():
    """On a perturbed, anisotropically scaled sphere (level 4) the
    preconditioned Neumann solve needs markedly fewer iterations."""
    from easybem_amd import _lib as L
    P, T, x0 = meshes.perturbed_icosphere(4, 77)
    m = L.Mesh(P, T)
    D64 = L.Operator(m, "D", 1, 1, "fp64", stab=1.0)
    D32 = L.Operator(m, "D", 1, 1, "fp32", stab=1.0)
    t, Mt = _neumann_data(P, T, x0)
    b = 0.5 * Mt - L.Operator(m, "Kt", 1, 0, "fp64")(t)
    _, st0 = L.Solver(D64, D32).solve(b, tol=1e-8, mixed=True)
    s = L.Solver(D64, D32)
    s.set_precond(L.precond_calderon_hs(m, "fp64"), L.precond_calderon_hs(m, "fp32"))
    _, st1 = s.solve(b, tol=1e-8, mixed=True)
    print("iterations", st0["iters"], st1["iters"])
    assert st1["relres"] <= 1e-8 and st1["iters"] <= 0.4 * st0["iters"]


def test_dirichlet_p1_server_matches_preconditioned_direct_solve():
    from easybem_amd import _lib as L
    P, T, x0 = stream(1, seed=6, levels=(3,))[0]
    srv = L.Server(max_slots=1, dirichlet_p1=True)
    res, t = srv.wait(srv.submit(P, T, x0))
    srv.close()
    assert res["status"] == 0 and res["relres"] <= 1e-8 and t.size == P.shape[0]
    m = L.Mesh(P, T)
    V64 = L.Operator(m, "V", 1, 1, "fp64")
    V32 = L.Operator(m, "V", 1, 1, "fp32")
    K = L.Operator(m, "K", 1, 1, "fp64")
    u = 1.0 / np.linalg.norm(P - x0, axis=1)
    b = 0.5 * m.mass(u, 1) + K(u)
    s = L.Solver(V64, V32)
    s.set_precond(L.precond_calderon(m, "fp64", line_rule=2, tri_rule=0),
                  L.precond_calderon(m, "fp32", line_rule=2, tri_rule=0))
    # (the server's first FP32 pass of a Dirichlet solve goes to 5e-9)
    t2, st = s.solve(b, tol=1e-8, mixed=True, inner_tol=5e-9)
    # right-hand sides formed by the server and here differ in the last bits
    err = np.linalg.norm(t - t2) / np.linalg.norm(t2)
    print("server vs direct", err)
    assert err <= 1e-9


def test_neumann_solution_converges_to_analytic_trace():
    """Interior Neumann problem for u = 1/|x - x0|: the P1 trace matches u
    up to a constant, with an error that shrinks under refinement."""
    from easybem_amd import _lib as L
    x0 = np.array([0.3, -1.9, 0.8])
    srv = L.Server(max_slots=2)
    errs = []
    for lev in (2, 3, 4):
        P, T = meshes.icosphere(lev)
        res, uh = srv.wait(srv.submit(P, T, x0, problem="neumann"))
        assert res["status"] == 0 and res["relres"] <= 1e-8, res
        u = 1.0 / np.linalg.norm(P - x0, axis=1)
        a, b = _weighted_zero_mean(P, T, uh), _weighted_zero_mean(P, T, u)
        errs.append(np.linalg.norm(a - b) / np.linalg.norm(b))
    srv.close()
    print("relative trace errors", errs)
    assert errs[2] < 0.5 * errs[1] < 0.25 * errs[0] and errs[2] < 1e-2


def test_dirichlet_and_neumann_mix_identical_alone():
    from easybem_amd import _lib as L
    reqs = [(P, T, x0, ("dirichlet", "neumann")[k % 2])
            for k, (P, T, x0) in enumerate(stream(6, seed=9, levels=(2, 3)))]
    srv = L.Server(max_slots=1)
    alone = []
    for P, T, x0, pb in reqs:
        res, v = srv.wait(srv.submit(P, T, x0, problem=pb))
        assert res["status"] == 0, res
        alone.append(v)
    srv.close()
    srv = L.Server(max_slots=3)
    ids = [srv.submit(P, T, x0, problem=pb) for P, T, x0, pb in reqs]
    for rid, v0 in zip(ids, alone):
        res, v = srv.wait(rid)
        assert res["status"] == 0 and res["relres"] <= 1e-8
        assert np.array_equal(v, v0)
    srv.close()


@pytest.mark.parametrize("p1", [True, False])
def test_setup_error_on_a_worker_thread_is_reported(p1):
    """A Dirichlet source point on a mesh node makes the nodal data, and so
    the right-hand side, non-finite.  The error arises in the set-up job
    that forms it (a host thread of its own); the request fails with that
    message, the server goes on, and the same mesh then solves with a
    valid point."""
    from easybem_amd import _lib as L
    P, T, x0 = stream(1, seed=12, levels=(3,))[0]
    srv = L.Server(max_slots=2, dirichlet_p1=p1)
    bad = srv.submit(P, T, P[7].copy(), key=77)
    res, _ = srv.wait(bad)
    assert res["status"] != 0 and "on the surface" in res["error"]
    good = srv.submit(P, T, x0, key=77)
    res2, sol = srv.wait(good)
    st = srv.stats()
    srv.close()
    assert res2["status"] == 0 and res2["relres"] <= 1e-8
    assert np.all(np.isfinite(sol))
    assert st["failed"] == 1 and st["completed"] == 1


def test_key_resident_for_another_mesh_fails_that_request_only():
    """A mesh_key already resident for a different mesh: the request fails
    with EBH_ERR_ARG before any apply, the resident mesh stays cached, and
    a later request with the original mesh still hits it."""
    from easybem_amd import _lib as L
    (P, T, x0), (P2, T2, x2) = stream(2, seed=13, levels=(3, 2))
    srv = L.Server(max_slots=1)
    r1, t1 = srv.wait(srv.submit(P, T, x0, key=99))
    r2, t2 = srv.wait(srv.submit(P2, T2, x2, key=99))
    # same sizes, moved nodes: also another mesh
    r3, _ = srv.wait(srv.submit(P * 1.01, T, x0, key=99))
    r4, t4 = srv.wait(srv.submit(P, T, x0, key=99))
    st = srv.stats()
    srv.close()
    assert r1["status"] == 0 and r4["status"] == 0
    for r in (r2, r3):
        assert r["status"] == -1 and "different mesh" in r["error"]
    assert t2 is None and r4["cache_hit"] == 1 and np.array_equal(t1, t4)
    assert st["failed"] == 2 and st["completed"] == 2


def test_measured_device_bytes_within_estimate_and_budget():
    """Every device allocation made for a request is counted: the measured
    per-request peak stays within the request's admission estimate, and
    the server's measured peak within its budget."""
    from easybem_amd import _lib as L
    reqs = stream(6, seed=14, levels=(3, 4))
    budget = 600 << 20
    srv = L.Server(max_slots=3, hbm_budget=budget)
    ids = [srv.submit(P, T, x0, problem=("dirichlet", "neumann")[k % 2])
           for k, (P, T, x0) in enumerate(reqs)]
    for rid in ids:
        res, _ = srv.wait(rid)
        print(f"measured {res['device_peak']} of {res['footprint']} estimated")
        assert res["status"] == 0
        assert 0 < res["device_peak"] <= res["footprint"]
    st = srv.stats()
    srv.close()
    print(st)
    assert 0 < st["device_peak"] <= budget and st["device_bytes"] == 0


def test_served_solves_replay_graphs_with_identical_results():
    """Workers capture a resident mesh's Arnoldi steps as hipGraphs (the
    capture gate keeps other workers' allocations out of a capture) and
    replay them in later solves on that mesh: the results equal those of a
    server that launches eagerly, bit for bit, also with other requests
    running at the same time."""
    from easybem_amd import _lib as L
    reqs = stream(6, seed=15, levels=(3, 4))
    out = {}
    for graphs in (False, True):
        srv = L.Server(max_slots=3, graphs=graphs)
        res = []
        for rep in range(2):  # the second pass hits the resident meshes
            # (each pass ends before the next starts: a request whose mesh
            # is still in use by another request builds its own copy)
            ids = []
            for k, (P, T, x0) in enumerate(reqs):
                pb = ("dirichlet", "neumann")[k % 2]
                ids.append(srv.submit(P, T, x0 + 0.1 * rep, key=100 + k, problem=pb))
            res += [srv.wait(r) for r in ids]
        srv.close()
        for r, _ in res:
            assert r["status"] == 0 and r["relres"] <= 1e-8, r
        assert sum(r["cache_hit"] for r, _ in res) == len(reqs)
        # the resident meshes' second solves replay the captured steps
        hits = [r for r, _ in res if r["cache_hit"]]
        if graphs:
            assert all(r["graph_steps"] > 0 for r in hits)
        else:
            assert all(r["graph_steps"] == 0 for r, _ in res)
        out[graphs] = [t for _, t in res]
    for a, b in zip(out[False], out[True]):
        assert np.array_equal(a, b)
