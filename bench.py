"""Benchmark: Laplace Dirichlet solves on one MI355X.

Workload (one step = one solve): icosphere subdivided 5 times (20480
triangles, 10242 vertices), P1 single layer V and double layer K, interior
Dirichlet data u = 1/|x - (0,0,2)|, right-hand side b = (1/2 M + K) u,
solved for the P1 Neumann trace with GMRES(50) by mixed-precision iterative
refinement (FP32 far field inside, FP64 outer residual) to a relative FP64
residual of 1e-8.  Near field: 5-point Gauss Sauter-Schwab; far field:
7-point triangle rule.  GMRES is right-preconditioned by the Calderon
preconditioner (the hypersingular operator between inverse lumped masses,
with a 1-point far rule); --precond none runs it without.  The line also
reports the unpreconditioned solve rate, configs 1, 3 and 4 (matvecs and
the batched 64-RHS solves), a serving stream on icosphere levels 4-6 and the
CPU baseline (--no-secondary, --no-serve, --no-cpu-baseline skip them).

Prints one JSON line (see the keys below).  --gpus N with N > 1 exits: this
build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X vector FP32 (datasheet)
PEAK_FP64_TFLOPS = 78.6    # MI355X vector FP64 (datasheet)
PEAK_FP64_MFMA_TFLOPS = 78.6  # MI355X dense FP64 matrix (datasheet)
# device bytes the config-4 operators may keep their tiles in (the FP64
# operator's 26.9 GB at icosphere L6 fit; the card holds 288 GB)
RESIDENT_TILE_LIMIT = 48 << 30
PEAK_HBM_GBPS = 8000.0     # MI355X HBM3E (datasheet)
FLOP_SL = 11               # per far interaction, single layer, one RHS
# HBM bytes per launch of the FP32 single-layer far kernel on the level-5
# icosphere with the packaged schedule (pair-symmetric, one step range): two
# rocprofv3 passes (FETCH_SIZE, WRITE_SIZE) over `bench.py --pmc-probe`, each
# scaled by the same counter on a 512 MiB device copy
# (profiles/k_far_sym_fp32_sl7_icosphere5_traffic.md).  A counter cannot be
# read inside this run, so the line reports that profile's figure, with its
# source, and only at level 5 with that schedule.
TRAFFIC_FAR32_L5_BYTES = 206.1e6
TRAFFIC_SOURCE = ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes over "
                  "bench.py --pmc-probe, scaled by a 512 MiB copy: "
                  "profiles/k_far_sym_fp32_sl7_icosphere5_traffic.md")
TRAFFIC_SCHEDULE = ("symmetric", 1)


def near_pass_bytes(w, ne, partial_rows, fused=False):
    """Algorithmic HBM bytes of the pass after the far kernel of a P1 x P1
    scalar operator on the pair-symmetric schedule: every partial row's
    3 ne sums (8 B each) read once, the near CSR (8 B value + 4 B column per
    entry, 4 B row pointers), x and y once.  Separate passes (the default):
    k_reduce_wide also writes the 3 ne sums, which k_finish gathers through
    the node lists (4 + 4 B indices, 8 B value) with their row pointers.
    Fused (k_reduce_finish): the sums are read at the nodes' incidence
    positions, with the node row pointers and one 4 B row block bound per
    workgroup (about 5 rows a block)."""
    rows, cols, nnz = w["rows"], w["cols"], w["near_nnz"]
    csr = 12 * nnz + 4 * (rows + 1)
    vectors = 8 * cols + 8 * rows
    partials = 8 * 3 * ne * partial_rows
    if fused:
        far = partials + 4 * (rows + 1) + 4 * (rows // 5 + 1)
    else:
        far = partials + 8 * 3 * ne + 16 * 3 * ne + 4 * (rows + 1)
    return csr + vectors + far


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    """The host CPU's model name (lscpu, else /proc/cpuinfo)."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _sampled_matvec(o, P, T, threads, seconds, seed, op="V", basis=1, nu=0.3):
    """Seconds of one full matvec extrapolated from the oracle on random
    test rows (node rows for P1, triangle rows for P0; the far rule 7 points,
    the near rule 5-point Gauss), sized to about `seconds` of work: config 2's
    P1 V by default, config 3's Kelvin P1 V, config 4's P0 V."""
    rng = np.random.default_rng(seed)
    ncols = (P.shape[0] if basis else T.shape[0]) * (3 if op == "kelvin" else 1)
    ntest = P.shape[0] if basis else T.shape[0]
    x = rng.uniform(-1, 1, ncols)
    nrows = 8
    while True:
        rows = rng.choice(ntest, size=nrows, replace=False)
        t0 = time.perf_counter()
        o.apply(op, basis, basis, P, T, x, rows=rows, line_idx=4, tri_idx=4, nu=nu,
                nthreads=threads)
        elapsed = time.perf_counter() - t0
        if elapsed > seconds / 4 or nrows >= ntest:
            break
        nrows = min(ntest, int(nrows * max(2.0, seconds / 4 / max(elapsed, 1e-3))))
    return elapsed / nrows * ntest, nrows, elapsed


def _quiet_fd1(fn):
    """Runs fn with the process's file descriptor 1 on /dev/null (the
    reference's C++ code prints its assembly progress there)."""
    import ctypes
    libc = ctypes.CDLL(None)
    sys.stdout.flush()
    saved = os.dup(1)
    null = os.open(os.devnull, os.O_WRONLY)
    try:
        os.dup2(null, 1)
        return fn()
    finally:
        libc.fflush(None)
        os.dup2(saved, 1)
        os.close(null)
        os.close(saved)


def reference_cpu():
    """The reference's own CPU path (getLaplace3d compiled from its sources
    into oracle/_ref; single-threaded, as written): dense assembly of V
    (P0), K, D and M at order 4 measured whole on icosphere L3 and
    extrapolated to config 2's mesh (L5) by the triangle-pair count, plus
    a dense LU solve of an L5-sized P0 system extrapolated from one of
    L4's size.  Its CPU code has no P1 single layer, so this is the P0
    formulation of config 2's Dirichlet problem (its CPU example's)."""
    from oracle import Reference
    from easybem_amd import meshes
    if not Reference.available():
        return {"available": False,
                "note": "oracle/_ref not built (the reference sources were absent)"}
    r = Reference()
    P3, T3 = meshes.icosphere(3)
    t0 = time.perf_counter()
    _quiet_fd1(lambda: r.laplace3d(P3, T3, 4))
    t_asm3 = time.perf_counter() - t0
    ne3, ne4, ne5 = T3.shape[0], 4 * T3.shape[0], 16 * T3.shape[0]
    A = np.random.default_rng(5).standard_normal((ne4, ne4)) + ne4 * np.eye(ne4)
    t0 = time.perf_counter()
    np.linalg.solve(A, np.ones(ne4))
    t_lu4 = time.perf_counter() - t0
    t_asm5 = t_asm3 * (ne5 / ne3) ** 2
    t_lu5 = t_lu4 * (ne5 / ne4) ** 3
    return {"kind": "reference", "cores": 1, "value": 1.0 / (t_asm5 + t_lu5),
            "unit": "solves/s", "assembly_s_ico3_measured": t_asm3,
            "assembly_s_ico5_extrapolated": t_asm5, "lu_s_ico5_extrapolated": t_lu5,
            "sample": f"getLaplace3d (V P0, K, D, M, order 4) on icosphere L3 "
                      f"({ne3} triangles, {t_asm3:.2f} s), times {(ne5 / ne3) ** 2:.0f} "
                      f"pairs for L5; numpy LU of a {ne4}-square system "
                      f"({t_lu4:.2f} s), times {(ne5 / ne4) ** 3:.0f} for L5's P0 system"}


def cpu_baseline(P, T, n_matvec, seconds=10.0):
    """The oracle (FP64 C restatement of the reference CPU operator) on the
    host cores, 1 thread (as the single-threaded reference) and all cores
    (OpenMP over targets).  The headline value is config 2's solve rate at
    all cores: a full P1 V matvec extrapolated from sampled node rows, times
    the FP64 GPU solve's matvec count (preconditioner applies counted by
    their far-field work).  Config 1: full matvecs at both thread counts and
    a dense solve (oracle assembly at order 4 + LAPACK)."""
    from oracle import Oracle
    from easybem_amd import meshes
    o = Oracle()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    t_all, rows_all, el_all = _sampled_matvec(o, P, T, threads, seconds, 7)
    t_one, rows_one, el_one = _sampled_matvec(o, P, T, 1, seconds / 2, 8)
    # config 1: icosphere L3, P0 V, 4-point Gauss near / 7-point far
    P1, T1 = meshes.icosphere(3)
    x1 = np.random.default_rng(20261203).uniform(-1, 1, T1.shape[0])
    c1 = {}
    for name, th in (("1_thread", 1), ("all_cores", threads)):
        t0 = time.perf_counter()
        o.apply("V", 0, 0, P1, T1, x1, line_idx=3, tri_idx=4, nthreads=th)
        c1[f"matvec_s_{name}"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    V, _, _ = o.laplace_dense(P1, T1, 4)
    t_asm = time.perf_counter() - t0
    t0 = time.perf_counter()
    np.linalg.solve(V, x1)
    c1["dense_solve_s"] = {"assembly_all_cores": t_asm,
                           "lapack_solve": time.perf_counter() - t0}
    # configs 3 and 4: sampled rows on all cores, extrapolated
    P3, T3 = meshes.torus(1.0, 0.4, 256, 128)
    t3, r3, e3 = _sampled_matvec(o, P3, T3, threads, 4.0, 9, op="kelvin", basis=1)
    P4, T4 = meshes.icosphere(6)
    t4, r4, e4 = _sampled_matvec(o, P4, T4, threads, 4.0, 10, op="V", basis=0)
    c34 = {"config3_torus_p1_kelvin": {
               "matvec_s_all_cores": t3, "extrapolated": True,
               "sample": f"{r3} random node rows ({e3:.1f} s)"},
           "config4_ico6_p0_V": {
               "matvec_s_all_cores": t4, "extrapolated": True,
               "block_matvec_64rhs_s": 64 * t4,
               "sample": f"{r4} random triangle rows ({e4:.1f} s), one right-hand "
                         "side per apply as the reference's operator, times 64"}}
    return {"value": 1.0 / (t_all * n_matvec), "unit": "solves/s",
            "cores": threads, "kind": "port",
            "value_1_thread": 1.0 / (t_one * n_matvec),
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"config 2 (icosphere L5 P1 V): oracle apply on {rows_all} "
                      f"random node rows on {threads} threads ({el_all:.1f} s) and "
                      f"{rows_one} rows on 1 thread ({el_one:.1f} s), extrapolated "
                      f"to full matvecs ({t_all:.1f} s / {t_one:.0f} s) x "
                      f"{n_matvec:.1f} FP64 matvec-equivalents per solve (the FP64 "
                      "GPU solve's operator and preconditioner applies); config 1 "
                      "full matvecs and dense solve measured whole; configs 3 "
                      "and 4 row-sampled on all cores, extrapolated",
            "config1_ico3_p0_V": c1, **c34, "reference_cpu_path": reference_cpu()}


def secondary(L, meshes, solve_batched=False):
    """Configs 1, 3 and 4 at full size: matvec times (HIP events) and, for
    config 4, optionally the batched 64-RHS mixed-precision solve."""
    out = {}
    rng = np.random.default_rng(20261203)
    # config 1: icosphere L3, P0 V, 7-point far / 4-point near
    P, T = meshes.icosphere(3)
    m = L.Mesh(P, T)
    c = meshes.centroids(P, T)
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    t = 3 * c[:, 2] ** 2 - 1
    r1 = {}
    for prec in ("fp64", "fp32"):
        op = L.Operator(m, "V", 0, 0, prec, line_rule=3, tri_rule=4)
        x = L.DeviceArray.from_host(t)
        y = L.DeviceArray(op.nrows)
        tm = op.time(x, y, reps=20)
        Vt = y.to_host()
        A = meshes.areas(P, T)
        tf = 11 * op.work()["far_interactions"] / tm["far"] / 1e9
        r1[prec] = {"matvec_ms": tm["total"], "far_ms": tm["far"],
                    "far_tflops": tf,
                    "far_frac": tf / (PEAK_FP32_TFLOPS if prec == "fp32"
                                      else PEAK_FP64_TFLOPS),
                    "rayleigh_Y20": float(t @ Vt / np.sum(A * t * t))}
    out["cfg1_ico3_p0_V"] = r1
    # config 3: torus 256x128, Kelvin P1, FP32 far field
    P, T = meshes.torus(1.0, 0.4, 256, 128)
    m = L.Mesh(P, T)
    r3 = {}
    for prec in ("fp32", "fp64"):
        op = L.Operator(m, "kelvin", 1, 1, prec, line_rule=4, tri_rule=4, nu=0.3)
        x = L.DeviceArray.from_host(rng.standard_normal(op.ncols))
        y = L.DeviceArray(op.nrows)
        tm = op.time(x, y, reps=2)
        tf = 30 * op.work()["far_interactions"] / tm["far"] / 1e9
        r3[prec] = {"matvec_ms": tm["total"], "far_tflops": tf,
                    "far_frac": tf / (PEAK_FP32_TFLOPS if prec == "fp32"
                                      else PEAK_FP64_TFLOPS),
                    "schedule": op.schedule, "partial_passes": op.partial["passes"]}
    out["cfg3_torus_p1_kelvin"] = r3
    # config 4: icosphere L6, P0 V, 64 right-hand sides in one launch
    P, T = meshes.icosphere(6)
    m = L.Mesh(P, T)
    r4 = {}
    ops = {}
    for prec in ("fp32", "fp64"):
        op = L.Operator(m, "V", 0, 0, prec, line_rule=4, tri_rule=4, max_rhs=64)
        ops[prec] = op
        B = rng.uniform(-1, 1, 64 * op.ncols)
        x = L.DeviceArray.from_host(B)
        y = L.DeviceArray(op.nrows * 64)
        tm = op.time(x, y, reps=1, nrhs=64)
        # the tile launches alone (far-kernel profiling records them)
        L.op_profile(op, True, 64)
        op.apply(x, y, nrhs=64)
        tiles_ms, tiles_n = L.op_profile(op, False)
        sched, passes = op.mrhs_schedule
        pts = op.work()["far_interactions"]   # ordered far point pairs
        tri_pairs = pts / 49                  # ordered far triangle pairs
        # FLOP the pair-symmetric schedule executes: each unordered
        # triangle pair's 7 x 7 point pairs once (11 FLOP each: r^2 and the
        # source weight) plus 2 per target point for the target weights
        # (k_mrhs_tiles, FP32 or FP64 VALU), then 2 FLOP per right-hand side
        # per ordered pair (k_mrhs_tile_mfma, FP64 matrix cores)
        flop_tiles = tri_pairs / 2 * (49 * 11 + 2 * 7)
        flop_mfma = tri_pairs * 2 * 64
        mfma_ms = tm["far"] - tiles_ms
        valu_peak = PEAK_FP32_TFLOPS if prec == "fp32" else PEAK_FP64_TFLOPS
        r4[prec] = {
            "block_matvec_ms": tm["total"], "schedule": sched, "passes": passes,
            "far_ms": tm["far"],
            "k_mrhs_tiles": {"ms": tiles_ms, "launches": tiles_n, "bound": "valu",
                             "tflops": flop_tiles / tiles_ms / 1e9,
                             "frac": flop_tiles / tiles_ms / 1e9 / valu_peak},
            "k_mrhs_tile_mfma_and_reduce": {
                "ms": mfma_ms, "bound": "mfma",
                "tflops": flop_mfma / mfma_ms / 1e9,
                "frac": flop_mfma / mfma_ms / 1e9 / PEAK_FP64_MFMA_TFLOPS},
            # the multi-RHS counting convention (9 + 2 nrhs FLOP per ordered
            # point pair)
            # divided by the far time: above the peaks because the P0
            # densities factor per triangle pair and each unordered pair's
            # point sums are formed once
            "far_tflops_convention": pts * (9 + 2 * 64) / tm["far"] / 1e9}
    if solve_batched:
        s = L.Solver(ops["fp64"], ops["fp32"], restart=50, batch=64)
        # near-field sparse approximate inverse on 3 triangle rings (at L6
        # with 64 RHS: 10 block steps against 12 with 2 rings and 9 with 4;
        # built in about 0.35 s against 0.19 and 1.15 s)
        t0 = time.perf_counter()
        S = L.precond_spai(ops["fp64"], rings=3, max_rhs=64)
        r4["spai_build_s"] = time.perf_counter() - t0
        s.set_precond(S, S)
        B = rng.uniform(-1, 1, 64 * ops["fp64"].ncols)
        db = L.DeviceArray.from_host(B)
        dx = L.DeviceArray.from_host(np.zeros_like(B))
        t0 = time.perf_counter()
        st = s.solve_device(db, dx, tol=1e-8, mixed=True, nb=64, max_iters=3000)
        dt = time.perf_counter() - t0
        r4["batched_mixed_solve"] = {"seconds": dt, "solves_per_s": 64 / dt,
                                     "iters_max": max(v["iters"] for v in st),
                                     "relres_max": max(v["relres"] for v in st)}
        # the same 64 systems in plain FP64 (FP64 64-RHS applies inside)
        s64 = L.Solver(ops["fp64"], None, restart=50, batch=64)
        s64.set_precond(S, S)
        dx = L.DeviceArray.from_host(np.zeros_like(B))
        t0 = time.perf_counter()
        st = s64.solve_device(db, dx, tol=1e-8, mixed=False, nb=64, max_iters=3000)
        dt64 = time.perf_counter() - t0
        r4["batched_fp64_solve"] = {"seconds": dt64, "solves_per_s": 64 / dt64,
                                    "iters_max": max(v["iters"] for v in st),
                                    "relres_max": max(v["relres"] for v in st)}
        r4["mixed_speedup_vs_fp64"] = dt64 / dt
        s64.close()
        # block GMRES: one shared block Krylov space of 64 columns per cycle
        for name, a32 in (("block_mixed_solve", ops["fp32"]), ("block_fp64_solve", None)):
            sb = L.Solver(ops["fp64"], a32, restart=50, batch=64)
            sb.set_precond(S, S)
            sb.set_block(True)
            dx = L.DeviceArray.from_host(np.zeros_like(B))
            sb.solve_device(db, dx, tol=1e-8, mixed=a32 is not None, nb=64,
                            max_iters=3000)
            dx = L.DeviceArray.from_host(np.zeros_like(B))
            t0 = time.perf_counter()
            st = sb.solve_device(db, dx, tol=1e-8, mixed=a32 is not None, nb=64,
                                 max_iters=3000)
            dtb = time.perf_counter() - t0
            r4[name] = {"seconds": dtb, "solves_per_s": 64 / dtb,
                        "block_steps": max(v["iters"] for v in st),
                        "relres_max": max(v["relres"] for v in st),
                        "note": "second solve on the solver (captured block steps "
                                "replayed)"}
            sb.close()
        r4["block_mixed_speedup_vs_block_fp64"] = (r4["block_fp64_solve"]["seconds"] /
                                                   r4["block_mixed_solve"]["seconds"])
        r4["block_mixed_speedup_vs_lockstep_fp64"] = dt64 / r4["block_mixed_solve"]["seconds"]
        # the plain FP64 path: the same systems one at a time, a single-RHS
        # FP64 operator (pair-symmetric single-RHS schedule) with the sparse
        # inverse built for one right-hand side; one untimed solve, then
        # `nseq` of the 64 systems timed
        n = ops["fp64"].ncols
        V1 = L.Operator(m, "V", 0, 0, "fp64", line_rule=4, tri_rule=4)
        S1 = L.precond_spai(V1, rings=3)
        s1 = L.Solver(V1, None, restart=50)
        s1.set_precond(S1, S1)
        nseq = 4
        its = []
        for r in range(nseq + 1):
            b1 = L.DeviceArray.from_host(B[r * n:(r + 1) * n])
            x1 = L.DeviceArray.from_host(np.zeros(n))
            if r == 1:
                t0 = time.perf_counter()
            st1 = s1.solve_device(b1, x1, tol=1e-8, mixed=False, max_iters=3000)
            if r >= 1:
                its.append(st1[0]["iters"] if isinstance(st1, list) else st1["iters"])
        dts = (time.perf_counter() - t0) / nseq
        s1.close()
        r4["sequential_fp64_solve"] = {
            "seconds_per_solve": dts, "solves_per_s": 1.0 / dts, "systems_timed": nseq,
            "iters_max": max(its),
            "note": "plain FP64 path: one system at a time, single-RHS FP64 operator"}
        r4["block_mixed_speedup_vs_sequential_fp64"] = (
            r4["block_mixed_solve"]["solves_per_s"] * dts)
        # resident tiles: each operator's tiles formed once and kept in HBM
        # (a stored far field, 8 or 16 bytes per unordered triangle pair);
        # every apply is the FP64 matrix-core products and their reduction
        res = {}
        B = rng.uniform(-1, 1, 64 * ops["fp64"].ncols)
        x = L.DeviceArray.from_host(B)
        y = L.DeviceArray(ops["fp64"].nrows * 64)
        for prec in ("fp32", "fp64"):
            op = ops[prec]
            t0 = time.perf_counter()
            op.set_resident_tiles(RESIDENT_TILE_LIMIT)
            build_s = time.perf_counter() - t0
            ok, nbytes = op.resident_tiles
            tm = op.time(x, y, reps=3, nrhs=64)
            tri_pairs = op.work()["far_interactions"] / 49
            res[prec] = {"resident": ok, "tile_bytes": nbytes, "build_s": build_s,
                         "block_matvec_ms": tm["total"], "products_ms": tm["far"],
                         "products_mfma_frac": tri_pairs * 2 * 64 / tm["far"] / 1e9
                         / PEAK_FP64_MFMA_TFLOPS}
        sb = L.Solver(ops["fp64"], None, restart=50, batch=64)
        sb.set_precond(S, S)
        sb.set_block(True)
        for k in range(2):
            dx = L.DeviceArray.from_host(np.zeros_like(B))
            t0 = time.perf_counter()
            st = sb.solve_device(db, dx, tol=1e-8, mixed=False, nb=64, max_iters=3000)
            dtr = time.perf_counter() - t0
        sb.close()
        res["block_fp64_solve"] = {
            "seconds": dtr, "solves_per_s": 64 / dtr,
            "block_steps": max(v["iters"] for v in st),
            "relres_max": max(v["relres"] for v in st),
            "note": "second solve on the solver; tiles formed before it"}
        res["speedup_vs_matrix_free_block_mixed"] = r4["block_mixed_solve"]["seconds"] / dtr
        r4["resident_tiles"] = res
    out["cfg4_ico6_p0_V_64rhs"] = r4
    out["multi_mesh"] = multi_mesh(L, meshes)
    return out


def _hip_stream():
    """A non-blocking HIP stream from the runtime the library already
    loaded (for issuing several operators' applies on one stream)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    s = ctypes.c_void_p()
    if hip.hipStreamCreateWithFlags(ctypes.byref(s), ctypes.c_uint(1)) != 0:
        raise RuntimeError("hipStreamCreateWithFlags failed")
    return s, hip


def multi_mesh(L, meshes, reps=20):
    """Many meshes of uneven size in one persistent far-field launch
    (ebh_batch) against their own applies issued back to back, each
    operator on its own stream (they overlap on the hardware queues) and
    all on one stream (in sequence); tuned launch shapes."""
    rng = np.random.default_rng(7)
    res = {}
    one, hip = _hip_stream()
    for name, levels in (("32 x icosphere L3", [3] * 32),
                         ("16 x icosphere L2-L4", [2, 2, 3, 3, 3, 4, 4, 4] * 2)):
        for prec in ("fp32", "fp64"):
            ops, xs, ys, work = [], [], [], 0
            for lev in levels:
                P, T, _ = meshes.perturbed_icosphere(lev, int(rng.integers(1 << 30)))
                op = L.Operator(L.Mesh(P, T), "V", 0, 0, prec)
                ops.append(op)
                xs.append(L.DeviceArray.from_host(rng.uniform(-1, 1, T.shape[0])))
                ys.append(L.DeviceArray(T.shape[0]))
                work += op.work()["far_interactions"]

            def timed(fn):
                for _ in range(3):
                    fn()
                L.check(L.lib().ebh_synchronize())
                t0 = time.perf_counter()
                for _ in range(reps):
                    fn()
                L.check(L.lib().ebh_synchronize())
                return (time.perf_counter() - t0) / reps * 1e3

            t_seq = timed(lambda: [op.apply(x, y) for op, x, y in zip(ops, xs, ys)])
            t_one = timed(lambda: [op.apply(x, y, stream=one.value)
                                   for op, x, y in zip(ops, xs, ys)])
            b = L.Batch(ops)
            t_b = timed(lambda: b.apply(xs, ys))
            res[f"{name} {prec}"] = {
                "separate_ms": t_seq, "separate_one_stream_ms": t_one,
                "batched_ms": t_b, "speedup": t_seq / t_b,
                "speedup_vs_one_stream": t_one / t_b,
                "batched_far_tflops": 11 * work / t_b / 1e9,
                "work_items": b.info()["items"]}
            b.close()
    hip.hipStreamDestroy(one)
    return res


# single-stream service time (s) of one solve, by (problem, level): sets the
# offered load of the serving stream
SERVICE_S = {("dirichlet", 4): 0.03, ("dirichlet", 5): 0.1,
             ("dirichlet", 6): 1.0, ("neumann", 4): 0.05, ("neumann", 5): 0.2,
             ("neumann", 6): 1.6}


def serving(L, meshes, n_requests=24, levels=(4, 5), rate=None, slots=4,
            seed=20261203, neumann_share=1 / 3):
    """Config 5: a stream of Dirichlet and Neumann solves with Poisson
    arrivals.  Each mesh is a seeded icosphere (level drawn from `levels`)
    with radius 1 + 0.1 sum_{l<=4} c_lm Y_lm and axes scaled by
    U(0.5, 1.5); the boundary data comes from a point source outside the
    mesh.  Latency is submit -> result."""
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(n_requests):
        lev = int(rng.choice(levels))
        pb = "neumann" if rng.random() < neumann_share else "dirichlet"
        P, T, x0 = meshes.perturbed_icosphere(lev, int(rng.integers(1 << 30)))
        reqs.append((lev, pb, P, T, x0))
    # P1 Dirichlet flux and Neumann trace, both Calderon-preconditioned
    srv = L.Server(max_slots=slots, tol=1e-8, mixed=True, dirichlet_p1=True)
    # warm the runtime (kernels loaded, streams created) on a small request
    P, T, x0 = meshes.perturbed_icosphere(2, 1)
    srv.wait(srv.submit(P, T, x0))
    if rate is None:
        # offered load about 0.7 of the single-stream service rate
        rate = 0.7 / np.mean([SERVICE_S[(pb, lv)] for lv, pb, *_ in reqs])
    gaps = rng.exponential(1.0 / rate, n_requests)
    t0 = time.perf_counter()
    ids, arrivals = [], []
    t_next = t0
    for (lev, pb, P, T, x0), gap in zip(reqs, gaps):
        t_next += gap
        now = time.perf_counter()
        if t_next > now:
            time.sleep(t_next - now)
        arrivals.append(time.perf_counter() - t0)
        ids.append((lev, pb, srv.submit(P, T, x0, problem=pb)))
    res = []
    for lev, pb, rid in ids:
        r, _ = srv.wait(rid)
        r["level"] = lev
        r["problem"] = pb
        res.append(r)
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    ok = [r for r in res if r["status"] == 0 and r["relres"] <= 1e-8]
    lat = np.array([r["total_ms"] for r in ok])
    by_level = {}
    for pb in ("dirichlet", "neumann"):
        for lev in sorted(set(levels)):
            sel = [r for r in ok if r["level"] == lev and r["problem"] == pb]
            if sel:
                v = np.array([r["total_ms"] for r in sel])
                by_level[f"{pb} L{lev}"] = {
                    "n": len(sel), "p50_ms": float(np.percentile(v, 50)),
                    "max_ms": float(v.max()),
                    "setup_ms_p50": float(np.median([r["setup_ms"] for r in sel])),
                    "solve_ms_p50": float(np.median([r["solve_ms"] for r in sel])),
                    "iters_max": int(max(r["iters"] for r in sel))}
    return {"requests": n_requests, "completed": len(ok), "levels": list(levels),
            "formulation": "Dirichlet: P1 flux, V with the Calderon "
                           "preconditioner Q D Q; Neumann: P1 trace, D + m m^T "
                           "with the dual Calderon preconditioner Q V Q (Q a "
                           "Chebyshev inverse of the P1 mass); mixed-precision "
                           "refinement to 1e-8",
            "arrival_rate_per_s": rate, "slots": slots,
            "solves_per_s": len(ok) / wall,
            "p50_ms": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_ms": float(np.percentile(lat, 99)) if len(lat) else None,
            "mean_queue_ms": float(np.mean([r["queue_ms"] for r in ok])) if ok else None,
            "by_level": by_level, "peak_device_bytes": st["peak_bytes"],
            "max_relres": max(r["relres"] for r in ok) if ok else None}


def resident_graphs(L, meshes, level=4, n=12):
    """Repeated Dirichlet requests on one resident mesh (the same mesh_key,
    moving point sources), one at a time, with the workers' Arnoldi steps
    replayed from captured hipGraphs or launched eagerly: median solve and
    total latency of the cache hits."""
    P, T, x0 = meshes.perturbed_icosphere(level, 4242)
    out = {}
    for graphs in (True, False):
        srv = L.Server(max_slots=1, tol=1e-8, mixed=True, dirichlet_p1=True,
                       graphs=graphs)
        res = []
        for k in range(n):
            r, _ = srv.wait(srv.submit(P, T, x0 * (1.0 + 0.01 * k), key=7))
            res.append(r)
        srv.close()
        hits = [r for r in res if r["cache_hit"]]
        out["graphs" if graphs else "eager"] = {
            "solve_ms_p50": float(np.median([r["solve_ms"] for r in hits])),
            "total_ms_p50": float(np.median([r["total_ms"] for r in hits])),
            "graph_steps": int(hits[-1]["graph_steps"]),
            "iters": int(hits[-1]["iters"])}
    out["level"] = level
    return out


def pmc_probe(L, meshes, level):
    """Fixed workload for hardware-counter passes: a plain device copy of
    512 MiB (the calibration of FETCH_SIZE / WRITE_SIZE) three times, then
    three FP32-far and three FP64 P1 single-layer applies at the bench
    mesh."""
    import torch
    n = 64 << 20
    src = torch.empty(n, dtype=torch.float64, device="cuda").uniform_()
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    P, T = meshes.icosphere(level)
    mesh = L.Mesh(P, T)
    V32 = L.Operator(mesh, "V", 1, 1, "fp32", line_rule=4, tri_rule=4)
    x = L.DeviceArray.from_host(np.random.default_rng(1).uniform(-1, 1, V32.ncols))
    y = L.DeviceArray(V32.nrows)
    for _ in range(3):
        V32.apply(x, y)
    L.check(L.lib().ebh_synchronize())
    V64 = L.Operator(mesh, "V", 1, 1, "fp64", line_rule=4, tri_rule=4)
    for _ in range(3):
        V64.apply(x, y)
    L.check(L.lib().ebh_synchronize())
    print(json.dumps({"probe": "pmc", "copy_bytes_each_way": n * 8,
                      "far_interactions": V32.work()["far_interactions"]}))


def mrhs_compare(L, meshes, level=6, reps=2):
    """Config 4's 64-RHS block matvec (icosphere `level`, P0 V, 7-point far
    rule) under both multi-RHS schedules and with the pair schedule's tiles
    resident, FP32 and FP64: per-apply times (HIP events, ebh_op_time) and
    the pair schedule's largest deviation from the target schedule, one JSON
    line each."""
    P, T = meshes.icosphere(level)
    m = L.Mesh(P, T)
    n = T.shape[0]
    X = np.random.default_rng(1).uniform(-1, 1, 64 * n)
    for prec in ("fp32", "fp64"):
        ys = {}
        for sched in ("target", "pairs", "resident"):
            op = L.Operator(m, "V", 0, 0, prec, line_rule=4, tri_rule=4, max_rhs=64)
            op.set_mrhs_schedule("target" if sched == "target" else "pairs")
            if sched == "resident":
                op.set_resident_tiles(RESIDENT_TILE_LIMIT)
            x = L.DeviceArray.from_host(X)
            y = L.DeviceArray(64 * n)
            tm = op.time(x, y, reps=reps, nrhs=64)
            ys[sched] = y.to_host().reshape(64, n)
            print(json.dumps({"precision": prec, "schedule": sched,
                              "passes": op.mrhs_schedule[1], "ms": tm}), flush=True)
            del op
        a, b = ys["target"], ys["pairs"]
        e = max(np.linalg.norm(a[r] - b[r]) / np.linalg.norm(a[r]) for r in range(64))
        print(json.dumps({"precision": prec, "pairs_vs_target_max_rel": e,
                          "resident_bitwise_equal": bool(np.array_equal(
                              ys["resident"], ys["pairs"]))}), flush=True)


def operator_probe(L, meshes, level=6, reps=3):
    """The double layer, its adjoint, the P0 single layer and the
    hypersingular operator (a P0 single layer on three curl densities) in
    both precisions on the icosphere: reps applies each after one warm-up,
    wall time per apply from host timers around synchronized loops, one
    JSON line per operator (the kernel trace of this workload is the
    profile of the symmetric double-layer and P0 kernels)."""
    P, T = meshes.icosphere(level)
    m = L.Mesh(P, T)
    rng = np.random.default_rng(3)
    for name, kern, te, tr in (("K P1xP1", "K", 1, 1), ("Kt P1xP0", "Kt", 1, 0),
                               ("V P0", "V", 0, 0), ("D P1", "D", 1, 1)):
        for prec in ("fp32", "fp64"):
            op = L.Operator(m, kern, te, tr, prec, line_rule=4, tri_rule=4,
                            stab=1.0 if kern == "D" else 0.0)
            x = L.DeviceArray.from_host(rng.uniform(-1, 1, op.ncols))
            y = L.DeviceArray(op.nrows)
            op.apply(x, y)
            L.check(L.lib().ebh_synchronize())
            t0 = time.perf_counter()
            for _ in range(reps):
                op.apply(x, y)
            L.check(L.lib().ebh_synchronize())
            ms = (time.perf_counter() - t0) / reps * 1e3
            print(json.dumps({"operator": name, "precision": prec, "level": level,
                              "schedule": op.schedule, "ms_per_apply": ms}), flush=True)
            del op


def kelvin_probe(L, meshes, reps=3):
    """The Kelvin single layer on the 256 x 128 torus (65536 triangles, P1,
    7-point far rule, nu 0.3) in both precisions on both far schedules: the
    default (pair-symmetric) and the chunked kernel at the packaged table's
    chunked shape (256-thread blocks, 32-triangle target tiles, 8 source
    chunks: 2048 tiles x 8 = 16384 workgroups).
    One warm-up and reps applies each, one JSON line per case (the kernel
    trace of this workload is the profile of the Kelvin kernels)."""
    P, T = meshes.torus(1.0, 0.4, 256, 128)
    m = L.Mesh(P, T)
    rng = np.random.default_rng(3)
    for prec in ("fp32", "fp64"):
        for sched in ("default", "chunked"):
            op = L.Operator(m, "kelvin", 1, 1, prec, line_rule=4, tri_rule=4, nu=0.3)
            if sched == "chunked":
                op.set_launch(256, 32, 8)
            x = L.DeviceArray.from_host(rng.uniform(-1, 1, op.ncols))
            y = L.DeviceArray(op.nrows)
            op.apply(x, y)
            L.check(L.lib().ebh_synchronize())
            t0 = time.perf_counter()
            for _ in range(reps):
                op.apply(x, y)
            L.check(L.lib().ebh_synchronize())
            ms = (time.perf_counter() - t0) / reps * 1e3
            print(json.dumps({"operator": "kelvin P1", "precision": prec,
                              "schedule": op.schedule, "partial": op.partial,
                              "ms_per_apply": ms}), flush=True)
            del op


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--level", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64-compare", action="store_true")
    ap.add_argument("--inner-tol", type=float, default=1e-8,
                    help="tolerance of the first FP32 inner solve (default: the "
                         "target itself; its FP64 residual certifies the result "
                         "or starts a second pass)")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip configs 1, 3 and 4 (matvecs, batched solves), "
                         "which the default run times")
    ap.add_argument("--serve", type=int, default=24,
                    help="serving stream of N requests on icosphere levels 4-6")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--precond", choices=("calderon", "spai", "none"), default="calderon",
                    help="right preconditioner of the GMRES solves (default: the "
                         "hypersingular operator with a 1-point far rule between "
                         "Chebyshev inverses of the P1 mass; spai: the near-field "
                         "sparse approximate inverse of V)")
    ap.add_argument("--retune", action="store_true",
                    help="run the launch-shape autotuner on the bench operators "
                         "instead of using the packaged table")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="run only the fixed copy + apply workload profiled "
                         "with hardware counters (no JSON bench line)")
    ap.add_argument("--mrhs-compare", action="store_true",
                    help="run only config 4's 64-RHS block matvec under both "
                         "multi-RHS schedules (no JSON bench line)")
    ap.add_argument("--operator-probe", action="store_true",
                    help="run only the double layer, adjoint, P0 single layer "
                         "and hypersingular applies at icosphere L6 in both "
                         "precisions (no JSON bench line)")
    ap.add_argument("--kelvin-probe", action="store_true",
                    help="run only the Kelvin single layer on the 256 x 128 "
                         "torus, both precisions and far schedules (no JSON "
                         "bench line)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "easybem-hip runs on one GPU; --gpus must "
                          "be 1 (multi-GPU sharding is a design note only)"}))
        sys.exit(2)

    from easybem_amd import _lib as L, meshes

    if args.pmc_probe:
        # same launch shapes as the bench line (the packaged table)
        pmc_probe(L, meshes, args.level)
        return
    if args.mrhs_compare:
        mrhs_compare(L, meshes)
        return
    if args.operator_probe:
        operator_probe(L, meshes)
        return
    if args.kelvin_probe:
        kelvin_probe(L, meshes)
        return
    P, T = meshes.icosphere(args.level)
    x0 = np.array([0.0, 0.0, 2.0])
    u = 1.0 / np.linalg.norm(P - x0, axis=1)
    mesh = L.Mesh(P, T)
    V64 = L.Operator(mesh, "V", 1, 1, "fp64", line_rule=4, tri_rule=4)
    V32 = L.Operator(mesh, "V", 1, 1, "fp32", line_rule=4, tri_rule=4)
    K64 = L.Operator(mesh, "K", 1, 1, "fp64", line_rule=4, tri_rule=4)
    # launch shapes: the autotuner's packaged table (loaded with the
    # library), or a fresh search with --retune (untimed setup)
    tune = {"source": "autotuner table " + os.path.basename(L.TUNE_FILE)
            if os.path.exists(L.TUNE_FILE) else "default shapes"}
    if args.retune:
        tune["source"] = "autotuner run"
        for name, op in (("fp32", V32), ("fp64", V64)):
            r = op.autotune(budget_ms=3000.0)
            tune[name + "_search"] = {
                "apply_ms": r["ms"], "apply_ms_default_shape": r["ms_default"],
                "shapes_timed": r["tried"]}
    for name, op in (("fp32", V32), ("fp64", V64)):
        blk, tile, chunks = op.launch()
        if op.schedule == "symmetric":
            ks = op.split
            ng = chunks // ks
            wg = ks * sum(ng - 4 * m for m in range(-(-ng // 4)))
            tune[name] = {"schedule": "symmetric", "split": ks,
                          "groups": ng, "workgroups": wg}
        else:
            tune[name] = {"schedule": "chunked", "block": blk, "tile": tile,
                          "chunks": chunks,
                          "workgroups": chunks * -(-int(T.shape[0]) // blk)}
    b = 0.5 * mesh.mass(u, 1) + K64(u)
    db = L.DeviceArray.from_host(b)
    dx = L.DeviceArray(b.size)
    zeros = np.zeros(b.size)
    solver = L.Solver(V64, V32, restart=50)
    pc = None
    if args.precond == "spai":
        # rows of the inverse of V on 3-ring node patches (about 37 entries
        # per row), applied in FP64 for both operators
        S = L.precond_spai(V64, rings=3)
        solver.set_precond(S, S)
        pc = ("near-field sparse approximate inverse, right: row c of M is row c "
              "of the inverse of V on the nodes within three edges of node c "
              "(V's entries there exact: near-field Sauter-Schwab plus far "
              "quadrature), applied in FP64; the residual is that of the "
              "unpreconditioned system")
    elif args.precond == "calderon":
        # Q (D + m m^T / (4|Gamma|)) Q: D through curls, 3-point Gauss near
        # field, 1-point far rule; Q five Chebyshev steps for the P1 mass
        # inverse.  The FP32 inner solves take D's far field in FP32: GMRES
        # keeps each preconditioned vector (flexible), so P need not be
        # linear beyond its rounding
        P64 = L.precond_calderon(mesh, "fp64", line_rule=2, tri_rule=0)
        P32 = L.precond_calderon(mesh, "fp32", line_rule=2, tri_rule=0)
        solver.set_precond(P64, P32)
        pc = ("Calderon, right: Q (D + m m^T/(4|Gamma|)) Q, D the hypersingular "
              "operator through curls with a 1-point far rule and 3-point Gauss "
              "near rule, Q five Chebyshev steps for the consistent P1 mass "
              "inverse (lumped-mass preconditioned); D's far field in FP32 for the "
              "FP32 inner solves (flexible GMRES); the residual is that of the "
              "unpreconditioned system")

    def solve(mixed=True, s=None):
        L.check(L.lib().ebh_memcpy_h2d(dx.ptr, zeros.ctypes.data, zeros.nbytes))
        return (s or solver).solve_device(db, dx, tol=1e-8, mixed=mixed,
                                          max_iters=3000, inner_tol=args.inner_tol)

    for _ in range(args.warmup):
        solve()
    L.check(L.lib().ebh_synchronize())
    lat, stats = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        st = solve()
        lat.append((time.perf_counter() - ts) * 1e3)
        stats.append(st)
    L.check(L.lib().ebh_synchronize())
    wall = time.perf_counter() - t0
    if not all(s["converged"] and s["relres"] <= 1e-8 for s in stats):
        log("solve did not converge:", stats)
        sys.exit(1)
    solves_per_s = args.steps / wall

    # matvec throughput of both far-field precisions
    xv = L.DeviceArray.from_host(np.random.default_rng(1).uniform(-1, 1, b.size))
    yv = L.DeviceArray(b.size)
    w = V32.work()
    t32 = V32.time(xv, yv, reps=5)
    t64 = V64.time(xv, yv, reps=3)
    # far-kernel and near-field-pass launch times with HIP events on the
    # operator's stream
    L.op_profile(V32, True, 64)
    for _ in range(10):
        V32.apply(xv, yv)
    near_ms, near_launches = L.op_profile_near(V32)
    far_ms, far_launches = L.op_profile(V32, False)
    far_flop = FLOP_SL * w["far_interactions"]
    # the pass after the far kernel (partial reduction and near field),
    # with HIP events around it on the operator's stream: the default
    # separate passes, then the fused pass on the same operator
    near = {}
    if V32.schedule == "symmetric":
        prow = V32.launch()[2]
        for fz in (False, True):
            V32.set_reduction(fz)
            tn = V32.time(xv, yv, reps=10)["near"]
            b_ = near_pass_bytes(w, int(T.shape[0]), prow, fz)
            near[fz] = {"ms_per_apply": tn, "bytes_per_apply": b_,
                        "achieved": b_ / (tn * 1e-3) / 1e9}
        V32.set_reduction(False)

    unpc = {}
    if pc is not None and not args.no_fp64_compare:
        # the same solve without the preconditioner, for comparison
        plain = L.Solver(V64, V32, restart=50)
        solve(s=plain)
        ts = time.perf_counter()
        for _ in range(3):
            stp = solve(s=plain)
        tp = (time.perf_counter() - ts) / 3
        unpc = {"unpreconditioned": {"solves_per_s": 1.0 / tp, "ms": tp * 1e3,
                                     "iters": stp["iters"],
                                     "matvec32": stp["matvec32"],
                                     "matvec64": stp["matvec64"]}}
        plain.close()

    out_fp64 = {}
    if not args.no_fp64_compare:
        solve(mixed=False)
        n64 = 2
        ts = time.perf_counter()
        for _ in range(n64):
            st64 = solve(mixed=False)
        t64s = (time.perf_counter() - ts) / n64
        out_fp64 = {"fp64_solves_per_s": 1.0 / t64s,
                    "fp64_iters": st64["iters"],
                    "mixed_speedup_vs_fp64": solves_per_s * t64s}
        # CPU-equivalent operator applies of the FP64 solve: a Calderon
        # preconditioner apply is 3 single-layer applies with 1 of the 49 far
        # interactions (the sparse inverse is negligible)
        if args.precond == "calderon":
            st64["matvec64"] += st64.get("precond", 0) * 3.0 / 49.0
    else:
        st64 = {"matvec64": stats[-1]["iters"]}

    avg_far_ms = far_ms / max(far_launches, 1)
    sym32 = V32.schedule == "symmetric"
    achieved = far_flop / (avg_far_ms * 1e-3) / 1e12
    lat_sorted = np.sort(lat)
    res = {
        "metric": "solves/sec to 1e-8 FP64 residual (GMRES(50), mixed-precision "
                  "iterative refinement)",
        "value": solves_per_s,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 far field / f64 near field, Krylov and residual",
        "data": "synthetic",
        "config": {"workload": f"icosphere L{args.level} P1 Laplace Dirichlet "
                               "solve, u=1/|x-(0,0,2)|, b=(M/2+K)u, "
                               "GMRES(50) mixed IR to 1e-8",
                   "preconditioner": pc,
                   "triangles": int(T.shape[0]), "vertices": int(P.shape[0]),
                   "near_rule": "5-point Gauss (line index 4)",
                   "far_rule": "7-point triangle (index 4)"},
        "p50_ms": float(np.percentile(lat_sorted, 50)),
        "p99_ms": float(np.percentile(lat_sorted, 99)),
        "iters": stats[-1]["iters"], "outer": stats[-1]["outer"],
        "matvec32_per_solve": stats[-1]["matvec32"],
        "matvec64_per_solve": stats[-1]["matvec64"],
        "precond_applies_per_solve": stats[-1]["precond"],
        **out_fp64,
        **unpc,
        "matvec_ms": {"fp32": t32["total"], "fp64": t64["total"]},
        "autotune": tune,
        "matvec_far_tflops": {"fp32": far_flop / (t32["far"] * 1e-3) / 1e12,
                              "fp64": far_flop / (t64["far"] * 1e-3) / 1e12},
        "roofline": {"kernel": ("k_far_sym<float, 7 points>" if sym32 else
                                "k_far<float, single layer, 7 points>"),
                     "bound": "valu", "achieved": achieved,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": (TRAFFIC_FAR32_L5_BYTES if args.level == 5
                                 and (V32.schedule, V32.split) == TRAFFIC_SCHEDULE
                                 else None),
                     "traffic_source": TRAFFIC_SOURCE,
                     "flop_per_launch": far_flop,
                     "avg_launch_ms": avg_far_ms, "launches": far_launches},
        "near_field": ({
            "kernel": "k_reduce_wide + k_finish (partial sums reduced, then near CSR x "
                      "plus the gathered sums; FP64)",
            "bound": "hbm", "achieved": near[False]["achieved"], "peak": PEAK_HBM_GBPS,
            "unit": "GB/s", "frac": near[False]["achieved"] / PEAK_HBM_GBPS,
            "bytes_per_apply": near[False]["bytes_per_apply"],
            "ms_per_apply": near[False]["ms_per_apply"],
            "k_finish_ms": near_ms / max(near_launches, 1),
            "fused": {"kernel": "k_reduce_finish (one launch)",
                      "ms_per_apply": near[True]["ms_per_apply"],
                      "bytes_per_apply": near[True]["bytes_per_apply"],
                      "frac": near[True]["achieved"] / PEAK_HBM_GBPS}}
            if near else None),
    }
    if not args.no_cpu_baseline:
        log("CPU baseline")
        res["cpu_baseline"] = cpu_baseline(P, T, st64.get("matvec64", 1))
    if not args.no_secondary:
        log("secondary configurations")
        res["secondary"] = secondary(L, meshes, solve_batched=True)
    if not args.no_serve and args.serve > 0:
        log("serving")
        res["serving"] = serving(L, meshes, n_requests=args.serve, levels=(4, 5, 6))
        res["serving"]["resident_mesh_graphs"] = resident_graphs(L, meshes)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
