"""Host-side logic through the C-ABI (no GPU): library symbols, mesh
validation and pair classification, bench contract."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from easybem_amd import meshes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def L():
    from easybem_amd import _lib
    _lib.lib()
    return _lib


def declared_symbols():
    syms = []
    for h in ("easybem_hip.h",):
        txt = open(os.path.join(ROOT, "include", h)).read()
        syms += re.findall(r"\b(ebh_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(syms))


def test_library_exports_every_declared_symbol(L):
    lib = L.lib()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_legacy_cpp_surface_exported():
    out = subprocess.run(["nm", "-DC", os.path.join(ROOT, "easybem_amd",
                                                     "libeasybem_hip.so")],
                         capture_output=True, text=True).stdout
    for sym in ["initGPU(int, double const*, int, int const*, int, int, int)",
                "closeGPU()",
                "applyGPULaplace3d1Layer(double const*, double*, int)",
                "applyGPULaplace3d2Layer(double const*, double*)",
                "applyGPULaplace3dHyperSing(double const*, double*)"]:
        assert sym in out, sym


@pytest.mark.parametrize("lev,far,nnz", [
    (3, 1621820, (16580, 15300, 15300, 12102)),
    (5, 419164220, (266180, 245700, 245700, 194502)),
])
def test_icosphere_pair_counts(L, lev, far, nnz):
    P, T = meshes.icosphere(lev)
    d = L.mesh_check(P, T)
    ne = T.shape[0]
    assert d["coincident"] == ne and d["edge"] == 3 * ne
    # far = ne (ne - 4) - sum_v d_v (d_v - 3)  (src/meshTopology.cpp:258-260)
    deg = np.bincount(T.ravel() - 1)
    assert d["vertex"] == np.sum(deg * (deg - 3))
    assert d["far"] == far == ne * (ne - 4) - d["vertex"]
    assert tuple(d["nnz"].values()) == nnz


def test_torus_pair_counts(L):
    P, T = meshes.torus()
    d = L.mesh_check(P, T)
    assert (d["coincident"], d["edge"], d["vertex"], d["far"]) == \
        (65536, 196608, 589824, 4294115328)
    assert d["max_touch"] == 13


@pytest.mark.parametrize("mutate,msg", [
    (lambda P, T: (P, T[:-1]), "closed 2-manifold"),
    (lambda P, T: (P, np.vstack([T[:1, [0, 2, 1]], T[1:]])), "orientation"),
    (lambda P, T: (P, np.where(T == 5, 999, T)), "outside"),
    (lambda P, T: (np.where(np.arange(P.size).reshape(P.shape) == 4, np.nan, P), T),
     "non-finite"),
    (lambda P, T: (P, np.vstack([[[1, 1, 2]], T[1:]])), "repeats"),
])
def test_mesh_rejections(L, mutate, msg):
    P, T = meshes.icosphere(1)
    P2, T2 = mutate(P.copy(), T.copy())
    with pytest.raises(L.EbhError) as e:
        L.mesh_check(P2, T2)
    assert e.value.code == -2 and msg in str(e.value)


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                        "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stdout


def test_tuning_table_round_trip(L, tmp_path):
    """The autotuner's table persists as text and rejects malformed lines
    (host-only: no device call)."""
    saved = tmp_path / "saved.txt"
    L.tune_save(str(saved))
    L.tune_clear()
    src = tmp_path / "in.txt"
    src.write_text("gfx950x256/k0/p1/nt7/test1/rb0/ne29 256 64 4096 4096 3.2 2\n"
                   "gfx950x256/k0/p0/nt7/test1/rb0/ne29 128 32 8192 0 10.5\n")
    L.tune_load(str(src))
    out = tmp_path / "out.txt"
    L.tune_save(str(out))
    lines = sorted(out.read_text().split("\n"))
    # a six-column line (no schedule column) reads as the chunked schedule
    assert lines[1:] == sorted(
        ["gfx950x256/k0/p1/nt7/test1/rb0/ne29 256 64 4096 4096 3.2 2",
         "gfx950x256/k0/p0/nt7/test1/rb0/ne29 128 32 8192 0 10.5 0"])
    for text in ["key 96 64 4096 0 1.0\n", "key 256 64 4096 0 1.0 3\n",
                 "key 256 64 4096 0 1.0 128\n"]:
        bad = tmp_path / "bad.txt"
        bad.write_text(text)
        with pytest.raises(L.EbhError):
            L.tune_load(str(bad))
    L.tune_clear()
    L.tune_save(str(out))
    assert out.read_text() == ""
    L.tune_load(str(saved))


def test_packaged_tuning_table_is_well_formed(L, tmp_path):
    from easybem_amd import _lib
    if not os.path.exists(_lib.TUNE_FILE):
        pytest.skip("no packaged table")
    saved = tmp_path / "saved.txt"
    L.tune_save(str(saved))
    L.tune_clear()
    L.tune_load(_lib.TUNE_FILE)
    out = tmp_path / "out.txt"
    L.tune_save(str(out))
    assert len(out.read_text().split()) >= 6
    L.tune_clear()
    L.tune_load(str(saved))


def test_dropin_module_curls_match_reference():
    """easybem.get_laplace3d_curl (host only) against getLaplace3dCurl."""
    sys.path.insert(0, os.path.join(ROOT, "easybem_amd"))
    import easybem
    z = np.load(os.path.join(ROOT, "tests", "golden", "laplace_torus8x4_o4.npz"))
    Cs = easybem.get_laplace3d_curl(z["nodes"], z["elements"].astype(np.int32))
    assert len(Cs) == 3
    for c in range(3):
        assert Cs[c].shape == z["C1"].shape
        assert np.array_equal(Cs[c], z[f"C{c + 1}"])
    with pytest.raises(ValueError):
        easybem.get_laplace3d_curl(z["nodes"], z["elements"].astype(np.int32) + 1000)
