"""The allocation-free evaluation entry point (ebh_evaluate_laplace_noalloc)
against ebh_evaluate_laplace, both through the C-ABI: identical results on
two calls in a row, the second with fewer points (the shared workspace keeps
the first call's contents), in both precisions; sizes beyond the workspace
are rejected on the host."""
import ctypes as C

import numpy as np
import pytest

from easybem_amd import meshes

pytestmark = pytest.mark.gpu


def _entry_points(L):
    lib = L.lib()
    vp = C.c_void_p
    sig = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp]
    lib.ebh_evaluate_laplace.argtypes = sig
    lib.ebh_evaluate_laplace_noalloc.argtypes = sig
    return lib.ebh_evaluate_laplace, lib.ebh_evaluate_laplace_noalloc


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_evaluation_matches_allocating_entry_point(prec):
    from easybem_amd import _lib as L
    ev, ev0 = _entry_points(L)
    P, T = meshes.icosphere(3)
    m = L.Mesh(P, T)
    rng = np.random.default_rng(30 + prec)
    t = L.DeviceArray.from_host(rng.standard_normal(T.shape[0]))
    u = L.DeviceArray.from_host(rng.standard_normal(P.shape[0]))
    for npts, seed in ((3000, 1), (700, 2)):
        print(f"case noalloc prec={prec} npts={npts}", flush=True)
        d = np.random.default_rng(seed).standard_normal((npts, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        pts = d * np.random.default_rng(seed + 9).uniform(1.2, 3.0, (npts, 1))
        dp = L.DeviceArray.from_host(pts)
        y_ref, y_new = L.DeviceArray(npts), L.DeviceArray(npts)
        L.check(ev(m.h, 4, prec, npts, dp.ptr, t.ptr, u.ptr, y_ref.ptr, None))
        L.check(ev0(m.h, 4, prec, npts, dp.ptr, t.ptr, u.ptr, y_new.ptr, None))
        L.check(L.lib().ebh_synchronize())
        a, b = y_ref.to_host(), y_new.to_host()
        assert np.all(np.isfinite(b))
        assert np.array_equal(a, b)


def test_noalloc_rejects_sizes_beyond_the_workspace():
    from easybem_amd import _lib as L
    _, ev0 = _entry_points(L)
    P, T = meshes.icosphere(3)
    m = L.Mesh(P, T)
    npts = 9_000_000  # 72 MB of partial sums alone: more than the workspace
    t = L.DeviceArray(T.shape[0])
    u = L.DeviceArray(P.shape[0])
    # the size check comes before any launch; the buffers are full size all
    # the same
    dp, y = L.DeviceArray(3 * npts), L.DeviceArray(npts)
    rc = ev0(m.h, 4, 0, npts, dp.ptr, t.ptr, u.ptr, y.ptr, None)
    assert rc == -1 and b"workspace" in L.lib().ebh_last_error()
