"""Parity at the benchmark configurations' full sizes.

The oracle is too slow for whole matvecs at these sizes, so the FP64 HIP
apply is checked against the oracle on seeded random rows (<= 1e-12
relative per row set), and the FP32-far apply against the FP64 apply over
the whole vector (<= 1e-6, 2-norm).  Configurations: icosphere L3 P0 with
the hybrid rule (7-point far, 4-point near), icosphere L5 P1 V and K,
torus 256x128 Kelvin P1 (nu = 0.3), icosphere L6 P0 V.
"""
import numpy as np
import pytest

from easybem_amd import meshes

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


CONFIGS = {
    "ico3_p0_V_hybrid": (lambda: meshes.icosphere(3), "V", 0, 0, 3, 4, 128),
    "ico5_p1_V": (lambda: meshes.icosphere(5), "V", 1, 1, 4, 4, 32),
    "ico5_p1_K": (lambda: meshes.icosphere(5), "K", 1, 1, 4, 4, 32),
    "ico5_p0p1_K": (lambda: meshes.icosphere(5), "K", 0, 1, 4, 4, 64),
    "torus_p1_kelvin": (lambda: meshes.torus(1.0, 0.4, 256, 128), "kelvin", 1, 1,
                        4, 4, 8),
    "ico6_p0_V": (lambda: meshes.icosphere(6), "V", 0, 0, 4, 4, 48),
}


@pytest.mark.parametrize("name", list(CONFIGS))
def test_fullsize_parity(oracle, name):
    from easybem_amd import _lib as L
    mk, kern, tb, trb, li, ti, nrows = CONFIGS[name]
    P, T = mk()
    print(f"case {name}: {T.shape[0]} triangles", flush=True)
    rng = np.random.default_rng(20261203)
    nc = 3 if kern == "kelvin" else 1
    ncols = P.shape[0] if trb else T.shape[0]
    x = rng.standard_normal(nc * ncols) if kern == "kelvin" else \
        rng.uniform(-1, 1, ncols)
    m = L.Mesh(P, T)
    y64 = L.Operator(m, kern, tb, trb, "fp64", line_rule=li, tri_rule=ti,
                     nu=0.3)(x)
    y32 = L.Operator(m, kern, tb, trb, "fp32", line_rule=li, tri_rule=ti,
                     nu=0.3)(x)
    e32 = rel(y32, y64)
    nt = P.shape[0] if tb else T.shape[0]
    rows = np.sort(rng.choice(nt, size=nrows, replace=False)).astype(np.int32)
    yo = oracle.apply(kern, tb, trb, P, T, x, rows=rows, line_idx=li,
                      tri_idx=ti, nu=0.3)
    if nc == 3:
        y64r = np.stack([y64[c * nt + rows] for c in range(3)])
    else:
        y64r = y64[rows]
    e64 = rel(y64r.ravel(), np.asarray(yo).ravel())
    print(f"{name}: fp64 vs oracle (rows) {e64:.2e}, fp32 vs fp64 {e32:.2e}")
    assert e64 <= 1e-12
    assert e32 <= 1e-6
