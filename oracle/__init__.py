"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker (never as the measured or shipped path).

``Oracle`` wraps oracle/liboracle.so (the C restatement of the reference
operator, oracle/oracle.c).  ``Reference`` wraps oracle/_ref/libeasybem_ref.so,
the reference's own CPU sources compiled by oracle/Makefile; it exists only
where the reference sources were present at build time.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REFLIB = os.path.join(HERE, "_ref", "libeasybem_ref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build(quiet: bool = True) -> None:
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True,
                   stdout=subprocess.DEVNULL if quiet else None)


def _mesh(nodes, elems):
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    elems = np.ascontiguousarray(elems, dtype=np.int32)
    return nodes, elems, nodes.shape[0], elems.shape[0]


def lumped_areas(nodes, elems):
    """m_n = sum of |T|/3 over the triangles at node n."""
    P = np.asarray(nodes, dtype=np.float64)
    T = np.asarray(elems) - 1
    a = 0.5 * np.linalg.norm(np.cross(P[T[:, 1]] - P[T[:, 0]],
                                      P[T[:, 2]] - P[T[:, 0]]), axis=1)
    m = np.zeros(P.shape[0])
    np.add.at(m, T, a[:, None] / 3)
    return m


class Oracle:
    def __init__(self, path: str = LIB):
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_line_rule.argtypes = [C.c_int, _dp, _dp]
        L.orc_tri_rule.argtypes = [C.c_int, _dp, _dp]
        L.orc_gauss01.argtypes = [C.c_int, _dp, _dp]
        L.orc_laplace_dense.argtypes = [C.c_int, _dp, C.c_int, _ip, C.c_int,
                                        C.c_int, C.c_void_p, C.c_void_p,
                                        C.c_void_p]
        L.orc_kelvin_dense.argtypes = [C.c_int, _dp, C.c_int, _ip, C.c_double,
                                       C.c_int, C.c_int, _dp]
        L.orc_apply.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp,
                                C.c_int, _ip, C.c_int, C.c_int, C.c_double,
                                _dp, C.c_int, _ip, _dp, C.c_int]
        L.orc_pair_block.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _ip,
                                     C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_int, _dp, _ip, _ip]
        L.orc_ss_full.argtypes = [C.c_int, C.c_int, _ip, _dp, C.c_int]
        L.orc_ss_full.restype = C.c_double
        L.orc_ss_table.argtypes = [C.c_int, C.c_int, C.c_int, _dp]
        L.orc_surface_curls.argtypes = [C.c_int, _dp, C.c_int, _ip, _dp]
        L.orc_evaluate.argtypes = [C.c_int, _dp, C.c_int, _ip, C.c_int, _dp,
                                   _dp, _dp, C.c_int, _dp, C.c_int]
        self.L = L

    # quadrature -------------------------------------------------------
    def line_rule(self, idx):
        p = np.zeros(128); w = np.zeros(128)
        n = self.L.orc_line_rule(idx, p, w)
        if n < 0:
            raise ValueError(idx)
        return p[:n].copy(), w[:n].copy()

    def tri_rule(self, idx):
        p = np.zeros(256); w = np.zeros(128)
        n = self.L.orc_tri_rule(idx, p, w)
        if n < 0:
            raise ValueError(idx)
        return p[:2 * n].reshape(n, 2).copy(), w[:n].copy()

    def gauss01(self, n):
        p = np.zeros(n); w = np.zeros(n)
        self.L.orc_gauss01(n, p, w)
        return p, w

    # operators ----------------------------------------------------------
    def laplace_dense(self, nodes, elems, line_idx, tri_idx=None):
        """Dense (V[test,trial], K[test,node], M[test,node]) in P0 x P0 / P0 x P1."""
        tri_idx = line_idx if tri_idx is None else tri_idx
        nodes, elems, np_, ne = _mesh(nodes, elems)
        V = np.zeros((ne, ne)); K = np.zeros((ne, np_)); M = np.zeros((ne, np_))
        rc = self.L.orc_laplace_dense(np_, nodes, ne, elems, line_idx, tri_idx,
                                      V.ctypes.data, K.ctypes.data,
                                      M.ctypes.data)
        if rc:
            raise RuntimeError(rc)
        return V, K, M

    def kelvin_dense(self, nodes, elems, nu, line_idx, tri_idx=None):
        tri_idx = line_idx if tri_idx is None else tri_idx
        nodes, elems, np_, ne = _mesh(nodes, elems)
        V = np.zeros((3 * ne, 3 * ne))
        if self.L.orc_kelvin_dense(np_, nodes, ne, elems, nu, line_idx,
                                   tri_idx, V):
            raise RuntimeError("kelvin_dense")
        return V

    def apply(self, op, test_basis, trial_basis, nodes, elems, x, rows=None,
              line_idx=4, tri_idx=None, nu=0.3, nthreads=0):
        """Row-sampled apply.  op: 'V', 'K' or 'kelvin'.  Returns y[rows]
        (shape (nrows,) or (3, nrows) for Kelvin)."""
        opc = {"V": 0, "K": 1, "kelvin": 2}[op]
        tri_idx = line_idx if tri_idx is None else tri_idx
        nodes, elems, np_, ne = _mesh(nodes, elems)
        ntest = np_ if test_basis else ne
        rows = np.arange(ntest, dtype=np.int32) if rows is None else \
            np.ascontiguousarray(rows, dtype=np.int32)
        nc = 3 if opc == 2 else 1
        y = np.zeros(nc * len(rows))
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        rc = self.L.orc_apply(opc, test_basis, trial_basis, np_, nodes, ne,
                              elems, line_idx, tri_idx, nu, x, len(rows), rows,
                              y, nthreads)
        if rc:
            raise RuntimeError(rc)
        return y.reshape(nc, len(rows)) if nc == 3 else y

    def pair_block(self, op, nodes, elems, i, j, line_idx=4, tri_idx=None,
                   use_tables=True):
        opc = {"V": 0, "K": 1, "kelvin": 2}[op]
        tri_idx = line_idx if tri_idx is None else tri_idx
        nodes, elems, np_, ne = _mesh(nodes, elems)
        out = np.zeros(63)
        oi = np.zeros(3, np.int32); oj = np.zeros(3, np.int32)
        cls = self.L.orc_pair_block(opc, np_, nodes, ne, elems, i, j, line_idx,
                                    tri_idx, int(use_tables), out, oi, oj)
        return cls, out, oi, oj

    def ss_full(self, cls, kind, exps, geom, n):
        return self.L.orc_ss_full(cls, kind, np.asarray(exps, np.int32),
                                  np.ascontiguousarray(geom, np.float64), n)

    def curl(self, nodes, elems):
        """(C1, C2, C3), each dense [triangle, node], restating
        getLaplace3dCurl (src/BEM3dLaplace.cpp:489-558)."""
        nodes, elems, np_, ne = _mesh(nodes, elems)
        cu = np.zeros(9 * ne)
        self.L.orc_surface_curls(np_, nodes, ne, elems, cu)
        cu = cu.reshape(ne, 3, 3)
        Cs = [np.zeros((ne, np_)) for _ in range(3)]
        rows = np.arange(ne)
        for c in range(3):
            for k in range(3):
                Cs[c][rows, elems[:, k] - 1] = cu[:, c, k]
        return tuple(Cs)

    def evaluate(self, nodes, elems, points, t, u, tri_idx=4, nthreads=0):
        """evaluateLaplace3d (src/BEM3dLaplace.cpp:784-843)."""
        nodes, elems, np_, ne = _mesh(nodes, elems)
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(pts.shape[0])
        rc = self.L.orc_evaluate(np_, nodes, ne, elems, pts.shape[0], pts,
                                 np.ascontiguousarray(t, dtype=np.float64),
                                 np.ascontiguousarray(u, dtype=np.float64),
                                 tri_idx, out, nthreads)
        if rc:
            raise RuntimeError(rc)
        return out

    def hypersingular_apply(self, nodes, elems, x, line_idx=4, tri_idx=None,
                            stab=0.0):
        """y = sum_c C_c^T V C_c x (+ stab m (m.x)): the dense D of
        getLaplace3d (src/BEM3dLaplace.cpp:652,769) applied through three P0
        single-layer applies; m = lumped node areas as in
        python/examples/laplace3dNeumannGPU.py:36-49."""
        Cs = self.curl(nodes, elems)
        y = np.zeros(np.asarray(nodes).shape[0])
        for Cc in Cs:
            y += Cc.T @ self.apply("V", 0, 0, nodes, elems, Cc @ x,
                                   line_idx=line_idx, tri_idx=tri_idx).ravel()
        if stab:
            m = lumped_areas(nodes, elems)
            y += stab * m * (m @ x)
        return y

    def ss_table(self, cls, kdeg, nq):
        nreg = {1: 2, 2: 5, 3: 3}[cls]
        nout = {1: nq ** 3, 2: nq ** 2, 3: nq}[cls]
        W = np.zeros(9 * nreg * nout)
        self.L.orc_ss_table(cls, kdeg, nq, W)
        return W.reshape(nreg, nout, 3, 3)


class Reference:
    """The reference's own CPU operator, compiled from its sources."""

    def __init__(self, path: str = REFLIB):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = C.CDLL(path)
        L.ebref_laplace3d.argtypes = [C.c_int, _dp, C.c_int, _ip, C.c_int,
                                      _dp, _dp, _dp, _dp]
        L.ebref_laplace3d_mass.argtypes = [C.c_int, _dp, C.c_int, _ip, _dp]
        L.ebref_laplace3d_curl.argtypes = [C.c_int, _dp, C.c_int, _ip, _dp,
                                           _dp, _dp]
        L.ebref_evaluate_laplace3d.argtypes = [C.c_int, _dp, C.c_int, _ip,
                                               C.c_int, _dp, _dp, _dp, C.c_int,
                                               _dp]
        L.ebref_elasticity3d.argtypes = [C.c_int, _dp, C.c_int, _ip,
                                         C.c_double, C.c_int, _dp, C.c_void_p,
                                         C.c_void_p, C.c_void_p]
        L.ebref_line_rule.argtypes = [C.c_int, _dp, _dp]
        L.ebref_tri_rule.argtypes = [C.c_int, _dp, _dp]
        self.L = L

    @staticmethod
    def available(path: str = REFLIB) -> bool:
        return os.path.exists(path)

    def laplace3d(self, nodes, elems, order):
        """(V[test,trial], K[test,node], D, M[test,node]) like getLaplace3d."""
        nodes, elems, np_, ne = _mesh(nodes, elems)
        V = np.zeros(ne * ne); K = np.zeros(ne * np_); D = np.zeros(np_ * np_)
        M = np.zeros(ne * np_)
        self.L.ebref_laplace3d(np_, nodes, ne, elems, order, V, K, D, M)
        # V is written row-major [test*ne+trial]; K, D, M column-major
        return (V.reshape(ne, ne), K.reshape(np_, ne).T.copy(),
                D.reshape(np_, np_).T.copy(), M.reshape(np_, ne).T.copy())

    def mass(self, nodes, elems):
        nodes, elems, np_, ne = _mesh(nodes, elems)
        M = np.zeros(ne * np_)
        self.L.ebref_laplace3d_mass(np_, nodes, ne, elems, M)
        return M.reshape(np_, ne).T.copy()

    def evaluate(self, nodes, elems, points, t, u, order):
        nodes, elems, np_, ne = _mesh(nodes, elems)
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(pts.shape[0])
        self.L.ebref_evaluate_laplace3d(
            np_, nodes, ne, elems, pts.shape[0], pts,
            np.ascontiguousarray(t, dtype=np.float64),
            np.ascontiguousarray(u, dtype=np.float64), order, out)
        return out

    def curl(self, nodes, elems):
        """(C1, C2, C3), each [triangle, node], like getLaplace3dCurl."""
        nodes, elems, np_, ne = _mesh(nodes, elems)
        Cs = [np.zeros(ne * np_) for _ in range(3)]
        self.L.ebref_laplace3d_curl(np_, nodes, ne, elems, *Cs)
        return tuple(c.reshape(np_, ne).T.copy() for c in Cs)

    def elasticity3d_V(self, nodes, elems, nu, order):
        nodes, elems, np_, ne = _mesh(nodes, elems)
        V = np.zeros(9 * ne * ne)
        self.L.ebref_elasticity3d(np_, nodes, ne, elems, nu, order, V, None,
                                  None, None)
        return V.reshape(3 * ne, 3 * ne).T.copy()  # column-major -> [row, col]

    def elasticity3d(self, nodes, elems, nu, order):
        """(V, K, D, M) like getElasticity3d, each [row, col]."""
        nodes, elems, np_, ne = _mesh(nodes, elems)
        V = np.zeros(9 * ne * ne); K = np.zeros(9 * ne * np_)
        D = np.zeros(9 * np_ * np_); M = np.zeros(9 * ne * np_)
        self.L.ebref_elasticity3d(np_, nodes, ne, elems, nu, order, V,
                                  K.ctypes.data, D.ctypes.data, M.ctypes.data)
        return (V.reshape(3 * ne, 3 * ne).T.copy(),
                K.reshape(3 * np_, 3 * ne).T.copy(),
                D.reshape(3 * np_, 3 * np_).T.copy(),
                M.reshape(3 * np_, 3 * ne).T.copy())

    def line_rule(self, idx):
        p = np.zeros(128); w = np.zeros(128)
        n = self.L.ebref_line_rule(idx, p, w)
        return p[:n].copy(), w[:n].copy()

    def tri_rule(self, idx):
        p = np.zeros(256); w = np.zeros(128)
        n = self.L.ebref_tri_rule(idx, p, w)
        return p[:2 * n].reshape(n, 2).copy(), w[:n].copy()
