"""ctypes binding of the C-ABI in include/easybem_hip.h.

The HIP library is required: importing this module raises if
libeasybem_hip.so has not been built (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(PKG, "libeasybem_hip.so")
TUNE_FILE = os.path.join(PKG, "tuning_mi355x.txt")

KERNELS = {"V": 0, "K": 1, "kelvin": 2, "D": 3, "Kt": 4, "kelvin_K": 5,
           "kelvin_D": 6}
EBH_OK = 0


class EbhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"easybem-hip error {code}: {msg}")
        self.code = code


class OpDesc(C.Structure):
    _fields_ = [("kernel", C.c_int), ("test_basis", C.c_int),
                ("trial_basis", C.c_int), ("far_precision", C.c_int),
                ("line_rule", C.c_int), ("tri_rule", C.c_int),
                ("nu", C.c_double), ("max_rhs", C.c_int), ("stab", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(
                f"{LIBPATH} is missing: build it with `python -m easybem_amd.build`")
        L = C.CDLL(LIBPATH)
        vp, i64p = C.c_void_p, C.POINTER(C.c_int64)
        L.ebh_last_error.restype = C.c_char_p
        L.ebh_malloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.ebh_free.argtypes = [vp]
        L.ebh_memcpy_h2d.argtypes = [vp, vp, C.c_size_t]
        L.ebh_memcpy_d2h.argtypes = [vp, vp, C.c_size_t]
        L.ebh_memset.argtypes = [vp, C.c_int, C.c_size_t]
        L.ebh_device_info.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_size_t),
                                      C.POINTER(C.c_size_t)]
        L.ebh_mesh_create.argtypes = [C.c_int, vp, C.c_int, vp,
                                      C.POINTER(C.c_void_p)]
        L.ebh_mesh_info.argtypes = [vp, i64p]
        L.ebh_mesh_destroy.argtypes = [vp]
        L.ebh_mesh_destroy.restype = None
        L.ebh_op_create.argtypes = [vp, C.POINTER(OpDesc), C.POINTER(C.c_void_p)]
        L.ebh_op_dims.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.ebh_op_apply.argtypes = [vp, vp, vp, C.c_int, vp]
        L.ebh_op_apply_host.argtypes = [vp, vp, vp, C.c_int]
        L.ebh_op_set_launch.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.ebh_op_get_launch.argtypes = [vp, C.POINTER(C.c_int),
                                        C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.ebh_op_set_schedule.argtypes = [vp, C.c_int, C.c_int]
        L.ebh_op_get_schedule.argtypes = [vp, C.POINTER(C.c_int),
                                          C.POINTER(C.c_int)]
        L.ebh_op_set_reduction.argtypes = [vp, C.c_int]
        L.ebh_op_set_partial_limit.argtypes = [vp, C.c_int64]
        L.ebh_op_get_partial.argtypes = [vp, i64p, C.POINTER(C.c_int), i64p]
        L.ebh_op_get_reduction.argtypes = [vp, C.POINTER(C.c_int)]
        L.ebh_op_set_mrhs_schedule.argtypes = [vp, C.c_int]
        L.ebh_op_get_mrhs_schedule.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.ebh_op_set_resident_tiles.argtypes = [vp, C.c_int64]
        L.ebh_op_get_resident_tiles.argtypes = [vp, C.POINTER(C.c_int), i64p]
        L.ebh_op_time.argtypes = [vp, vp, vp, C.c_int, C.c_int,
                                  C.POINTER(C.c_double)]
        L.ebh_op_work.argtypes = [vp, i64p]
        L.ebh_op_destroy.argtypes = [vp]
        L.ebh_op_destroy.restype = None
        L.ebh_mass_apply.argtypes = [vp, C.c_int, vp, vp]
        _lib = L
        # packaged launch shapes from the autotuner (EBH_TUNE_FILE overrides
        # the path; an empty value starts with an empty table)
        tf = os.environ.get("EBH_TUNE_FILE", TUNE_FILE)
        if tf and os.path.exists(tf):
            L.ebh_tune_load.argtypes = [C.c_char_p]
            check(L.ebh_tune_load(os.fsencode(tf)))
    return _lib


def check(rc):
    if rc != EBH_OK:
        raise EbhError(rc, lib().ebh_last_error().decode())
    return rc


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class DeviceArray:
    """A float64 device buffer (ebh_malloc) with host copies."""

    def __init__(self, n):
        self.n = int(n)
        p = C.c_void_p()
        check(lib().ebh_malloc(C.byref(p), max(1, self.n) * 8))
        self.ptr = p

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        d = cls(a.size)
        check(lib().ebh_memcpy_h2d(d.ptr, _ptr(a), a.nbytes))
        return d

    def to_host(self):
        out = np.empty(self.n)
        check(lib().ebh_memcpy_d2h(_ptr(out), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().ebh_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_info():
    n, cu, fr, tot = C.c_int(), C.c_int(), C.c_size_t(), C.c_size_t()
    check(lib().ebh_device_info(C.byref(n), C.byref(cu), C.byref(fr),
                                C.byref(tot)))
    return {"devices": n.value, "cu": cu.value, "free": fr.value,
            "total": tot.value}


class Mesh:
    """Device-resident mesh (ebh_mesh).  elements are 1-based int32."""

    def __init__(self, nodes, elements):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self.elements = np.ascontiguousarray(elements, dtype=np.int32)
        h = C.c_void_p()
        check(lib().ebh_mesh_create(self.nodes.shape[0], _ptr(self.nodes),
                                    self.elements.shape[0], _ptr(self.elements),
                                    C.byref(h)))
        self.h = h

    def info(self):
        a = (C.c_int64 * 8)()
        check(lib().ebh_mesh_info(self.h, a))
        keys = ["np", "ne", "coincident", "edge", "vertex", "far",
                "max_touch", "bytes"]
        return dict(zip(keys, list(a)))

    def mass(self, u, test_basis=0):
        u = np.ascontiguousarray(u, dtype=np.float64)
        n = self.nodes.shape[0] if test_basis else self.elements.shape[0]
        y = np.empty(n)
        check(lib().ebh_mass_apply(self.h, test_basis, _ptr(u), _ptr(y)))
        return y

    def close(self):
        if getattr(self, "h", None):
            lib().ebh_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Operator:
    """Matrix-free layer-potential operator (ebh_op) on a Mesh."""

    def __init__(self, mesh: Mesh, kernel="V", test_basis=0, trial_basis=0,
                 precision="fp64", line_rule=4, tri_rule=None, nu=0.3,
                 max_rhs=1, stab=0.0):
        self.mesh = mesh
        d = OpDesc(KERNELS[kernel], int(test_basis), int(trial_basis),
                   0 if precision == "fp64" else 1, int(line_rule),
                   int(line_rule if tri_rule is None else tri_rule), float(nu),
                   int(max_rhs), float(stab))
        self.desc = d
        h = C.c_void_p()
        check(lib().ebh_op_create(mesh.h, C.byref(d), C.byref(h)))
        self.h = h
        r, c = C.c_int(), C.c_int()
        check(lib().ebh_op_dims(h, C.byref(r), C.byref(c)))
        self.nrows, self.ncols = r.value, c.value

    def apply_host(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.size == 0 or x.size % self.ncols:
            raise ValueError(f"input of {x.size} entries is not a whole number "
                             f"of {self.ncols}-entry vectors")
        nrhs = x.size // self.ncols
        y = np.empty(nrhs * self.nrows)
        check(lib().ebh_op_apply_host(self.h, _ptr(x), _ptr(y), nrhs))
        return y if nrhs > 1 or x.ndim == 1 else y

    __call__ = apply_host

    def apply(self, x: DeviceArray, y: DeviceArray, nrhs=1, stream=None):
        if nrhs < 1 or x.n < nrhs * self.ncols or y.n < nrhs * self.nrows:
            raise ValueError(f"apply of {nrhs} vectors needs x of "
                             f"{nrhs * self.ncols} and y of {nrhs * self.nrows} "
                             f"entries (got {x.n} and {y.n})")
        check(lib().ebh_op_apply(self.h, x.ptr, y.ptr, nrhs, stream))

    def set_launch(self, block=0, tile=0, chunks=0):
        check(lib().ebh_op_set_launch(self.h, block, tile, chunks))

    def launch(self):
        b, t, c = C.c_int(), C.c_int(), C.c_int()
        check(lib().ebh_op_get_launch(self.h, C.byref(b), C.byref(t),
                                      C.byref(c)))
        return b.value, t.value, c.value

    SCHEDULES = {"chunked": 0, "symmetric": 1}

    def set_schedule(self, schedule, split=0):
        """Far-field schedule: "chunked" (every ordered triangle pair, the
        shape of set_launch) or "symmetric" (each unordered pair of 64-triangle
        groups once; V with equal test and trial bases only).  split: step
        ranges per group pair of the symmetric schedule (0: default)."""
        check(lib().ebh_op_set_schedule(self.h, self.SCHEDULES[schedule],
                                        int(split)))

    @property
    def schedule(self):
        v = C.c_int()
        check(lib().ebh_op_get_schedule(self.h, C.byref(v), None))
        return {0: "chunked", 1: "symmetric"}[v.value]

    @property
    def split(self):
        v, k = C.c_int(), C.c_int()
        check(lib().ebh_op_get_schedule(self.h, C.byref(v), C.byref(k)))
        return k.value

    def set_reduction(self, fused=False):
        """Symmetric schedule: reduce the group partial sums in a kernel of
        their own before the near-field pass (the default) or inside it
        (fused); y is bitwise the same."""
        check(lib().ebh_op_set_reduction(self.h, int(bool(fused))))

    def set_partial_limit(self, nbytes=0):
        """Bound on the symmetric schedule's partial sums in bytes (0: the
        default 4 GiB); above it the schedule runs in passes."""
        check(lib().ebh_op_set_partial_limit(self.h, int(nbytes)))

    @property
    def partial(self):
        """(limit, passes, partial buffer bytes) of the far schedule."""
        lim, nb = C.c_int64(), C.c_int64()
        npass = C.c_int()
        check(lib().ebh_op_get_partial(self.h, C.byref(lim), C.byref(npass), C.byref(nb)))
        return {"limit": lim.value, "passes": npass.value, "bytes": nb.value}

    MRHS_SCHEDULES = {"auto": 0, "target": 1, "pairs": 2}

    def set_mrhs_schedule(self, schedule="auto"):
        """Multi-RHS schedule of a P0 single layer: "target" (every ordered
        triangle pair), "pairs" (each unordered pair's point sums once, in
        64 x 64 tiles applied in both directions) or "auto" (pairs where it
        applies)."""
        check(lib().ebh_op_set_mrhs_schedule(self.h, self.MRHS_SCHEDULES[schedule]))

    @property
    def mrhs_schedule(self):
        """(schedule in effect, its passes) of multi-RHS applies."""
        v, k = C.c_int(), C.c_int()
        check(lib().ebh_op_get_mrhs_schedule(self.h, C.byref(v), C.byref(k)))
        return {1: "target", 2: "pairs"}[v.value], k.value

    def set_resident_tiles(self, max_bytes):
        """Keep the "pairs" schedule's tiles in device memory (formed once,
        now) when they and the item sums fit in max_bytes; 0 turns it off.
        The far field is then stored, not recomputed per apply."""
        check(lib().ebh_op_set_resident_tiles(self.h, int(max_bytes)))

    @property
    def resident_tiles(self):
        """(tiles resident, their bytes)."""
        v, nb = C.c_int(), C.c_int64()
        check(lib().ebh_op_get_resident_tiles(self.h, C.byref(v), C.byref(nb)))
        return bool(v.value), nb.value

    @property
    def fused(self):
        """Whether the fused reduction and near-field pass runs."""
        v = C.c_int()
        check(lib().ebh_op_get_reduction(self.h, C.byref(v)))
        return bool(v.value)

    def time(self, x: DeviceArray, y: DeviceArray, reps=10, nrhs=1):
        ms = (C.c_double * 4)()
        check(lib().ebh_op_time(self.h, x.ptr, y.ptr, nrhs, reps, ms))
        return {"fold": ms[0], "far": ms[1], "near": ms[2], "total": ms[3]}

    def autotune(self, budget_ms=2000.0):
        """Search far-field launch shapes (ebh_op_autotune); the fastest is
        kept on this operator and recorded for later operators of its key."""
        L = lib()
        L.ebh_op_autotune.argtypes = [C.c_void_p, C.c_double,
                                      C.POINTER(TuneResult)]
        r = TuneResult()
        check(L.ebh_op_autotune(self.h, float(budget_ms), C.byref(r)))
        return {k: getattr(r, k) for k, _ in TuneResult._fields_}

    def work(self):
        w = (C.c_int64 * 4)()
        check(lib().ebh_op_work(self.h, w))
        return {"far_interactions": w[0], "near_nnz": w[1], "rows": w[2],
                "cols": w[3]}

    def close(self):
        if getattr(self, "h", None):
            lib().ebh_op_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SolveOpts(C.Structure):
    _fields_ = [("restart", C.c_int), ("max_iters", C.c_int),
                ("tol", C.c_double), ("mixed", C.c_int),
                ("inner_tol", C.c_double), ("max_outer", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("outer", C.c_int), ("matvec64", C.c_int),
                ("matvec32", C.c_int), ("relres", C.c_double),
                ("ms", C.c_double), ("converged", C.c_int),
                ("precond", C.c_int), ("graph_steps", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _solver_sigs(L):
    if getattr(L, "_solver_sigs", False):
        return
    vp = C.c_void_p
    L.ebh_solver_create.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_void_p)]
    L.ebh_solver_create_batched.argtypes = [vp, vp, C.c_int, C.c_int,
                                            C.POINTER(C.c_void_p)]
    L.ebh_solver_set.argtypes = [vp, C.c_int, C.c_int]
    L.ebh_solver_solve.argtypes = [vp, vp, vp, C.POINTER(SolveOpts),
                                   C.POINTER(SolveStats)]
    L.ebh_solver_solve_batched.argtypes = [vp, C.c_int, vp, vp,
                                           C.POINTER(SolveOpts),
                                           C.POINTER(SolveStats)]
    L.ebh_solver_destroy.argtypes = [vp]
    L.ebh_solver_destroy.restype = None
    L._solver_sigs = True


class Solver:
    """GMRES(restart) on device for up to `batch` systems sharing the
    operator; mixed-precision refinement when a32 is given."""

    def __init__(self, a64: Operator, a32: Operator = None, restart=50,
                 batch=1, graphs=True, check_every=0):
        L = lib()
        _solver_sigs(L)
        h = C.c_void_p()
        check(L.ebh_solver_create_batched(a64.h, a32.h if a32 else None,
                                          restart, batch, C.byref(h)))
        self.h = h
        self.n = a64.nrows
        self.a64, self.a32 = a64, a32
        self.restart = restart
        self.batch = batch
        check(L.ebh_solver_set(h, int(graphs), check_every))

    def set(self, graphs=True, check_every=0):
        check(lib().ebh_solver_set(self.h, int(graphs), check_every))

    def set_block(self, block=True):
        """Batched solves in one shared block Krylov space (block GMRES)
        instead of independent Krylov spaces in lockstep."""
        L = lib()
        L.ebh_solver_set_block.argtypes = [C.c_void_p, C.c_int]
        check(L.ebh_solver_set_block(self.h, int(block)))

    def set_precond(self, p64=None, p32=None):
        """Right preconditioners for the FP64 / FP32 operator (None: off)."""
        L = lib()
        L.ebh_solver_set_precond.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        check(L.ebh_solver_set_precond(self.h, p64.h if p64 else None,
                                       p32.h if p32 else None))
        self.p64, self.p32 = p64, p32

    def solve_device(self, b: DeviceArray, x: DeviceArray, tol=1e-8,
                     mixed=False, max_iters=2000, inner_tol=1e-5,
                     max_outer=10, raise_on_noconv=True, nb=1):
        if nb < 1 or nb > self.batch or b.n < nb * self.n or x.n < nb * self.n:
            raise ValueError(f"solve of {nb} systems (batch {self.batch}) needs "
                             f"b and x of {nb * self.n} entries (got {b.n} and "
                             f"{x.n})")
        o = SolveOpts(self.restart, max_iters, tol, int(mixed), inner_tol,
                      max_outer)
        st = (SolveStats * nb)()
        rc = lib().ebh_solver_solve_batched(self.h, nb, b.ptr, x.ptr,
                                            C.byref(o), st)
        if rc != EBH_OK and (raise_on_noconv or rc != -6):
            check(rc)
        out = [s.as_dict() for s in st]
        return out[0] if nb == 1 else out

    def solve(self, b, x0=None, **kw):
        b = np.ascontiguousarray(b, dtype=np.float64)
        if b.size == 0 or b.size % self.n:
            raise ValueError(f"right-hand side of {b.size} entries is not a "
                             f"whole number of {self.n}-entry vectors")
        nb = b.size // self.n
        db = DeviceArray.from_host(b)
        dx = DeviceArray.from_host(np.zeros_like(b) if x0 is None else x0)
        st = self.solve_device(db, dx, nb=nb, **kw)
        return dx.to_host().reshape(b.shape), st

    def close(self):
        if getattr(self, "h", None):
            lib().ebh_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def precond_calderon(mesh: Mesh, precision="fp32", line_rule=2, tri_rule=0,
                     max_rhs=1):
    """Calderon preconditioner Q (D + m m^T/(4|Gamma|)) Q for the P1 single
    layer, Q a 5-step Chebyshev approximation of the consistent P1 mass
    inverse (ebh_precond_calderon), as an Operator."""
    L = lib()
    L.ebh_precond_calderon.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int,
                                       C.c_int, C.POINTER(C.c_void_p)]
    h = C.c_void_p()
    check(L.ebh_precond_calderon(mesh.h, 0 if precision == "fp64" else 1,
                                 int(line_rule), int(tri_rule), int(max_rhs),
                                 C.byref(h)))
    op = Operator.__new__(Operator)
    op.mesh = mesh
    op.h = h
    op.desc = None
    r, c = C.c_int(), C.c_int()
    check(L.ebh_op_dims(h, C.byref(r), C.byref(c)))
    op.nrows, op.ncols = r.value, c.value
    return op


def precond_spai(V: Operator, rings=None, max_rhs=1):
    """Near-field sparse approximate inverse of the single layer V
    (ebh_precond_spai) on patches of `rings` adjacency steps (default: 3
    for P1, 2 for P0), as an Operator applying y = M x in FP64; use it as
    both preconditioners of a solver."""
    L = lib()
    L.ebh_precond_spai.argtypes = [C.c_void_p, C.c_int, C.c_int,
                                   C.POINTER(C.c_void_p)]
    if rings is None:
        rings = 3 if V.nrows == V.mesh.nodes.shape[0] else 2
    h = C.c_void_p()
    check(L.ebh_precond_spai(V.h, int(rings), int(max_rhs), C.byref(h)))
    op = Operator.__new__(Operator)
    op.mesh = V.mesh
    op.desc = None
    op.h = h
    op.nrows = op.ncols = V.nrows
    return op


def precond_calderon_hs(mesh: Mesh, precision="fp32", line_rule=3, tri_rule=1):
    """Dual Calderon preconditioner Q V Q for the P1 hypersingular
    operator, Q a 5-step Chebyshev approximation of the consistent P1 mass
    inverse (ebh_precond_calderon_hs), as an Operator."""
    L = lib()
    L.ebh_precond_calderon_hs.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int,
                                          C.POINTER(C.c_void_p)]
    h = C.c_void_p()
    check(L.ebh_precond_calderon_hs(mesh.h, 0 if precision == "fp64" else 1,
                                    int(line_rule), int(tri_rule), C.byref(h)))
    op = Operator.__new__(Operator)
    op.mesh = mesh
    op.h = h
    op.desc = None
    op.nrows = op.ncols = mesh.nodes.shape[0]
    return op


class Batch:
    """Operators of one kind on several meshes applied with one persistent
    far-field launch (ebh_batch_*)."""

    def __init__(self, ops):
        L = lib()
        L.ebh_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int,
                                       C.POINTER(C.c_void_p)]
        L.ebh_batch_apply.argtypes = [C.c_void_p, C.POINTER(C.c_void_p),
                                      C.POINTER(C.c_void_p), C.c_void_p]
        L.ebh_batch_info.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        L.ebh_batch_destroy.argtypes = [C.c_void_p]
        L.ebh_batch_destroy.restype = None
        self.ops = list(ops)
        arr = (C.c_void_p * len(self.ops))(*[o.h.value for o in self.ops])
        h = C.c_void_p()
        check(L.ebh_batch_create(arr, len(self.ops), C.byref(h)))
        self.h = h

    def apply(self, xs, ys, stream=None):
        n = len(self.ops)
        xa = (C.c_void_p * n)(*[x.ptr for x in xs])
        ya = (C.c_void_p * n)(*[y.ptr for y in ys])
        check(lib().ebh_batch_apply(self.h, xa, ya, stream))

    def info(self):
        a = (C.c_int64 * 4)()
        check(lib().ebh_batch_info(self.h, a))
        return {"ops": a[0], "items": a[1], "workgroups": a[2], "lds": a[3]}

    def close(self):
        if getattr(self, "h", None):
            lib().ebh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TuneResult(C.Structure):
    _fields_ = [("block", C.c_int), ("tile", C.c_int), ("chunks", C.c_int),
                ("grid", C.c_int), ("mgrid", C.c_int), ("ms", C.c_double),
                ("ms_default", C.c_double), ("mms", C.c_double),
                ("mms_default", C.c_double), ("tried", C.c_int),
                ("schedule", C.c_int), ("split", C.c_int)]


def evaluate(mesh: Mesh, points, t, u, tri_rule=4, precision="fp64"):
    """Representation formula at off-surface points (evaluateLaplace3d):
    single layer of the P0 data t minus double layer of the P1 data u."""
    L = lib()
    L.ebh_evaluate_laplace_host.argtypes = [C.c_void_p, C.c_int, C.c_int,
                                            C.c_int, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_void_p]
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(t, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.empty(pts.shape[0])
    check(L.ebh_evaluate_laplace_host(mesh.h, int(tri_rule),
                                      0 if precision == "fp64" else 1,
                                      pts.shape[0], _ptr(pts), _ptr(t), _ptr(u),
                                      _ptr(out)))
    return out


def assemble(mesh: Mesh, order=4):
    """Dense (V, K, D, M) like getLaplace3d, assembled on the device."""
    L = lib()
    L.ebh_assemble_laplace.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 4
    ne, np_ = mesh.elements.shape[0], mesh.nodes.shape[0]
    V, K = np.empty((ne, ne)), np.empty((ne, np_))
    D, M = np.empty((np_, np_)), np.empty((ne, np_))
    check(L.ebh_assemble_laplace(mesh.h, int(order), _ptr(V), _ptr(K), _ptr(D),
                                 _ptr(M)))
    return V, K, D, M


def tune_save(path):
    lib().ebh_tune_save.argtypes = [C.c_char_p]
    check(lib().ebh_tune_save(os.fsencode(path)))


def tune_load(path):
    lib().ebh_tune_load.argtypes = [C.c_char_p]
    check(lib().ebh_tune_load(os.fsencode(path)))


def tune_clear():
    check(lib().ebh_tune_clear())


def op_profile(op: Operator, enable: bool, max_launches=4096):
    """Start/stop far-kernel event timing; returns (total_ms, launches)
    recorded since the previous call."""
    L = lib()
    L.ebh_op_profile.argtypes = [C.c_void_p, C.c_int, C.c_int,
                                 C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    ms, n = C.c_double(), C.c_int64()
    check(L.ebh_op_profile(op.h, int(enable), max_launches, C.byref(ms),
                           C.byref(n)))
    return ms.value, n.value


def op_profile_near(op: Operator):
    """Near-field pass timing recorded since op_profile(op, True):
    (total_ms, launches)."""
    L = lib()
    L.ebh_op_profile_near.argtypes = [C.c_void_p, C.POINTER(C.c_double),
                                      C.POINTER(C.c_int64)]
    ms, n = C.c_double(), C.c_int64()
    check(L.ebh_op_profile_near(op.h, C.byref(ms), C.byref(n)))
    return ms.value, n.value


def mesh_check(nodes, elements):
    """Host-only validation and classification (no GPU needed)."""
    L = lib()
    L.ebh_mesh_check.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                 C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    elements = np.ascontiguousarray(elements, dtype=np.int32)
    info = (C.c_int64 * 8)()
    nnz = (C.c_int64 * 4)()
    check(L.ebh_mesh_check(nodes.shape[0], _ptr(nodes), elements.shape[0],
                           _ptr(elements), info, nnz))
    keys = ["np", "ne", "coincident", "edge", "vertex", "far", "max_touch"]
    d = dict(zip(keys, list(info)[:7]))
    d["nnz"] = {"P0xP0": nnz[0], "P0xP1": nnz[1], "P1xP0": nnz[2],
                "P1xP1": nnz[3]}
    return d


class ServerOpts(C.Structure):
    _fields_ = [("max_slots", C.c_int), ("hbm_budget", C.c_size_t),
                ("restart", C.c_int), ("tol", C.c_double), ("mixed", C.c_int),
                ("line_rule", C.c_int), ("tri_rule", C.c_int),
                ("max_iters", C.c_int), ("dirichlet_p1", C.c_int),
                ("no_precond", C.c_int), ("no_graphs", C.c_int),
                ("partial_limit", C.c_size_t), ("inner_tol", C.c_double)]


class RequestResult(C.Structure):
    _fields_ = [("status", C.c_int), ("iters", C.c_int), ("relres", C.c_double),
                ("queue_ms", C.c_double), ("setup_ms", C.c_double),
                ("solve_ms", C.c_double), ("total_ms", C.c_double),
                ("ne", C.c_int), ("footprint", C.c_size_t),
                ("cache_hit", C.c_int), ("device_peak", C.c_size_t),
                ("graph_steps", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Server:
    """Continuous-batching Dirichlet / Neumann solve server (ebh_server_*)."""

    def __init__(self, max_slots=4, hbm_budget=0, restart=50, tol=1e-8,
                 mixed=True, line_rule=4, tri_rule=4, max_iters=2000,
                 dirichlet_p1=False, precond=True, graphs=True, partial_limit=0,
                 inner_tol=0.0):
        L = lib()
        vp = C.c_void_p
        L.ebh_server_create.argtypes = [C.POINTER(ServerOpts), C.POINTER(vp)]
        L.ebh_server_submit.argtypes = [vp, C.c_int, vp, C.c_int, vp, vp,
                                        C.c_uint64, C.POINTER(C.c_int64)]
        L.ebh_server_submit_bvp.argtypes = [vp, C.c_int, C.c_int, vp, C.c_int,
                                            vp, vp, C.c_uint64,
                                            C.POINTER(C.c_int64)]
        L.ebh_server_wait.argtypes = [vp, C.c_int64, C.c_double,
                                      C.POINTER(RequestResult), vp]
        L.ebh_server_stats.argtypes = [vp, C.POINTER(C.c_int64)]
        L.ebh_server_destroy.argtypes = [vp]
        L.ebh_server_destroy.restype = None
        o = ServerOpts(max_slots, hbm_budget, restart, tol, int(mixed),
                       line_rule, tri_rule, max_iters, int(dirichlet_p1),
                       int(not precond), int(not graphs), int(partial_limit),
                       float(inner_tol))
        self.dirichlet_p1 = bool(dirichlet_p1)
        h = vp()
        check(L.ebh_server_create(C.byref(o), C.byref(h)))
        self.h = h
        self.ne = {}

    def submit(self, nodes, elements, x0, key=0, problem="dirichlet"):
        """problem: "dirichlet" (returns the P0 flux, ne entries) or
        "neumann" (returns the P1 trace, np entries)."""
        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        elements = np.ascontiguousarray(elements, dtype=np.int32)
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        kind = {"dirichlet": 0, "neumann": 1}[problem]
        rid = C.c_int64()
        check(lib().ebh_server_submit_bvp(self.h, kind, nodes.shape[0],
                                          _ptr(nodes), elements.shape[0],
                                          _ptr(elements), _ptr(x0), key,
                                          C.byref(rid)))
        self.ne[rid.value] = (nodes.shape[0] if kind or self.dirichlet_p1
                              else elements.shape[0])
        return rid.value

    def wait(self, rid, timeout=0.0, raise_on_error=False):
        res = RequestResult()
        t = np.empty(self.ne[rid])
        rc = lib().ebh_server_wait(self.h, rid, timeout, C.byref(res), _ptr(t))
        if rc != -7:  # EBH_ERR_STATE: timed out, the request is still pending
            self.ne.pop(rid)
        if rc != EBH_OK and raise_on_error:
            check(rc)
        d = res.as_dict()
        d["error"] = "" if rc == EBH_OK else lib().ebh_last_error().decode()
        return d, (t if rc == EBH_OK else None)

    def stats(self):
        a = (C.c_int64 * 10)()
        check(lib().ebh_server_stats(self.h, a))
        keys = ["submitted", "completed", "failed", "active", "queued",
                "bytes_held", "peak_bytes", "cache_hits", "device_bytes",
                "device_peak"]
        return dict(zip(keys, list(a)))

    def close(self):
        if getattr(self, "h", None):
            lib().ebh_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
