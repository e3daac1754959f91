"""Builds the packaged launch-shape table (``tuning_mi355x.txt``) by running
the library's autotuner (``ebh_op_autotune``) on the operators of the
standard workloads, on the GPU this runs on:

    python -m easybem_amd.tune_db [--out PATH] [--budget-ms MS] [--only NAME]

--only retunes the workloads whose name contains NAME and keeps the other
entries of the packaged table.

Workloads: icosphere levels 3-6 (single layer P0 and P1 with the 7-point
and the 1-point far rule -- the latter is the Calderon preconditioners' --,
both far-field precisions; the FP64 double layer with P0 and P1 tests and
its adjoint), the 256x128 torus with the
Kelvin single layer, and the 64-RHS single layer on level 6.  Each line of the output keys a
shape on device, operator kind, precision, rule size, test basis, RHS block
and mesh size class; ``easybem_amd._lib`` loads the file when the library is
first opened.

One entry of the packaged table is set from longer timings than the
tuner's budget allows: the FP64 Kelvin 7-point torus key records the
pair-symmetric schedule (143.9 ms against 154.6 ms chunked, three runs
each), where the tuner's short runs time the two within noise.  Rerunning
``--only torus`` rewrites it; restore that line afterwards.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

from . import _lib as L, meshes


def workloads():
    for lev in (3, 4, 5, 6):
        P, T = meshes.icosphere(lev)
        yield f"icosphere L{lev}", P, T, [
            ("V", b, b, p, 1, r) for r in (4, 0) for b in (0, 1)
            for p in ("fp32", "fp64")] + [
            ("K", 1, 1, "fp64", 1, 4), ("K", 0, 1, "fp64", 1, 4),
            ("Kt", 1, 0, "fp64", 1, 4)]
    P, T = meshes.icosphere(6)
    yield "icosphere L6, 64 RHS", P, T, [("V", 0, 0, p, 64, 4)
                                         for p in ("fp32", "fp64")]
    P, T = meshes.torus(1.0, 0.4, 256, 128)
    yield "torus 256x128", P, T, [("kelvin", 1, 1, p, 1, 4)
                                  for p in ("fp32", "fp64")]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=L.TUNE_FILE)
    ap.add_argument("--budget-ms", type=float, default=4000.0)
    ap.add_argument("--only", default="", help="retune only these workloads")
    args = ap.parse_args()
    if not args.only:
        os.environ["EBH_TUNE_FILE"] = ""  # start from default shapes
    L.lib()
    if not args.only:
        L.tune_clear()
    for name, P, T, ops in workloads():
        if args.only not in name:
            continue
        m = L.Mesh(P, T)
        for kern, test, trial, prec, nrhs, rule in ops:
            t0 = time.time()
            op = L.Operator(m, kern, test, trial, prec,
                            line_rule=4 if rule else 2, tri_rule=rule,
                            nu=0.3, max_rhs=nrhs)
            r = op.autotune(budget_ms=args.budget_ms)
            sched = (f"symmetric split {r['split']}" if r["schedule"] else
                     f"block {r['block']} tile {r['tile']} workgroups "
                     f"{r['grid']}")
            print(f"{name} {kern}{test}{trial} {prec} rule {rule} "
                  f"rhs={nrhs}: {sched} "
                  f"mgrid {r['mgrid']}  {r['ms_default']:.3f} -> "
                  f"{r['ms']:.3f} ms"
                  + (f", multi {r['mms_default']:.2f} -> {r['mms']:.2f} ms"
                     if r["mgrid"] else "")
                  + f"  ({r['tried']} shapes, {time.time() - t0:.1f} s)",
                  flush=True)
            op.close()
        m.close()
    L.tune_save(args.out)
    print("wrote", args.out)


if __name__ == "__main__":
    sys.exit(main())
