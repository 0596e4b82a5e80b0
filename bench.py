"""Benchmark of the FEMOCS field-solve hot path on one MI355X.

One step = one Poisson solve of the nanotip problem (Neumann face load + FP64 Dirichlet
lifting + CG from phi = 0 to |b - Ax| / |b| <= 1e-9 + export of the nodal potential), on the
mesh of a nanotip r = 10 nm, h = 60 nm on a 120 x 120 nm slab (tet-split Q1 hexahedra).
The headline value is the mixed-precision solver (FP32 operator and CG inside FP64 defect
correction); the plain FP64 CG is measured on the same mesh beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--small]
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

METRIC = "Poisson solves/s + p50/p99 latency at 1e-9 residual: mixed vs FP64 CG, 1 GPU"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


USE_GRAPH = True


def one_solve(solver, sys_, precision, tol):
    from femocs_amd import fgpu
    sys_.reset()
    sys_.add_dirichlet(fgpu.BID_COPPER_SURFACE, 0.0)
    sys_.add_neumann(fgpu.BID_VACUUM_TOP, None, 1.0)   # applied_field = -E0, E0 = -1 V/A
    sys_.apply_dirichlet()
    sys_.fill_solution(0.0)
    st = sys_.solve(fgpu.opts(tol=tol, relative=True, precision=precision, inner_tol=0.1, use_graph=USE_GRAPH))
    phi = sys_.solution()
    return st, phi


def time_mode(sys_, precision, steps, warmup, tol):
    for _ in range(warmup):
        one_solve(None, sys_, precision, tol)
    lat = []
    its = []
    sys_.ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        t = time.perf_counter()
        st, phi = one_solve(None, sys_, precision, tol)
        lat.append(time.perf_counter() - t)
        its.append(st.iterations)
        if st.iterations <= 0:
            raise RuntimeError(f"solve did not converge: {st.as_dict()}")
    sys_.ctx.sync()
    total = time.perf_counter() - t0
    lat = np.array(lat) * 1e3
    return {
        "solves_per_s": steps / total,
        "ms_per_solve": total * 1e3 / steps,
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "iterations": int(np.median(its)),
        "outer": int(st.outer),
        "rel_residual": st.residual / st.rhs_norm,
    }, phi


def cpu_baseline(mesh, budget_s=60.0):
    """CPU oracle (restatement of the reference algorithm: assembled CSR, QGauss(2), deal.II
    lifting, SSOR(1.2)-PCG), one full solve of the same problem on one host core."""
    from oracle import oracle as O
    t0 = time.perf_counter()
    om = O.Mesh(mesh.nodes, mesh.vacuum_cells(), 8, 2, 4, 2)
    om.assemble()
    b = om.face_rhs(8, constant=1.0)
    x = np.zeros(om.n_dofs)
    mask = om.dirichlet_mask(2)
    om.apply_dirichlet(b, x, mask, None)
    t1 = time.perf_counter()
    it, res = om.pcg(b, x, tol=1e-9, relative=True, precond="ssor", omega=1.2, max_it=10000)
    t2 = time.perf_counter()
    return {"value": 1.0 / (t2 - t0), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"one full solve of the same mesh ({om.n_dofs} dofs): import+assembly+lifting "
                      f"{t1 - t0:.2f} s, SSOR(1.2)-PCG {it} iterations {t2 - t1:.2f} s, 1 thread"}, x


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--small", action="store_true", help="nanotip r=3nm h=15nm mesh instead (quick check)")
    ap.add_argument("--tol", type=float, default=1e-9)
    ap.add_argument("--no-graph", action="store_true", help="launch kernels directly (for profilers)")
    args = ap.parse_args()
    global USE_GRAPH
    USE_GRAPH = not args.no_graph
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the solver has no multi-GPU path in this build "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    from femocs_amd import fgpu
    from femocs_amd import mesh as M

    g = M.tip_r3nm_geometry() if args.small else M.tip_r10nm_geometry()
    t = time.perf_counter()
    mesh = M.make_tip_mesh(g, with_atoms=False)
    log(f"mesh: {mesh.n_tets} tets, {mesh.n_hexes} hexes ({time.perf_counter() - t:.1f} s)")
    ctx = fgpu.Context(0)
    t = time.perf_counter()
    gm = fgpu.Mesh(ctx, mesh.nodes, mesh.vacuum_cells(), 8, 2, 4, 2)
    sys_ = fgpu.System(gm)
    setup_ms = (time.perf_counter() - t) * 1e3
    log(f"device mesh: {gm.n_dofs} dofs, {gm.n_cells} cells, import {setup_ms:.0f} ms")

    from femocs_amd import autotune
    tuned = {}
    best, _ = autotune.autotune(sys_, fgpu.FGPU_FP64)
    tuned["fp64"] = {"cells_per_patch": best[0], "patches_per_block": best[1]}
    fp64, phi64 = time_mode(sys_, fgpu.FGPU_FP64, args.steps, args.warmup, args.tol)
    log("fp64", fp64, tuned["fp64"])
    best, _ = autotune.autotune(sys_, fgpu.FGPU_MIXED)
    tuned["mixed"] = {"cells_per_patch": best[0], "patches_per_block": best[1]}
    mixed, phimx = time_mode(sys_, fgpu.FGPU_MIXED, args.steps, args.warmup, args.tol)
    log("mixed", mixed, tuned["mixed"])
    err_mixed_vs_fp64 = float(np.linalg.norm(phimx - phi64) / np.linalg.norm(phi64))

    # dominant kernel: the element apply, FP32 in the mixed solver's iterations (the system
    # is left tuned for mixed precision).  Algorithmic bytes per launch = compulsory traffic
    # of a matrix-free Q1 apply: connectivity 32 B/cell (8 x int32) + coordinates 3*vb B/dof
    # + the two input vectors (z, p) 2*vb B/dof + one output per dof vb B/dof.
    H, N = gm.n_cells, gm.n_dofs
    kern = {}
    for prec, name, vb in ((fgpu.FGPU_MIXED, "k_patch<float>", 4), (fgpu.FGPU_FP64, "k_patch<double>", 8)):
        ms = sys_.time_kernel(0, prec, reps=50)
        byts = 32 * H + 3 * vb * N + 2 * vb * N + vb * N
        kern[name] = {"ms": ms, "bytes": byts, "GBs": byts / (ms * 1e-3) / 1e9}
    log("kernels", kern)
    k = kern["k_patch<float>"]
    roofline = {"bound": "hbm", "achieved": round(k["GBs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(k["GBs"] / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_patch<float>",
                "avg_us": round(k["ms"] * 1e3, 2), "algorithmic_bytes": k["bytes"],
                "fp64_kernel": {"avg_us": round(kern["k_patch<double>"]["ms"] * 1e3, 2),
                                "GBs": round(kern["k_patch<double>"]["GBs"], 1)}}

    cpu = None
    err_vs_cpu = None
    if not args.no_cpu:
        cpu, xcpu = cpu_baseline(mesh)
        err_vs_cpu = float(np.linalg.norm(phi64 - xcpu) / np.linalg.norm(xcpu))
        cpu["value"] = round(cpu["value"], 5)
        log("cpu", cpu)

    out = {
        "metric": METRIC,
        "value": round(mixed["solves_per_s"], 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["ms_per_solve"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 outer",
        "data": "synthetic",
        "config": {"workload": ("Cu nanotip r=3nm h=15nm on 30x30x3 nm slab" if args.small else
                                "Cu nanotip r=10nm h=60nm on 120x120 nm slab") +
                               ", Q1 tet-split hexes, 1 Poisson solve per step",
                   "n_dofs": N, "n_cells": H, "tol_rel": args.tol, "precond": "jacobi",
                   "parallelism": "single GPU"},
        "latency_ms": {"p50": round(mixed["p50_ms"], 3), "p99": round(mixed["p99_ms"], 3)},
        "mixed": mixed,
        "fp64": fp64,
        "speedup_mixed_vs_fp64": round(mixed["solves_per_s"] / fp64["solves_per_s"], 3),
        "err_mixed_vs_fp64": err_mixed_vs_fp64,
        "err_fp64_vs_cpu_oracle": err_vs_cpu,
        "mesh_import_ms": round(setup_ms, 1),
        "autotuned": tuned,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
