"""Benchmark of the FEMOCS field-solve hot path on one MI355X: Poisson solves served to
concurrent molecular-dynamics clients.

Workload: 16 MD clients, each with its own Cu nanotip (radius 3-12 nm, aspect ratio 4-7,
112k-1.6M vacuum Q1 hexes from the tet split, ~10M dofs in all), whose surface atoms vibrate
(0.1 A Gaussian noise per step).  The clients submit the Laplace solves of their next
timesteps (Dirichlet 0 on the copper surface, Neumann field on the top; |b - Ax| / |b| <=
1e-9) as a Poisson process of 200 requests/s in all (--closed-loop: each client waits an
exponential think time after its previous result instead).  A client's solves run in order.
One step = one served solve; `value` is solves/s over the K timed requests, from the first
timed arrival to the last timed result; latency runs from arrival to result.

* headline: mixed precision (FP32 operator and CG inside FP64 defect correction), two-level
  preconditioner (Jacobi + a coarse correction on 1024-dof blocks, built once per resident
  mesh), continuous batching, meshes and last solutions resident on the GPU (warm starts);
* `plain_fp64`: the same request stream through the plain form: FP64 CG, one solve at a time,
  from phi = 0, in arrival order.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--small] [--jacobi]
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
N_CLIENTS = 16
RATE = 200.0


def measured_traffic():
    """HBM bytes per full-batch launch of the element apply from the committed counter summary
    (rocprofv3 FETCH_SIZE and WRITE_SIZE passes over `bench.py --traffic-probe`), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_k_patch_serving_16_clients.txt")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def client_geometry(i: int, n: int = N_CLIENTS, small: bool = False):
    """Tip i of the client set: radii spread over 3-12 nm, aspect ratios 4-7 interleaved."""
    from femocs_amd import mesh as M
    if small:
        return M.small_geometry(radius=10.0 + 2.0 * i, height=24.0 + 4.0 * i, seed=100 + i)
    r = 3.0 + 9.0 * i / (n - 1)
    a = 4.0 + 3.0 * ((5 * i) % n) / (n - 1)
    return M.TipGeometry(radius=10 * r, height=10 * r * a, slab_width=120 * r,
                         slab_thickness=30.0 if r < 6 else 60.0, seed=1000 + i)


def _make_client(args):
    i, n, small = args
    from femocs_amd import mesh as M
    from femocs_amd import serving
    tip = M.make_tip_mesh(client_geometry(i, n, small), with_atoms=False)
    return serving.Client(i, tip, field=-1.0, sigma=0.1, seed=i)


def make_clients(n: int, small: bool):
    from multiprocessing import get_context
    workers = max(1, min(n, 16, os.cpu_count() or 1))
    with get_context("fork").Pool(workers) as pool:
        return pool.map(_make_client, [(i, n, small) for i in range(n)])


def serve(ctx, clients, precision, batched, warm, steps, warmup, tuned, seed, coarse=False, arrivals="poisson"):
    from femocs_amd import fgpu, serving
    # inner_tol: the FP32 CG runs until |r| dropped 100-fold, then the FP64 residual replaces it
    o = fgpu.opts(tol=1e-9, relative=True, precision=precision, inner_tol=0.01, use_graph=False)
    cells, ppb = tuned

    def configure(mesh, system):
        mesh.set_patch_cells(cells)
        system.set_patches_per_block(ppb)

    for c in clients:
        c.seq = 0
        c.rng = np.random.default_rng(1000 + c.cid)
    srv = serving.Server(ctx, clients, o, batched=batched, warm=warm, configure=configure, coarse=coarse)
    srv.make_resident()                      # every tip imported and assembled before the clock runs
    serving.run_clients(srv, rate=RATE, duration=1e9, max_requests=warmup, seed=seed, arrivals=arrivals)
    srv.stats = serving.ServeStats()
    ctx.sync()
    serving.run_clients(srv, rate=RATE, duration=1e9, max_requests=steps, seed=seed + 1, arrivals=arrivals)
    ctx.sync()
    summ = srv.stats.summary()
    bad = [r for r in srv.stats.requests if r.iterations <= 0]
    if bad:
        raise RuntimeError(f"{len(bad)} solves did not converge")
    return srv, summ


def fp64_errors(srv):
    """Error of each client's last mixed result against an FP64 solve of the same system."""
    from femocs_amd import fgpu
    o = fgpu.opts(tol=1e-12, relative=True, precision=fgpu.FGPU_FP64, use_graph=False)
    errs = []
    for cid, r in srv.resident.items():
        xm = r.sys.solution()
        r.sys.fill_solution(0.0)
        st = r.sys.solve(o)
        if st.iterations <= 0:
            raise RuntimeError("FP64 reference solve did not converge")
        x64 = r.sys.solution()
        errs.append(float(np.linalg.norm(xm - x64) / np.linalg.norm(x64)))
    return errs


def cpu_baseline(client):
    """CPU oracle (restatement of the reference algorithm: assembled CSR, QGauss(2), deal.II
    lifting, SSOR(1.2)-PCG), one cold solve of one client's request on one host core."""
    from oracle import oracle as O
    X = client.next_positions()
    t0 = time.perf_counter()
    om = O.Mesh(X, client.cells, 8, 2, 4, 2)
    om.assemble()
    b = om.face_rhs(8, constant=-client.field)
    x = np.zeros(om.n_dofs)
    om.apply_dirichlet(b, x, om.dirichlet_mask(2), None)
    t1 = time.perf_counter()
    it, _ = om.pcg(b, x, tol=1e-9, relative=True, precond="ssor", omega=1.2, max_it=10000)
    t2 = time.perf_counter()
    if it <= 0:
        raise RuntimeError("CPU oracle did not converge")
    return {"value": round(1.0 / (t2 - t0), 5), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"one cold solve of the median client's request ({om.n_dofs} dofs, "
                      f"{len(client.cells)} hexes): import+assembly+lifting {t1 - t0:.2f} s, "
                      f"SSOR(1.2)-PCG {it} iterations {t2 - t1:.2f} s, 1 thread"}


def traffic_probe(n_launches: int = 5) -> None:
    """Counter target for the roofline's `traffic`: every client's system resident in one
    mixed batch (as served), `n_launches` element applies over the whole batch, then two device
    copies of 1 GiB (the calibration kernel of known size).  Run under
    rocprofv3 --pmc FETCH_SIZE (then WRITE_SIZE)."""
    import ctypes as C
    from femocs_amd import fgpu, serving
    # in this process, no worker pool: forked workers would inherit the profiler's state
    clients = []
    for i in range(N_CLIENTS):
        clients.append(_make_client((i, N_CLIENTS, False)))
        log(f"client {i} ready ({clients[-1].n_dofs} dofs)")
    ctx = fgpu.Context(0)
    o = fgpu.opts(tol=1e-9, relative=True, precision=fgpu.FGPU_MIXED, inner_tol=0.01, use_graph=False)

    def configure(mesh, system):
        mesh.set_patch_cells(256)
        system.set_patches_per_block(1)

    srv = serving.Server(ctx, clients, o, configure=configure, coarse=True)
    for i in range(N_CLIENTS):
        srv.make_resident([i])
        log(f"client {i} resident")
    b = fgpu.Batch(ctx, o)
    for r in srv.resident.values():
        b.add(r.sys)
    log("element apply avg us", round(b.time_kernel(0, reps=n_launches) * 1e3, 2))
    hip = C.CDLL("libamdhip64.so")
    n = 1 << 30
    src, dst = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), C.c_size_t(n)) or hip.hipMalloc(C.byref(dst), C.c_size_t(n)):
        raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(src, 1, C.c_size_t(n)):
        raise RuntimeError("hipMemset failed")
    for _ in range(2):
        if hip.hipMemcpy(dst, src, C.c_size_t(n), 3):          # device to device
            raise RuntimeError("hipMemcpy failed")
    if hip.hipDeviceSynchronize():
        raise RuntimeError("hipDeviceSynchronize failed")
    hip.hipFree(src)
    hip.hipFree(dst)
    log("copies done")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=480, help="timed requests (steady state: the queue of the "
                    "busiest clients drains for a while after the last arrival)")
    ap.add_argument("--warmup", type=int, default=16, help="untimed requests (meshes become resident)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 run")
    ap.add_argument("--small", action="store_true", help="4 small tips instead (quick check)")
    ap.add_argument("--jacobi", action="store_true", help="mixed run without the two-level preconditioner")
    ap.add_argument("--closed-loop", action="store_true",
                    help="clients submit an exponential think time after each result instead of Poisson arrivals")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only the counter target of roofline.traffic (see traffic_probe)")
    args = ap.parse_args()
    if args.traffic_probe:
        traffic_probe()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the field solver has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    from femocs_amd import autotune, fgpu

    n = 4 if args.small else N_CLIENTS
    t = time.perf_counter()
    clients = make_clients(n, args.small)
    hexes = [len(c.cells) for c in clients]
    dofs = [c.n_dofs for c in clients]
    log(f"clients: {n} tips, hexes {min(hexes)}..{max(hexes)}, {sum(dofs)} dofs ({time.perf_counter() - t:.1f} s)")

    ctx = fgpu.Context(0)
    big = int(np.argmax(hexes))
    gm = fgpu.Mesh(ctx, clients[big].X0, clients[big].cells, 8, 2, 4, 2)
    gs = fgpu.System(gm)
    tuned = {}
    for name, prec in (("mixed", fgpu.FGPU_MIXED), ("fp64", fgpu.FGPU_FP64)):
        best, _ = autotune.autotune(gs, prec)
        tuned[name] = best
    del gs, gm
    log("autotuned (cells per patch, patches per workgroup):", tuned)

    arrivals = "closed" if args.closed_loop else "poisson"
    plain = None
    if not args.no_plain:
        _, plain = serve(ctx, clients, fgpu.FGPU_FP64, False, False, args.steps, args.warmup, tuned["fp64"], 1,
                         arrivals=arrivals)
        log("plain fp64", plain)
    srv, mixed = serve(ctx, clients, fgpu.FGPU_MIXED, True, True, args.steps, args.warmup, tuned["mixed"], 1,
                       coarse=not args.jacobi, arrivals=arrivals)
    log("mixed batched", mixed)

    # dominant kernel: the FP32 element apply over the whole batch (all clients resident).
    # Algorithmic bytes per launch = compulsory traffic of a matrix-free Q1 apply: cell
    # connectivity 32 B/hex (8 x int32), the geometry (tet-split form: the 6-float metric,
    # 24 B/hex; otherwise coordinates 3 x 4 B/dof), the two input vectors (z, p) 2 x 4 B/dof
    # and one output 4 B/dof.
    b = fgpu.Batch(ctx, srv.opts)
    affine = True
    for r in srv.resident.values():
        affine = r.mesh.use_affine(True) and affine
        b.add(r.sys)
    ms = b.time_kernel(0, reps=30)
    H = sum(r.mesh.n_cells for r in srv.resident.values())
    N = sum(r.mesh.n_dofs for r in srv.resident.values())
    byts = 32 * H + (24 * H if affine else 12 * N) + 8 * N + 4 * N
    gbs = byts / (ms * 1e-3) / 1e9
    gather_ms = b.time_kernel(1, reps=30)
    coarse_ms = b.time_kernel(2, reps=30) if not args.jacobi else None
    del b
    roofline = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": measured_traffic(), "kernel": "k_patch<float>",
                "avg_us": round(ms * 1e3, 2), "algorithmic_bytes": byts, "cells": H, "dofs": N,
                "tet_split_form": affine,
                "gather_avg_us": round(gather_ms * 1e3, 2),
                "coarse_avg_us": round(coarse_ms * 1e3, 2) if coarse_ms is not None else None}
    log("roofline", roofline)

    errs = fp64_errors(srv)
    log("max rel error mixed vs fp64", max(errs))

    cpu = None
    if not args.no_cpu:
        order = np.argsort(hexes)
        cpu = cpu_baseline(clients[int(order[len(order) // 2])])
        log("cpu", cpu)

    ms_per = 1e3 / mixed["solves_per_s"]
    out = {
        "metric": METRIC,
        "value": round(mixed["solves_per_s"], 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 outer",
        "data": "synthetic",
        "config": {"workload": f"{n} MD clients, Cu nanotips r=3-12 nm, {min(hexes)}-{max(hexes)} Q1 hexes each, "
                               f"Poisson solve per timestep, "
                               + (f"closed-loop arrivals {RATE:.0f} req/s offered" if args.closed_loop else
                                  f"Poisson arrivals at {RATE:.0f} req/s total"),
                   "clients": n, "total_dofs": int(sum(dofs)), "total_hexes": int(sum(hexes)),
                   "tol_rel": 1e-9,
                   "precond": "jacobi" if args.jacobi else "two-level: jacobi + 1024-dof block aggregates",
                   "precond_plain_fp64": "jacobi", "parallelism": "single GPU"},
        "latency_ms": {"p50": round(mixed["p50_ms"], 2), "p99": round(mixed["p99_ms"], 2)},
        "mixed_batched": mixed,
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": round(mixed["solves_per_s"] / plain["solves_per_s"], 3) if plain else None,
        "err_mixed_vs_fp64_max": max(errs),
        "autotuned": {k: {"cells_per_patch": v[0], "patches_per_block": v[1]} for k, v in tuned.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
