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
  preconditioner (Jacobi + a coarse correction on blocks of 256-1024 dofs by mesh size, built once per resident
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


def _best_of(f, reps=3):
    """Smallest wall time of reps calls of f (each synchronises before it returns)."""
    best = float("inf")
    out = None
    for _ in range(reps):
        t = time.perf_counter()
        out = f()
        best = min(best, time.perf_counter() - t)
    return best, out


def paths_bench(no_cpu: bool):
    """The paths around the field solve, measured on the larger tip (r = 10 nm, h = 60 nm on a
    120 x 120 nm slab): the plain FP64 solve, the transfer of the solution to every surface
    atom (ranks 1-3, located and mapped), the PIC pieces (space-charge deposit, particle cell
    update, particle field) for 1M superparticles, and one coupled sigma(T) current +
    Crank-Nicolson heat step on the r = 3 nm tip's bulk.  Host arrays cross PCIe inside these
    timings (the calls take and return host buffers).  CPU baselines: the oracle on bounded
    samples, one core.  One JSON line."""
    from femocs_amd import fgpu, interp
    from femocs_amd import mesh as M
    out = {"metric": "secondary paths: transfer to atoms, PIC pieces, coupled heat step (1 GPU)",
           "data": "synthetic"}
    tip = M.make_tip_mesh(M.tip_r10nm_geometry())
    ctx = fgpu.Context(0)
    gm = fgpu.Mesh(ctx, tip.nodes, tip.vacuum_cells(), 8, 2, 4, 2)
    s = fgpu.System(gm)
    s.reset()
    s.add_dirichlet(2, 0.0)
    s.add_neumann(8, None, -1.0)
    s.apply_dirichlet()

    # direct launches, as in the serving loop (a rocprofv3 kernel trace of graph-captured
    # solves ended the host process)
    def cold(precision):
        s.fill_solution(0.0)
        return s.solve(fgpu.opts(tol=1e-9, relative=True, precision=precision, inner_tol=0.01, use_graph=False))

    t64, st64 = _best_of(lambda: cold(fgpu.FGPU_FP64))
    tmx, stmx = _best_of(lambda: cold(fgpu.FGPU_MIXED))
    out["solve_r10nm"] = {"hexes": gm.n_cells, "dofs": gm.n_dofs, "fp64_ms": round(t64 * 1e3, 2),
                          "fp64_iterations": st64.iterations, "mixed_ms": round(tmx * 1e3, 2),
                          "mixed_iterations": stmx.iterations}
    log("solve", out["solve_r10nm"])
    # mesh import (DealSolver::import_mesh + mark_boundary and the kernels' internal order):
    # tables built on the device against the same tables built on the host (host arrays in)
    vc = tip.vacuum_cells()

    def imp(host):
        m = fgpu.Mesh(ctx, tip.nodes, vc, 8, 2, 4, 2, host_import=host)
        ctx.sync()
        return m

    tid, md = _best_of(lambda: imp(False))
    tih, mh = _best_of(lambda: imp(True))
    out["mesh_import"] = {"hexes": md.n_cells, "device_ms": round(tid * 1e3, 2), "host_ms": round(tih * 1e3, 2),
                          "same_tables": md.digest() == mh.digest()}
    del md, mh
    log("mesh import", out["mesh_import"])
    # interpolator tables (tet tables, neighbour lists, search tree, node lists): device and host
    def icreate(host):
        r = interp.Interpolator(ctx, tip, interp.TYPES_VACUUM, host_build=host)
        ctx.sync()
        return r

    tcd, ipd = _best_of(lambda: icreate(False))
    tch, iph = _best_of(lambda: icreate(True), reps=1)
    out["interpolator_create"] = {"tets": len(tip.tets), "device_ms": round(tcd * 1e3, 1),
                                  "host_ms": round(tch * 1e3, 1), "same_tables": ipd.digest() == iph.digest()}
    del iph
    # the first located transfer of a new interpolator also fetches the host copy of its tet
    # tables (the chained guess reads them)
    ipf = icreate(False)
    t = time.perf_counter()
    ipf.points(tip.atoms, 1)
    out["interpolator_create"]["first_locate_ms"] = round((time.perf_counter() - t) * 1e3, 1)
    del ipf
    log("interpolator create", out["interpolator_create"])
    ip = ipd
    # the reference's tet -> 4-hex split with its numbering, on the device (host arrays in/out)
    tris, _ = tip.surface_triangles()
    tsp, sp_out = _best_of(lambda: fgpu.tethex_split(ctx, tip.nodes[:tip.n_tetnodes], tip.tets, tris))
    out["tet_split"] = {"tets": len(tip.tets), "boundary_triangles": len(tris), "ms": round(tsp * 1e3, 2),
                        "nodes": len(sp_out[0])}
    log("tet split", out["tet_split"])
    te, _ = _best_of(lambda: ip.extract_solution(s, smoothen=True))
    atoms = tip.atoms
    tr = {"atoms": len(atoms), "extract_ms": round(te * 1e3, 2)}
    for rank in (1, 2, 3):
        tl, (cells, _) = _best_of(lambda: ip.points(atoms, rank))
        tm, _ = _best_of(lambda: ip.in_cells(atoms, rank, cells))
        tr[f"rank{rank}_locate_atoms_per_s"] = round(len(atoms) / tl)
        tr[f"rank{rank}_mapped_atoms_per_s"] = round(len(atoms) / tm)
    out["transfer"] = tr
    log("transfer", tr)
    # PIC pieces: superparticles at random natural coordinates of random solver cells
    rng = np.random.default_rng(7)
    npart = 1_000_000
    hx = np.nonzero(tip.hex_marker > 0)[0]
    pc = rng.integers(0, len(hx), npart).astype(np.int32)
    UVW = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]])
    uvw = rng.uniform(-0.95, 0.95, (npart, 3))
    w = np.prod(1 + UVW[None] * uvw[:, None, :], axis=2) / 8.0
    pts = np.einsum("nk,nkd->nd", w, tip.nodes[tip.hexes[hx[pc]]])
    sp = fgpu.System(gm)

    def deposit():
        sp.reset()
        sp.add_space_charge(ip, pts, pc, -1e-3)

    td, _ = _best_of(deposit)
    moved = pts + rng.normal(0.0, 0.5, pts.shape)
    tu, _ = _best_of(lambda: ip.update_cells(moved, pc))
    tg, _ = _best_of(lambda: ip.hex_gradient(pts, hx[pc].astype(np.int32)))
    pic = {"particles": npart, "deposit_per_s": round(npart / td), "cell_update_per_s": round(npart / tu),
           "field_per_s": round(npart / tg)}
    out["pic"] = pic
    log("pic", pic)
    # one coupled step on the r = 3 nm tip's bulk (mixed precision, as served)
    small = M.make_tip_mesh(M.tip_r3nm_geometry())
    bm = fgpu.Mesh(ctx, small.nodes, small.bulk_cells(), 2, 7, 4, 2)
    cen = bm.face_centroids(2)
    J = 3e-5 * (1.0 + 0.5 * np.sin(cen[:, 0] / 7.0) * np.cos(cen[:, 2] / 5.0))
    Nq = 1e-9 * np.ones(len(cen))
    cur, heat = fgpu.System(bm), fgpu.System(bm)
    xyz, _ = bm.dofs()
    heat.set_solution(300.0 + 400.0 * np.exp(-((xyz[:, 2] - xyz[:, 2].max()) / 20.0) ** 2))
    cur.fill_solution(0.0)
    o = fgpu.opts(tol=1e-9, relative=True, precision=fgpu.FGPU_MIXED, inner_tol=0.01, use_graph=False)

    def coupled_step():
        cur.reset()
        cur.sigma_terms(heat, 1.0, 3.9e-3)
        cur.add_neumann(2, J, 0.0)
        cur.add_dirichlet(7, 0.0)
        cur.apply_dirichlet()
        cur.solve(o)
        heat.reset()
        heat.heat_terms(cur, 3.4496e-24 / 4e-15, 2.44e-8, theta=0.5)
        heat.add_neumann(2, Nq, 0.0)
        heat.add_dirichlet(7, 300.0)
        heat.apply_dirichlet()
        return heat.solve(o)

    th, _ = _best_of(coupled_step, reps=5)
    out["coupled_step_r3nm"] = {"bulk_hexes": bm.n_cells, "dofs": bm.n_dofs, "ms": round(th * 1e3, 2)}
    log("coupled", out["coupled_step_r3nm"])
    if not no_cpu:
        from oracle import oracle as O
        cpu = {"cores": 1, "kind": "port"}
        L = O.LinTet(tip.nodes, tip.tets, tip.tet_nbr, tip.tet_marker, 3)
        q10 = O.quadtet_cells(tip.tets, tip.hexes, tip.node_kind)
        sub = atoms[:: max(1, len(atoms) // 2000)]
        sol = ip.nodal()
        t = time.perf_counter()
        O.interp_points(L, tip.nodes, tip.tets, q10, tip.hexes, sol, sub, 2)
        cpu["rank2_locate_atoms_per_s"] = round(len(sub) / (time.perf_counter() - t), 1)
        om = O.Mesh(tip.nodes, tip.vacuum_cells(), 8, 2, 4, 2)
        d2f = hx.astype(np.int32)
        f2d = -np.ones(len(tip.hexes), np.int32)
        f2d[d2f] = np.arange(len(d2f))
        k = 20000
        t = time.perf_counter()
        O.space_charge(om, tip.nodes, tip.hexes, d2f, pts[:k], pc[:k], -1e-3)
        cpu["deposit_per_s"] = round(k / (time.perf_counter() - t))
        t = time.perf_counter()
        O.update_cells(L, d2f, f2d, moved[:k], pc[:k])
        cpu["cell_update_per_s"] = round(k / (time.perf_counter() - t))
        t = time.perf_counter()
        M.tet_to_hex(tip.nodes[:tip.n_tetnodes], tip.tets)
        cpu["tet_split_numpy_ms"] = round((time.perf_counter() - t) * 1e3, 1)
        cpu["sample"] = (f"{len(sub)} atoms (every {max(1, len(atoms) // 2000)}th), {k} particles; the tet split "
                         "by the generator's numpy version (its own numbering)")
        out["cpu_baseline"] = cpu
        log("cpu", cpu)
    print(json.dumps(out), flush=True)


def setup_bench(reps: int):
    """One remesh of the r = 10 nm tip on the device, `reps` times: the Tethex split of its tets
    (host arrays in and out), the import of the vacuum hexes and the interpolator tables; then
    the same import and tables once on the host, and whether both give the same tables.  One
    JSON line; host buffers cross PCIe inside these timings."""
    from femocs_amd import fgpu, interp
    from femocs_amd import mesh as M
    tip = M.make_tip_mesh(M.tip_r10nm_geometry(), with_atoms=False)
    ctx = fgpu.Context(0)
    tris, _ = tip.surface_triangles()
    vc = tip.vacuum_cells()
    steps = {"tet_split": [], "mesh_import": [], "interpolator": []}
    md = ipd = None
    for r in range(reps + 1):          # the first pass loads the kernels and is not kept
        t0 = time.perf_counter()
        fgpu.tethex_split(ctx, tip.nodes[:tip.n_tetnodes], tip.tets, tris)
        t1 = time.perf_counter()
        md = fgpu.Mesh(ctx, tip.nodes, vc, 8, 2, 4, 2)
        ctx.sync()
        t2 = time.perf_counter()
        ipd = interp.Interpolator(ctx, tip, interp.TYPES_VACUUM)
        ctx.sync()
        t3 = time.perf_counter()
        if r > 0:
            for k, v in zip(steps, (t1 - t0, t2 - t1, t3 - t2)):
                steps[k].append(v * 1e3)
        log("setup pass", r, [round(v * 1e3, 1) for v in (t1 - t0, t2 - t1, t3 - t2)])
    t = time.perf_counter()
    mh = fgpu.Mesh(ctx, tip.nodes, vc, 8, 2, 4, 2, host_import=True)
    t_imp = time.perf_counter() - t
    t = time.perf_counter()
    iph = interp.Interpolator(ctx, tip, interp.TYPES_VACUUM, host_build=True)
    t_int = time.perf_counter() - t
    out = {"metric": "remesh setup on the device: tet split, mesh import, interpolator tables (1 GPU)",
           "unit": "ms", "data": "synthetic", "reps": reps,
           "config": {"workload": "r = 10 nm tip", "tets": len(tip.tets), "vacuum_hexes": md.n_cells,
                      "dofs": md.n_dofs},
           "device_ms": {k: {"best": round(min(v), 2), "mean": round(sum(v) / len(v), 2)} for k, v in steps.items()},
           "host_ms": {"mesh_import": round(t_imp * 1e3, 1), "interpolator": round(t_int * 1e3, 1)},
           "same_tables": {"mesh_import": md.digest() == mh.digest(), "interpolator": ipd.digest() == iph.digest()}}
    out["device_ms"]["total_best"] = round(sum(v["best"] for v in out["device_ms"].values()), 2)
    print(json.dumps(out), flush=True)


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
    ap.add_argument("--paths", action="store_true",
                    help="the secondary paths instead of the headline (transfer, PIC pieces, coupled heat)")
    ap.add_argument("--setup", action="store_true",
                    help="only the remesh setup on the device (tet split, import, interpolator tables)")
    args = ap.parse_args()
    if args.setup:
        setup_bench(max(1, min(args.steps, 20)))
        return
    if args.traffic_probe:
        traffic_probe()
        return
    if args.paths:
        paths_bench(args.no_cpu)
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
                   "precond": "jacobi" if args.jacobi else "two-level: jacobi + aggregates of 256-1024 consecutive dofs by mesh size",
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
