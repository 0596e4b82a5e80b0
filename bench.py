"""Benchmark of the FEMOCS field-solve hot path on one MI355X: field + heat timesteps served to
concurrent molecular-dynamics clients.

Workload: 16 MD clients, each with its own Cu nanotip (radius 2.6-8.4 nm, aspect ratio 4-7;
about 0.1M-1.5M hexes per client mesh, vacuum and bulk together, as the tet split makes them)
whose surface atoms vibrate (0.1 A Gaussian noise per step).  One request is one timestep of
the reference's runaway loop (src/ProjectRunaway.cpp:422-447, 535-571): the Laplace solve on the
vacuum mesh (Dirichlet 0 on the copper surface, Neumann field on the top), then the current
solve and the implicit-Euler heat step (dt = 4 fs) on the bulk mesh, each to
|b - Ax| / |b| <= 1e-9.  A client's requests run in order; requests of different clients
overlap on the GPU.

* headline (`value`): mixed precision (FP32 operator and CG inside FP64 defect correction),
  two-level preconditioner (Jacobi + a coarse correction, built once per resident mesh),
  continuous batching, meshes and last solutions resident on the GPU (warm starts): the
  capacity, i.e. timesteps/s with every client submitting its next timestep as soon as the
  previous one returns.  One step = 24 served timesteps; --warmup steps are not timed;
* `latency_200`: the same server under Poisson arrivals of 200 timesteps/s in all (p50/p99 from
  arrival to the heat result); above the capacity, so the queue grows during the phase;
* `latency_sustained`: the same at 80 % of the measured capacity, a load the server keeps up
  with (stable p50/p99);
* `plain_fp64`: the plain form, FP64 CG with the Chebyshev-Jacobi preconditioner, one timestep
  at a time, Laplace and current from zero (the reference rebuilds its mesh every step), heat
  from the last temperature;
* `fp64_two_level_batched`: the ablation, FP64 CG with the headline's preconditioner, warm
  starts and batching: what mixed precision itself buys;
* every timed mixed timestep is checked, stage by stage, against an FP64 solve (|r|/|b| <=
  1e-12) of the same request from the same inputs (moved mesh, current, last temperature);
  the bench fails when an error exceeds 1e-8.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--small] [--no-plain]
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
FP64_PEAK_TFS = 78.6      # MI355X FP64 (vector and matrix), dense
HBM_PEAK_GBS = 8000.0
N_CLIENTS = 16
RATE = 200.0
PER_STEP = 24                 # served timesteps per bench step
ERR_LIMIT = 1e-8


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
    """Tip i of the client set: radii spread over 2.6-8.4 nm, aspect ratios 4-7 interleaved."""
    from femocs_amd import mesh as M
    if small:
        return M.small_geometry(radius=10.0 + 2.0 * i, height=24.0 + 4.0 * i, seed=100 + i)
    r = 2.6 + 5.8 * i / (n - 1)
    a = 4.0 + 3.0 * ((5 * i) % n) / (n - 1)
    return M.TipGeometry(radius=10 * r, height=10 * r * a, slab_width=120 * r,
                         slab_thickness=30.0 if r < 6 else 60.0, seed=1000 + i)


def _make_client(args):
    i, n, small = args
    from femocs_amd import mesh as M
    from femocs_amd import serving
    tip = M.make_tip_mesh(client_geometry(i, n, small), with_atoms=False)
    return serving.Client(i, tip, field=-1.0, sigma=0.1, seed=i, heat=True)


def make_clients(n: int, small: bool):
    from multiprocessing import get_context
    workers = max(1, min(n, 16, os.cpu_count() or 1))
    with get_context("fork").Pool(workers) as pool:
        return pool.map(_make_client, [(i, n, small) for i in range(n)])


class Snapshot:
    """A served result kept on the device: address of n FP64 values in a pool."""

    def __init__(self, pool, ptr: int):
        self.pool = pool
        self.ptr = ptr

    def data_ptr(self) -> int:
        return self.ptr


class Snapshots:
    """Device checkpoints of served results (one FP64 vector per stage and timestep): slices of
    one device pool allocated before the clock runs; kept until the timesteps are verified."""

    def __init__(self, ctx, doubles: int):
        self.ctx = ctx
        self.pool = ctx.alloc(8 * int(doubles))
        self.cap = int(doubles)
        self.off = 0
        self.extra = []
        self.overflow = 0

    def __call__(self, system):
        n = system.mesh.n_dofs
        if self.off + n <= self.cap:
            snap = Snapshot(self.pool, self.pool.ptr + 8 * self.off)
            self.off += n
        else:
            buf = self.ctx.alloc(8 * n)
            self.extra.append(buf)
            snap = Snapshot(buf, buf.ptr)
            self.overflow += 1
        system.solution_to_device(snap.ptr)
        return snap

    def clear(self):
        self.pool = None
        self.extra = []


def set_mode(srv, precision, batched, warm, coarse_in_cg, plain_precond=None):
    """Reconfigures an idle server (same resident meshes and systems) for another form."""
    from femocs_amd import fgpu
    if not srv.idle():
        raise RuntimeError("server busy")
    pc = plain_precond if plain_precond is not None else fgpu.FGPU_PC_JACOBI
    # inner_tol: the FP32 CG runs until |r| dropped 100-fold, then the FP64 residual replaces it
    o = fgpu.opts(tol=1e-9, relative=True, precision=precision, inner_tol=0.01, use_graph=False, precond=pc)
    srv.opts = o
    srv.heat_opts = o
    srv.batched = batched
    srv.warm = warm
    srv.make_batches()
    if not coarse_in_cg and pc == fgpu.FGPU_PC_JACOBI:
        raise ValueError("the two-level correction is on whenever the systems have one")


def run_phase(srv, n_warm: int, n_timed: int, arrivals: str, seed: int, keep=None, rate: float = RATE):
    """n_warm untimed then n_timed timed timesteps; returns (summary, timed requests).
    keep(requests) sees the untimed ones.  rate: total Poisson arrival rate."""
    from femocs_amd import serving
    rate = rate if arrivals == "poisson" else 1e12          # closed loop: no think time
    if n_warm:
        serving.run_clients(srv, rate=rate, duration=1e9, max_requests=n_warm, seed=seed, arrivals=arrivals)
        if keep is not None:
            keep(srv.stats.requests)
    srv.stats = serving.ServeStats()
    srv.ctx.sync()
    serving.run_clients(srv, rate=rate, duration=1e9, max_requests=n_timed, seed=seed + 1, arrivals=arrivals)
    srv.ctx.sync()
    summ = srv.stats.summary()
    bad = [r for r in srv.stats.requests if min(r.stage_iterations or [0]) <= 0]
    if bad:
        raise RuntimeError(f"{len(bad)} timesteps did not converge")
    return summ, list(srv.stats.requests)


def verify(srv, requests, history):
    """Error of every stage of every given timestep against an FP64 solve (|r|/|b| <= 1e-12,
    Jacobi + the two-level correction, warm from the served result) of the same request: the
    meshes moved to that timestep's displacement, the heat step from the served current of
    that timestep and the served temperature of the client's previous one (its initial
    temperature for its first).  history: (client, seq) -> snapshots of every served timestep."""
    from femocs_amd import fgpu
    o64 = fgpu.opts(tol=1e-12, relative=True, precision=fgpu.FGPU_FP64, use_graph=False)
    errs = {"poisson": 0.0, "current": 0.0, "heat": 0.0}
    rounds = {}
    for r in requests:
        rounds.setdefault(r.client, []).append(r)
    for v in rounds.values():
        v.sort(key=lambda q: q.seq)
    n_round = max(len(v) for v in rounds.values())
    count = 0
    for j in range(n_round):
        batch = [(cid, v[j]) for cid, v in rounds.items() if j < len(v)]
        for stage, name in enumerate(("poisson", "current", "heat")):
            systems = []
            for cid, req in batch:
                res = srv.resident[cid]
                s = res.stage_system(stage)
                if stage == 0:
                    res.mesh.displace(req.disp)
                    res.bmesh.displace(req.disp)
                    s.reassemble()
                elif stage == 1:
                    s.reassemble()
                else:
                    res.cur.solution_from_device(req.snaps[1].data_ptr())
                    prev = history.get((cid, req.seq - 1))
                    if prev is not None:
                        s.solution_from_device(prev[2].data_ptr())
                    else:
                        s.set_solution(res.T0)
                    s.reassemble()
                s.solution_from_device(req.snaps[stage].data_ptr())
                systems.append(s)
            stats = srv.ctx.solve_batch(systems, o64)
            for (cid, req), s, st in zip(batch, systems, stats):
                if st.iterations <= 0:
                    raise RuntimeError(f"FP64 reference solve did not converge (client {cid}, {name})")
                x64 = s.solution()
                s.solution_from_device(req.snaps[stage].data_ptr())
                xm = s.solution()
                errs[name] = max(errs[name], float(np.linalg.norm(xm - x64) / np.linalg.norm(x64)))
        count += len(batch)
    errs["max"] = max(errs.values())
    errs["timesteps"] = count
    return errs


def cpu_baseline(client):
    """CPU reference timing: the oracle's restatement of the reference algorithm (assembled
    CSR, QGauss(2), deal.II-style lifting, SSOR(1.2)-PCG with sequential sweeps and the matrix
    products and vector operations on OpenMP threads), one timestep of the smallest client
    (Laplace, current, implicit-Euler heat step, each to 1e-9 relative), on every host core
    of this process and on one; median of 5 runs after one warm-up."""
    from oracle import oracle as O
    from femocs_amd import serving
    cores = len(os.sched_getaffinity(0))
    X = client.next_positions()
    J, N = client.surface_loads(O.Mesh(client.X0, client.bulk_cells, *serving.HEAT_MARKS).face_centroids(2))

    def one(threads):
        t0 = time.perf_counter()
        _, _, itp, _, _ = O.poisson_laplace(X, client.cells, client.field, tol=1e-9, relative=True, threads=threads)
        m, phic, itc = O.current_solve(X, client.bulk_cells, J, tol=1e-9, threads=threads)
        xyz = m.x.reshape(-1, 3)
        T0 = client.initial_temperature(xyz)
        _, ith = O.heat_step(m, T0, phic, client.dt, N, serving.T_AMBIENT, tol=1e-9, threads=threads)
        if min(itp, itc, ith) <= 0:
            raise RuntimeError("CPU oracle did not converge")
        return time.perf_counter() - t0, (itp, itc, ith)

    out = {"unit": "timesteps/s", "kind": "port", "cores": cores}
    for label, th in (("all_cores", cores), ("one_core", 1)):
        one(th)
        runs = [one(th) for _ in range(5)]
        med = float(np.median([t for t, _ in runs]))
        out[label] = {"threads": th, "median_s": round(med, 3), "timesteps_per_s": round(1.0 / med, 4),
                      "iterations": list(runs[0][1])}
    out["value"] = out["all_cores"]["timesteps_per_s"]
    try:
        model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][0]
    except (OSError, IndexError):
        model = "unknown"
    out["cpu_model"] = model
    out["host_cpus"] = os.cpu_count()          # the host's CPUs; `cores`: those this process may use
    out["sample"] = (f"one timestep of the smallest client ({client.n_dofs} vacuum dofs, {len(client.cells)} vacuum + "
                     f"{len(client.bulk_cells)} bulk hexes): assemble + SSOR(1.2)-PCG for Laplace, current and heat; "
                     f"median of 5 after a warm-up, {cores} threads (value) and 1 thread")
    return out


def traffic_probe(n_launches: int = 5) -> None:
    """Counter target for the roofline's `traffic`: the Laplace and current systems of every
    client in one mixed batch (as served), `n_launches` element applies over the whole batch,
    then two device copies of 1 GiB (the calibration kernel of known size).  Run under
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
        b.add(r.cur)
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
    if not out["mesh_import"]["same_tables"]:
        raise RuntimeError("device and host mesh imports differ")
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
    if not out["interpolator_create"]["same_tables"]:
        raise RuntimeError("device and host interpolator tables differ")
    del iph
    # the first located transfer of a new interpolator (it allocates the point and chain buffers)
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
    # the same located transfer with the chained guess resolved on the host, point after point
    # (its first call also fetches the host copy of the tet tables)
    iph = interp.Interpolator(ctx, tip, interp.TYPES_VACUUM, host_chain=True)
    iph.set_nodal(ip.nodal())
    iph.points(atoms, 1)
    th, (cells_h, _) = _best_of(lambda: iph.points(atoms, 1), reps=2)
    if not np.array_equal(cells_h, ip.points(atoms, 1)[0]):
        raise RuntimeError("device and host chains differ")
    tr["rank1_locate_host_chain_atoms_per_s"] = round(len(atoms) / th)
    # dim 2: the surface faces (quadrangles for rank 3), then the tet or hex each one picks
    tris, t2t = tip.surface_triangles()
    quads = tip.surface_quads(tris)
    ip.set_triangles(tris, t2t, quads)
    iph.set_triangles(tris, t2t, quads)
    for rank in (1, 3):
        ts, (cs, vs) = _best_of(lambda: ip.surface_points(atoms, rank))
        tsh, (csh, vsh) = _best_of(lambda: iph.surface_points(atoms, rank), reps=1)
        if not (np.array_equal(cs, csh) and np.array_equal(vs, vsh)):
            raise RuntimeError("device and host surface chains differ")
        tr[f"surface_rank{rank}_locate_atoms_per_s"] = round(len(atoms) / ts)
        tr[f"surface_rank{rank}_host_chain_atoms_per_s"] = round(len(atoms) / tsh)
    del iph
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
    if not all(out["same_tables"].values()):
        log("device and host builds differ:", out["same_tables"])
        sys.exit(3)


N_BATCH = 64


def _make_batch_tip(i):
    from femocs_amd import mesh as M
    tip = M.make_tip_mesh(M.batched_tip_geometry(i), with_atoms=False)
    return tip.nodes, np.ascontiguousarray(tip.vacuum_cells(), np.int32)


def make_batch_tips(n: int = N_BATCH):
    from multiprocessing import get_context
    workers = max(1, min(n, 16, os.cpu_count() or 1))
    with get_context("fork").Pool(workers) as pool:
        return pool.map(_make_batch_tip, range(n))


def batch_systems(ctx, tips, coarse: bool, configure=None):
    """Laplace systems of the tips (phi = 0 on the copper, field -1 V/A on the top), from 0."""
    from femocs_amd import fgpu, serving
    out = []
    for X, cells in tips:
        m = fgpu.Mesh(ctx, X, cells, *serving.POISSON_MARKS)
        s = fgpu.System(m)
        if configure is not None:
            configure(m, s)
        s.reset()
        s.add_dirichlet(2, 0.0)
        s.add_neumann(8, None, -1.0)
        s.apply_dirichlet()
        s.fill_solution(0.0)
        if coarse:
            s.set_coarse(True)
        out.append(s)
    return out


def batched_bench(steps: int):
    """The batched workload: 64 tips of uneven size solved together, mixed precision with the
    two-level preconditioner, every CG iteration one launch per kernel over all 64 systems,
    `unroll` iterations captured per hipGraph (autotuned with the patch size); beside it the
    same batch in FP64 (two-level, graphs) and the plain FP64 form one tip at a time
    (Chebyshev-Jacobi).  Each run starts from phi = 0; the best of `steps` runs is kept."""
    from femocs_amd import autotune, fgpu
    t = time.perf_counter()
    tips = make_batch_tips()
    hexes = [len(c) for _, c in tips]
    log(f"{len(tips)} tips, {min(hexes)}..{max(hexes)} vacuum hexes, {sum(hexes)} in all "
        f"({time.perf_counter() - t:.1f} s)")
    ctx = fgpu.Context(0)
    big = int(np.argmax(hexes))
    from femocs_amd import serving
    gm = fgpu.Mesh(ctx, tips[big][0], tips[big][1], *serving.POISSON_MARKS)
    tuned, _ = autotune.autotune(fgpu.System(gm), fgpu.FGPU_MIXED)
    del gm

    def configure(mesh, system):
        mesh.set_patch_cells(tuned[0])
        system.set_patches_per_block(tuned[1])

    t = time.perf_counter()
    systems = batch_systems(ctx, tips, True, configure)
    ctx.sync()
    n_dofs = sum(s.mesh.n_dofs for s in systems)
    log(f"assembled with coarse levels ({time.perf_counter() - t:.1f} s), {n_dofs} dofs")

    def run(precision, unroll, reps=1, precond=fgpu.FGPU_PC_JACOBI):
        o = fgpu.opts(tol=1e-9, relative=True, precision=precision, inner_tol=0.01, use_graph=True,
                      check_every=unroll, precond=precond)
        best, stats = float("inf"), None
        for _ in range(reps):
            for s in systems:
                s.fill_solution(0.0)
            ctx.sync()
            t0 = time.perf_counter()
            st = ctx.solve_batch(systems, o)
            dt = time.perf_counter() - t0
            if dt < best:
                best, stats = dt, st
        if min(x.iterations for x in stats) <= 0:
            raise RuntimeError("a batched solve did not converge")
        return best, stats

    # unroll: CG iterations per captured graph (host convergence checks in between)
    unroll, unrolls = autotune.autotune_unroll(ctx, systems, fgpu.opts(tol=1e-9, relative=True,
                                                                      precision=fgpu.FGPU_MIXED, inner_tol=0.01))
    log("unroll (iterations per graph) seconds:", {k: round(v, 3) for k, v in unrolls.items()}, "->", unroll)
    t_mx, st_mx = run(fgpu.FGPU_MIXED, unroll, reps=max(1, steps))
    rel = [x.residual / x.rhs_norm for x in st_mx]
    x_mx = [s.solution() for s in systems]
    t_64, st_64 = run(fgpu.FGPU_FP64, unroll, reps=max(1, steps))
    err = max(float(np.linalg.norm(a - s.solution()) / np.linalg.norm(s.solution())) for a, s in zip(x_mx, systems))
    # plain FP64: one tip at a time, Chebyshev-Jacobi, no coarse level
    plain_o = fgpu.opts(tol=1e-9, relative=True, precision=fgpu.FGPU_FP64, use_graph=True,
                        precond=fgpu.FGPU_PC_CHEBYSHEV)
    t0 = time.perf_counter()
    for s in systems:
        s.fill_solution(0.0)
        if s.solve(plain_o).iterations <= 0:
            raise RuntimeError("a plain solve did not converge")
    t_plain = time.perf_counter() - t0
    out = {"metric": "batched tips: solves/s of 64 tips solved together (1 GPU)", "unit": "solves/s",
           "value": round(len(systems) / t_mx, 3), "data": "synthetic", "dtype": "f32 inner / f64 outer",
           "config": {"workload": f"{len(systems)} tips, radius U[2,12] nm, aspect U[3,10], "
                                  f"{min(hexes)}-{max(hexes)} vacuum Q1 hexes each ({sum(hexes)} in all, {n_dofs} dofs)",
                      "tol_rel": 1e-9, "cells_per_patch": tuned[0], "patches_per_block": tuned[1],
                      "unroll": unroll},
           "mixed_batched_s": round(t_mx, 4), "mixed_iterations_max": max(x.iterations for x in st_mx),
           "mixed_fp64_residual_max": max(rel),
           "fp64_two_level_batched_s": round(t_64, 4), "fp64_batched_solves_per_s": round(len(systems) / t_64, 3),
           "plain_fp64_sequential_s": round(t_plain, 4), "plain_fp64_solves_per_s": round(len(systems) / t_plain, 3),
           "speedup_vs_plain_fp64": round(t_plain / t_mx, 3), "err_mixed_vs_fp64_max": err,
           "unroll_seconds": {str(k): round(v, 4) for k, v in unrolls.items()}}
    print(json.dumps(out), flush=True)
    if not err <= ERR_LIMIT or not max(rel) <= 1e-9:
        log("batched results out of tolerance")
        sys.exit(3)


def q2_bench(steps: int):
    """The r = 10 nm tip (h = 60 nm on a 120 x 120 nm slab) with Q2 elements: the FP64
    Laplace solve to 1e-9 (Chebyshev-Jacobi and Jacobi), the Q1 solve of the same mesh beside
    it, and the roofline of the Q2 element apply (HIP events over `reps` launches).  Algorithmic
    bytes per apply: the 27 int32 dofs of every cell, the vertex coordinates (3 x 8 B), the two
    inputs z, p and the output (3 x 8 B per dof)."""
    from femocs_amd import fgpu, serving
    from femocs_amd import mesh as M
    tip = M.make_tip_mesh(M.tip_r10nm_geometry(), with_atoms=False)
    ctx = fgpu.Context(0)
    gm = fgpu.Mesh(ctx, tip.nodes, tip.vacuum_cells(), *serving.POISSON_MARKS)
    t = time.perf_counter()
    q = fgpu.Q2System(gm)
    t_setup = time.perf_counter() - t
    q.reset()
    q.add_dirichlet(2, 0.0)
    q.add_neumann(8, -1.0)
    q.apply_dirichlet()
    res = {}
    for name, pc in (("chebyshev", fgpu.FGPU_PC_CHEBYSHEV), ("jacobi", fgpu.FGPU_PC_JACOBI)):
        best, st = float("inf"), None
        for _ in range(max(1, steps)):
            q.fill_solution(0.0)
            ctx.sync()
            t0 = time.perf_counter()
            st = q.solve(fgpu.opts(tol=1e-9, relative=True, precond=pc, use_graph=False))
            best = min(best, time.perf_counter() - t0)
        if st.iterations <= 0:
            raise RuntimeError("Q2 solve did not converge")
        res[name] = {"seconds": round(best, 4), "iterations": st.iterations,
                     "rel_residual": st.residual / st.rhs_norm}
    s1 = fgpu.System(gm)
    s1.reset()
    s1.add_dirichlet(2, 0.0)
    s1.add_neumann(8, None, -1.0)
    s1.apply_dirichlet()
    best = float("inf")
    for _ in range(max(1, steps)):
        s1.fill_solution(0.0)
        ctx.sync()
        t0 = time.perf_counter()
        st1 = s1.solve(fgpu.opts(tol=1e-9, relative=True, precond=fgpu.FGPU_PC_CHEBYSHEV))
        best = min(best, time.perf_counter() - t0)
    tet_split = q.uses_tet_split()
    ms = q.time_apply(20)
    N, H, Nv = q.n_dofs, q.n_cells, gm.n_dofs
    # tet-split form: the FP64 per-cell metric (6 x 8 B) instead of the vertex coordinates
    byts = 108 * H + (48 * H if tet_split else 24 * Nv) + 24 * N
    gbs = byts / (ms * 1e-3) / 1e9
    # FP64 work of the tet-split form's apply per cell: 2025 fused multiply-adds (the five
    # contraction sweeps 162 + 243 + 1215 + 243 + 162, the metric 27 per Gauss point included)
    flops = 4050 * H
    tfs = flops / (ms * 1e-3) / 1e12
    general_ms = None
    if tet_split:
        gm.use_affine(False)
        general_ms = q.time_apply(20)
        gm.use_affine(True)
    out = {"metric": "Q2 Laplace solve, FP64 (1 GPU)", "unit": "solves/s", "data": "synthetic", "dtype": "f64",
           "value": round(1.0 / res["chebyshev"]["seconds"], 3),
           "config": {"workload": "r = 10 nm tip, vacuum hexes with Q2 elements (QGauss(3))", "hexes": H,
                      "q2_dofs": N, "q1_dofs": Nv, "tol_rel": 1e-9},
           "q2": res, "q2_setup_s": round(t_setup, 2),
           "q1_chebyshev": {"seconds": round(best, 4), "iterations": st1.iterations},
           "roofline": {"bound": "hbm", "kernel": "k_q2_apply" + ("<tet-split>" if tet_split else "<general>"),
                        "avg_us": round(ms * 1e3, 2), "algorithmic_bytes": byts, "achieved": round(gbs, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                        "fp64": {"flops": flops, "achieved_tflops": round(tfs, 2), "peak_tflops": FP64_PEAK_TFS,
                                 "frac": round(tfs / FP64_PEAK_TFS, 4)} if tet_split else None,
                        "general_form_avg_us": round(general_ms * 1e3, 2) if general_ms else None}}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help=f"timed steps of {PER_STEP} served timesteps each")
    ap.add_argument("--warmup", type=int, default=5, help="untimed steps (meshes already resident)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 and ablation runs")
    ap.add_argument("--no-verify", action="store_true", help="skip the per-timestep FP64 checks")
    ap.add_argument("--small", action="store_true", help="4 small tips instead (quick check)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only the counter target of roofline.traffic (see traffic_probe)")
    ap.add_argument("--paths", action="store_true",
                    help="the secondary paths instead of the headline (transfer, PIC pieces, coupled heat)")
    ap.add_argument("--setup", action="store_true",
                    help="only the remesh setup on the device (tet split, import, interpolator tables)")
    ap.add_argument("--batched", action="store_true",
                    help="the batched workload instead: 64 tips solved together (see batched_bench)")
    ap.add_argument("--q2", action="store_true", help="the Q2 Laplace solve of the r = 10 nm tip instead (q2_bench)")
    args = ap.parse_args()
    if args.q2:
        q2_bench(max(1, min(args.steps, 5)))
        return
    if args.batched:
        batched_bench(max(1, min(args.steps, 5)))
        return
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

    from femocs_amd import autotune, fgpu, serving

    n = 4 if args.small else N_CLIENTS
    t = time.perf_counter()
    clients = make_clients(n, args.small)
    vac = [len(c.cells) for c in clients]
    bulk = [len(c.bulk_cells) for c in clients]
    tot = [a + b for a, b in zip(vac, bulk)]
    log(f"clients: {n} tips, hexes per client {min(tot)}..{max(tot)} (vacuum {sum(vac)}, bulk {sum(bulk)}), "
        f"{sum(c.n_dofs for c in clients)} + {sum(c.n_bulk_dofs for c in clients)} dofs "
        f"({time.perf_counter() - t:.1f} s)")

    ctx = fgpu.Context(0)
    big = int(np.argmax(vac))
    gm = fgpu.Mesh(ctx, clients[big].X0, clients[big].cells, *serving.POISSON_MARKS)
    gs = fgpu.System(gm)
    tuned, _ = autotune.autotune(gs, fgpu.FGPU_MIXED)
    del gs, gm
    log("autotuned (cells per patch, patches per workgroup):", tuned)

    def configure(mesh, system):
        mesh.set_patch_cells(tuned[0])
        system.set_patches_per_block(tuned[1])

    history = {}
    n_w, n_t = args.warmup * PER_STEP, args.steps * PER_STEP
    # checkpoints of every served timestep of both mixed phases: about n / clients timesteps
    # per client and phase, with room for uneven arrivals
    per_client = sum(c.n_dofs + 2 * c.n_bulk_dofs for c in clients) / n
    snaps = Snapshots(ctx, 1.5 * (2 * n_w + 2 * n_t) * per_client)

    def snapshot(system):
        return snaps(system)

    o = fgpu.opts(tol=1e-9, relative=True, precision=fgpu.FGPU_MIXED, inner_tol=0.01, use_graph=False)
    srv = serving.Server(ctx, clients, o, configure=configure, coarse=True, snapshot=snapshot)
    t = time.perf_counter()
    srv.make_resident()                      # every tip imported and assembled before the clock runs
    ctx.sync()
    log(f"resident: {len(srv.resident)} clients ({time.perf_counter() - t:.1f} s, "
        f"{srv.used_bytes() / 2**30:.1f} GiB)")
    for c in clients:
        c.seq = 0
        c.rng = np.random.default_rng(1000 + c.cid)

    def keep(reqs):
        for r in reqs:
            history[(r.client, r.seq)] = r.snaps

    # headline: capacity, closed loop (every client's next timestep right after its result)
    serving.run_clients(srv, rate=1e12, duration=1e9, max_requests=n_w, seed=1, arrivals="closed")
    keep(srv.stats.requests)
    cap, cap_reqs = run_phase(srv, 0, n_t, "closed", 3)
    keep(cap_reqs)
    log("mixed capacity", cap)
    # latency at 200 timesteps/s (Poisson arrivals), same server and clients
    lat, lat_reqs = run_phase(srv, n_w, n_t, "poisson", 5, keep=keep)
    keep(lat_reqs)
    log("latency at 200/s", lat)
    # latency at a load the server sustains: Poisson arrivals at 80 % of the measured capacity
    sus_rate = 0.8 * cap["solves_per_s"]
    sus, sus_reqs = run_phase(srv, n_w, n_t, "poisson", 11, keep=keep, rate=sus_rate)
    keep(sus_reqs)
    log(f"latency at {sus_rate:.1f}/s", sus)

    # dominant kernel: the FP32 element apply over every Laplace and current system (all
    # clients resident, as served).  Algorithmic bytes per launch = compulsory traffic of a
    # matrix-free Q1 apply: cell connectivity 32 B/hex (8 x int32), the geometry (tet-split
    # form: the 6-float metric, 24 B/hex; otherwise coordinates 3 x 4 B/dof), the two input
    # vectors (z, p) 2 x 4 B/dof and one output 4 B/dof.
    b = fgpu.Batch(ctx, srv.opts)
    affine = True
    for r in srv.resident.values():
        for m, s in ((r.mesh, r.sys), (r.bmesh, r.cur)):
            affine = m.use_affine(True) and affine
            b.add(s)
    ms = b.time_kernel(0, reps=30)
    H = sum(r.mesh.n_cells + r.bmesh.n_cells for r in srv.resident.values())
    N = sum(r.mesh.n_dofs + r.bmesh.n_dofs for r in srv.resident.values())
    byts = 32 * H + (24 * H if affine else 12 * N) + 8 * N + 4 * N
    gbs = byts / (ms * 1e-3) / 1e9
    gather_ms = b.time_kernel(1, reps=30)
    coarse_ms = b.time_kernel(2, reps=30)
    del b
    # the same systems through the FP64 element apply (the plain form's operator): coordinates
    # 3 x 8 B/dof instead of the metric, vectors 3 x 8 B/dof
    b64 = fgpu.Batch(ctx, fgpu.opts(precision=fgpu.FGPU_FP64))
    for r in srv.resident.values():
        b64.add(r.sys)
        b64.add(r.cur)
    ms64 = b64.time_kernel(0, reps=30)
    del b64
    byts64 = 32 * H + 24 * N + 24 * N
    gbs64 = byts64 / (ms64 * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": measured_traffic(), "kernel": "k_patch<float>",
                "avg_us": round(ms * 1e3, 2), "algorithmic_bytes": byts, "cells": H, "dofs": N,
                "tet_split_form": affine, "gather_avg_us": round(gather_ms * 1e3, 2),
                "coarse_avg_us": round(coarse_ms * 1e3, 2),
                "fp64": {"kernel": "k_patch<double>", "avg_us": round(ms64 * 1e3, 2), "algorithmic_bytes": byts64,
                         "achieved": round(gbs64, 1), "frac": round(gbs64 / HBM_PEAK_GBS, 4)}}
    log("roofline", roofline)
    for r in srv.resident.values():
        for sysm in (r.sys, r.cur, r.heat):
            sysm.set_patches_per_block(tuned[1])    # (re)builds nothing: the batch left

    errs = None
    if not args.no_verify:
        t = time.perf_counter()
        errs = {"capacity": verify(srv, cap_reqs, history), "latency_200": verify(srv, lat_reqs, history),
                "latency_sustained": verify(srv, sus_reqs, history)}
        errs["max"] = max(errs["capacity"]["max"], errs["latency_200"]["max"], errs["latency_sustained"]["max"])
        errs["seconds"] = round(time.perf_counter() - t, 1)
        errs["snapshots_outside_pool"] = snaps.overflow
        log("mixed vs FP64 per stage", errs)
    history.clear()
    snaps.clear()
    for r in cap_reqs + lat_reqs + sus_reqs:
        r.snaps = []
    srv.snapshot = None

    plain = ablation = None
    if not args.no_plain:
        # ablation: FP64 CG with the same two-level preconditioner, warm starts and batching
        set_mode(srv, fgpu.FGPU_FP64, True, True, True)
        ablation, _ = run_phase(srv, PER_STEP, max(2 * PER_STEP, n_t // 4), "closed", 7)
        log("fp64 two-level batched", ablation)
        # plain form: FP64 Chebyshev-Jacobi, one timestep at a time, Laplace and current cold
        set_mode(srv, fgpu.FGPU_FP64, False, False, False, plain_precond=fgpu.FGPU_PC_CHEBYSHEV)
        plain, _ = run_phase(srv, 2, max(16, n_t // 8), "closed", 9)
        log("plain fp64", plain)

    cpu = None
    if not args.no_cpu:
        order = np.argsort(tot)
        cpu = cpu_baseline(clients[int(order[0])])
        log("cpu", cpu)

    ms_per = 1e3 * cap["seconds"] / args.steps
    out = {
        "metric": METRIC,
        "value": round(cap["solves_per_s"], 3),
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
        "config": {"workload": f"{n} MD clients, Cu nanotips r=2.6-8.4 nm, {min(tot)}-{max(tot)} Q1 hexes per client "
                               f"(vacuum + bulk); one solve = one timestep: Laplace + current + implicit-Euler heat "
                               f"(dt 4 fs); capacity with every client resubmitting at once, latency at Poisson "
                               f"arrivals of {RATE:.0f}/s",
                   "clients": n, "timesteps_per_step": PER_STEP,
                   "total_dofs": int(sum(c.n_dofs + c.n_bulk_dofs for c in clients)),
                   "total_hexes": int(sum(tot)), "tol_rel": 1e-9,
                   "precond": "jacobi + two-level aggregates of 256-1024 consecutive dofs by mesh size",
                   "precond_plain_fp64": "chebyshev-jacobi (lanczos lambda_max)", "parallelism": "single GPU"},
        "latency_ms": {"p50": round(lat["p50_ms"], 2), "p99": round(lat["p99_ms"], 2),
                       "offered_per_s": RATE, "served_per_s": round(lat["solves_per_s"], 2),
                       "sustained": {"p50": round(sus["p50_ms"], 2), "p99": round(sus["p99_ms"], 2),
                                     "offered_per_s": round(sus_rate, 2),
                                     "served_per_s": round(sus["solves_per_s"], 2)}},
        "mixed_capacity": cap,
        "latency_200": lat,
        "latency_sustained": sus,
        "plain_fp64": plain,
        "fp64_two_level_batched": ablation,
        "speedup_vs_plain_fp64": round(cap["solves_per_s"] / plain["solves_per_s"], 3) if plain else None,
        "speedup_vs_fp64_two_level_batched": round(cap["solves_per_s"] / ablation["solves_per_s"], 3) if ablation else None,
        "err_mixed_vs_fp64": errs,
        "err_mixed_vs_fp64_max": errs["max"] if errs else None,
        "autotuned": {"cells_per_patch": tuned[0], "patches_per_block": tuned[1]},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    if errs is not None and not errs["max"] <= ERR_LIMIT:
        log(f"mixed results differ from the FP64 solves by {errs['max']:.3g} > {ERR_LIMIT}")
        sys.exit(3)


if __name__ == "__main__":
    main()
