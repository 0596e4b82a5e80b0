"""Continuous-batching field-solve server for concurrent molecular-dynamics clients.

The reference solves one Laplace problem per MD timestep, inline, with the previous solution as
the initial guess only inside a PIC run (src/ProjectRunaway.cpp:422-431, 455-498; the solution
vector survives between `PoissonSolver::assemble` calls, src/DealSolver.cpp:369-387 resets it
only in `setup_system`), then, with heating on, the current and the heat step on the bulk mesh
(ProjectRunaway::solve_heat, src/ProjectRunaway.cpp:535-571: current.assemble + solve,
heat.assemble(dt) + solve).  Here many MD clients share one GPU; a request is one timestep:
the Laplace solve, and with `heat` the current solve and the implicit-Euler heat step after it,
in that order (the current density of the reference comes from the field: the stages of one
request run one after another, requests of different clients overlap):

* every client keeps its tip mesh and its last solution resident in device memory, so a new
  timestep only uploads moved vertex coordinates and starts CG from the previous potential;
* a request is admitted into the running batch (`fgpu.Batch`) as soon as its client is idle
  and the device-memory budget holds the client's mesh and system; idle clients are evicted
  least-recently-used first when the budget is short;
* the batch runs a fixed number of CG iterations per step, finished solves leave and new ones
  join between steps, so short solves are not held back by long ones.

`Server(..., batched=False, warm=False)` is the plain form: the same requests solved one at a
time, each from the Dirichlet value, in arrival order.
"""
from __future__ import annotations

import heapq
import time
from collections import OrderedDict, deque
from dataclasses import dataclass, field

import numpy as np

from . import fgpu
from . import mesh as M

# Poisson boundary ids for the vacuum mesh: top 8 (Neumann field), bottom/sides/other
# as src/PoissonSolver.cpp:52-55
POISSON_MARKS = (8, 2, 4, 2)
# bulk mesh: top and interface 2 (current and heat fluxes), bottom 7 (Dirichlet), sides 4, as
# CurrentHeatSolver::mark_mesh (src/CurrentHeatSolver.cpp:535-538)
HEAT_MARKS = (2, 7, 4, 2)
BYTES_PER_DOF_ESTIMATE = 480          # mesh + system device bytes per dof before the first import
CU_RHO_CP = 3.4496e-24                # J/(K A^3), include/CurrentHeatSolver.h:110
LORENTZ = 2.44e-8                     # src/Config.cpp:75
T_AMBIENT = 300.0
STAGES = ("poisson", "current", "heat")


@dataclass
class Request:
    client: int
    seq: int
    t_submit: float
    t_admit: float = -1.0
    t_done: float = -1.0
    iterations: int = 0                     # CG iterations of the Laplace solve
    residual: float = 0.0
    stage: int = 0                          # 0 Laplace, 1 current, 2 heat (in flight)
    stage_iterations: list = field(default_factory=list)
    disp: np.ndarray | None = None          # the surface displacement of this timestep
    moved: bool = True                      # the displacement was accepted (else the last mesh)
    snaps: list = field(default_factory=list)   # per stage: the solution's checkpoint handle
    x_init: np.ndarray | None = None        # kept only when the server records inputs
    positions: np.ndarray | None = None
    bpositions: np.ndarray | None = None    # bulk mesh positions (heat on)
    x_final: np.ndarray | None = None

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit


def vertex_altitudes(nodes: np.ndarray, tets: np.ndarray) -> np.ndarray:
    """Per tet vertex: the smallest altitude (distance to the opposite face) over its tets."""
    P = nodes[tets]
    v6 = np.abs(np.einsum("ij,ij->i", P[:, 1] - P[:, 0], np.cross(P[:, 2] - P[:, 0], P[:, 3] - P[:, 0])))
    h = np.full(len(nodes), np.inf)
    for k in range(4):
        a, b, c = [P[:, j] for j in range(4) if j != k]
        area2 = np.linalg.norm(np.cross(b - a, c - a), axis=1)          # twice the face area
        np.minimum.at(h, tets[:, k], v6 / np.maximum(area2, 1e-300))
    return h


class Client:
    """An MD client: a tip whose surface atoms vibrate about their lattice sites with Gaussian
    displacements of `sigma` (Angstrom) per timestep; the rest of the mesh follows through the
    tet-to-hex averaging of node positions.

    A resident mesh keeps its topology, so a move may not invert a tet: each surface node's
    displacement is capped at `cap` times its smallest tet altitude (the first-order volume
    change of a tet is then below 4 * cap of its volume).  Without the cap, Gaussian noise of
    0.1 A turns the thinnest Delaunay slivers inside out."""

    def __init__(self, cid: int, tip: M.TipMesh, field: float = -1.0, sigma: float = 0.1, seed: int = 0,
                 cap: float = 0.15, heat: bool = False, dt_fs: float = 4.0):
        self.cid = cid
        self.tip = tip
        self.field = field
        self.sigma = sigma
        self.rng = np.random.default_rng(seed)
        self.cells = np.ascontiguousarray(tip.vacuum_cells(), np.int32)
        self.X0 = np.ascontiguousarray(tip.nodes, np.float64)
        self.n_dofs = int(np.unique(self.cells).size)
        # current and heat on the bulk mesh (solve_heat), time step dt_fs femtoseconds
        self.heat = heat
        self.dt = dt_fs * 1e-15
        self.bulk_cells = np.ascontiguousarray(tip.bulk_cells(), np.int32) if heat else None
        self.n_bulk_dofs = int(np.unique(self.bulk_cells).size) if heat else 0
        W = M.hex_node_weights(tip.n_tetnodes, tip.tets)
        ns = tip.n_surface_nodes
        self.W_surf = W[:, :ns].tocsr()
        # the nodes a surface move reaches (the surface vertices and the edge, face and tet
        # centres around them): only these are sent to the server
        self.moved_ids = np.flatnonzero(np.diff(self.W_surf.indptr) > 0).astype(np.int32)
        self.W_moved = self.W_surf[self.moved_ids]
        self.limit = cap * vertex_altitudes(tip.nodes[: tip.n_tetnodes], tip.tets)[:ns]
        self.seq = 0
        self.last_disp = None                 # displacement the resident meshes hold (None: 0)

    def displacements(self) -> np.ndarray:
        d = self.rng.normal(0.0, self.sigma, size=(self.tip.n_surface_nodes, 3))
        norm = np.linalg.norm(d, axis=1)
        scale = np.minimum(1.0, self.limit / np.maximum(norm, 1e-300))
        return d * scale[:, None]

    def next_positions(self) -> np.ndarray:
        return self.X0 + self.W_surf @ self.displacements()

    def next_move(self):
        """(vertex ids, their new positions) of the next timestep."""
        return self.moved_ids, self.X0[self.moved_ids] + self.W_moved @ self.displacements()

    def positions_after(self, move) -> np.ndarray:
        X = self.X0.copy()
        X[move[0]] = move[1]
        return X

    def surface_loads(self, centroids: np.ndarray):
        """Per copper-surface face of the bulk mesh (face order of DealSolver::assemble_rhs):
        the current density and the Nottingham heat flux.  GETELEC computes both from the
        field in the reference (src/EmissionReader.cpp:199-228); it is absent here, so they are
        seeded smooth fields over the face centroids of the unmoved tip."""
        rng = np.random.default_rng(5000 + self.cid)
        c = centroids
        shape = 1.0 + 0.5 * np.sin(c[:, 0] / 7.0) * np.cos(c[:, 2] / 5.0) + 0.1 * rng.uniform(size=len(c))
        J = 1e-3 * shape
        N = 1e-9 * (1.0 + 0.3 * np.cos(c[:, 1] / 5.0))
        return J, N

    def initial_temperature(self, xyz: np.ndarray) -> np.ndarray:
        return T_AMBIENT + 100.0 * np.exp(-((xyz[:, 2] - xyz[:, 2].max()) / 20.0) ** 2)


@dataclass
class _Resident:
    mesh: fgpu.Mesh
    sys: fgpu.System
    bytes: int
    bmesh: fgpu.Mesh | None = None          # bulk mesh, current and heat systems (heat on)
    cur: fgpu.System | None = None
    heat: fgpu.System | None = None
    T0: np.ndarray | None = None            # temperature before the first step (dof order)

    def stage_system(self, stage: int) -> fgpu.System:
        return (self.sys, self.cur, self.heat)[stage]


@dataclass
class ServeStats:
    requests: list = field(default_factory=list)
    evictions: int = 0
    rejected_moves: int = 0
    prep_seconds: float = 0.0                # host time assembling requests (GPU idle meanwhile)
    step_seconds: float = 0.0                # time in batch steps and retirement
    t_start: float = 0.0
    t_end: float = 0.0

    def summary(self) -> dict:
        """Throughput counts the solves finished inside [t_start, t_end]; latency percentiles
        cover every finished request, including those drained after the window."""
        done = [r for r in self.requests if r.t_done >= 0]
        lat = np.array([r.latency for r in done]) * 1e3
        span = max(self.t_end - self.t_start, 1e-12)
        its = np.array([r.iterations for r in done]) if done else np.zeros(1)
        in_window = sum(1 for r in done if r.t_done <= self.t_end)
        # mean number of solves on the GPU at once (admission to result, over the span)
        busy = sum(r.t_done - r.t_admit for r in done if r.t_admit >= 0)
        stage_its = {}
        for k, name in enumerate(STAGES):
            v = [r.stage_iterations[k] for r in done if len(r.stage_iterations) > k]
            if v:
                stage_its[name] = round(float(np.mean(v)), 1)
        return {
            "solves": in_window,
            "solves_per_s": in_window / span,
            "p50_ms": float(np.percentile(lat, 50)) if done else None,
            "p99_ms": float(np.percentile(lat, 99)) if done else None,
            "mean_iterations": float(its.mean()),
            "mean_stage_iterations": stage_its,
            "mean_in_flight": round(busy / span, 2),
            "evictions": self.evictions,
            "rejected_moves": self.rejected_moves,
            "prep_seconds": round(self.prep_seconds, 4),
            "step_seconds": round(self.step_seconds, 4),
            "seconds": span,
        }


class Server:
    def __init__(self, ctx: fgpu.Context, clients: list[Client], opts: fgpu.SolveOpts,
                 budget_bytes: int = 64 << 30, batched: bool = True, warm: bool = True,
                 step_iterations: int = 32, max_batch: int = 64, record_inputs: bool = False, configure=None,
                 coarse: bool = False, pipelined: bool = False, heat_opts: fgpu.SolveOpts | None = None,
                 snapshot=None, bulk_iterations: int | None = None):
        self.ctx = ctx
        self.clients = clients
        self.opts = opts
        self.heat_opts = heat_opts if heat_opts is not None else opts
        self.budget = int(budget_bytes)
        self.batched = batched
        self.warm = warm
        self.step_iterations = step_iterations
        # CG iterations per step of the bulk-mesh batches (current, heat): their systems are
        # smaller, so their steps are shorter than the Laplace batch's they run beside
        self.bulk_iterations = bulk_iterations if bulk_iterations is not None else step_iterations
        self.max_batch = max_batch
        self.record_inputs = record_inputs
        self.configure = configure                    # called as configure(mesh, system) on import
        # snapshot(system) -> handle: called when a stage's solve ends (checkpoints of results)
        self.snapshot = snapshot
        # two-level preconditioner, built at residency: False, True (aggregate size by mesh
        # size), 256 / 512 / 1024, or a function of the dof count returning one of those
        self.coarse = coarse
        self.heat = any(c.heat for c in clients)
        if self.heat and not all(c.heat for c in clients):
            raise ValueError("either every client or none runs the heat stages")
        self.resident: OrderedDict[int, _Resident] = OrderedDict()   # LRU order
        self.known_bytes: dict[int, int] = {}
        self.pending: deque[Request] = deque()
        self.busy: dict[int, Request] = {}                            # client -> in-flight request
        self.by_sys: dict[int, Request] = {}
        self.stats = ServeStats()
        self.make_batches()
        # pipelined steps (Laplace only): the next step is queued before the host retires
        # (requests join one step later).  Off by default: admission synchronises with the
        # stream (the validity check of a displacement, the metric of the moved mesh), so the
        # queued step only delays it; measured 154 against 160 solves/s on the serving workload
        if pipelined and self.heat:
            raise ValueError("pipelined steps serve Laplace-only requests")
        self.pipelined = pipelined
        self._inflight = False

    def make_batches(self):
        """One continuous batch per stage: Laplace (vacuum meshes), current and heat (bulk
        meshes; heat systems carry per-cell coefficients, kappa(T) and the mass term).  Each
        batch has its own stream, so the three advance side by side on the GPU."""
        self.batch = fgpu.Batch(self.ctx, self.opts) if self.batched else None
        self.cbatch = fgpu.Batch(self.ctx, self.opts) if self.batched and self.heat else None
        self.hbatch = fgpu.Batch(self.ctx, self.heat_opts) if self.batched and self.heat else None

    # ---- memory-budget admission
    def used_bytes(self) -> int:
        return sum(r.bytes for r in self.resident.values())

    def _estimate(self, c: Client) -> int:
        return BYTES_PER_DOF_ESTIMATE * (c.n_dofs + 2 * c.n_bulk_dofs)

    def _make_resident(self, c: Client) -> _Resident | None:
        r = self.resident.get(c.cid)
        if r is not None:
            self.resident.move_to_end(c.cid)
            return r
        need = self.known_bytes.get(c.cid, self._estimate(c))
        if need > self.budget:
            raise ValueError(f"client {c.cid} needs {need} bytes, more than the whole budget")
        for cid in list(self.resident):
            if self.used_bytes() + need <= self.budget:
                break
            if cid not in self.busy:                  # idle clients leave, oldest first
                del self.resident[cid]
                self.stats.evictions += 1
        if self.used_bytes() + need > self.budget:
            return None
        m = fgpu.Mesh(self.ctx, c.X0, c.cells, *POISSON_MARKS)
        m.set_motion(c.moved_ids, c.W_moved)
        c.last_disp = None                            # a fresh import holds the unmoved tip
        s = fgpu.System(m)
        if self.configure is not None:
            self.configure(m, s)
        # the system is assembled once here, on the unmoved tip; requests move it and replay
        # the same terms on the device
        s.reset()
        s.add_dirichlet(2, 0.0)
        s.add_neumann(8, None, c.field)
        s.apply_dirichlet()
        s.fill_solution(0.0)
        if self.coarse:
            # built once per residency: it stays a valid preconditioner as the mesh moves
            s.set_coarse(self._aggregate(m.n_dofs))
        b = m.device_bytes() + s.device_bytes() + self._coarse_bytes(m.n_dofs)
        r = _Resident(m, s, 0)
        if c.heat:
            bm = fgpu.Mesh(self.ctx, c.X0, c.bulk_cells, *HEAT_MARKS)
            bm.set_motion(c.moved_ids, c.W_moved)             # rows outside the bulk are dropped
            cur, heat = fgpu.System(bm), fgpu.System(bm)
            if self.configure is not None:
                self.configure(bm, cur)
                self.configure(bm, heat)
            J, N = c.surface_loads(bm.face_centroids(2))
            # CurrentSolver::assemble (src/CurrentHeatSolver.cpp:421-484): sigma = 1, the
            # current density on the copper surface, phi = 0 on the bottom
            cur.reset()
            cur.add_neumann(2, J, 0.0)
            cur.add_dirichlet(7, 0.0)
            cur.apply_dirichlet()
            cur.fill_solution(0.0)
            if self.coarse:
                cur.set_coarse(self._aggregate(bm.n_dofs))
            # HeatSolver::assemble (src/CurrentHeatSolver.cpp:116-152): implicit Euler from the
            # present temperature, Joule heat of the present current, Nottingham flux, T_amb on
            # the bottom; every request replays it with the new temperature and current
            xyz, _ = bm.dofs()
            r.T0 = c.initial_temperature(xyz)
            heat.set_solution(r.T0)
            heat.reset()
            heat.heat_terms(cur, CU_RHO_CP / c.dt, LORENTZ)
            heat.add_neumann(2, N, 0.0)
            heat.add_dirichlet(7, T_AMBIENT)
            heat.apply_dirichlet()
            r.T0 = heat.solution()        # the first step's temperature: T_amb on the bottom
            if self.coarse:
                # from the first step's kappa(T); kappa drifts slowly and the mass term stays, so
                # the level remains a good preconditioner over the steps
                heat.set_coarse(self._aggregate(bm.n_dofs))
            r.bmesh, r.cur, r.heat = bm, cur, heat
            b += bm.device_bytes() + cur.device_bytes() + heat.device_bytes() + 2 * self._coarse_bytes(bm.n_dofs)
        r.bytes = b
        self.known_bytes[c.cid] = b
        self.resident[c.cid] = r
        return r

    def _coarse_bytes(self, n_dofs: int) -> int:
        agg = self._aggregate(n_dofs)
        if not agg:
            return 0
        nc = (n_dofs + agg - 1) // agg
        return nc * nc * 2 + nc * 12             # FP16 inverse, FP32 and FP64 corrections

    def _assemble(self, c: Client, r: _Resident, req: Request):
        t0 = time.perf_counter()
        try:
            self._assemble_timed(c, r, req)
        finally:
            self.stats.prep_seconds += time.perf_counter() - t0

    def _assemble_timed(self, c: Client, r: _Resident, req: Request):
        d = c.displacements()
        try:
            r.mesh.displace_noalloc(d)
            if r.bmesh is not None:
                r.bmesh.displace_noalloc(d)
            c.last_disp = d
        except fgpu.MeshInvertedError:
            # the step would need a new mesh; this request is solved on the last valid one
            self.stats.rejected_moves += 1
            req.moved = False
        # the displacement the meshes now hold (relative to the unmoved tip)
        req.disp = c.last_disp if c.last_disp is not None else np.zeros_like(d)
        s = r.sys
        # same boundary terms, new geometry: rebuilt on the device
        s.reassemble()
        if not self.warm:
            s.fill_solution(0.0)
        if self.record_inputs:
            xyz, vod = r.mesh.dofs()
            req.positions = c.X0.copy()
            req.positions[vod] = xyz
            req.x_init = s.solution()
            if r.bmesh is not None:
                xyz, vod = r.bmesh.dofs()
                req.bpositions = c.X0.copy()
                req.bpositions[vod] = xyz

    def _aggregate(self, n_dofs: int) -> int:
        """Aggregate size of the coarse level for a mesh of n_dofs (0: none)."""
        c = self.coarse(n_dofs) if callable(self.coarse) else self.coarse
        if c is True:
            return 256 if n_dofs <= 700000 else (512 if n_dofs <= 1500000 else 1024)
        return int(c) if c else 0

    def make_resident(self, cids=None) -> int:
        """Imports and assembles the clients' meshes ahead of their requests (while the budget
        allows); returns how many are resident."""
        for cid in (range(len(self.clients)) if cids is None else cids):
            if self._make_resident(self.clients[cid]) is None:
                break
        return len(self.resident)

    # ---- request flow
    def submit(self, req: Request):
        self.stats.requests.append(req)
        self.pending.append(req)

    def _admit(self, now: float) -> int:
        admitted = 0
        waiting = deque()
        while self.pending:
            req = self.pending.popleft()
            if req.client in self.busy or len(self.busy) >= self.max_batch:
                waiting.append(req)
                continue
            c = self.clients[req.client]
            r = self._make_resident(c)
            if r is None:
                waiting.append(req)
                continue
            self._assemble(c, r, req)
            req.t_admit = now
            self.busy[req.client] = req
            self.by_sys[r.sys.h.value] = req
            self.batch.add(r.sys)
            admitted += 1
        self.pending = waiting
        return admitted

    def _next_stage(self, req: Request) -> fgpu.System:
        """Prepares the next stage of a request on the device: the current solve on the moved
        bulk mesh, or the heat step from the new current and the last temperature."""
        t0 = time.perf_counter()
        r = self.resident[req.client]
        req.stage += 1
        nxt = r.stage_system(req.stage)
        nxt.reassemble()
        if req.stage == 1 and not self.warm:
            nxt.fill_solution(0.0)
        self.stats.prep_seconds += time.perf_counter() - t0
        return nxt

    def _stage_done(self, req: Request, s: fgpu.System, st: fgpu.SolveStats):
        req.stage_iterations.append(st.iterations)
        if req.stage == 0:
            req.iterations = st.iterations
            req.residual = st.residual
            if self.record_inputs:
                req.x_final = s.solution()
        if self.snapshot is not None:
            req.snaps.append(self.snapshot(s))

    def _finish(self, req: Request, now: float):
        req.t_done = now
        del self.busy[req.client]

    def step(self, clock=time.perf_counter) -> list[Request]:
        """Admits what fits, advances the GPU work by one step and returns finished requests."""
        done = []
        if not self.batched:
            # plain form: the oldest request is solved to the end, alone
            if not self.pending:
                return done
            req = self.pending.popleft()
            c = self.clients[req.client]
            r = self._make_resident(c)
            if r is None:
                raise RuntimeError("budget too small for a single client")
            self._assemble(c, r, req)
            req.t_admit = clock()
            self.busy[req.client] = req
            t0 = time.perf_counter()
            self._stage_done(req, r.sys, r.sys.solve(self.opts))
            while c.heat and req.stage < 2:
                s = self._next_stage(req)
                self._stage_done(req, s, s.solve(self.heat_opts if req.stage == 2 else self.opts))
            self._finish(req, clock())
            self.stats.step_seconds += time.perf_counter() - t0
            return [req]
        self._admit(clock())
        if not self.pipelined:
            active = []
            for b in (self.batch, self.cbatch, self.hbatch):
                if b is not None and len(b) > 0 and all(b is not a for a in active):
                    active.append(b)
            if not active:
                return done
            t0 = time.perf_counter()
            for b in active:
                b.step_async(self.step_iterations if b is self.batch else self.bulk_iterations)
            for b in active:
                b.wait()
            now = clock()
            retired = []
            for b in active:
                retired += b.retire()
        else:
            if len(self.batch) == 0 and not self._inflight:
                return done
            t0 = time.perf_counter()
            waited = self._inflight
            if waited:
                self.batch.wait()
                self._inflight = False
            now = clock()
            if len(self.batch) > 0:
                self.batch.step_async(self.step_iterations)
                self._inflight = True
            retired = self.batch.retire() if waited else []
        for s, st in retired:
            req = self.by_sys.pop(s.h.value)
            self._stage_done(req, s, st)
            if self.heat and req.stage < 2:
                nxt = self._next_stage(req)
                self.by_sys[nxt.h.value] = req
                (self.hbatch if req.stage == 2 else self.cbatch).add(nxt)
                continue
            self._finish(req, now)
            done.append(req)
        self.stats.step_seconds += time.perf_counter() - t0
        return done

    def idle(self) -> bool:
        return not self.pending and not self.busy and not self._inflight

    def solution(self, cid: int) -> np.ndarray:
        return self.resident[cid].sys.solution()


def run_clients(server: Server, rate: float, duration: float, seed: int = 0, max_requests: int | None = None,
                clock=time.perf_counter, arrivals: str = "poisson") -> ServeStats:
    """Drives the server with MD clients for `duration` seconds (or until `max_requests` were
    submitted), then drains what is in flight; throughput counts the solves finished inside the
    window.

    arrivals "poisson": the clients' requests form one Poisson process of total rate `rate`
    (each client a Poisson process of rate / n); a request's latency runs from its arrival
    time, and a client's requests are solved in arrival order.
    arrivals "closed": each client submits its next request an exponential think time (mean
    n / rate) after its previous result came back."""
    if arrivals not in ("poisson", "closed"):
        raise ValueError("arrivals must be 'poisson' or 'closed'")
    rng = np.random.default_rng(seed)
    n = len(server.clients)
    mean_think = n / rate
    t0 = clock()
    server.stats.t_start = t0
    if arrivals == "closed":
        events = [(t0 + rng.exponential(mean_think), cid) for cid in range(n)]
    else:
        events = [(t0 + rng.exponential(1.0 / rate), int(rng.integers(n)))]
    heapq.heapify(events)
    submitted = 0
    while True:
        now = clock()
        in_window = now - t0 < duration and (max_requests is None or submitted < max_requests)
        while in_window and events and events[0][0] <= now:
            ta, cid = heapq.heappop(events)
            c = server.clients[cid]
            server.submit(Request(cid, c.seq, ta if arrivals == "poisson" else now))
            c.seq += 1
            submitted += 1
            if arrivals == "poisson":
                heapq.heappush(events, (ta + rng.exponential(1.0 / rate), int(rng.integers(n))))
            if max_requests is not None and submitted >= max_requests:
                break
        finished = server.step(clock)
        now = clock()
        if arrivals == "closed":
            for req in finished:
                heapq.heappush(events, (now + rng.exponential(mean_think), req.client))
        if not in_window and server.idle():
            break
        if not finished and server.idle() and events:
            wait = events[0][0] - clock()
            if wait > 0:
                time.sleep(min(wait, 0.01))
    server.stats.t_end = min(clock(), t0 + duration) if max_requests is None else clock()
    return server.stats
