"""Host-side mirror of the reference's solver classes, running on libfemocs_gpu.

``PoissonSolver`` keeps the method contract of the reference's PoissonSolver<3>
(include/PoissonSolver.h:24-105, src/PoissonSolver.cpp): import_mesh, setup, assemble, solve,
export_solution, check_limits/stat, size, get_n_cells.  The matrix never exists: the
operator is applied matrix-free on the GPU.  Solver differences, stated where they arise:
Jacobi instead of SSOR (SSOR's sweeps are sequential) and an optional FP32-inner mixed
precision mode; the converged solution is the same discrete solution.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import fgpu
from .fgpu import BID_COPPER_BOTTOM, BID_COPPER_SURFACE, BID_SIDES, BID_VACUUM_TOP


@dataclass
class FieldConfig:
    """Config::Field with the reference defaults (src/Config.cpp:63-72)."""
    E0: float = 0.0
    ssor_param: float = 1.2
    cg_tolerance: float = 1e-9
    n_cg: int = 10000
    V0: float = 0.0
    anode_BC: str = "neumann"
    mode: str = "laplace"
    V_min: float = -1.0
    V_max: float = 1e4


_default_ctx = None


def default_context() -> fgpu.Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = fgpu.Context(0)
    return _default_ctx


class Stat:
    def __init__(self):
        self.sol_min = 0.0
        self.sol_max = 0.0

    def __repr__(self):
        return f"min={self.sol_min}, max={self.sol_max}"


class DealSolver:
    """Shared part of the reference's DealSolver<3> (include/DealSolver.h:33-242)."""

    #: boundary ids passed to mark_boundary(top, bottom, sides, other)
    MARKS = (BID_VACUUM_TOP, BID_COPPER_SURFACE, BID_SIDES, BID_COPPER_SURFACE)

    def __init__(self, ctx: fgpu.Context | None = None):
        self.ctx = ctx or default_context()
        self.mesh = None
        self.system = None
        self.stat = Stat()
        self.dirichlet_bc_value = 0.0
        self.precision = fgpu.FGPU_FP64
        self.relative = False
        self.inner_tol = 1e-4
        self.use_graph = True
        # GPU preconditioner standing in for SSOR: FGPU_PC_JACOBI (with the two-level
        # correction of a system that has one) or FGPU_PC_CHEBYSHEV (FP64)
        self.preconditioner = fgpu.FGPU_PC_JACOBI
        self.last_stats = None
        #: FileWriter::write_time (src/FileWriter.cpp:23-26) is the caller's clock here: set
        #: True before an assembly whose write-step quantities are wanted
        self.write_step = False

    def write_time(self) -> bool:
        return self.write_step

    def import_mesh(self, vertices, cells) -> bool:
        """DealSolver::import_mesh (src/DealSolver.cpp:192-209): False on a bad mesh."""
        try:
            self.mesh = fgpu.Mesh(self.ctx, vertices, cells, *self.MARKS)
        except fgpu.FgpuError:
            self.mesh = None
            return False
        self.system = fgpu.System(self.mesh)
        return True

    def size(self) -> int:
        return self.mesh.n_dofs

    def get_n_cells(self) -> int:
        return self.mesh.n_cells

    def setup_system(self):
        """DealSolver::setup_system (src/DealSolver.cpp:369-387): solution = dirichlet value."""
        self.system.reset()
        self.system.fill_solution(self.dirichlet_bc_value)

    def export_solution(self) -> np.ndarray:
        """Solution per mesh vertex (DealSolver::export_solution, src/DealSolver.cpp:270-278)."""
        return self.system.solution()

    def import_solution(self, sol):
        self.system.set_solution(sol)

    def export_solution_grad(self) -> np.ndarray:
        return self.system.solution_grad()

    def check_limits(self, low_limit: float, high_limit: float) -> bool:
        """DealSolver::check_limits (src/DealSolver.cpp:158-167): True if out of limits."""
        self.stat.sol_min, self.stat.sol_max = self.system.minmax()
        return self.stat.sol_min < low_limit or self.stat.sol_max > high_limit

    def solve_cg(self, n_steps: int, tolerance: float, ssor_param: float = 1.2) -> int:
        """DealSolver::solve_cg (src/DealSolver.cpp:443-458): iterations, negative on failure.
        As the reference, ssor_param <= 0 solves without a preconditioner
        (src/DealSolver.cpp:447-452); otherwise the GPU preconditioner `self.preconditioner`
        (Jacobi by default, FGPU_PC_CHEBYSHEV for FP64) stands in for SSOR(ssor_param)."""
        precond = self.preconditioner if ssor_param > 0 else fgpu.FGPU_PC_NONE
        o = fgpu.opts(max_it=n_steps, tol=tolerance, relative=self.relative, precond=precond,
                      precision=self.precision, inner_tol=self.inner_tol, use_graph=self.use_graph)
        st = self.system.solve(o)
        self.last_stats = st
        return st.iterations


@dataclass
class ParticleSpecies:
    """The part of ParticleSpecies the field solve reads (include/ParticleSpecies.h): super-
    particle positions and solver cells, charge over eps0 and the super-particle weight."""
    pos: np.ndarray
    cell: np.ndarray
    q_over_eps0: float
    Wsp: float


class PoissonSolver(DealSolver):
    """PoissonSolver<3> (src/PoissonSolver.cpp)."""

    MARKS = (BID_VACUUM_TOP, BID_COPPER_SURFACE, BID_SIDES, BID_COPPER_SURFACE)

    def __init__(self, particles=None, conf: FieldConfig | None = None, interpolator=None,
                 ctx: fgpu.Context | None = None):
        super().__init__(ctx)
        self.particles = particles
        self.conf = conf or FieldConfig()
        self.interpolator = interpolator
        self.applied_field = 0.0
        self.applied_potential = 0.0

    def setup(self, field: float, potential: float):
        """PoissonSolver::setup (src/PoissonSolver.cpp:162-167)."""
        self.setup_system()
        self.applied_field = field
        self.applied_potential = potential

    def assemble(self, full_run: bool = True):
        """PoissonSolver::assemble (src/PoissonSolver.cpp:170-210) for laplace mode: the
        matrix is implicit, so full_run only matters for the cached diagonal."""
        if self.conf.anode_BC not in ("neumann", "dirichlet"):
            raise ValueError("Unimplemented anode BC: " + self.conf.anode_BC)
        s = self.system
        s.reset()
        s.add_dirichlet(BID_COPPER_SURFACE, self.dirichlet_bc_value)
        if self.conf.anode_BC == "neumann":
            s.add_neumann(BID_VACUUM_TOP, None, self.applied_field)
        else:
            s.add_dirichlet(BID_VACUUM_TOP, self.applied_potential)
        if self.conf.mode != "laplace":
            self.assemble_space_charge()
        # charge density of a write step, before the Dirichlet conditions; zeros otherwise
        if self.write_time():
            s.save_charge_density()
        s.apply_dirichlet()

    def assemble_space_charge(self):
        """PoissonSolver::assemble_space_charge(_fast) (src/PoissonSolver.cpp:267-319): each
        particle's charge q/eps0 * W_sp spread over its cell's dofs by the shape functions.
        ``particles``: pos (n, 3), cell (n) solver cells, q_over_eps0, Wsp; ``interpolator``:
        the Interpolator over this mesh's region hexes.  Nothing without particles."""
        P = self.particles
        if P is None or len(P.cell) == 0:
            return
        if self.interpolator is None:
            raise ValueError("space charge needs the interpolator of the solver's hexes")
        self.system.add_space_charge(self.interpolator, P.pos, P.cell, P.q_over_eps0 * P.Wsp)

    def export_charge_dens(self) -> np.ndarray:
        """PoissonSolver::export_charge_dens (src/PoissonSolver.cpp:141-149), per vertex."""
        return self.system.charge_density()

    def solve(self) -> int:
        """PoissonSolver::solve (include/PoissonSolver.h:54)."""
        return self.solve_cg(self.conf.n_cg, self.conf.cg_tolerance, self.conf.ssor_param)


@dataclass
class HeatingConfig:
    """Config::Heating with the reference defaults (src/Config.cpp:73-84)."""
    t_ambient: float = 300.0
    lorentz: float = 2.44e-8
    n_cg: int = 2000
    cg_tolerance: float = 1e-9
    ssor_param: float = 1.2
    delta_time: float = 10.0          # fs
    #: extensions (not in the reference): the heat step's theta (1 implicit Euler as the
    #: reference, 1/2 Crank-Nicolson) and sigma(T) = sigma0 / (1 + alpha (T - 300 K)) in the
    #: current solve instead of sigma = 1 (None)
    theta: float = 1.0
    sigma_T: tuple | None = None      # (sigma0, alpha)


CU_RHO_CP = 3.4496e-24                # J/(K A^3), include/CurrentHeatSolver.h:110


class _BulkSolver(DealSolver):
    """A solver on the bulk mesh of a CurrentHeatSolver, sharing its mesh."""

    def __init__(self, owner: "CurrentHeatSolver"):
        super().__init__(owner.ctx)
        self.owner = owner
        self.conf = owner.conf

    def solve(self) -> int:
        """EmissionSolver::solve (include/CurrentHeatSolver.h:35)."""
        return self.solve_cg(self.conf.n_cg, self.conf.cg_tolerance, self.conf.ssor_param)


class CurrentSolver(_BulkSolver):
    """CurrentSolver<3> (src/CurrentHeatSolver.cpp:400-484): sigma = 1, the per-face current
    densities as Neumann flux on the copper surface, phi = 0 on the copper bottom."""

    def __init__(self, owner):
        super().__init__(owner)
        self.bc_values = None            # one current density per copper-surface face

    def assemble(self):
        s = self.system
        s.reset()
        if self.conf.sigma_T is not None:
            s.sigma_terms(self.owner.heat.system, *self.conf.sigma_T)
        s.add_neumann(BID_COPPER_SURFACE, self.bc_values, 0.0)
        s.add_dirichlet(BID_COPPER_BOTTOM, self.dirichlet_bc_value)
        s.apply_dirichlet()


class HeatSolver(_BulkSolver):
    """HeatSolver<3> (src/CurrentHeatSolver.cpp:106-152, 347-393): implicit Euler with kappa and
    sigma at the previous temperature, Joule heat from the current solution, Nottingham heat
    per copper-surface face, T = T_ambient on the copper bottom."""

    def __init__(self, owner):
        super().__init__(owner)
        self.bc_values = None            # Nottingham heat per copper-surface face

    def assemble(self, delta_time: float):
        """delta_time in seconds (the caller converts fs, src/ProjectRunaway.cpp:549)."""
        if not delta_time > 0:
            raise ValueError(f"Invalid delta time: {delta_time}")
        s = self.system
        s.reset()
        s.heat_terms(self.owner.current.system, CU_RHO_CP / delta_time, self.conf.lorentz, self.conf.theta)
        if self.bc_values is not None:
            s.add_neumann(BID_COPPER_SURFACE, self.bc_values, 0.0)
        s.add_dirichlet(BID_COPPER_BOTTOM, self.dirichlet_bc_value)
        s.apply_dirichlet()


class CurrentHeatSolver:
    """CurrentHeatSolver<3> (src/CurrentHeatSolver.cpp:487-538): current and heat solvers on one
    bulk mesh, marked (copper surface, copper bottom, sides, copper surface)."""

    MARKS = (BID_COPPER_SURFACE, BID_COPPER_BOTTOM, BID_SIDES, BID_COPPER_SURFACE)

    def __init__(self, conf: HeatingConfig | None = None, ctx: fgpu.Context | None = None):
        self.ctx = ctx or default_context()
        self.conf = conf or HeatingConfig()
        self.mesh = None
        self.current = CurrentSolver(self)
        self.heat = HeatSolver(self)

    def import_mesh(self, vertices, cells) -> bool:
        try:
            self.mesh = fgpu.Mesh(self.ctx, vertices, cells, *self.MARKS)
        except fgpu.FgpuError:
            self.mesh = None
            return False
        for sub in (self.current, self.heat):
            sub.mesh = self.mesh
            sub.system = fgpu.System(self.mesh)
        return True

    def setup(self, temperature: float):
        """CurrentHeatSolver::setup (src/CurrentHeatSolver.cpp:510-514)."""
        self.heat.dirichlet_bc_value = temperature
        self.current.setup_system()
        self.heat.setup_system()

    def export_temp_rho(self):
        """(temperature per dof, current density -grad phi_c per dof) (src/CurrentHeatSolver.cpp:525-532)."""
        return self.heat.export_solution(), self.current.export_solution_grad()
