"""Fused MD propagators (md/integrators.py).

``make_velocity_verlet_step`` and ``make_baoab_langevin_step`` return
propagators with the reference's step order (273-317: B(dt/2) A(dt) F B(dt/2);
405-453: B(dt/2) A(dt/2) O A(dt/2) F B(dt/2); positions wrapped after every A,
104-141).  Where the reference composes MomentumStep/PositionStep/
StochasticStep around a Potential, here one persistent HIP launch
(``kupsmi_md_lj``) runs any number of steps of every system, each system
resident in LDS.  The neighbor set of every step is the reference's (every
pair within the cutoff, rebuilt every step); the kernel finds it through
Verlet lists of radius cutoff + skin kept in HBM and rebuilt when an atom may
have moved by skin/2 (``skin=0``: a cell list in LDS every step).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from .. import _lib
from ..core.verlet import neighbor_capacity
from ..potential.lennard_jones import LennardJonesParameters
from .data import MDState

VELOCITY_VERLET = 0
BAOAB_LANGEVIN = 1
CSVR = 2
CSVR_NPT = 3


@dataclass
class MDObservables:
    potential_energy: torch.Tensor  # (n_sys, n_steps) f64
    kinetic_energy: torch.Tensor  # (n_sys, n_steps) f64
    virial: torch.Tensor  # (n_sys, 3, 3) f64 at the last step


@dataclass
class FusedLJPropagator:
    """n MD steps of every system of an MDState in one launch.

    skin: Verlet-list skin (length units of the positions); None picks 0.3 x
    the largest LJ sigma, 0 rebuilds a cell list every step.  It changes
    the cost, not the pair set.  max_neighbors: list entries per atom (None:
    from the densest system); a system whose lists overflow falls back to the
    per-step cell list."""

    params: LennardJonesParameters
    integrator: int = VELOCITY_VERLET
    seed: int = 0
    skin: float | None = None
    max_neighbors: int | None = None

    def skin_length(self) -> float:
        if self.skin is None:
            return 0.3 * float(np.max(self.params.sigma))
        if self.skin < 0:
            raise ValueError("skin must be >= 0")
        return float(self.skin)

    def _capacity(self, state: MDState, skin: float) -> int:
        if self.max_neighbors is not None:
            return int(self.max_neighbors)
        key = (id(state), state.n_sys, skin)
        if getattr(self, "_cap_key", None) != key:
            cut = float(state.cutoff.max().item()) if state.n_sys else 0.0
            cap = neighbor_capacity(state.sizes, state.tril.cpu().numpy(), cut, skin)
            self._cap_key, self._cap = key, min(cap, 32767)
        return self._cap

    def _tables(self, dev):
        return (torch.as_tensor(self.params.sigma, dtype=torch.float64, device=dev).contiguous(),
                torch.as_tensor(self.params.epsilon, dtype=torch.float64, device=dev).contiguous())

    def evaluate_forces(self, state: MDState) -> tuple[torch.Tensor, torch.Tensor]:
        """F, per-system energy and virial at the current positions (updates state.forces)."""
        obs = self._launch(state, 0, record=True)
        return obs.potential_energy[:, 0], obs.virial

    def __call__(self, state: MDState, n_steps: int = 1, record: bool = True) -> MDObservables | None:
        return self._launch(state, n_steps, record)

    def _launch(self, state: MDState, n_steps: int, record: bool):
        if state.sizes.size and state.sizes.max() > 4096:
            raise ValueError("the fused MD kernel holds systems of at most 4096 atoms")
        dev = state.positions.device
        sig, eps = self._tables(dev)
        n_rec = max(n_steps, 1)
        pot = kin = vir = None
        if record:
            pot = torch.empty((state.n_sys, n_rec), dtype=torch.float64, device=dev)
            kin = torch.empty((state.n_sys, n_rec), dtype=torch.float64, device=dev) if n_steps else None
            vir = torch.empty((state.n_sys, 3, 3), dtype=torch.float64, device=dev)
        kT = state.kT.contiguous()
        skin = self.skin_length() if n_steps >= 2 else 0.0
        cap = self._capacity(state, skin) if skin > 0 else 0
        _lib.check(_lib.load().kupsmi_md_lj(
            _lib.ctx(), _lib.stream_ptr(), int(self.integrator), int(n_steps), int(state.step),
            int(self.seed) & (2**64 - 1), state.n_sys, _lib.ptr(state.sys_off), state.n,
            _lib.ptr(state.positions), _lib.ptr(state.momenta), _lib.ptr(state.forces),
            _lib.ptr(state.masses), _lib.ptr(state.species), self.params.n_species,
            _lib.ptr(sig), _lib.ptr(eps), _lib.ptr(state.tril), _lib.ptr(state.cutoff),
            skin, cap, _lib.ptr(state.time_step), _lib.ptr(kT), _lib.ptr(state.friction),
            _lib.ptr(state.thermostat_time_constant),
            _lib.ptr(state.barostat) if self.integrator == CSVR_NPT else None,
            _lib.ptr(state.tril) if self.integrator == CSVR_NPT else None,
            _lib.ptr(pot), _lib.ptr(kin), _lib.ptr(vir)))
        state.step += n_steps
        if not record:
            return None
        if kin is None:
            kin = torch.zeros_like(pot)
        return MDObservables(pot, kin, vir)


def make_velocity_verlet_step(params: LennardJonesParameters, seed: int = 0,
                              skin: float | None = None) -> FusedLJPropagator:
    """NVE velocity Verlet (md/integrators.py:273-317)."""
    return FusedLJPropagator(params, VELOCITY_VERLET, seed, skin)


def make_baoab_langevin_step(params: LennardJonesParameters, seed: int = 0,
                             skin: float | None = None) -> FusedLJPropagator:
    """NVT BAOAB Langevin (md/integrators.py:405-453); O step noise from Philox4x32-10."""
    return FusedLJPropagator(params, BAOAB_LANGEVIN, seed, skin)


def make_csvr_step(params: LennardJonesParameters, seed: int = 0,
                   skin: float | None = None) -> FusedLJPropagator:
    """NVT velocity Verlet with the CSVR thermostat before every step
    (md/integrators.py:470-635; Bussi, Donadio and Parrinello 2007): one
    R1 ~ N(0, 1) for the batch and R2 ~ chi^2(3N - 4) per system, from
    Philox4x32-10 (seed, step); coupling time from the state."""
    return FusedLJPropagator(params, CSVR, seed, skin)


def make_csvr_npt_step(params: LennardJonesParameters, seed: int = 0) -> FusedLJPropagator:
    """NPT: CSVR, velocity Verlet, stochastic cell rescaling of cell and
    positions, forces again (md/integrators.py:829-888; Bernetti and Bussi
    2020); one R ~ N(0, 1) per step for the batch.  The cell of the state
    (state.tril) is updated in place; the cell list is rebuilt every step."""
    return FusedLJPropagator(params, CSVR_NPT, seed, 0.0)


def normals(seed: int, step: int, n_atoms: int, device=None) -> torch.Tensor:
    """The (n_atoms, 3) standard normals of the O step at global step ``step``."""
    dev = device or _lib.require_gpu()
    out = torch.empty((n_atoms, 3), dtype=torch.float32, device=dev)
    _lib.check(_lib.load().kupsmi_normals(_lib.stream_ptr(), int(seed) & (2**64 - 1), int(step),
                                          n_atoms, _lib.ptr(out)))
    return out


@dataclass
class EwaldObservables:
    lj: torch.Tensor          # (n_sys, n_steps)
    short_range: torch.Tensor
    long_range: torch.Tensor
    self_energy: torch.Tensor  # (n_sys,)
    kinetic_energy: torch.Tensor

    @property
    def potential_energy(self) -> torch.Tensor:
        return self.lj + self.short_range + self.long_range + self.self_energy[:, None]


class EwaldLJPropagator:
    """MD of charged LJ systems (LJ + Ewald) on the HBM-resident path, one
    force evaluation per step: integrate(phase 0) -> kupsmi_pair_forces (LJ and
    real-space Ewald, one pair loop) -> kupsmi_ewald_recip (k-space, added) ->
    integrate(phase 1).  Same step order as the fused LJ kernel.

    integrator CSVR puts the thermostat rescaling (kupsmi_thermostat_barostat
    phase 0) before the velocity Verlet step; CSVR_NPT also rescales cell and
    positions after the closing kick (phase 1, from that step's virial and
    kinetic energy) and evaluates the forces again at the rescaled state
    (md/integrators.py:826-888).  The cell changes every step, so NPT runs
    with a cell list per evaluation (no Verlet reuse)."""

    def __init__(self, params: LennardJonesParameters, ewald, integrator: int = BAOAB_LANGEVIN,
                 seed: int = 0, skin: float | None = None, max_neighbors: int | None = None):
        """skin: reuse a Verlet list of radius cutoff + skin across steps
        (rebuilt on the GPU per system when an atom has moved more than
        skin / 2); None evaluates with a fresh cell list every step."""
        self.params = params
        self.ewald = ewald
        self.integrator = integrator
        self.seed = seed
        self.skin = skin
        self.max_neighbors = max_neighbors
        self.verlet = None
        self._dev = None
        # library context of the force calls (None: the device's shared one)
        self.context = None
        # reciprocal-space kernels run on a second stream beside the pair
        # kernels, into buffers of their own, joined before the closing kick
        # (overlap = False: both on the current stream, one after the other)
        self.overlap = True
        # with overlap and Verlet lists, True holds the pair passes (not the
        # list upkeep) until the structure factors are done, so the list
        # upkeep runs beside them and the pair passes beside the force pass
        # (measured 1-5 % faster than False on the 512-melt workload at skins
        # 1.0 to 2.0 A); False lets the pair passes follow the list upkeep at
        # once
        self.gate_pairs = True
        self._side = None
        self._f2 = None
        self._vir2 = None

    def _prep(self, state: MDState):
        dev = state.positions.device
        if self._dev != dev:
            self._sig = torch.as_tensor(self.params.sigma, dtype=torch.float64, device=dev).contiguous()
            self._eps = torch.as_tensor(self.params.epsilon, dtype=torch.float64, device=dev).contiguous()
            self._alpha = torch.as_tensor(self.ewald.alpha, dtype=torch.float64, device=dev).contiguous()
            self._rc = torch.as_tensor(self.ewald.cutoff, dtype=torch.float64, device=dev).contiguous()
            self._kmax = torch.as_tensor(self.ewald.kmax, dtype=torch.float64, device=dev).contiguous()
            self._dev = dev
        self._maxn = int(state.sizes.max()) if state.sizes.size else 0
        if self.integrator == CSVR_NPT and (state.barostat is None or
                                            state.thermostat_time_constant is None):
            raise ValueError("NPT needs the state's barostat and thermostat coupling")
        if self.integrator == CSVR and state.thermostat_time_constant is None:
            raise ValueError("CSVR needs the state's thermostat coupling time")
        if self.skin is not None and self.verlet is None and self.integrator != CSVR_NPT:
            from ..core.verlet import VerletList, neighbor_capacity

            rc = max(float(np.max(self.ewald.cutoff)), float(state.cutoff.max().item()))
            cap = self.max_neighbors or neighbor_capacity(state.sizes, state.tril.cpu().numpy(), rc,
                                                          self.skin)
            self.verlet = VerletList(self.skin, cap)

    def _side_stream(self, state: MDState):
        dev = state.positions.device
        if self._side is None or self._side.device != dev:
            self._side = torch.cuda.Stream(device=dev)
        if self._f2 is None or self._f2.shape != state.forces.shape or self._f2.device != dev:
            self._f2 = torch.empty_like(state.forces)
        if self._vir2 is None or self._vir2.shape[0] != state.n_sys or self._vir2.device != dev:
            self._vir2 = torch.empty((state.n_sys, 3, 3), dtype=torch.float64, device=dev)
        return self._side

    def forces(self, state: MDState, e2, elr, vir=None):
        """Pair forces (LJ + real-space Ewald) on the current stream and the
        reciprocal-space Ewald on a second stream, concurrently: the k-space
        kernels write forces and virial into buffers of their own, added to
        the pair results after the join (the same additions, in the same
        order, as accumulating in place).  elr is set to the k-space energy."""
        lib = _lib.load()
        ctx = self.context.handle if self.context is not None else _lib.ctx()
        main = torch.cuda.current_stream(state.positions.device)
        side = self._side_stream(state) if self.overlap else main
        side.wait_stream(main)
        mx, my, mz = self.ewald.grid
        with torch.cuda.stream(side):
            _lib.check(lib.kupsmi_ewald_recip(
                ctx, _lib.stream_ptr(), state.n_sys, _lib.ptr(state.sys_off), state.n, self._maxn,
                _lib.ptr(state.positions), _lib.ptr(state.charges), _lib.ptr(state.tril),
                _lib.ptr(self._alpha), _lib.ptr(self._kmax), mx, my, mz, 0, _lib.ptr(self._f2),
                _lib.ptr(elr), _lib.ptr(self._vir2) if vir is not None else None, None))
        if self.overlap and self.gate_pairs and self.verlet is not None:
            # the pair call's list upkeep runs beside the structure factors,
            # its pair passes after them, beside the force pass
            ev = C.c_void_p()
            _lib.check(lib.kupsmi_ewald_phase_event(ctx, C.byref(ev)))
            _lib.check(lib.kupsmi_verlet_gate(self.verlet.handle, ev))
        if self.verlet is not None:
            _lib.check(lib.kupsmi_pair_forces_verlet(
                self.verlet.handle, _lib.stream_ptr(), state.n_sys, _lib.ptr(state.sys_off), state.n,
                self._maxn, _lib.ptr(state.positions), _lib.ptr(state.charges),
                _lib.ptr(state.species), self.params.n_species, _lib.ptr(self._sig),
                _lib.ptr(self._eps), _lib.ptr(state.cutoff), None, _lib.ptr(self._alpha),
                _lib.ptr(self._rc), None, _lib.ptr(state.tril), self.verlet.skin,
                self.verlet.max_neighbors, 0, _lib.ptr(state.forces), _lib.ptr(e2), _lib.ptr(vir)))
        else:
            _lib.check(lib.kupsmi_pair_forces(
                ctx, _lib.stream_ptr(), state.n_sys, _lib.ptr(state.sys_off), state.n,
                self._maxn, _lib.ptr(state.positions), _lib.ptr(state.charges),
                _lib.ptr(state.species), self.params.n_species, _lib.ptr(self._sig),
                _lib.ptr(self._eps), _lib.ptr(state.cutoff), None, _lib.ptr(self._alpha),
                _lib.ptr(self._rc), None, _lib.ptr(state.tril), 0, _lib.ptr(state.forces),
                _lib.ptr(e2), _lib.ptr(vir), None))
        main.wait_stream(side)
        state.forces.add_(self._f2)
        if vir is not None:
            vir.add_(self._vir2)

    def evaluate_forces(self, state: MDState):
        self._prep(state)
        dev = state.positions.device
        e2 = torch.zeros((state.n_sys, 2), dtype=torch.float64, device=dev)
        elr = torch.zeros(state.n_sys, dtype=torch.float64, device=dev)
        vir = torch.zeros((state.n_sys, 3, 3), dtype=torch.float64, device=dev)
        self.forces(state, e2, elr, vir)
        return e2, elr, vir

    def __call__(self, state: MDState, n_steps: int = 1) -> EwaldObservables:
        self._prep(state)
        obs = self._steps(state, n_steps, None, self._self_energy(state))
        state.step += n_steps
        return obs

    def _self_energy(self, state: MDState) -> torch.Tensor:
        from ..potential.ewald import self_energy

        dev = state.positions.device
        system = torch.repeat_interleave(torch.arange(state.n_sys, device=dev),
                                         torch.as_tensor(state.sizes, device=dev))
        return self_energy(state.charges, system, self._alpha, state.n_sys)

    def _steps(self, state: MDState, n_steps: int, step_dev, es) -> EwaldObservables:
        """n_steps launches sequences; step_dev: device int64 step counter read
        by the kernels (and advanced after each step) instead of state.step,
        which this method leaves alone."""
        dev = state.positions.device
        lib = _lib.load()
        e2 = torch.zeros((n_steps, state.n_sys, 2), dtype=torch.float64, device=dev)
        elr = torch.zeros((n_steps, state.n_sys), dtype=torch.float64, device=dev)
        kin = torch.zeros((n_steps, state.n_sys), dtype=torch.float64, device=dev)
        kT = state.kT.contiguous()
        seed = int(self.seed) & (2**64 - 1)
        csvr = self.integrator in (CSVR, CSVR_NPT)
        npt = self.integrator == CSVR_NPT
        base = VELOCITY_VERLET if csvr else int(self.integrator)
        vir = torch.zeros((state.n_sys, 3, 3), dtype=torch.float64, device=dev) if npt else None
        tau = _lib.ptr(state.thermostat_time_constant) if csvr else None
        sd = _lib.ptr(step_dev) if step_dev is not None else None
        for k in range(n_steps):
            t = int(state.step) + k
            if csvr:
                _lib.check(lib.kupsmi_thermostat_barostat(
                    _lib.stream_ptr(), 0, state.n_sys, _lib.ptr(state.sys_off), None,
                    _lib.ptr(state.momenta), _lib.ptr(state.masses), None,
                    _lib.ptr(state.time_step), _lib.ptr(kT), tau, None, None, None, seed, t, sd))
            _lib.check(lib.kupsmi_integrate(
                _lib.stream_ptr(), 0, base, state.n_sys, _lib.ptr(state.sys_off),
                self._maxn, _lib.ptr(state.positions), _lib.ptr(state.momenta),
                _lib.ptr(state.forces), _lib.ptr(state.masses), _lib.ptr(state.tril),
                _lib.ptr(state.time_step), _lib.ptr(kT), _lib.ptr(state.friction), seed,
                t, None, sd))
            self.forces(state, e2[k], elr[k], vir)
            _lib.check(lib.kupsmi_integrate(
                _lib.stream_ptr(), 1, base, state.n_sys, _lib.ptr(state.sys_off),
                self._maxn, _lib.ptr(state.positions), _lib.ptr(state.momenta),
                _lib.ptr(state.forces), _lib.ptr(state.masses), _lib.ptr(state.tril),
                _lib.ptr(state.time_step), _lib.ptr(kT), _lib.ptr(state.friction), seed,
                t, _lib.ptr(kin[k]), sd))
            if npt:
                _lib.check(lib.kupsmi_thermostat_barostat(
                    _lib.stream_ptr(), 1, state.n_sys, _lib.ptr(state.sys_off),
                    _lib.ptr(state.positions), None, None, _lib.ptr(state.tril),
                    _lib.ptr(state.time_step), _lib.ptr(kT), None, _lib.ptr(state.barostat),
                    _lib.ptr(vir), _lib.ptr(kin[k]), seed, t, sd))
                self.forces(state, e2[k], elr[k], vir)
            if step_dev is not None:
                _lib.check(lib.kupsmi_step_advance(_lib.stream_ptr(), sd, 1))
        return EwaldObservables(e2[:, :, 0].T, e2[:, :, 1].T, elr.T, es, kin.T)

    def capture(self, state: MDState, n_steps: int) -> "GraphedSteps":
        """Record n_steps of this propagator on ``state`` in a HIP graph
        (torch.cuda.CUDAGraph); ``replay()`` then runs them with one launch
        from the host.  The state's tensors are bound by address."""
        return GraphedSteps(self, state, n_steps)


class GraphedSteps:
    """n_steps of an EwaldLJPropagator recorded in one HIP graph.

    Every kernel of the steps (integrator phases, thermostat and barostat,
    pair and reciprocal-space forces, Verlet-list maintenance) is recorded;
    the step counter of the random streams is a device int64 that the
    recorded kernels read and advance, set from ``state.step`` before each
    replay, so replays draw the same noise as eager steps would.  The
    recorded launches use scratch, a Verlet list and k-space force and
    virial buffers of their own (a private library context and Verlet
    handle, frozen after capture) and hold the propagator's parameter
    tables: no eager call, of this propagator or any other, can reallocate or
    free memory the graph reads.
    Capture first evaluates the forces once eagerly with those so they are
    sized before recording (the recorded launches allocate nothing).  The
    observables of the last replay are in ``obs``."""

    def __init__(self, prop: EwaldLJPropagator, state: MDState, n_steps: int):
        from ..core.verlet import VerletList

        self.prop, self.state, self.n = prop, state, int(n_steps)
        prop._prep(state)
        self.context = _lib.Context()
        self.verlet = (VerletList(prop.verlet.skin, prop.verlet.max_neighbors)
                       if prop.verlet is not None else None)
        saved = prop.context, prop.verlet, prop._f2, prop._vir2
        # fresh k-space force / virial buffers for the recorded launches
        prop.context, prop.verlet, prop._f2, prop._vir2 = self.context, self.verlet, None, None
        try:
            f0 = state.forces.clone()
            prop.evaluate_forces(state)
            state.forces.copy_(f0)
            es = prop._self_energy(state)
            dev = state.positions.device
            self.step_dev = torch.zeros(1, dtype=torch.int64, device=dev)
            torch.cuda.synchronize()
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                self.obs = prop._steps(state, self.n, self.step_dev, es)
            torch.cuda.synchronize()
            self._buffers = (prop._f2, prop._vir2)
        finally:
            prop.context, prop.verlet, prop._f2, prop._vir2 = saved
        self.context.freeze()
        if self.verlet is not None:
            self.verlet.freeze()
        # the parameter tables the recorded launches read: held here, so a
        # later _prep of the propagator (another device or state) that
        # replaces them cannot free memory the graph uses
        self._tables = (prop._sig, prop._eps, prop._alpha, prop._rc, prop._kmax)

    def replay(self) -> EwaldObservables:
        self.step_dev.fill_(int(self.state.step))
        self.graph.replay()
        self.state.step += self.n
        return self.obs
