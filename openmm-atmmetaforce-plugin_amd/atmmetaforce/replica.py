"""Lambda-replica molecular dynamics with replica exchange for ATM on one GPU.

The reference runs one MD loop per alchemical state (example/abfe/abfe.py:
101-102, 149-160: a LangevinIntegrator over the force groups {0, ATM group},
one sample of (temperature, lambda, lambda1, lambda2, alpha, u0, w0, potential
energy, perturbation energy) every few thousand steps) and leaves the exchange
of states between replicas to an external asynchronous driver
(README.md:179-201).  Here all replicas of a lambda schedule live in one
batched context: every MD step integrates all of them in one launch sequence
(atm_md_step: the ATM force of every replica in one HIP graph, then the
Langevin update), and between cycles the driver proposes exchanges of
lambda-states between replicas.

Exchange rule.  The ATM energy of replica a in state i is
W_i(x_a) = e0(x_a) + B_i(u_a), with e0 independent of the state and the bias
B_i(u) = [alpha_i > 0] (l2_i - l1_i)/alpha_i ln(1 + exp(-alpha_i (u - u0_i)))
         + l2_i u + w0_i
(platforms/reference/src/ReferenceATMMetaForceKernels.cpp:86-98) of the
soft-core perturbation energy u.  Swapping states i and j between replicas a
and b is accepted with probability min(1, exp(-beta Delta)),
Delta = B_i(u_b) + B_j(u_a) - B_i(u_a) - B_j(u_b), which leaves the joint
distribution prod_a exp(-beta W_{s(a)}(x_a)) invariant (Metropolis).  Swaps
are proposed between neighbouring states of the schedule, in a random order,
several sweeps per cycle.  All states must share the direction (e0 must not
depend on the state) and the soft-core parameters umax, ubcore and acore (u is
the soft-core perturbation energy of the replica's own state).

Units: nm, ps, K, kJ/mol (the sample lines use the reference's kcal/mol
columns).
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
from typing import Iterable, List, Sequence

import numpy as np

from . import _capi as C
from .batch import ATMBatch
from .workloads import ATMParams, System

KB = 0.00831446261815324   # kJ/mol/K
KCAL = 4.184               # kJ per kcal


def softplus_bias(p: ATMParams, u: float) -> float:
    """B(u): the state-dependent part of W as a function of the soft-core
    perturbation energy (stable ln(1 + e^x))."""
    b = p.lambda2 * u + p.w0
    if p.alpha > 0.0:
        x = -p.alpha * (u - p.u0)
        lnee = x + math.log1p(math.exp(-x)) if x > 0.0 else math.log1p(math.exp(x))
        b += (p.lambda2 - p.lambda1) / p.alpha * lnee
    return b


def exchange_log_ratio(beta: float, pi: ATMParams, pj: ATMParams, ua: float, ub: float) -> float:
    """log of the Metropolis ratio for giving state j to replica a and state i
    to replica b (a in i, b in j before): -beta Delta."""
    delta = softplus_bias(pi, ub) + softplus_bias(pj, ua) - softplus_bias(pi, ua) - softplus_bias(pj, ub)
    return -beta * delta


def _check_states(states):
    """The exchange rule evaluates each state's softplus bias at the other replica's
    soft-core perturbation energy u_sc (computed with that replica's own umax,
    ubcore, acore): it is exact only when every state shares those three
    parameters and the direction, so schedules that vary them are rejected."""
    if len({p.direction > 0 for p in states}) != 1:
        raise ValueError("all states must have the same direction")
    if len({(float(p.umax), float(p.ubcore), float(p.acore)) for p in states}) != 1:
        raise ValueError("all states must share the soft-core parameters umax, ubcore and acore")


@dataclasses.dataclass
class Sample:
    cycle: int
    replica: int
    state: int
    temperature: float
    params: ATMParams
    lam: float
    potential_energy: float      # W, kJ/mol
    perturbation_energy: float   # u_sc, kJ/mol
    kinetic_energy: float        # kJ/mol

    def out_line(self) -> str:
        """The reference's sample line (example/abfe/abfe.py:158): temperature,
        lambda, lambda1, lambda2, alpha (1/(kcal/mol)), u0, w0, potential energy,
        perturbation energy (kcal/mol)."""
        p = self.params
        return "%f %f %f %f %f %f %f %f %f" % (self.temperature, self.lam, p.lambda1, p.lambda2, p.alpha * KCAL,
                                              p.u0 / KCAL, p.w0 / KCAL, self.potential_energy / KCAL,
                                              self.perturbation_energy / KCAL)


class ATMReplicaExchange:
    """Replicas of one system, one per lambda-state, integrated together.

    system: the system (positions, parameters, displacements, box, cutoff);
    states: the ATMParams of the schedule (one replica per state, replica k
        starts in state k); lambdas: the label printed in the lambda column
        (default (lambda1 + lambda2)/2);
    masses: per-atom masses (amu); positions: one (N, 3) array per replica or
        None (all start at system.positions)."""

    def __init__(self, system: System, states: Sequence[ATMParams], masses, temperature: float = 300.0,
                 friction: float = 1.0, dt: float = 0.002, seed: int = 1, positions=None,
                 lambdas: Sequence[float] | None = None, device: int = 0):
        if len(states) < 1:
            raise ValueError("need at least one state")
        _check_states(states)
        self.states = list(states)
        self.lambdas = list(lambdas) if lambdas is not None else [0.5 * (p.lambda1 + p.lambda2) for p in states]
        self.temperature, self.friction, self.dt = float(temperature), float(friction), float(dt)
        self.beta = 1.0 / (KB * self.temperature)
        R = len(states)
        reps = [dataclasses.replace(system, params=p) for p in states]
        if positions is not None:
            reps = [dataclasses.replace(r, positions=np.asarray(x, dtype=np.float64)) for r, x in zip(reps, positions)]
        self.batch = ATMBatch(reps, device=device)
        self.n = system.num_atoms
        self.state_of = list(range(R))          # replica -> state
        self.rng = np.random.default_rng(seed)
        self.attempts = np.zeros(max(R - 1, 1), dtype=np.int64)
        self.accepts = np.zeros(max(R - 1, 1), dtype=np.int64)
        m = np.ascontiguousarray(np.tile(np.asarray(masses, dtype=np.float64), R))
        x0 = np.ascontiguousarray(np.concatenate([r.positions for r in reps]).reshape(-1))
        lib = self.batch._lib
        self._lib = lib
        C.check(lib.atm_md_init(self.batch.handle, m.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                x0.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), self.temperature,
                                ctypes.c_uint64(int(seed) & 0xFFFFFFFFFFFFFFFF)))
        self.batch.positions_changed_on_device()
        self.cycle = 0

    # ------------------------------------------------------------ dynamics
    def step(self, steps: int):
        C.check(self._lib.atm_md_step(self.batch.handle, int(steps), self.dt, self.friction, self.temperature))
        self.batch.positions_changed_on_device()

    def state(self, replica: int):
        """(positions, velocities, kinetic energy) of a replica."""
        x = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        ke = ctypes.c_double()
        dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        C.check(self._lib.atm_md_get(self.batch.handle, replica, dp(x), dp(v), ctypes.byref(ke)))
        return x, v, ke.value

    # ------------------------------------------------------------ exchange
    def exchange(self, sweeps: int = 1):
        """Metropolis exchanges of neighbouring states (random order); the
        forces are re-evaluated for the replicas' new states."""
        R = len(self.states)
        if R < 2:
            return 0
        u = [self.batch.result(a)["perturbation_energy"] for a in range(R)]
        replica_in = {s: a for a, s in enumerate(self.state_of)}
        swapped = 0
        for _ in range(sweeps):
            for i in self.rng.permutation(R - 1):
                a, b = replica_in[i], replica_in[i + 1]
                self.attempts[i] += 1
                lr = exchange_log_ratio(self.beta, self.states[i], self.states[i + 1], u[a], u[b])
                if lr >= 0.0 or self.rng.random() < math.exp(lr):
                    self.accepts[i] += 1
                    replica_in[i], replica_in[i + 1] = b, a
                    swapped += 1
        for s, a in replica_in.items():
            if self.state_of[a] != s:
                self.state_of[a] = s
                self.batch.set_params(a, self.states[s])
        if swapped:
            self.batch.compute(forces=True)   # forces and results of the new states
        return swapped

    # ------------------------------------------------------------ driver
    def run(self, cycles: int, steps_per_cycle: int, exchange: bool = True, sweeps: int = 1,
            out=None) -> List[Sample]:
        """`cycles` x (`steps_per_cycle` MD steps, one sample per replica,
        exchanges); sample lines are written to `out` (a file object) if given."""
        samples = []
        for _ in range(cycles):
            self.step(steps_per_cycle)
            self.cycle += 1
            for a in range(len(self.states)):
                r = self.batch.result(a)
                s = self.state_of[a]
                ke = ctypes.c_double()
                C.check(self._lib.atm_md_get(self.batch.handle, a, None, None, ctypes.byref(ke)))
                smp = Sample(self.cycle, a, s, self.temperature, self.states[s], self.lambdas[s], r["energy"],
                             r["perturbation_energy"], ke.value)
                samples.append(smp)
                if out is not None:
                    print(smp.out_line(), file=out)
            if exchange:
                self.exchange(sweeps)
        if out is not None:
            out.flush()
        return samples

    def acceptance(self) -> np.ndarray:
        """Accepted fraction of the proposed swaps per neighbouring state pair."""
        return np.where(self.attempts > 0, self.accepts / np.maximum(self.attempts, 1), 0.0)

    def close(self):
        self.batch.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class ContextReplicaExchange:
    """Replica exchange over lambda-states for replicas that are separate
    Contexts of the OpenMM-compatible API (one per replica), as the
    reference's asynchronous driver runs them (README.md:179-201): any
    variable group works, e.g. the PME NonbondedForce of an AMBER system.

    replicas: [(context, integrator, atm_force)], each force the ATMMetaForce of
    that context's System (all in force group `atm_group`, default the first
    force's group); states: ATMParams (or dicts) of the schedule, replica k
    starting in state k; `lambdas` label the sample lines.  Exchanges use the
    rule of ATMReplicaExchange (Metropolis on the softplus bias of u_sc,
    neighbouring states, random order)."""

    _NAMES = ("Lambda1", "Lambda2", "Alpha", "U0", "W0", "Umax", "Ubcore", "Acore", "Direction")

    def __init__(self, replicas, states, temperature: float = 300.0, atm_group: int | None = None,
                 lambdas=None, seed: int = 1):
        if len(replicas) != len(states) or len(states) < 1:
            raise ValueError("one replica per state")
        self.replicas = [(c, i) for c, i, _ in replicas]
        self.forces = [f for _, _, f in replicas]
        force = self.forces[0]
        self.states = [p if isinstance(p, ATMParams) else ATMParams(**p) for p in states]
        _check_states(self.states)
        self.lambdas = list(lambdas) if lambdas is not None else [0.5 * (p.lambda1 + p.lambda2) for p in self.states]
        self.temperature = float(temperature)
        self.beta = 1.0 / (KB * self.temperature)
        self.atm_group = force.getForceGroup() if atm_group is None else atm_group
        self.state_of = list(range(len(self.states)))
        self.rng = np.random.default_rng(seed)
        n = max(len(self.states) - 1, 1)
        self.attempts = np.zeros(n, dtype=np.int64)
        self.accepts = np.zeros(n, dtype=np.int64)
        self.cycle = 0
        for a, (ctx, _) in enumerate(self.replicas):
            self._apply(ctx, self.states[a])

    def _apply(self, ctx, p: ATMParams):
        for name, v in zip(self._NAMES, p.as_tuple()):
            ctx.setParameter(getattr(self.forces[0], name)(), float(v))

    def _perturbation_energies(self):
        u = []
        for (ctx, _), f in zip(self.replicas, self.forces):
            ctx.getState(getEnergy=True, groups={self.atm_group})   # evaluates at the current positions
            u.append(f.getPerturbationEnergy(ctx))
        return u

    def step(self, steps: int):
        for _, integ in self.replicas:
            integ.step(int(steps))

    def exchange(self, sweeps: int = 1) -> int:
        R = len(self.states)
        if R < 2:
            return 0
        u = self._perturbation_energies()
        replica_in = {s: a for a, s in enumerate(self.state_of)}
        swapped = 0
        for _ in range(sweeps):
            for i in self.rng.permutation(R - 1):
                a, b = replica_in[i], replica_in[i + 1]
                self.attempts[i] += 1
                lr = exchange_log_ratio(self.beta, self.states[i], self.states[i + 1], u[a], u[b])
                if lr >= 0.0 or self.rng.random() < math.exp(lr):
                    self.accepts[i] += 1
                    replica_in[i], replica_in[i + 1] = b, a
                    swapped += 1
        for s, a in replica_in.items():
            if self.state_of[a] != s:
                self.state_of[a] = s
                self._apply(self.replicas[a][0], self.states[s])
        return swapped

    def run(self, cycles: int, steps_per_cycle: int, exchange: bool = True, sweeps: int = 1, out=None):
        """`cycles` x (MD steps, one sample line per replica, exchanges)."""
        samples = []
        for _ in range(cycles):
            self.step(steps_per_cycle)
            self.cycle += 1
            for a, (ctx, _) in enumerate(self.replicas):
                st = ctx.getState(getEnergy=True, groups={0, self.atm_group})
                s = self.state_of[a]
                smp = Sample(self.cycle, a, s, self.temperature, self.states[s], self.lambdas[s],
                             st.getPotentialEnergy(), self.forces[a].getPerturbationEnergy(ctx),
                             st.getKineticEnergy())
                samples.append(smp)
                if out is not None:
                    print(smp.out_line(), file=out)
            if exchange:
                self.exchange(sweeps)
        if out is not None:
            out.flush()
        return samples

    def acceptance(self) -> np.ndarray:
        return np.where(self.attempts > 0, self.accepts / np.maximum(self.attempts, 1), 0.0)
