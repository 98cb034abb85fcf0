"""Seeded synthetic ATM workloads: jittered-lattice LJ + reaction-field systems.

Every coordinate (positions, displacements, box edges) is rounded to a 2^-20 nm
grid.  Below 16 nm such values are exact in FP32, so are their differences and
minimum-image shifts, and the squared distance of two of them is exact in FP64.
That makes the in-cutoff pair set of a system a well-defined integer object that
the FP32 device path and the FP64 CPU path must reproduce bit for bit.

Generator (one system):
  * n = ceil(N^(1/3)) lattice sites per edge, spacing a = L/n; site (ix, iy, iz)
    sits at ((ix+0.5)a, (iy+0.5)a, (iz+0.5)a), sites ordered with ix slowest and
    iz fastest, and the first N sites are used;
  * every coordinate jittered by U(-0.03, 0.03) nm;
  * sigma ~ U[0.25, 0.35] nm, epsilon ~ U[0.1, 0.8] kJ/mol,
    q ~ N(0, 0.4) e shifted to zero net charge;
  * the ligand is the first n_lig atoms; they carry the displacement d, every
    other atom carries (0, 0, 0);
  * draws, in order, from numpy.random.default_rng(seed): jitter (N x 3),
    sigma (N), epsilon (N), q (N).

The ATM parameter names and meanings follow the reference's global parameters
(openmmapi/include/ATMMetaForce.h:151-218); the direction convention
(> 0 means "from the reference state to the displaced state") follows
platforms/reference/src/ReferenceATMMetaForceKernels.cpp:84-85.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List

import numpy as np

GRID = 2.0 ** -20  # nm


def quantize(v):
    """Round to the 2^-20 nm grid (exact in FP32 for |v| < 16 nm)."""
    return np.round(np.asarray(v, dtype=np.float64) / GRID) * GRID


@dataclasses.dataclass
class ATMParams:
    """The nine ATM global parameters (ATMMetaForce.h:151-218)."""
    lambda1: float = 0.5
    lambda2: float = 0.5
    alpha: float = 0.1      # (kJ/mol)^-1
    u0: float = 110.0       # kJ/mol
    w0: float = 0.0         # kJ/mol
    umax: float = 200.0     # kJ/mol
    ubcore: float = 100.0   # kJ/mol
    acore: float = 0.0625
    direction: float = 1.0

    def as_tuple(self):
        return (self.lambda1, self.lambda2, self.alpha, self.u0, self.w0,
                self.umax, self.ubcore, self.acore, self.direction)


@dataclasses.dataclass
class Exceptions:
    """NonbondedForce exceptions: atom pairs (m, 2) with their charge product (e^2),
    sigma (nm) and epsilon (kJ/mol); charge product 0 and epsilon 0 is an exclusion."""
    pairs: np.ndarray
    charge_prod: np.ndarray
    sigma: np.ndarray
    epsilon: np.ndarray

    @property
    def count(self) -> int:
        return int(self.pairs.shape[0])


@dataclasses.dataclass
class System:
    """One ATM system: N atoms in a periodic box, orthorhombic (edges `box`) or, with
    `tilt` = (b_x, c_x, c_y), triclinic with vectors a = (box[0], 0, 0),
    b = (b_x, box[1], 0), c = (c_x, c_y, box[2]) (OpenMM's reduced form)."""
    positions: np.ndarray      # (N, 3) float64, nm
    charges: np.ndarray        # (N,) float64, e
    sigmas: np.ndarray         # (N,) float64, nm
    epsilons: np.ndarray       # (N,) float64, kJ/mol
    displacements: np.ndarray  # (N, 3) float64, nm
    box: np.ndarray            # (3,) float64, nm
    cutoff: float              # nm
    rf_dielectric: float = 78.5
    params: ATMParams = dataclasses.field(default_factory=ATMParams)
    seed: int = 0
    num_ligand: int = 0
    exceptions: Exceptions | None = None
    tilt: np.ndarray | None = None  # (3,) float64, nm: (b_x, c_x, c_y); None: orthorhombic

    @property
    def num_atoms(self) -> int:
        return int(self.positions.shape[0])


def lattice_system(num_atoms: int, box: float, cutoff: float, num_ligand: int,
                   displacement, seed: int, params: ATMParams | None = None,
                   rf_dielectric: float = 78.5) -> System:
    rng = np.random.default_rng(seed)
    n = int(math.ceil(num_atoms ** (1.0 / 3.0) - 1e-9))
    while n ** 3 < num_atoms:
        n += 1
    L = float(quantize(box))
    a = L / n
    idx = np.arange(num_atoms)
    ix, iy, iz = idx // (n * n), (idx // n) % n, idx % n
    sites = (np.stack([ix, iy, iz], axis=1).astype(np.float64) + 0.5) * a
    jitter = rng.uniform(-0.03, 0.03, size=(num_atoms, 3))
    sig = rng.uniform(0.25, 0.35, size=num_atoms)
    eps = rng.uniform(0.1, 0.8, size=num_atoms)
    q = rng.normal(0.0, 0.4, size=num_atoms)
    q = q - q.mean()
    pos = quantize(sites + jitter)
    d = np.zeros((num_atoms, 3))
    d[:num_ligand] = np.asarray(displacement, dtype=np.float64)
    d = quantize(d)
    return System(positions=pos, charges=q, sigmas=sig, epsilons=eps,
                  displacements=d, box=np.array([L, L, L]), cutoff=float(cutoff),
                  rf_dielectric=rf_dielectric,
                  params=params if params is not None else ATMParams(),
                  seed=seed, num_ligand=num_ligand)


def config1(seed: int = 1) -> System:
    """3000 atoms, 3.1 nm box, 0.9 nm cutoff, 24-atom ligand, d=(1.5,1.5,1.5).

    The lattice spacing is L/ceil(N^(1/3)) = 3.1/15 nm: a 0.31 nm lattice in a
    3.1 nm box holds only 1000 sites, so N, L and the cutoff are kept.
    """
    return lattice_system(3000, 3.1, 0.9, 24, (1.5, 1.5, 1.5), seed,
                          ATMParams(0.5, 0.5, 0.1, 110.0, 0.0, 200.0, 100.0, 0.0625, 1.0))


def config2(seed: int = 2) -> System:
    """23,000 atoms, 8.8 nm box, 1.0 nm cutoff, 40-atom ligand, d=(2.2,2.2,2.2)."""
    return lattice_system(23000, 8.8, 1.0, 40, (2.2, 2.2, 2.2), seed,
                          ATMParams(0.5, 0.5, 0.1, 110.0, 0.0, 200.0, 100.0, 0.0625, 1.0))


def lambda_schedule(num_replicas: int = 22):
    """lambda2_k = 0.5 k/(K-1); lambda1_k = lambda2_k for even k, lambda2_{k-1} for odd k."""
    lam2 = [0.5 * k / (num_replicas - 1) for k in range(num_replicas)]
    lam1 = [lam2[k] if k % 2 == 0 else lam2[k - 1] for k in range(num_replicas)]
    return lam1, lam2


def replica_batch(num_replicas: int = 22, base: System | None = None,
                  noise: float = 0.01) -> List[System]:
    """Copies of the 23k-atom system with N(0, noise) nm jitter and a lambda schedule.

    Replica k draws its noise from default_rng(1000 + k).
    """
    base = base if base is not None else config2()
    lam1, lam2 = lambda_schedule(num_replicas)
    out = []
    for k in range(num_replicas):
        rng = np.random.default_rng(1000 + k)
        pos = quantize(base.positions + rng.normal(0.0, noise, size=base.positions.shape))
        p = dataclasses.replace(base.params, lambda1=lam1[k], lambda2=lam2[k])
        out.append(dataclasses.replace(base, positions=pos, params=p, seed=1000 + k))
    return out


def uneven_batch(num_systems: int = 64, min_atoms: int = 5000, max_atoms: int = 60000,
                 cutoff: float = 1.0, seed: int = 4000) -> List[System]:
    """Systems with log-uniform atom counts at the 23k-system density (33.75 nm^-3).

    Batch-level draws from default_rng(seed): per system, in order, the atom
    count, the ligand size U{10..80} and two U[0, 0.5] lambdas (sorted).
    System s's atoms are drawn from default_rng(seed + 1 + s).  d = (L/4)(1,1,1).
    """
    rho = 23000.0 / 8.8 ** 3
    rng = np.random.default_rng(seed)
    out = []
    for s in range(num_systems):
        n = int(round(math.exp(rng.uniform(math.log(min_atoms), math.log(max_atoms)))))
        nlig = int(rng.integers(10, 81))
        l1, l2 = sorted(rng.uniform(0.0, 0.5, size=2))
        L = (n / rho) ** (1.0 / 3.0)
        p = ATMParams(float(l1), float(l2), 0.1, 110.0, 0.0, 200.0, 100.0, 0.0625, 1.0)
        out.append(lattice_system(n, L, cutoff, nlig, (L / 4, L / 4, L / 4), seed + 1 + s, p))
    return out


def tiny_system(num_atoms: int, box: float, cutoff: float, num_ligand: int,
                displacement, seed: int, params: ATMParams | None = None) -> System:
    """Small systems for exhaustive tests (same generator, any size)."""
    return lattice_system(num_atoms, box, cutoff, num_ligand, displacement, seed, params)


def box_vectors(box, tilt=None) -> np.ndarray:
    """Rows a, b, c (3, 3) of the box with edges `box` and tilts (b_x, c_x, c_y)."""
    t = np.zeros(3) if tilt is None else np.asarray(tilt, dtype=np.float64)
    return np.array([[box[0], 0.0, 0.0], [t[0], box[1], 0.0], [t[1], t[2], box[2]]], dtype=np.float64)


def triclinic(system: System, sites=(5, -4, 7)) -> System:
    """The same atoms in a triclinic box with the same edges and tilts
    (b_x, c_x, c_y) = sites x the lattice spacing L/n, rounded to the grid.  The
    brick [0, L)^3 holding the lattice is a unit cell of the triclinic lattice
    too, and whole-site tilts map lattice sites onto lattice sites, so the atoms
    still form a jittered simple-cubic lattice across every face of the box."""
    n = int(math.ceil(system.num_atoms ** (1.0 / 3.0) - 1e-9))
    while n ** 3 < system.num_atoms:
        n += 1
    a = float(system.box[0]) / n
    tilt = quantize(np.asarray(sites, dtype=np.float64) * a)
    if abs(tilt[0]) > 0.5 * system.box[0] or abs(tilt[1]) > 0.5 * system.box[0] or abs(tilt[2]) > 0.5 * system.box[1]:
        raise ValueError("tilts beyond half an edge: not OpenMM's reduced form")
    return dataclasses.replace(system, tilt=tilt)


def with_molecules(system: System, size: int = 3, scale14: float = 0.5, coulomb14: float = 0.8333,
                   chain_every: int = 4, crosslink_every: int = 0) -> System:
    """The same atoms grouped into molecules of `size` consecutive atoms (on the
    lattice they are z-neighbours ~0.3 nm apart), with the exceptions a force field
    would create from their bonds: every pair of a molecule is excluded, and every
    `chain_every`-th molecule is instead a chain whose 1-4 pair (first, last atom)
    keeps a scaled interaction (charge product coulomb14 q_i q_j, sigma the mixed
    sigma, epsilon scale14 sqrt(eps_i eps_j)), like AMBER's 1-4 scaling.

    crosslink_every > 0 adds, from every crosslink_every-th atom a, a pair with
    its lattice x-neighbour a + n^2 (n = lattice sites per edge; spatially close,
    far apart in index, like a disulfide bond), alternately an exclusion and a
    scaled pair."""
    n = system.num_atoms
    pairs, qq, sg, ep = [], [], [], []

    def scaled(i, j):
        qq.append(coulomb14 * system.charges[i] * system.charges[j])
        sg.append(0.5 * (system.sigmas[i] + system.sigmas[j]))
        ep.append(scale14 * math.sqrt(system.epsilons[i] * system.epsilons[j]))

    for m, a in enumerate(range(0, n - size + 1, size)):
        chain = size >= 4 and chain_every > 0 and m % chain_every == 0
        for i in range(a, a + size):
            for j in range(i + 1, a + size):
                pairs.append((i, j))
                if chain and i == a and j == a + size - 1:
                    scaled(i, j)
                else:
                    qq.append(0.0); sg.append(0.0); ep.append(0.0)
    if crosslink_every > 0:
        side = int(math.ceil(n ** (1.0 / 3.0) - 1e-9))
        while side ** 3 < n:
            side += 1
        stride = side * side
        seen = set(pairs)
        for k, a in enumerate(range(0, n - stride, crosslink_every)):
            if (a, a + stride) in seen:
                continue
            pairs.append((a, a + stride))
            if k % 2 == 0:
                qq.append(0.0); sg.append(0.0); ep.append(0.0)
            else:
                scaled(a, a + stride)
    ex = Exceptions(np.asarray(pairs, dtype=np.int32).reshape(-1, 2), np.asarray(qq), np.asarray(sg), np.asarray(ep))
    return dataclasses.replace(system, exceptions=ex)
