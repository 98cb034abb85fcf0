"""AMBER topology input for the PME variable-group path.

The reference's end-to-end test (python/tests/test_abfe.py:38-43) builds its
system with OpenMM's AmberPrmtopFile.createSystem(nonbondedMethod=PME,
nonbondedCutoff=1 nm, constraints=HBonds) and loads the coordinates from a
State XML checkpoint (:127).  This module reads the same files and produces the
force-field terms that system holds, in OpenMM units (nm, kJ/mol, radians, e):

  NonbondedForce   per-atom (charge, sigma, epsilon) with Lorentz-Berthelot
                   combining; exceptions = the scaled 1-4 pairs (charge product
                   / SCEE, epsilon / SCNB) plus every other pair of the
                   excluded-atoms list (zero charge product and epsilon);
  HarmonicBondForce  the bonds without hydrogen (bonds to hydrogen are
                   constraints under HBonds), k = 2 K_amber;
  HarmonicAngleForce every angle, k = 2 K_amber;
  PeriodicTorsionForce every dihedral term (propers and impropers).

The prmtop layout (%FLAG / %FORMAT sections, pointer-style atom indices 3*(i),
a negative third index marking a dihedral without 1-4 term and a negative
fourth index an improper) is AMBER's published file format.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field

import numpy as np

KCAL = 4.184                  # kJ per kcal
AMBER_CHARGE_SCALE = 18.2223  # prmtop charges are e * 18.2223


def read_prmtop_sections(path: str) -> dict:
    """All %FLAG sections of a prmtop file as lists of str / int / float."""
    with open(path) as fh:
        text = fh.read()
    out = {}
    for blk in text.split("%FLAG ")[1:]:
        lines = blk.split("\n")
        name = lines[0].split()[0]
        k = 1
        while k < len(lines) and not lines[k].startswith("%FORMAT"):
            k += 1
        m = re.match(r"%FORMAT\((\d+)([aAIiEeFf])(\d+)", lines[k].strip()) if k < len(lines) else None
        if m is None:
            raise ValueError(f"prmtop section {name}: no %FORMAT line")
        typ, width = m.group(2).lower(), int(m.group(3))
        vals = []
        for line in lines[k + 1:]:
            if line.startswith("%"):
                continue
            for c in range(0, len(line), width):
                f = line[c:c + width]
                if not f.strip():
                    continue
                vals.append(f.strip() if typ == "a" else (int(f) if typ == "i" else float(f)))
        out[name] = vals
    return out


def read_state_xml(path: str):
    """(box edges (3,), positions (N,3)) of an OpenMM State XML file with a
    rectangular box (nm)."""
    with open(path) as fh:
        text = fh.read()
    box = []
    for ax in "ABC":
        m = re.search(r"<%s x=\"([^\"]+)\" y=\"([^\"]+)\" z=\"([^\"]+)\"" % ax, text)
        if m is None:
            raise ValueError(f"{path}: no periodic box vector {ax}")
        box.append([float(m.group(g)) for g in (1, 2, 3)])
    box = np.array(box)
    if np.any(np.abs(box - np.diag(np.diag(box))) > 0):
        raise NotImplementedError("triclinic boxes are not supported")
    blk = text[text.index("<Positions>"):text.index("</Positions>")]
    pos = np.array([[float(a), float(b), float(c)] for a, b, c in
                    re.findall(r"<Position x=\"([^\"]+)\" y=\"([^\"]+)\" z=\"([^\"]+)\"", blk)])
    return np.diag(box).copy(), pos


def read_inpcrd(path: str) -> np.ndarray:
    """Coordinates (N,3, nm) of an AMBER ASCII restart file."""
    with open(path) as fh:
        lines = fh.read().split("\n")
    n = int(lines[1].split()[0])
    vals = []
    for line in lines[2:]:
        for c in range(0, len(line), 12):
            f = line[c:c + 12]
            if f.strip():
                vals.append(float(f))
            if len(vals) >= 3 * n:
                break
        if len(vals) >= 3 * n:
            break
    return np.array(vals[:3 * n]).reshape(n, 3) / 10.0


@dataclass
class AmberSystem:
    """Force-field terms of a prmtop under PME with HBonds constraints."""
    charges: np.ndarray
    sigmas: np.ndarray
    epsilons: np.ndarray
    masses: np.ndarray
    atom_names: list
    residue_ids: np.ndarray          # 1-based residue number per atom
    residue_names: list              # per residue
    exception_pairs: np.ndarray      # (E, 2)
    exception_qq: np.ndarray         # charge products (e^2)
    exception_sigma: np.ndarray
    exception_eps: np.ndarray
    bonds: np.ndarray                # (B, 2)
    bond_params: np.ndarray          # (B, 2): r0 (nm), k (kJ/mol/nm^2)
    angles: np.ndarray               # (A, 3)
    angle_params: np.ndarray         # (A, 2): theta0 (rad), k (kJ/mol/rad^2)
    torsions: np.ndarray             # (T, 4)
    torsion_params: np.ndarray       # (T, 3): periodicity, phase (rad), k (kJ/mol)
    constraints: np.ndarray          # (C, 2) atom pairs of the HBonds constraints
    constraint_lengths: np.ndarray
    box: np.ndarray | None = None
    extra: dict = field(default_factory=dict)

    @property
    def num_atoms(self) -> int:
        return len(self.charges)

    def residue_atoms(self, resid: int) -> np.ndarray:
        return np.nonzero(self.residue_ids == resid)[0]


def system_from_sections(s: dict) -> AmberSystem:
    """AmberSystem from prmtop sections (read_prmtop_sections)."""
    ptr = s["POINTERS"]
    n, ntypes = ptr[0], ptr[1]
    q = np.asarray(s["CHARGE"], dtype=np.float64) / AMBER_CHARGE_SCALE
    types = np.asarray(s["ATOM_TYPE_INDEX"], dtype=np.int64) - 1
    nbidx = np.asarray(s["NONBONDED_PARM_INDEX"], dtype=np.int64)
    A = np.asarray(s["LENNARD_JONES_ACOEF"], dtype=np.float64)
    B = np.asarray(s["LENNARD_JONES_BCOEF"], dtype=np.float64)
    sig_t = np.full(ntypes, 0.1)
    eps_t = np.zeros(ntypes)
    for t in range(ntypes):
        k = nbidx[ntypes * t + t] - 1
        if k < 0:
            raise NotImplementedError("10-12 hydrogen-bond terms are not supported")
        if A[k] != 0.0 and B[k] != 0.0:
            sig_t[t] = (A[k] / B[k]) ** (1.0 / 6.0) / 10.0
            eps_t[t] = B[k] * B[k] / (4.0 * A[k]) * KCAL
    for t in range(ntypes):
        for u in range(ntypes):
            k = nbidx[ntypes * t + u] - 1
            if eps_t[t] == 0.0 or eps_t[u] == 0.0:
                if A[k] != 0.0 or B[k] != 0.0:
                    raise NotImplementedError("LJ pair table is not Lorentz-Berthelot (NBFIX)")
                continue
            sg = 5.0 * (sig_t[t] + sig_t[u])
            ep = math.sqrt(eps_t[t] * eps_t[u]) / KCAL
            if abs(4.0 * ep * sg ** 12 / A[k] - 1.0) > 1e-4 or abs(4.0 * ep * sg ** 6 / B[k] - 1.0) > 1e-4:
                raise NotImplementedError("LJ pair table is not Lorentz-Berthelot (NBFIX)")
    sig, eps = sig_t[types], eps_t[types]

    rptr = np.asarray(s["RESIDUE_POINTER"], dtype=np.int64) - 1
    resid = np.zeros(n, dtype=np.int64)
    for r, start in enumerate(rptr):
        resid[start:] = r + 1

    def rows(name, k):
        return np.asarray(s.get(name, []), dtype=np.int64).reshape(-1, k)

    bh, bn = rows("BONDS_INC_HYDROGEN", 3), rows("BONDS_WITHOUT_HYDROGEN", 3)
    bk = np.asarray(s["BOND_FORCE_CONSTANT"], dtype=np.float64)
    br = np.asarray(s["BOND_EQUIL_VALUE"], dtype=np.float64)
    bonds = np.stack([bn[:, 0] // 3, bn[:, 1] // 3], 1)
    bond_params = np.stack([br[bn[:, 2] - 1] / 10.0, 2.0 * bk[bn[:, 2] - 1] * KCAL * 100.0], 1)
    constraints = np.stack([bh[:, 0] // 3, bh[:, 1] // 3], 1)
    constraint_lengths = br[bh[:, 2] - 1] / 10.0

    an = np.concatenate([rows("ANGLES_INC_HYDROGEN", 4), rows("ANGLES_WITHOUT_HYDROGEN", 4)])
    ak = np.asarray(s["ANGLE_FORCE_CONSTANT"], dtype=np.float64)
    a0 = np.asarray(s["ANGLE_EQUIL_VALUE"], dtype=np.float64)
    angles = np.stack([an[:, 0] // 3, an[:, 1] // 3, an[:, 2] // 3], 1)
    angle_params = np.stack([a0[an[:, 3] - 1], 2.0 * ak[an[:, 3] - 1] * KCAL], 1)

    dh = np.concatenate([rows("DIHEDRALS_INC_HYDROGEN", 5), rows("DIHEDRALS_WITHOUT_HYDROGEN", 5)])
    dk = np.asarray(s["DIHEDRAL_FORCE_CONSTANT"], dtype=np.float64)
    dn = np.asarray(s["DIHEDRAL_PERIODICITY"], dtype=np.float64)
    dp = np.asarray(s["DIHEDRAL_PHASE"], dtype=np.float64)
    ntyp = len(dk)
    scee = np.asarray(s.get("SCEE_SCALE_FACTOR", [1.2] * ntyp), dtype=np.float64)
    scnb = np.asarray(s.get("SCNB_SCALE_FACTOR", [2.0] * ntyp), dtype=np.float64)
    ti = dh[:, 4] - 1
    torsions = np.stack([dh[:, 0] // 3, dh[:, 1] // 3, np.abs(dh[:, 2]) // 3, np.abs(dh[:, 3]) // 3], 1)
    torsion_params = np.stack([np.abs(dn[ti]), dp[ti], dk[ti] * KCAL], 1)

    # exceptions: scaled 1-4 pairs first, then the rest of the excluded-atoms list
    ex = {}
    for row in dh:
        if row[2] < 0 or row[3] < 0:
            continue   # no 1-4 term (multi-term dihedral or ring closure) / improper
        i, l, t = row[0] // 3, row[3] // 3, row[4] - 1
        key = (min(i, l), max(i, l))
        if key in ex:
            continue
        ex[key] = (q[i] * q[l] / scee[t], 0.5 * (sig[i] + sig[l]), math.sqrt(eps[i] * eps[l]) / scnb[t])
    nex = s["NUMBER_EXCLUDED_ATOMS"]
    lst = s["EXCLUDED_ATOMS_LIST"]
    pos = 0
    for i in range(n):
        for c in range(nex[i]):
            j = lst[pos + c] - 1
            if j >= 0:
                key = (min(i, j), max(i, j))
                if key not in ex:
                    ex[key] = (0.0, 0.1, 0.0)
        pos += nex[i]
    keys = sorted(ex)
    ex_pairs = np.array(keys, dtype=np.int64).reshape(-1, 2)
    vals = np.array([ex[k] for k in keys], dtype=np.float64).reshape(-1, 3)
    box = None
    if "BOX_DIMENSIONS" in s:
        bd = s["BOX_DIMENSIONS"]
        box = np.array(bd[1:4]) / 10.0
    return AmberSystem(
        charges=q, sigmas=sig, epsilons=eps, masses=np.asarray(s["MASS"], dtype=np.float64),
        atom_names=list(s["ATOM_NAME"]), residue_ids=resid, residue_names=list(s["RESIDUE_LABEL"]),
        exception_pairs=ex_pairs, exception_qq=vals[:, 0], exception_sigma=vals[:, 1], exception_eps=vals[:, 2],
        bonds=bonds, bond_params=bond_params, angles=angles, angle_params=angle_params,
        torsions=torsions, torsion_params=torsion_params, constraints=constraints,
        constraint_lengths=constraint_lengths, box=box)


def read_prmtop(path: str) -> AmberSystem:
    return system_from_sections(read_prmtop_sections(path))


def system_from_npz(path: str):
    """(AmberSystem, arrays) from an archive of prmtop sections: entries
    "prmtop/<FLAG>" hold the sections, every other entry is returned in
    `arrays` (e.g. the box and positions of a State checkpoint)."""
    with np.load(path, allow_pickle=False) as z:
        sec = {k[len("prmtop/"):]: z[k] for k in z.files if k.startswith("prmtop/")}
        arrays = {k: z[k].copy() for k in z.files if not k.startswith("prmtop/")}
    sec = {k: ([str(v) for v in a] if a.dtype.kind == "U" else a.tolist()) for k, a in sec.items()}
    return system_from_sections(sec), arrays


def create_system(system: AmberSystem, box=None, nonbonded_cutoff: float = 1.0, ewald_tolerance: float = 5e-4,
                  constraints: bool = True):
    """An atmmetaforce.openmm.System holding what
    AmberPrmtopFile.createSystem(nonbondedMethod=PME, nonbondedCutoff, constraints=HBonds)
    holds: masses, the hydrogen-bond constraints, a PME NonbondedForce with the
    exceptions, HarmonicBondForce, HarmonicAngleForce and PeriodicTorsionForce
    (force group 0 each).  box: periodic box edges (nm); default the prmtop's."""
    from . import openmm as mm
    out = mm.System()
    for m in system.masses:
        out.addParticle(float(m))
    if constraints:
        for (i, j), d in zip(system.constraints, system.constraint_lengths):
            out.addConstraint(int(i), int(j), float(d))
    b = system.box if box is None else np.asarray(box, dtype=np.float64)
    if b is None:
        raise ValueError("no periodic box: pass box=")
    out.setDefaultPeriodicBoxVectors((float(b[0]), 0.0, 0.0), (0.0, float(b[1]), 0.0), (0.0, 0.0, float(b[2])))
    nb = mm.NonbondedForce()
    nb.setNonbondedMethod(mm.NonbondedForce.PME)
    nb.setCutoffDistance(nonbonded_cutoff)
    nb.setEwaldErrorTolerance(ewald_tolerance)
    nb.setUseDispersionCorrection(True)
    for q, s, e in zip(system.charges, system.sigmas, system.epsilons):
        nb.addParticle(float(q), float(s), float(e))
    for (i, j), qq, s, e in zip(system.exception_pairs, system.exception_qq, system.exception_sigma,
                                system.exception_eps):
        nb.addException(int(i), int(j), float(qq), float(s), float(e))
    out.addForce(nb)
    bf = mm.HarmonicBondForce()
    for (i, j), (r0, k) in zip(system.bonds, system.bond_params):
        bf.addBond(int(i), int(j), float(r0), float(k))
    out.addForce(bf)
    af = mm.HarmonicAngleForce()
    for (i, j, k), (t0, kk) in zip(system.angles, system.angle_params):
        af.addAngle(int(i), int(j), int(k), float(t0), float(kk))
    out.addForce(af)
    tf = mm.PeriodicTorsionForce()
    for (i, j, k, l), (n, ph, kk) in zip(system.torsions, system.torsion_params):
        tf.addTorsion(int(i), int(j), int(k), int(l), int(n), float(ph), float(kk))
    out.addForce(tf)
    return out
