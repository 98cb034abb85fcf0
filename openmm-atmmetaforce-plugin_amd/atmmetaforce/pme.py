"""ATM evaluation with a PME NonbondedForce as the variable group, and the
bonded terms beside it, on one MI355X (atm_pme_* / atm_bonded_* of
include/atm_mi355x.h).

This is the setup of the reference's end-to-end test
(python/tests/test_abfe.py): an AMBER system under PME with a 1 nm cutoff
whose NonbondedForce is the ATM variable group (:86-96), bonded terms and
restraints in group 0.  ATMPMEForce evaluates U1 (x), U2 (x + d), the
soft-core/softplus combine and the hybrid force in one device pass;
BondedForces evaluates the harmonic bonds, angles and periodic torsions.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi as C


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _params(p) -> C.Params:
    if isinstance(p, C.Params):
        return p
    if hasattr(p, "as_tuple"):
        return C.Params(*p.as_tuple())
    if isinstance(p, dict):
        return C.Params(**p)
    return C.Params(*p)


class ATMPMEForce:
    """ATM force whose variable group is a PME NonbondedForce.

    charges (e), sigmas (nm), epsilons (kJ/mol), displacements (N,3 nm);
    exceptions = (pairs (E,2), charge products, sigmas, epsilons) as
    NonbondedForce.addException; box (3,) nm; cutoff nm."""

    def __init__(self, charges, sigmas, epsilons, displacements, box, cutoff=1.0, exceptions=None, params=None,
                 ewald_tolerance=5e-4, dispersion_correction=True, grid=None, alpha=0.0, device=0):
        self._lib = C.load()
        q = np.ascontiguousarray(charges, dtype=np.float64)
        self.n = len(q)
        d = C.PmeDesc()
        d.num_atoms = self.n
        for k in range(3):
            d.box[k] = float(box[k])
            d.grid[k] = int(grid[k]) if grid is not None else 0
        d.cutoff = float(cutoff)
        d.ewald_tolerance = float(ewald_tolerance)
        d.alpha = float(alpha)
        d.dispersion_correction = 1 if dispersion_correction else 0
        h = ctypes.c_void_p()
        C.check(self._lib.atm_pme_create(int(device), ctypes.byref(d), ctypes.byref(h)))
        self._h = h
        self.box = np.array(box, dtype=np.float64)
        s = np.ascontiguousarray(sigmas, dtype=np.float64)
        e = np.ascontiguousarray(epsilons, dtype=np.float64)
        dd = np.ascontiguousarray(displacements, dtype=np.float64).reshape(self.n, 3)
        if s.shape != (self.n,) or e.shape != (self.n,):
            raise ValueError("sigmas and epsilons must have one entry per atom")
        C.check(self._lib.atm_pme_set_particles(self._h, _dp(q), _dp(s), _dp(e), _dp(dd)))
        if exceptions is not None:
            self.set_exceptions(*exceptions)
        if params is not None:
            self.set_params(params)
        self._F = np.zeros((self.n, 3))

    def set_exceptions(self, pairs, charge_prod, sigma, epsilon):
        p = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        qq = np.ascontiguousarray(charge_prod, dtype=np.float64)
        sg = np.ascontiguousarray(sigma, dtype=np.float64)
        ep = np.ascontiguousarray(epsilon, dtype=np.float64)
        C.check(self._lib.atm_pme_set_exceptions(self._h, len(p), _ip(p), _dp(qq), _dp(sg), _dp(ep)))

    def set_energy_offsets(self, state1, state2):
        """Energies (kJ/mol) of the other variable-group forces at x and at x + d
        (atm_pme_set_energy_offsets): added to U1 and U2 before the soft-core combine."""
        C.check(self._lib.atm_pme_set_energy_offsets(self._h, float(state1), float(state2)))

    def set_params(self, params):
        p = _params(params)
        C.check(self._lib.atm_pme_set_params(self._h, ctypes.byref(p)))

    def info(self):
        a = ctypes.c_double()
        g = (ctypes.c_int32 * 3)()
        C.check(self._lib.atm_pme_get_info(self._h, ctypes.byref(a), g))
        return a.value, tuple(g)

    def evaluate(self, positions, box=None, forces=True, state1_only=False):
        """(result dict, hybrid force (N,3) or None, energy parts dict);
        state1_only: the NonbondedForce alone at x (energy = U1, force F1)."""
        x = np.ascontiguousarray(positions, dtype=np.float64).reshape(self.n, 3)
        b = np.ascontiguousarray(box if box is not None else self.box, dtype=np.float64)
        r = C.Result()
        parts = np.zeros(10)
        F = None
        if forces:
            F = np.zeros((self.n, 3))
        C.check(self._lib.atm_pme_eval(self._h, _dp(x), _dp(b), C.ATM_EVAL_STATE1_ONLY if state1_only else 0,
                                       _dp(F) if forces else None, ctypes.byref(r), _dp(parts)))
        names = ("direct", "exceptions", "reciprocal", "self", "dispersion")
        p = {f"{k}{s + 1}": parts[5 * s + i] for s in range(2) for i, k in enumerate(names)}
        return r.as_dict(), F, p

    def close(self):
        if getattr(self, "_h", None):
            C.check(self._lib.atm_pme_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class BondedForces:
    """HarmonicBondForce + HarmonicAngleForce + PeriodicTorsionForce on the device."""

    def __init__(self, num_atoms, bonds=None, bond_params=None, angles=None, angle_params=None, torsions=None,
                 torsion_params=None, device=0):
        self._lib = C.load()
        self.n = int(num_atoms)
        h = ctypes.c_void_p()
        C.check(self._lib.atm_bonded_create(int(device), ctypes.byref(h)))
        self._h = h

        def arr(a, k, dt):
            return np.ascontiguousarray(a if a is not None else np.zeros((0, k)), dtype=dt).reshape(-1, k)

        b, bp = arr(bonds, 2, np.int32), arr(bond_params, 2, np.float64)
        a, ap = arr(angles, 3, np.int32), arr(angle_params, 2, np.float64)
        t, tp = arr(torsions, 4, np.int32), arr(torsion_params, 3, np.float64)
        C.check(self._lib.atm_bonded_set(self._h, self.n, len(b), _ip(b), _dp(bp), len(a), _ip(a), _dp(ap),
                                         len(t), _ip(t), _dp(tp)))

    def evaluate(self, positions, forces=True):
        """({"bond", "angle", "torsion"} energies, forces (N,3) or None)."""
        x = np.ascontiguousarray(positions, dtype=np.float64).reshape(self.n, 3)
        e = np.zeros(3)
        F = np.zeros((self.n, 3)) if forces else None
        C.check(self._lib.atm_bonded_eval(self._h, _dp(x), _dp(F) if forces else None, _dp(e)))
        return {"bond": e[0], "angle": e[1], "torsion": e[2]}, F

    def close(self):
        if getattr(self, "_h", None):
            C.check(self._lib.atm_bonded_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def from_amber(system, displacements, box, params=None, cutoff=1.0, ewald_tolerance=5e-4, device=0):
    """(ATMPMEForce, BondedForces) of an atmmetaforce.amber.AmberSystem."""
    pme = ATMPMEForce(system.charges, system.sigmas, system.epsilons, displacements, box, cutoff,
                      exceptions=(system.exception_pairs, system.exception_qq, system.exception_sigma,
                                  system.exception_eps),
                      params=params, ewald_tolerance=ewald_tolerance, device=device)
    bonded = BondedForces(system.num_atoms, system.bonds, system.bond_params, system.angles, system.angle_params,
                          system.torsions, system.torsion_params, device=device)
    return pme, bonded


def _main(argv=None):
    """Time the PME ATM evaluation of an AMBER system:
    python -m atmmetaforce.pme --npz tests/golden/temoa-g1.npz [--reps N]
    (or --prmtop F --state F); ligand = residue --ligand, displaced by --displ nm."""
    import argparse
    import time
    from . import amber
    ap = argparse.ArgumentParser(prog="python -m atmmetaforce.pme")
    ap.add_argument("--npz")
    ap.add_argument("--prmtop")
    ap.add_argument("--state")
    ap.add_argument("--ligand", type=int, default=2)
    ap.add_argument("--displ", type=float, default=2.2)
    ap.add_argument("--reps", type=int, default=20)
    a = ap.parse_args(argv)
    if a.npz:
        system, arr = amber.system_from_npz(a.npz)
        box, x = arr["state/box"], arr["state/positions"]
    else:
        system = amber.read_prmtop(a.prmtop)
        box, x = amber.read_state_xml(a.state)
    d = np.zeros((system.num_atoms, 3))
    d[system.residue_atoms(a.ligand)] = a.displ
    params = dict(lambda1=0.5, lambda2=0.5, alpha=0.0, u0=0.0, w0=0.0, umax=836.8, ubcore=418.4, acore=0.0625,
                  direction=1.0)
    f, b = from_amber(system, d, box, params=params)
    try:
        for _ in range(3):
            f.evaluate(x, box)
        t0 = time.perf_counter()
        for _ in range(a.reps):
            r, F, parts = f.evaluate(x, box)
        dt = (time.perf_counter() - t0) / a.reps
        eb, _ = b.evaluate(x)
    finally:
        f.close()
        b.close()
    print(f"atoms {system.num_atoms}  ms/eval {dt * 1e3:.3f}  W {r['energy']:.4f}  u_sc {r['perturbation_energy']:.4f}  "
          f"bonded {sum(eb.values()):.4f}", flush=True)


if __name__ == "__main__":
    _main()
