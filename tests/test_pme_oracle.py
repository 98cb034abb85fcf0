"""Pins the FP64 PME / AMBER oracle (oracle/pme_oracle.py): the reference's
goldens on temoa-g1 (python/tests/test_abfe.py:147-150), the Madelung
constant of rock salt, alpha independence of the Ewald sum, central
differences and Newton's third law; and the AMBER reader against the
reference's files where they are present."""
from __future__ import annotations

import math
import os

import numpy as np
import pytest

from atmmetaforce import amber
from oracle import oracle, pme_oracle as po

import temoa


def nb(s, x, box, **kw):
    return po.nonbonded_pme(x, s.charges, s.sigmas, s.epsilons, box, temoa.CUTOFF, s.exception_pairs,
                            s.exception_qq, s.exception_sigma, s.exception_eps, tol=temoa.EWALD_TOL, **kw)


def temoa_energies(forces=False):
    s, box, x, ref40 = temoa.load()
    rcpt, lig, posr = temoa.restraint_setup(s)
    x2 = x.copy()
    x2[lig] += temoa.DISPLACEMENT
    U1, _, p1 = nb(s, x, box, forces=forces)
    U2, _, p2 = nb(s, x2, box, forces=forces)
    eb, _ = po.bonded(x, s.bonds, s.bond_params, s.angles, s.angle_params, s.torsions, s.torsion_params)
    ecm, _ = oracle.centroid_restraint(x, lig, rcpt, s.masses[lig], s.masses[rcpt], temoa.RESTRAINT_K,
                                       temoa.CM_TOL, np.zeros(3))
    epos, _ = oracle.position_restraints(x, posr, ref40[posr], np.full(len(posr), temoa.RESTRAINT_K),
                                         np.full(len(posr), temoa.POS_TOL), box)
    r = oracle.combine(temoa.ATM_PARAMS, U1, U2)
    return {"U1": U1, "U2": U2, "W": r["W"], "usc": r["usc"], "bonded": sum(eb.values()), "restraints": ecm + epos,
            "parts1": p1, "parts2": p2}


def test_temoa_goldens():
    """test_abfe.py:147-150: potential energy of groups {0, ATM} and the
    perturbation energy, both within 0.1 kJ/mol."""
    e = temoa_energies()
    pot = e["bonded"] + e["restraints"] + e["W"]
    assert e["parts1"]["grid"] == (35, 40, 35)
    assert abs(pot - temoa.GOLDEN_POTENTIAL) <= temoa.GOLDEN_ATOL, pot
    assert abs(e["usc"] - temoa.GOLDEN_PERTURBATION) <= temoa.GOLDEN_ATOL, e["usc"]


def test_madelung_constant_of_rock_salt():
    """E = -N/2 k_e M / r0 for an NaCl lattice (M = 1.747564594633)."""
    a = 0.5          # cubic cell (nm): 8 ions, nearest neighbours a/2
    nc = 4
    pts, q = [], []
    for i in range(2 * nc):
        for j in range(2 * nc):
            for k in range(2 * nc):
                pts.append((i * a / 2, j * a / 2, k * a / 2))
                q.append(1.0 if (i + j + k) % 2 == 0 else -1.0)
    x = np.array(pts) + 0.01
    q = np.array(q)
    box = np.full(3, nc * a)
    E, F, p = po.nonbonded_pme(x, q, np.full(len(q), 0.3), np.zeros(len(q)), box, 0.95, np.zeros((0, 2)),
                               [], [], [], tol=1e-7, grid=(64, 64, 64), dispersion=False)
    expect = -len(q) / 2 * po.KE * 1.747564594633 / (a / 2)
    assert abs(E / expect - 1.0) < 2e-6, (E, expect)
    assert np.abs(F).max() < 1e-3 * po.KE   # every ion sits at a symmetric site


def _random_system(n=80, L=2.4, seed=3):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, L, (n, 3))
    # keep atoms apart: reject pairs closer than 0.2 nm by nudging on a lattice
    g = int(math.ceil(n ** (1 / 3)))
    lat = np.array([(i, j, k) for i in range(g) for j in range(g) for k in range(g)][:n], dtype=float)
    x = (lat + 0.5) * (L / g) + rng.uniform(-0.05, 0.05, (n, 3))
    q = rng.normal(0, 0.5, n)
    q -= q.mean()
    sig = rng.uniform(0.25, 0.35, n)
    eps = rng.uniform(0.1, 0.6, n)
    # bonded-like exceptions: near pairs (plain distance = minimum-image distance)
    pairs = np.array([(0, 1), (1, 2), (2, 3), (0, 2), (5, 6), (10, 11)])
    qq = np.array([0.0, 0.0, 0.0, q[0] * q[2] / 1.2, 0.0, q[10] * q[11] / 1.2])
    es = 0.5 * (sig[pairs[:, 0]] + sig[pairs[:, 1]])
    ee = np.array([0.0, 0.0, 0.0, math.sqrt(eps[0] * eps[2]) / 2, 0.0, math.sqrt(eps[10] * eps[11]) / 2])
    return x, q, sig, eps, np.full(3, L), pairs, qq, es, ee


def test_total_ewald_energy_independent_of_alpha():
    x, q, sig, eps, box, pairs, qq, es, ee = _random_system()
    Es = [po.nonbonded_pme(x, q, sig, eps, box, 1.15, pairs, qq, es, ee, alpha=a, grid=(72, 72, 72),
                           forces=False)[0] for a in (3.2, 4.0)]
    assert abs(Es[0] - Es[1]) < 2e-6 * abs(Es[0]), Es


def test_nonbonded_forces_are_minus_gradient():
    x, q, sig, eps, box, pairs, qq, es, ee = _random_system()
    kw = dict(alpha=3.0, grid=(32, 32, 32))
    E0, F, _ = po.nonbonded_pme(x, q, sig, eps, box, 1.0, pairs, qq, es, ee, **kw)
    h = 1e-5
    for (a, k) in [(0, 0), (3, 2), (10, 1), (40, 0)]:
        xp, xm = x.copy(), x.copy()
        xp[a, k] += h
        xm[a, k] -= h
        Ep = po.nonbonded_pme(xp, q, sig, eps, box, 1.0, pairs, qq, es, ee, forces=False, **kw)[0]
        Em = po.nonbonded_pme(xm, q, sig, eps, box, 1.0, pairs, qq, es, ee, forces=False, **kw)[0]
        fd = -(Ep - Em) / (2 * h)
        assert abs(fd - F[a, k]) < 2e-5 * max(1.0, abs(F[a, k])), (a, k, fd, F[a, k])


def test_bonded_forces_are_minus_gradient_and_sum_to_zero():
    rng = np.random.default_rng(7)
    x = np.cumsum(rng.normal(0, 0.1, (6, 3)), 0) + np.array([[0.15, 0, 0]]) * np.arange(6)[:, None]
    bonds = np.array([[0, 1], [1, 2], [2, 3], [3, 4]])
    bp = np.array([[0.15, 2.5e5], [0.14, 3.0e5], [0.16, 2.0e5], [0.15, 2.2e5]])
    angles = np.array([[0, 1, 2], [1, 2, 3], [2, 3, 4]])
    ap = np.array([[1.9, 400.0], [2.0, 500.0], [2.1, 350.0]])
    tors = np.array([[0, 1, 2, 3], [1, 2, 3, 4], [2, 3, 4, 5], [0, 1, 2, 3]])
    tp = np.array([[3, 0.0, 1.5], [2, math.pi, 4.0], [1, 0.3, 2.0], [1, 1.1, 0.7]])
    E, F = po.bonded(x, bonds, bp, angles, ap, tors, tp)
    assert np.abs(F.sum(0)).max() < 1e-9 * np.abs(F).max()
    h = 1e-6
    for a in range(6):
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[a, k] += h
            xm[a, k] -= h
            ep = sum(po.bonded(xp, bonds, bp, angles, ap, tors, tp)[0].values())
            em = sum(po.bonded(xm, bonds, bp, angles, ap, tors, tp)[0].values())
            fd = -(ep - em) / (2 * h)
            assert abs(fd - F[a, k]) < 1e-5 * max(1.0, np.abs(F).max()), (a, k, fd, F[a, k])


def test_torsion_sign_convention():
    """IUPAC sign: B at the origin, C on +x, A on +y, D above C on +z.  Seen
    along B->C, A-B turns clockwise by 90 degrees onto C-D, so phi = +90 deg:
    k (1 + cos(phi - pi/2)) = 2k, where phi = -90 deg would give 0."""
    x = np.array([[0.0, 1.0, 0.0], [0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 0.0, 1.0]]) * 0.15
    E, _ = po.bonded(x, torsions=np.array([[0, 1, 2, 3]]), torsion_params=np.array([[1, math.pi / 2, 1.0]]))
    assert abs(E["torsion"] - 2.0) < 1e-12


@pytest.mark.skipif(not os.path.exists("/root/reference/python/tests/temoa-g1.prmtop"),
                    reason="the reference's prmtop is not present on this machine")
def test_reader_on_reference_files_matches_fixture():
    s_ref = amber.read_prmtop("/root/reference/python/tests/temoa-g1.prmtop")
    s_fix, box, x, _ = temoa.load()
    for f in ("charges", "sigmas", "epsilons", "masses", "exception_pairs", "exception_qq", "bonds", "angles",
              "torsions", "torsion_params"):
        assert np.array_equal(getattr(s_ref, f), getattr(s_fix, f)), f
    b2, x2 = amber.read_state_xml("/root/reference/python/tests/temoa-g1-equil.xml")
    assert np.array_equal(b2, box) and np.array_equal(x2, x)


def test_amber_terms_counts():
    s, box, x, _ = temoa.load()
    assert s.num_atoms == 8479
    assert len(s.bonds) == 167 and len(s.constraints) == 8324 and len(s.angles) == 410 and len(s.torsions) == 741
    assert len(s.residue_atoms(1)) == 196 and len(s.residue_atoms(2)) == 21
    assert abs(s.charges.sum()) < 5e-3
