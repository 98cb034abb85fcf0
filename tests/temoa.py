"""The temoa-g1 host-guest system of the reference's end-to-end test
(python/tests/test_abfe.py), loaded from the committed fixture
tests/golden/temoa-g1.npz (made by tests/golden/make_temoa_fixture.py).

Setup restated from test_abfe.py:22-35 and :58-120: receptor = residue 1,
ligand = residue 2, displacement 22 A along each axis, centroid restraint
kf = 25 kcal/mol/A^2, tolerance 5 A, zero offset (masses as weights), position
restraints on the receptor carbons with index < 40 at their inpcrd positions
(25 kcal/mol/A^2, 0.5 A), ATM parameters lambda1 = lambda2 = 0.5, alpha = 0,
u0 = w0 = 0, umax = 200 kcal/mol, ubcore = 100 kcal/mol, acore = 1/16 and the
direction parameter set to acore (test_abfe.py:142, as written there).
Goldens (:147-150): potential energy of groups {0, ATM} -116071.0 kJ/mol and
perturbation energy 58.2 kJ/mol, both within 0.1 kJ/mol.
"""
from __future__ import annotations

import os

import numpy as np

from atmmetaforce import amber

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "temoa-g1.npz")

KCAL = 4.184
GOLDEN_POTENTIAL = -116071.0
GOLDEN_PERTURBATION = 58.2
GOLDEN_ATOL = 0.1
DISPLACEMENT = np.array([2.2, 2.2, 2.2])
CUTOFF = 1.0
EWALD_TOL = 5e-4
RESTRAINT_K = 25.0 * KCAL * 100.0     # kJ/mol/nm^2
CM_TOL = 0.5                            # nm
POS_TOL = 0.05                          # nm
ATM_PARAMS = dict(lambda1=0.5, lambda2=0.5, alpha=0.0, u0=0.0, w0=0.0, umax=200.0 * KCAL,
                  ubcore=100.0 * KCAL, acore=0.0625, direction=0.0625)


def load():
    """(AmberSystem, box, positions, restraint reference positions of atoms 0..39)."""
    s, arr = amber.system_from_npz(FIXTURE)
    return s, arr["state/box"], arr["state/positions"], arr["inpcrd/positions40"]


def restraint_setup(s: amber.AmberSystem):
    """(receptor atoms, ligand atoms, position-restrained atoms)."""
    rcpt = s.residue_atoms(1)
    lig = s.residue_atoms(2)
    posr = np.array([a for a in rcpt if s.atom_names[a].startswith("C") and a < 40])
    return rcpt, lig, posr
