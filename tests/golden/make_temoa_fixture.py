"""Writes tests/golden/temoa-g1.npz from the reference's own test fixtures
python/tests/temoa-g1.prmtop, temoa-g1.inpcrd and temoa-g1-equil.xml (data
files, read as text; nothing of the reference is imported or run).

The archive holds the prmtop sections the AMBER reader needs (one array per
%FLAG, prefixed "prmtop/"), the box and positions of the State checkpoint
(state/box, state/positions) and the inpcrd coordinates of the first 40
atoms, the reference positions of test_abfe.py's position restraints
(inpcrd/positions40).  Run from the repository root:

    python tests/golden/make_temoa_fixture.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "openmm-atmmetaforce-plugin_amd"))
from atmmetaforce import amber  # noqa: E402

SRC = "/root/reference/python/tests"
SECTIONS = ["POINTERS", "ATOM_NAME", "CHARGE", "MASS", "ATOM_TYPE_INDEX", "NUMBER_EXCLUDED_ATOMS",
            "NONBONDED_PARM_INDEX", "RESIDUE_LABEL", "RESIDUE_POINTER", "BOND_FORCE_CONSTANT",
            "BOND_EQUIL_VALUE", "ANGLE_FORCE_CONSTANT", "ANGLE_EQUIL_VALUE", "DIHEDRAL_FORCE_CONSTANT",
            "DIHEDRAL_PERIODICITY", "DIHEDRAL_PHASE", "SCEE_SCALE_FACTOR", "SCNB_SCALE_FACTOR",
            "LENNARD_JONES_ACOEF", "LENNARD_JONES_BCOEF", "BONDS_INC_HYDROGEN", "BONDS_WITHOUT_HYDROGEN",
            "ANGLES_INC_HYDROGEN", "ANGLES_WITHOUT_HYDROGEN", "DIHEDRALS_INC_HYDROGEN",
            "DIHEDRALS_WITHOUT_HYDROGEN", "EXCLUDED_ATOMS_LIST", "BOX_DIMENSIONS"]


def main():
    sec = amber.read_prmtop_sections(os.path.join(SRC, "temoa-g1.prmtop"))
    box, pos = amber.read_state_xml(os.path.join(SRC, "temoa-g1-equil.xml"))
    x0 = amber.read_inpcrd(os.path.join(SRC, "temoa-g1.inpcrd"))
    out = {}
    for name in SECTIONS:
        v = sec[name]
        if isinstance(v[0], str):
            out["prmtop/" + name] = np.array(v, dtype="<U4")
        elif isinstance(v[0], int):
            out["prmtop/" + name] = np.array(v, dtype=np.int32)
        else:
            out["prmtop/" + name] = np.array(v, dtype=np.float64)
    out["state/box"] = box
    out["state/positions"] = pos
    out["inpcrd/positions40"] = x0[:40]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "temoa-g1.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
