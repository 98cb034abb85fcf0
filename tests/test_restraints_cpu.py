"""ATMMetaForceUtils (python/ATMMetaForceUtils.py) on the CPU side: the
restraint forces it builds, their parameters and argument checks."""
import numpy as np
import pytest

import atmmetaforce as atm
from atmmetaforce import openmm as mm


def small_system(n=12):
    s = mm.System()
    for i in range(n):
        s.addParticle(12.0 + i)
    nb = mm.NonbondedForce()
    for i in range(n):
        nb.addParticle(0.1, 0.0 if i == 3 else 0.3, 0.0 if i == 3 else 0.5)
    s.addForce(nb)
    return s, nb


def test_utils_build_the_reference_restraints():
    s, nb = small_system()
    u = atm.ATMMetaForceUtils(s)
    # fixZeroLJParams (:756-785): sigma 0.1 A, epsilon 1e-4 kcal/mol for zero-LJ particles only
    assert nb.getParticleParameters(3) == pytest.approx((0.1, 0.01, 1e-4 * 4.184))
    assert nb.getParticleParameters(4) == pytest.approx((0.1, 0.3, 0.5))
    f = u.addVsiteRestraintForceCMCM([0, 1], [5, 6, 7], 100.0, 0.5, (0.1, 0.2, 0.3))
    assert isinstance(f, atm.CentroidRestraintForce) and s.getNumForces() == 2
    g = u.addVsiteRestraintForceCMCM([2], [8, 9], 50.0, 0.2, (0, 0, 0))
    assert g is f and f.getNumGroups() == 4 and f.getNumBonds() == 2  # one force per utility object
    assert f.getBondParameters(1) == (2, 3, 50.0, 0.2, (0.0, 0.0, 0.0))
    assert f.getGroupParameters(1) == ([5, 6, 7], [])  # empty weights: masses
    p = u.addPosRestraints([2, 4], np.arange(36.0).reshape(12, 3) / 10, 50.0, 0.05)
    assert p.getNumParticles() == 2 and p.getPeriodic()
    assert p.getParticleParameters(1) == (4, pytest.approx((1.2, 1.3, 1.4)), 50.0, 0.05)
    assert u.addPosRestraints([], None) is None
    u.setNonbondedForceGroup(1)
    assert nb.getForceGroup() == 1
    # the reference's default position-restraint constants: 25 kcal/mol/A^2, 0.5 A
    q = u.addPosRestraints([1], np.zeros((12, 3)), periodic=False)
    assert q.getParticleParameters(0)[2:] == pytest.approx((25 * 4.184 * 100, 0.05)) and not q.getPeriodic()


def test_restraint_argument_checks():
    f = atm.CentroidRestraintForce()
    with pytest.raises(Exception, match="at least one particle"):
        f.addGroup([])
    with pytest.raises(Exception, match="one weight per particle"):
        f.addGroup([1, 2], [1.0])
    with pytest.raises(Exception, match="positive"):
        f.addGroup([1, 2], [1.0, 0.0])
    with pytest.raises(Exception, match="out of range"):
        f.getBondParameters(0)
    p = atm.PositionRestraintForce(False)
    with pytest.raises(Exception, match="out of range"):
        p.getParticleParameters(0)
    with pytest.raises(AssertionError):
        atm.ATMMetaForceUtils(small_system()[0]).addVsiteRestraintForceCMCM([], [1], 1.0, 1.0)
