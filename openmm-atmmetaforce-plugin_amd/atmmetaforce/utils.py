"""ATMMetaForceUtils: the setup helpers of the reference's Python module
(python/ATMMetaForceUtils.py) that ATM binding free energy scripts call before
creating the ATM force (python/tests/test_abfe.py:58-87):

  ATMMetaForceUtils(system, fix_zero_LJparams=True)    :13-46
  setNonbondedForceGroup(group)                         :48-61
  addRestraintForce(...)  (deprecated alias)            :63-77
  addVsiteRestraintForceCMCM(lig, rcpt, kfcm, tolcm, offset)   :79-138
  addPosRestraints(particles, refpos, fc, tol, periodic)       :707-753
  fixZeroLJParams(force, minsigma, minepsilon)          :756-785

The reference builds the restraints as OpenMM custom forces with expression
strings; here they are CentroidRestraintForce and PositionRestraintForce,
evaluated by hand-written MI355X kernels with the same energy functions.
openmm.unit is not installed: arguments are plain floats in OpenMM's units
(nm, kJ/mol, kJ/mol/nm^2), which is what the reference converts its Quantity
arguments to before building the forces.  Boresch, torsional, alignment and
centroid-angle restraints (:140-705) are not provided.
"""
from __future__ import annotations

import numpy as np

from . import openmm as mm
from ._atmmetaforce import CentroidRestraintForce, PositionRestraintForce

ANGSTROM = 0.1                 # nm
KCAL = 4.184                   # kJ


class ATMMetaForceUtils:
    def __init__(self, system, fix_zero_LJparams: bool = True):
        self.system = system
        self.CMCMDistForce = None
        if fix_zero_LJparams:
            for force in self._nonbonded():
                self.fixZeroLJParams(force)

    def _nonbonded(self):
        return [f for f in (self.system.getForce(i) for i in range(self.system.getNumForces()))
                if isinstance(f, mm.NonbondedForce)]

    def setNonbondedForceGroup(self, group: int):
        """Places the non-bonded forces in `group` (:48-61)."""
        for f in self._nonbonded():
            f.setForceGroup(group)

    def addRestraintForce(self, lig_cm_particles=None, rcpt_cm_particles=None, kfcm=0.0, tolcm=0.0,
                          offset=(0.0, 0.0, 0.0)):
        """Deprecated name of addVsiteRestraintForceCMCM (:63-77)."""
        print("warning: AddRestraintForce() is deprecated. Use addVsiteRestraintForceCMCM()")
        return self.addVsiteRestraintForceCMCM(lig_cm_particles, rcpt_cm_particles, kfcm, tolcm, offset)

    def addVsiteRestraintForceCMCM(self, lig_cm_particles=None, rcpt_cm_particles=None, kfcm=0.0, tolcm=0.0,
                                   offset=(0.0, 0.0, 0.0)):
        """Flat-bottom harmonic restraint of the ligand's centre of mass to the
        receptor's plus `offset`: E = kfcm/2 (d - tolcm)^2 for d > tolcm,
        d = |c_lig - offset - c_rcpt|, masses as weights (:79-138).  Groups are
        added to one force per utility object, as the reference does."""
        assert lig_cm_particles is not None and len(lig_cm_particles) > 0
        assert rcpt_cm_particles is not None and len(rcpt_cm_particles) > 0
        if self.CMCMDistForce is None:
            self.CMCMDistForce = CentroidRestraintForce()
            self.system.addForce(self.CMCMDistForce)
        force = self.CMCMDistForce
        g1 = force.addGroup([int(i) for i in lig_cm_particles])
        g2 = force.addGroup([int(i) for i in rcpt_cm_particles])
        off = [float(v) for v in np.asarray(offset, dtype=np.float64).reshape(3)]
        force.addBond(g1, g2, float(kfcm), float(tolcm), off)
        return force

    def addPosRestraints(self, particles, refpos, fc=25.0 * KCAL / ANGSTROM ** 2, tol=0.5 * ANGSTROM,
                         periodic: bool = True):
        """Flat-bottom position restraints of `particles` to refpos[particle]
        (reference positions of all atoms): E = fc/2 (d - tol)^2 for d > tol,
        d the minimum-image distance when `periodic` (:707-753)."""
        if not particles or len(particles) == 0:
            return None
        force = PositionRestraintForce(periodic)
        self.system.addForce(force)
        ref = np.asarray(refpos, dtype=np.float64)
        for p in particles:
            force.addParticle(int(p), [float(v) for v in ref[int(p)]], float(fc), float(tol))
        return force

    def fixZeroLJParams(self, force, minsigma=0.1 * ANGSTROM, minepsilon=1.0e-4 * KCAL):
        """Gives particles with zero LJ parameters minimum values (:756-785)."""
        small = 1.0e-6
        if isinstance(force, mm.NonbondedForce):
            for i in range(force.getNumParticles()):
                charge, sigma, epsilon = force.getParticleParameters(i)
                if sigma < small and epsilon < small:
                    force.setParticleParameters(i, charge, minsigma, minepsilon)
