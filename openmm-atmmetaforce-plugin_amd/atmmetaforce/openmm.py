"""Minimal OpenMM-compatible classes (the subset the ATM force is used with).

Names follow the `openmm` package so scripts read the same: System,
NonbondedForce (CutoffPeriodic reaction field or PME, with exceptions),
HarmonicBondForce, HarmonicAngleForce, PeriodicTorsionForce,
MonteCarloBarostat (a holder of its two global parameters), Context, State,
Platform, VerletIntegrator, LangevinIntegrator (OpenMM's leapfrog Langevin),
LangevinMiddleIntegrator (BAOAB), XmlSerializer.  Positions and forces are
(N, 3) arrays or lists of 3-tuples in nm and kJ/mol/nm; `groups` arguments
accept a set of force-group indices or a bit mask.
"""
from ._atmmetaforce import (Context, Force, HarmonicAngleForce, HarmonicBondForce, Integrator,  # noqa: F401
                            LangevinIntegrator, LangevinMiddleIntegrator, MonteCarloBarostat, NonbondedForce,
                            OpenMMException, PeriodicTorsionForce, Platform, State, System, VerletIntegrator,
                            XmlSerializer)

__all__ = ["Context", "Force", "HarmonicAngleForce", "HarmonicBondForce", "Integrator", "LangevinIntegrator",
           "LangevinMiddleIntegrator", "MonteCarloBarostat", "NonbondedForce", "OpenMMException",
           "PeriodicTorsionForce", "Platform", "State", "System", "VerletIntegrator", "XmlSerializer"]
