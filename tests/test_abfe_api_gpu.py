"""The reference's end-to-end test, python/tests/test_abfe.py, through this
package's OpenMM-compatible API on the MI355X platform: the AMBER system with a
PME NonbondedForce (amber.create_system, the terms of
AmberPrmtopFile.createSystem(PME, 1 nm, HBonds)), ATMMetaForceUtils restraints,
the NonbondedForce moved to group 1 and made the ATM variable group, a
disabled barostat, a Langevin integrator over groups {0, 2}, and the goldens
(:147-150); then abfe.py's MD loop with the hydrogen-bond constraints.  Coordinates and box come from the reference's State checkpoint
(temoa-g1-equil.xml, through the committed fixture) instead of loadState."""
from __future__ import annotations

import numpy as np
import pytest

import atmmetaforce as atm
from atmmetaforce import amber, openmm as mm

import temoa

pytestmark = pytest.mark.gpu


def build_abfe_context(bonded_group=0):
    """bonded_group: force group of the bonded forces (1 puts them into the ATM
    variable group next to the NonbondedForce)."""
    s, box, x, ref40 = temoa.load()
    system = amber.create_system(s, box)
    for i in range(system.getNumForces()):
        fc = system.getForce(i)
        if isinstance(fc, (mm.HarmonicBondForce, mm.HarmonicAngleForce, mm.PeriodicTorsionForce)):
            fc.setForceGroup(bonded_group)
    utils = atm.ATMMetaForceUtils(system)
    rcpt, lig, posr = temoa.restraint_setup(s)
    utils.addVsiteRestraintForceCMCM(lig_cm_particles=list(lig), rcpt_cm_particles=list(rcpt),
                                     kfcm=temoa.RESTRAINT_K, tolcm=temoa.CM_TOL, offset=[0.0, 0.0, 0.0])
    refpos = np.zeros((s.num_atoms, 3))
    refpos[:40] = ref40
    utils.addPosRestraints(list(posr), refpos, temoa.RESTRAINT_K, temoa.POS_TOL)
    atm_group, nb_group = 2, 1
    utils.setNonbondedForceGroup(nb_group)
    p = temoa.ATM_PARAMS
    force = atm.ATMMetaForce(p["lambda1"], p["lambda2"], p["alpha"], p["u0"], p["w0"], p["umax"], p["ubcore"],
                             p["acore"], 1.0, [nb_group])
    for i in range(s.num_atoms):
        force.addParticle(i, 0.0, 0.0, 0.0)
    for i in lig:
        force.setParticleParameters(int(i), int(i), *temoa.DISPLACEMENT)
    force.setForceGroup(atm_group)
    system.addForce(force)
    barostat = mm.MonteCarloBarostat(1.0, 300.0)
    barostat.setFrequency(0)
    system.addForce(barostat)
    integrator = mm.LangevinIntegrator(300.0, 0.5, 0.001)
    integrator.setIntegrationForceGroups({0, atm_group})
    ctx = mm.Context(system, integrator, mm.Platform.getPlatformByName("MI355X"))
    ctx.setPositions(x)
    ctx.setPeriodicBoxVectors((box[0], 0, 0), (0, box[1], 0), (0, 0, box[2]))
    # test_abfe.py:134-142 (its Direction line sets acore: positive, so the forward direction)
    for name, v in ((force.Lambda1(), p["lambda1"]), (force.Lambda2(), p["lambda2"]), (force.Alpha(), p["alpha"]),
                    (force.U0(), p["u0"]), (force.W0(), p["w0"]), (force.Umax(), p["umax"]),
                    (force.Ubcore(), p["ubcore"]), (force.Acore(), p["acore"]), (force.Direction(), p["acore"])):
        ctx.setParameter(name, v)
    return ctx, force, integrator, s, x


def test_binding_energy_goldens_through_the_api(gpu):
    print("[abfe-api] temoa-g1 through the OpenMM-compatible API", flush=True)
    ctx, force, integ, s, x = build_abfe_context()
    state = ctx.getState(getEnergy=True, groups={0, 2})
    pot = state.getPotentialEnergy()
    pert = force.getPerturbationEnergy(ctx)
    assert abs(pot - temoa.GOLDEN_POTENTIAL) <= temoa.GOLDEN_ATOL, pot
    assert abs(pert - temoa.GOLDEN_PERTURBATION) <= temoa.GOLDEN_ATOL, pert


def test_pme_nonbonded_group_alone_and_forces(gpu):
    """Group 1 alone is the PME NonbondedForce at x (U1 of the ATM force, minus
    nothing); the bonded groups' forces match the oracle's."""
    from oracle import pme_oracle as po
    ctx, force, integ, s, x = build_abfe_context()
    st = ctx.getState(getEnergy=True, getForces=True, groups={1})
    box = temoa.load()[1]
    U1, F1, _ = po.nonbonded_pme(x, s.charges, s.sigmas, s.epsilons, box, temoa.CUTOFF, s.exception_pairs,
                                 s.exception_qq, s.exception_sigma, s.exception_eps, tol=temoa.EWALD_TOL)
    assert abs(st.getPotentialEnergy() - U1) < 1e-9 * abs(U1)
    assert np.abs(st.getForces(asNumpy=True) - F1).max() < 2e-5
    eb, Fb = po.bonded(x, s.bonds, s.bond_params, s.angles, s.angle_params, s.torsions, s.torsion_params)
    st0 = ctx.getState(getEnergy=True, getForces=True, groups={0})
    # group 0 = bonded terms + restraints (zero at this state: both within their tolerances)
    assert abs(st0.getPotentialEnergy() - sum(eb.values())) < 1e-8 * sum(eb.values())
    assert np.abs(st0.getForces(asNumpy=True) - Fb).max() < 1e-5


def test_abfe_md_with_hbond_constraints(gpu):
    """abfe.py's MD loop on the temoa-g1 system: LangevinIntegrator (300 K,
    0.5/ps, 1 fs) over groups {0, ATM} with createSystem's 8,324 hydrogen-bond
    constraints applied by SHAKE; a second run with the same seed retraces
    the first."""
    finals = []
    for _ in range(2):
        ctx, force, integ, s, x = build_abfe_context()
        integ.setRandomNumberSeed(5)
        for _ in range(4):
            integ.step(25)
            st = ctx.getState(getEnergy=True, groups={0, 2})
            assert np.isfinite(st.getPotentialEnergy())
            # the binding energy fluctuates by tens of kJ/mol in MD; a broken force would blow it up
            assert abs(force.getPerturbationEnergy(ctx) - temoa.GOLDEN_PERTURBATION) < 250.0
        ke = ctx.getState(getEnergy=True).getKineticEnergy()
        dof = 3 * s.num_atoms - len(s.constraints)
        temperature = 2.0 * ke / (dof * 0.00831446261815324)
        assert 10.0 < temperature < 600.0, temperature   # started at 0 K, thermostat at 300 K
        xs = ctx.getState(getPositions=True).getPositions(asNumpy=True)
        d = np.linalg.norm(xs[s.constraints[:, 0]] - xs[s.constraints[:, 1]], axis=1)
        assert np.abs(d / s.constraint_lengths - 1.0).max() < 2e-5
        assert not np.array_equal(xs, x)
        finals.append(xs)
    assert np.array_equal(finals[0], finals[1])


def test_replica_exchange_over_pme_contexts(gpu):
    """Three lambda-states of temoa-g1 as three contexts (the reference's
    asynchronous-driver layout), MD cycles with exchanges: every context
    ends up holding its assigned state's parameters and the sample lines
    follow abfe.py's .out format."""
    from atmmetaforce import replica
    reps = [build_abfe_context() for _ in range(3)]
    p = temoa.ATM_PARAMS
    states = [dict(p, lambda1=l, lambda2=l, direction=1.0) for l in (0.0, 0.25, 0.5)]
    re = replica.ContextReplicaExchange([(r[0], r[2], r[1]) for r in reps], states, temperature=300.0, seed=4)
    lines = re.run(cycles=2, steps_per_cycle=5)
    assert sorted(re.state_of) == [0, 1, 2]
    for a, (ctx, f, integ, s, x) in enumerate(reps):
        st = states[re.state_of[a]]
        assert ctx.getParameter(f.Lambda1()) == st["lambda1"] and ctx.getParameter(f.Lambda2()) == st["lambda2"]
    assert len(lines) == 6
    for smp in lines:
        cols = smp.out_line().split()
        assert len(cols) == 9 and np.isfinite([float(c) for c in cols]).all()
    assert re.attempts.sum() == 4
