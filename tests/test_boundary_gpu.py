"""The drop-in path on the GPU: ATMMetaForce in an OpenMM-style System and
Context on the MI355X platform, checked against the FP64 oracle.

Layout as in the reference's own test (python/tests/test_abfe.py:88-104): the
NonbondedForce in force group 1, ATMMetaForce in group 2 with variable groups
[1], every particle added in order with its displacement.
"""
import dataclasses

import numpy as np
import pytest

import atmmetaforce as atm
from atmmetaforce import openmm as mm, workloads as W
from oracle import oracle

pytestmark = pytest.mark.gpu

E_TOL = 1e-6
F_TOL = 1e-5


def build(s, atm_group=2, nb_group=1, dispersion=False, method=None, add_all=True):
    system = mm.System()
    for _ in range(s.num_atoms):
        system.addParticle(12.0)
    L = s.box
    system.setDefaultPeriodicBoxVectors((L[0], 0, 0), (0, L[1], 0), (0, 0, L[2]))
    nb = mm.NonbondedForce()
    nb.setNonbondedMethod(mm.NonbondedForce.CutoffPeriodic if method is None else method)
    nb.setCutoffDistance(s.cutoff)
    nb.setReactionFieldDielectric(s.rf_dielectric)
    nb.setUseDispersionCorrection(dispersion)
    for q, sg, e in zip(s.charges, s.sigmas, s.epsilons):
        nb.addParticle(float(q), float(sg), float(e))
    if s.exceptions is not None:
        ex = s.exceptions
        for (i, j), qq, sg, e in zip(ex.pairs, ex.charge_prod, ex.sigma, ex.epsilon):
            nb.addException(int(i), int(j), float(qq), float(sg), float(e))
    nb.setForceGroup(nb_group)
    system.addForce(nb)
    p = s.params
    f = atm.ATMMetaForce(p.lambda1, p.lambda2, p.alpha, p.u0, p.w0, p.umax, p.ubcore, p.acore, p.direction, [1])
    n = s.num_atoms if add_all else s.num_atoms - 1
    for i in range(n):
        d = s.displacements[i]
        f.addParticle(i, float(d[0]), float(d[1]), float(d[2]))
    f.setForceGroup(atm_group)
    system.addForce(f)
    return system, f


def context_for(s, properties=None, **kw):
    system, f = build(s, **kw)
    integ = mm.VerletIntegrator(0.001)
    plat = mm.Platform.getPlatformByName("MI355X")
    ctx = mm.Context(system, integ, plat, properties or {})
    ctx.setPositions(s.positions)
    return system, f, integ, ctx


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def frms(F, R):
    return float(np.linalg.norm(F - R) / np.linalg.norm(R))


SYS = W.tiny_system(3000, 3.1, 0.9, 30, (1.03, 0.77, 0.62), 1,
                    W.ATMParams(0.25, 0.45, 0.1, 110.0, 2.0, 200.0, 100.0, 0.0625, 1.0))


@pytest.mark.parametrize("props", [{}, {"ATMEvaluation": "plain"}])
def test_reaction_field_exceptions(gpu, props):
    """NonbondedForce.addException on the reaction-field path: exclusions and
    scaled pairs of molecules of 4 atoms and of index-distant crosslinks."""
    s = W.with_molecules(SYS, size=4, crosslink_every=9)
    ref, Fref = oracle.atm_eval(s)
    system, f, integ, ctx = context_for(s, props)
    st = ctx.getState(getEnergy=True, getForces=True, groups={2})
    print(props, st.getPotentialEnergy(), ref["W"], flush=True)
    assert rel(st.getPotentialEnergy(), ref["W"]) < E_TOL
    assert frms(st.getForces(asNumpy=True), Fref) < F_TOL
    assert rel(f.getPerturbationEnergy(ctx), ref["usc"]) < E_TOL
    st1 = ctx.getState(getEnergy=True, groups={1})  # the NonbondedForce alone at x
    e1 = oracle.nonbonded(s.positions, s.charges, s.sigmas, s.epsilons, s.box, s.cutoff, s.rf_dielectric,
                          forces=False, exceptions=s.exceptions)[0]
    assert rel(st1.getPotentialEnergy(), e1) < E_TOL


@pytest.mark.parametrize("props", [{}, {"ATMEvaluation": "plain"}, {"Precision": "double"}])
def test_atm_energy_forces_and_perturbation_energy(gpu, props):
    ref, Fref = oracle.atm_eval(SYS)
    system, f, integ, ctx = context_for(SYS, props)
    st = ctx.getState(getEnergy=True, getForces=True, groups={2})
    print(props, st.getPotentialEnergy(), ref["W"], flush=True)
    assert rel(st.getPotentialEnergy(), ref["W"]) < E_TOL
    assert frms(st.getForces(asNumpy=True), Fref) < F_TOL
    assert rel(f.getPerturbationEnergy(ctx), ref["usc"]) < E_TOL


def test_fp32_pair_precision_property(gpu):
    ref, Fref = oracle.atm_eval(SYS)
    system, f, integ, ctx = context_for(SYS, {"Precision": "single"})
    st = ctx.getState(getEnergy=True, getForces=True, groups={2})
    # FP32 pair arithmetic: the bound scales with the sum of |pair energies|
    w1, w2 = 1 - ref["sp"], ref["sp"]
    assert abs(st.getPotentialEnergy() - ref["W"]) <= E_TOL * (abs(ref["W"]) + w1 * ref["abs1"] + w2 * ref["abs2"])
    assert frms(st.getForces(asNumpy=True), Fref) < F_TOL


def test_group_gate_and_nonbonded_group(gpu):
    system, f, integ, ctx = context_for(SYS)
    assert ctx.getState(getEnergy=True, groups={0}).getPotentialEnergy() == 0.0
    E, F, _, _ = oracle.nonbonded(SYS.positions, SYS.charges, SYS.sigmas, SYS.epsilons, SYS.box, SYS.cutoff,
                                  SYS.rf_dielectric)
    st = ctx.getState(getEnergy=True, getForces=True, groups={1})
    assert rel(st.getPotentialEnergy(), E) < E_TOL
    assert frms(st.getForces(asNumpy=True), F) < F_TOL
    ref, _ = oracle.atm_eval(SYS)
    both = ctx.getState(getEnergy=True, groups={1, 2}).getPotentialEnergy()
    assert rel(both, E + ref["W"]) < E_TOL


def test_global_parameters_drive_the_combine(gpu):
    system, f, integ, ctx = context_for(SYS)
    for l1, l2, a, d in ((0.1, 0.3, 0.2, 1.0), (0.4, 0.4, 0.05, 1.0), (0.2, 0.35, 0.1, -1.0)):
        ctx.setParameter(atm.ATMMetaForce.Lambda1(), l1)
        ctx.setParameter(f.Lambda2(), l2)
        ctx.setParameter(f.Alpha(), a)
        ctx.setParameter(f.Direction(), d)
        p = dataclasses.replace(SYS.params, lambda1=l1, lambda2=l2, alpha=a, direction=d)
        ref, Fref = oracle.atm_eval(SYS, p)
        st = ctx.getState(getEnergy=True, getForces=True, groups={2})
        if np.isfinite(ref["W"]):
            assert rel(st.getPotentialEnergy(), ref["W"]) < E_TOL
        assert frms(st.getForces(asNumpy=True), Fref) < F_TOL
    assert ctx.getParameter("ATMLambda1") == 0.2


def test_update_displacements_in_context(gpu):
    system, f, integ, ctx = context_for(SYS)
    ctx.getState(getEnergy=True, groups={2})
    new = SYS.displacements.copy()
    new[:30] = W.quantize([0.5, -0.6, 0.7])
    for i in range(30):
        f.setParticleParameters(i, i, *map(float, new[i]))
    f.updateParametersInContext(ctx)
    s2 = dataclasses.replace(SYS, displacements=new)
    ref, Fref = oracle.atm_eval(s2)
    st = ctx.getState(getEnergy=True, getForces=True, groups={2})
    assert rel(st.getPotentialEnergy(), ref["W"]) < E_TOL
    assert frms(st.getForces(asNumpy=True), Fref) < F_TOL
    f.setParticleParameters(3, 4, 0.0, 0.0, 0.0)  # a particle index change is rejected
    with pytest.raises(Exception, match="particle index has changed"):
        f.updateParametersInContext(ctx)


def test_dispersion_correction_shifts_energies_not_forces(gpu):
    _, _, _, c0 = context_for(SYS)
    _, _, _, c1 = context_for(SYS, dispersion=True)
    e_nb0 = c0.getState(getEnergy=True, groups={1}).getPotentialEnergy()
    e_nb1 = c1.getState(getEnergy=True, groups={1}).getPotentialEnergy()
    s0 = c0.getState(getEnergy=True, getForces=True, groups={2})
    s1 = c1.getState(getEnergy=True, getForces=True, groups={2})
    shift = e_nb1 - e_nb0
    assert shift < 0.0                                           # attractive tail
    assert s1.getPotentialEnergy() - s0.getPotentialEnergy() == pytest.approx(shift, rel=1e-9)
    assert np.array_equal(s1.getForces(asNumpy=True), s0.getForces(asNumpy=True))
    L = SYS.box * 1.02                                           # E_disp ~ 1/V
    c1.setPeriodicBoxVectors((L[0], 0, 0), (0, L[1], 0), (0, 0, L[2]))
    c0.setPeriodicBoxVectors((L[0], 0, 0), (0, L[1], 0), (0, 0, L[2]))
    shift2 = c1.getState(getEnergy=True, groups={1}).getPotentialEnergy() - \
        c0.getState(getEnergy=True, groups={1}).getPotentialEnergy()
    assert shift2 == pytest.approx(shift / 1.02 ** 3, rel=1e-9)


def test_errors_match_reference_behaviour(gpu):
    with pytest.raises(Exception, match="cannot be one of the variable force groups"):
        context_for(SYS, atm_group=1)
    with pytest.raises(Exception, match="CutoffPeriodic .reaction-field. or PME"):
        context_for(SYS, method=mm.NonbondedForce.Ewald)
    with pytest.raises(Exception, match="every particle"):
        context_for(SYS, add_all=False)
    with pytest.raises(Exception, match="Illegal property name"):
        context_for(SYS, {"CudaPrecision": "mixed"})


def test_verlet_steps_with_atm_force(gpu):
    s = W.tiny_system(800, 2.8, 0.9, 8, (0.9, 0.8, 0.7), 3)
    system, f, integ, ctx = context_for(s)
    integ.setIntegrationForceGroups({0, 2})
    x0 = ctx.getState(getPositions=True).getPositions(asNumpy=True)
    integ.step(3)
    x1 = ctx.getState(getPositions=True).getPositions(asNumpy=True)
    assert np.isfinite(x1).all() and not np.array_equal(x0, x1)
    assert ctx.getState(getEnergy=True).getPotentialEnergy() != 0.0


def test_state_checkpoint_restores_atm_parameters(gpu):
    """Simulation.loadState's path (example/abfe/abfe.py:121): the reference's
    checkpoint (cut to 16 particles) restored with Context.setState sets the
    box, positions, velocities and the ATM global parameters; the evaluation
    equals one set up by hand with the same values, and a checkpoint written
    from the context restores it again."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "temoa-g1-equil-excerpt.xml")
    st = mm.XmlSerializer.deserialize(open(path).read())
    x = st.getPositions(asNumpy=True)
    n = len(x)
    rng = np.random.default_rng(3)
    q = rng.normal(0, 0.4, n); q -= q.mean()
    sg = rng.uniform(0.25, 0.35, n); ep = rng.uniform(0.1, 0.8, n)
    d = np.zeros((n, 3)); d[:4] = (1.1, 0.9, 0.7)

    def make():
        system = mm.System()
        for _ in range(n):
            system.addParticle(12.0)
        system.setDefaultPeriodicBoxVectors((4.0, 0, 0), (0, 4.0, 0), (0, 0, 4.0))
        nb = mm.NonbondedForce()
        nb.setNonbondedMethod(mm.NonbondedForce.CutoffPeriodic)
        nb.setCutoffDistance(1.2)
        nb.setUseDispersionCorrection(False)
        for i in range(n):
            nb.addParticle(float(q[i]), float(sg[i]), float(ep[i]))
        nb.setForceGroup(1)
        system.addForce(nb)
        f = atm.ATMMetaForce(0.2, 0.3, 0.1, 10.0, 0.0, 200.0, 100.0, 0.0625, 1.0, [1])
        for i in range(n):
            f.addParticle(i, *map(float, d[i]))
        f.setForceGroup(2)
        system.addForce(f)
        system.addForce(mm.MonteCarloBarostat(1.0, 300.0, 10000000))
        ctx = mm.Context(system, mm.VerletIntegrator(0.001), mm.Platform.getPlatformByName("MI355X"))
        return system, f, ctx

    _, f1, c1 = make()
    c1.setState(st)
    for k, v in st.getParameters().items():
        assert c1.getParameter(k) == v
    assert c1.getParameter("ATMDirection") == 1.0          # absent from the checkpoint: unchanged
    assert [tuple(v) for v in c1.getState(getPositions=True).getPeriodicBoxVectors()] == \
        [tuple(v) for v in st.getPeriodicBoxVectors()]
    s1 = c1.getState(getEnergy=True, getForces=True, groups={2})

    _, f2, c2 = make()
    a, b, c = st.getPeriodicBoxVectors()
    c2.setPeriodicBoxVectors(a, b, c)
    c2.setPositions(x)
    for k, v in st.getParameters().items():
        c2.setParameter(k, v)
    s2 = c2.getState(getEnergy=True, getForces=True, groups={2})
    assert s1.getPotentialEnergy() == s2.getPotentialEnergy()
    assert np.array_equal(s1.getForces(asNumpy=True), s2.getForces(asNumpy=True))
    assert f1.getPerturbationEnergy(c1) == f2.getPerturbationEnergy(c2)

    # write a checkpoint from c1, restore it into a fresh context
    c1.setParameter(f1.Lambda2(), 0.7)
    ck = mm.XmlSerializer.serialize(c1.getState(getPositions=True, getVelocities=True, getParameters=True))
    _, f3, c3 = make()
    c3.setState(mm.XmlSerializer.deserialize(ck))
    assert c3.getParameter("ATMLambda2") == 0.7
    assert np.array_equal(c3.getState(getVelocities=True).getVelocities(asNumpy=True), st.getVelocities(asNumpy=True))
    e1 = c1.getState(getEnergy=True, groups={2}).getPotentialEnergy()
    e3 = c3.getState(getEnergy=True, groups={2}).getPotentialEnergy()
    assert e1 == e3


@pytest.mark.parametrize("integrator", ["LangevinIntegrator", "LangevinMiddleIntegrator"])
def test_langevin_md_loop_like_abfe(gpu, integrator):
    """The MD loop of example/abfe/abfe.py:100-160 on the MI355X platform:
    Langevin dynamics over the force groups {0, ATM group}, samples of the
    potential energy of those groups and of the perturbation energy."""
    s = W.tiny_system(900, 3.3, 0.9, 9, (1.0, 0.9, 0.8), 13)
    system, f = build(s)
    plat = mm.Platform.getPlatformByName("MI355X")
    finals = []
    for _ in range(2):
        integ = getattr(mm, integrator)(300.0, 5.0, 0.001)
        integ.setIntegrationForceGroups({0, 2})
        integ.setRandomNumberSeed(17)
        ctx = mm.Context(system, integ, plat)
        ctx.setPositions(s.positions)
        lines = []
        for _ in range(3):
            integ.step(5)
            st = ctx.getState(getEnergy=True, groups={0, 2})
            pert = f.getPerturbationEnergy(ctx)
            lines.append("%f %f %f %f" % (300.0, ctx.getParameter(f.Lambda1()), st.getPotentialEnergy(), pert))
        x = ctx.getState(getPositions=True, getVelocities=True)
        finals.append((x.getPositions(asNumpy=True), x.getVelocities(asNumpy=True)))
        assert all(np.isfinite([float(t) for t in l.split()]).all() for l in lines)
        assert ctx.getState(getEnergy=True).getKineticEnergy() > 0.0
    assert np.array_equal(finals[0][0], finals[1][0]) and np.array_equal(finals[0][1], finals[1][1])
    assert not np.array_equal(finals[0][0], s.positions)
