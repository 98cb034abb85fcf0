"""Flat-bottom restraints on the MI355X (atm_restraints.hip) against the FP64
oracle: the C ABI directly, then the forces that ATMMetaForceUtils builds in a
Context beside an ATM force (layout of python/tests/test_abfe.py:58-104)."""
import ctypes

import numpy as np
import pytest

import atmmetaforce as atm
from atmmetaforce import _capi as C, openmm as mm, workloads as W
from oracle import oracle

pytestmark = pytest.mark.gpu

F_ABS = 2e-7   # forces are summed in 2^-24 kJ/mol/nm fixed point


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def test_restraints_c_abi_against_oracle(gpu):
    L = C.load()
    rng = np.random.default_rng(11)
    n = 300
    box = np.array([3.1, 3.3, 3.7])
    x = np.ascontiguousarray(rng.uniform(-1.0, 4.5, (n, 3)))   # not wrapped: centroids use raw positions
    groups = [rng.choice(n, 40, replace=False), rng.choice(n, 250, replace=False),
              rng.choice(n, 1, replace=False), rng.choice(n, 7, replace=False)]
    weights = [rng.uniform(1.0, 16.0, len(g)) for g in groups]
    members = np.ascontiguousarray(np.concatenate(groups).astype(np.int32))
    mw = np.ascontiguousarray(np.concatenate(weights))
    first = np.cumsum([0] + [len(g) for g in groups])
    bonds = [(0, 1, 300.0, 0.05, (0.2, -0.3, 0.1)), (2, 3, 80.0, 0.0, (0.0, 0.0, 0.0)), (1, 3, 50.0, 100.0, (0, 0, 0))]
    cent = (C.CentroidRestraint * len(bonds))()
    for k, (g1, g2, kf, tol, off) in enumerate(bonds):
        cent[k] = C.CentroidRestraint(int(first[g1]), len(groups[g1]), int(first[g2]), len(groups[g2]), kf, tol,
                                      (ctypes.c_double * 3)(*off))
    patoms = rng.choice(n, 60, replace=False)
    refs = rng.uniform(0, 3, (60, 3))
    pk = rng.uniform(10.0, 500.0, 60)
    posr = (C.PositionRestraint * 60)()
    for k in range(60):
        posr[k] = C.PositionRestraint(int(patoms[k]), int(k % 2), (ctypes.c_double * 3)(*refs[k]), pk[k], 0.1)
    h = ctypes.c_void_p()
    C.check(L.atm_restraints_create(0, ctypes.byref(h)))
    try:
        C.check(L.atm_restraints_set(h, n, len(members), members.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                     _dp(mw), len(bonds), cent, 60, posr))
        F = np.zeros((n, 3))
        e = ctypes.c_double()
        C.check(L.atm_restraints_eval(h, _dp(x), _dp(box), _dp(F), ctypes.byref(e)))
        Eref, Fref = 0.0, np.zeros((n, 3))
        for g1, g2, kf, tol, off in bonds:   # groups 1 and 3 share atoms, with their own weights
            E1, F1 = oracle.centroid_restraint(x, groups[g1], groups[g2], weights[g1], weights[g2], kf, tol, off)
            Eref += E1
            Fref += F1
        for per in (0, 1):
            sel = np.arange(60) % 2 == per
            E2, F2 = oracle.position_restraints(x, patoms[sel], refs[sel], pk[sel], 0.1, box=box if per else None)
            Eref += E2
            Fref += F2
        assert e.value == pytest.approx(Eref, rel=1e-12)
        assert np.abs(F - Fref).max() < F_ABS
        # accumulate into the caller's forces; repeated evaluations are bit-identical
        F2 = F.copy()
        C.check(L.atm_restraints_eval(h, _dp(x), _dp(box), _dp(F2), ctypes.byref(e)))
        assert np.array_equal(F2, 2 * F)
        # argument checks
        bad = (C.PositionRestraint * 1)(C.PositionRestraint(n, 0, (ctypes.c_double * 3)(0, 0, 0), 1.0, 0.0))
        assert L.atm_restraints_set(h, n, 0, None, None, 0, None, 1, bad) == C.ATM_ERR_INVALID_ARGUMENT
        assert b"out of range" in L.atm_last_error()
    finally:
        L.atm_restraints_destroy(h)


def _abfe_like():
    s = W.tiny_system(1200, 2.9, 0.9, 12, (1.0, 0.9, 0.8), 31,
                      W.ATMParams(0.3, 0.45, 0.1, 110.0, 1.5, 200.0, 100.0, 0.0625, 1.0))
    system = mm.System()
    masses = np.where(np.arange(s.num_atoms) % 3 == 0, 16.0, 12.0)
    for m in masses:
        system.addParticle(float(m))
    L = s.box
    system.setDefaultPeriodicBoxVectors((L[0], 0, 0), (0, L[1], 0), (0, 0, L[2]))
    nb = mm.NonbondedForce()
    nb.setNonbondedMethod(mm.NonbondedForce.CutoffPeriodic)
    nb.setCutoffDistance(s.cutoff)
    nb.setUseDispersionCorrection(False)
    for q, sg, e in zip(s.charges, s.sigmas, s.epsilons):
        nb.addParticle(float(q), float(sg), float(e))
    system.addForce(nb)
    utils = atm.ATMMetaForceUtils(system)
    lig = list(range(12))
    rcpt = list(range(100, 160))
    cm = utils.addVsiteRestraintForceCMCM(lig, rcpt, 25.0 * 4.184 * 100, 0.05, (0.3, 0.2, -0.1))
    refpos = s.positions + 0.07
    pr = utils.addPosRestraints(list(range(200, 240)), refpos, 25.0 * 4.184 * 100, 0.05)
    utils.setNonbondedForceGroup(1)
    p = s.params
    f = atm.ATMMetaForce(p.lambda1, p.lambda2, p.alpha, p.u0, p.w0, p.umax, p.ubcore, p.acore, p.direction, [1])
    for i in range(s.num_atoms):
        f.addParticle(i, *map(float, s.displacements[i]))
    f.setForceGroup(2)
    system.addForce(f)
    ctx = mm.Context(system, mm.VerletIntegrator(0.001), mm.Platform.getPlatformByName("MI355X"))
    ctx.setPositions(s.positions)
    return s, masses, lig, rcpt, refpos, system, cm, pr, f, ctx


def test_utils_restraints_in_a_context(gpu):
    s, masses, lig, rcpt, refpos, system, cm, pr, f, ctx = _abfe_like()
    E1, F1 = oracle.centroid_restraint(s.positions, lig, rcpt, masses[lig], masses[rcpt], 25.0 * 4.184 * 100, 0.05,
                                       (0.3, 0.2, -0.1))
    E2, F2 = oracle.position_restraints(s.positions, list(range(200, 240)), refpos[200:240], 25.0 * 4.184 * 100, 0.05,
                                        box=s.box)
    st = ctx.getState(getEnergy=True, getForces=True, groups={0})
    assert E1 > 0 and E2 > 0
    assert st.getPotentialEnergy() == pytest.approx(E1 + E2, rel=1e-12)
    assert np.abs(st.getForces(asNumpy=True) - (F1 + F2)).max() < F_ABS
    # group 0 beside the ATM force, as an ABFE script integrates {0, atm group}
    ref, Fatm = oracle.atm_eval(s)
    both = ctx.getState(getEnergy=True, getForces=True, groups={0, 2})
    e2 = ctx.getState(getEnergy=True, groups={2}).getPotentialEnergy()
    assert e2 == pytest.approx(ref["W"], rel=1e-6)
    assert both.getPotentialEnergy() == pytest.approx(st.getPotentialEnergy() + e2, rel=1e-12)
    Fb = both.getForces(asNumpy=True)
    assert np.linalg.norm(Fb - (F1 + F2 + Fatm)) / np.linalg.norm(F1 + F2 + Fatm) < 1e-5
    # parameters updated in the context
    cm.setBondParameters(0, 0, 1, 10.0, 0.0, (0.0, 0.0, 0.0))
    cm.updateParametersInContext(ctx)
    E1b, _ = oracle.centroid_restraint(s.positions, lig, rcpt, masses[lig], masses[rcpt], 10.0, 0.0, (0.0, 0.0, 0.0))
    assert ctx.getState(getEnergy=True, groups={0}).getPotentialEnergy() == pytest.approx(E1b + E2, rel=1e-12)
    # an MD step with the restraints and the ATM force
    ctx.getIntegrator().setIntegrationForceGroups({0, 2})
    ctx.getIntegrator().step(2)
    assert np.isfinite(ctx.getState(getEnergy=True, groups={0, 2}).getPotentialEnergy())
