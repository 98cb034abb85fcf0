"""Variable force groups holding bonded forces next to the NonbondedForce.

The reference evaluates every force of the variable groups at x and at x + d
(openmmapi/src/ATMMetaForceImpl.cpp:32-67 clones them into its two inner
contexts, :113,116 evaluates both): U1 and U2 are sums over all of them, the
softcore and the perturbation energy see the sums, and every force of those
groups is blended with the same weights.  Checked here against the oracle's
nonbonded (reaction field, PME) and bonded energies at x and x + d, combined
and blended by the oracle, and against the reference's test_abfe.py goldens
with the AMBER bonded forces moved into the variable group.
"""
import dataclasses

import numpy as np
import pytest

import atmmetaforce as atm
from atmmetaforce import openmm as mm, workloads as W
from oracle import oracle, pme_oracle as po

import temoa

pytestmark = pytest.mark.gpu

E_TOL = 1e-6   # RF pair energies in FP64 on the 2^-20 nm grid, bonded in FP64
F_TOL = 1e-5   # relative RMS of the blended force (fixed-point pair sums)

SYS = W.with_molecules(
    W.tiny_system(3000, 3.1, 0.9, 30, (1.03, 0.77, 0.62), 1,
                  W.ATMParams(0.25, 0.45, 0.1, 110.0, 2.0, 200.0, 100.0, 0.0625, 1.0)),
    size=4, crosslink_every=9)


def chain_terms(x, n, size=4):
    """Bonds (i, i+1), angles (i, i+1, i+2) and torsions (i..i+3) inside every
    molecule of `size` consecutive atoms, with equilibrium values a little off the
    current geometry so that every term carries force.  The molecule holding
    atoms 28-31 spans the 30-atom ligand and the environment: its bonds are
    stretched by the displacement in state 2."""
    bonds, bp, angles, ap, tors, tp = [], [], [], [], [], []
    for a in range(0, n - size + 1, size):
        for i in range(a, a + size - 1):
            r = np.linalg.norm(x[i + 1] - x[i])
            bonds.append((i, i + 1))
            bp.append((0.9 * r, 900.0 + i % 7))
        for i in range(a, a + size - 2):
            v1, v2 = x[i] - x[i + 1], x[i + 2] - x[i + 1]
            th = np.arccos(np.clip(v1 @ v2 / np.linalg.norm(v1) / np.linalg.norm(v2), -1, 1))
            angles.append((i, i + 1, i + 2))
            ap.append((th * 0.97 if th > 0.3 else 1.9, 80.0))
        tors.append((a, a + 1, a + 2, a + 3))
        tp.append((1 + a % 3, 0.3 * (a % 5), 2.5))
    return (np.array(bonds), np.array(bp), np.array(angles), np.array(ap), np.array(tors), np.array(tp))


def add_bonded(system, terms, groups):
    """HarmonicBondForce, HarmonicAngleForce, PeriodicTorsionForce in force groups `groups`."""
    bonds, bp, angles, ap, tors, tp = terms
    bf = mm.HarmonicBondForce()
    for (i, j), (r0, k) in zip(bonds, bp):
        bf.addBond(int(i), int(j), float(r0), float(k))
    af = mm.HarmonicAngleForce()
    for (i, j, k), (t0, kk) in zip(angles, ap):
        af.addAngle(int(i), int(j), int(k), float(t0), float(kk))
    tf = mm.PeriodicTorsionForce()
    for (i, j, k, l), (per, ph, kk) in zip(tors, tp):
        tf.addTorsion(int(i), int(j), int(k), int(l), int(per), float(ph), float(kk))
    for f, g in zip((bf, af, tf), groups):
        f.setForceGroup(g)
        system.addForce(f)


def rf_system(s, terms, params, variable=(1, 3), groups=(1, 1, 3)):
    system = mm.System()
    for _ in range(s.num_atoms):
        system.addParticle(12.0)
    L = s.box
    system.setDefaultPeriodicBoxVectors((L[0], 0, 0), (0, L[1], 0), (0, 0, L[2]))
    nb = mm.NonbondedForce()
    nb.setNonbondedMethod(mm.NonbondedForce.CutoffPeriodic)
    nb.setCutoffDistance(s.cutoff)
    nb.setReactionFieldDielectric(s.rf_dielectric)
    nb.setUseDispersionCorrection(False)
    for q, sg, e in zip(s.charges, s.sigmas, s.epsilons):
        nb.addParticle(float(q), float(sg), float(e))
    ex = s.exceptions
    for (i, j), qq, sg, e in zip(ex.pairs, ex.charge_prod, ex.sigma, ex.epsilon):
        nb.addException(int(i), int(j), float(qq), float(sg), float(e))
    nb.setForceGroup(1)
    system.addForce(nb)
    add_bonded(system, terms, groups)
    p = params
    f = atm.ATMMetaForce(p.lambda1, p.lambda2, p.alpha, p.u0, p.w0, p.umax, p.ubcore, p.acore, p.direction,
                         list(variable))
    for i in range(s.num_atoms):
        d = s.displacements[i]
        f.addParticle(i, float(d[0]), float(d[1]), float(d[2]))
    f.setForceGroup(2)
    system.addForce(f)
    return system, f


def oracle_atm(params, U1, U2, F1, F2):
    r = oracle.combine(params, U1, U2)
    return r, oracle.blend(r["sp"], params["direction"] if isinstance(params, dict) else params.direction, F1, F2)


# direction -1 is covered on the PME system: here the displaced ligand overlaps
# environment atoms, so the state-2 energy that direction -1 takes as its base is
# astronomically large
@pytest.mark.parametrize("props,direction", [({}, 1.0), ({"ATMEvaluation": "plain"}, 1.0),
                                             ({"Precision": "double"}, 1.0)])
def test_reaction_field_with_bonded_forces(gpu, props, direction):
    s = SYS
    params = dataclasses.replace(s.params, direction=direction)
    x, xd = s.positions, s.positions + s.displacements
    terms = chain_terms(x, s.num_atoms)
    print(f"[vargroups] RF {props} direction={direction} bonds={len(terms[0])}", flush=True)
    system, f = rf_system(s, terms, params)
    ctx = mm.Context(system, mm.VerletIntegrator(0.001), mm.Platform.getPlatformByName("MI355X"), props)
    ctx.setPositions(x)
    st = ctx.getState(getEnergy=True, getForces=True, groups={2})

    def state(y):
        E, F = oracle.nonbonded(y, s.charges, s.sigmas, s.epsilons, s.box, s.cutoff, s.rf_dielectric,
                                exceptions=s.exceptions)[:2]
        eb, Fb = po.bonded(y, *terms)
        return E + sum(eb.values()), F + Fb, sum(eb.values())

    U1, F1, eb1 = state(x)
    U2, F2, eb2 = state(xd)
    assert abs(eb2 - eb1) > 100.0          # the spanning molecule's bonds are stretched in state 2
    ref, Fref = oracle_atm(params, U1, U2, F1, F2)
    print(props, st.getPotentialEnergy(), ref["W"], flush=True)
    assert abs(st.getPotentialEnergy() - ref["W"]) < E_TOL * abs(ref["W"])
    assert abs(f.getPerturbationEnergy(ctx) - ref["usc"]) < E_TOL * max(1.0, abs(ref["usc"]))
    F = st.getForces(asNumpy=True)
    assert np.linalg.norm(F - Fref) / np.linalg.norm(Fref) < F_TOL
    # a parameter change re-blends the bonded forces with the new weights
    ctx.setParameter(f.Lambda2(), 0.7)
    st = ctx.getState(getEnergy=True, getForces=True, groups={2})
    p2 = dataclasses.replace(params, lambda2=0.7)
    ref, Fref = oracle_atm(p2, U1, U2, F1, F2)
    assert abs(st.getPotentialEnergy() - ref["W"]) < E_TOL * abs(ref["W"])
    assert np.linalg.norm(st.getForces(asNumpy=True) - Fref) / np.linalg.norm(Fref) < F_TOL


def test_bonded_forces_outside_the_variable_groups_stay_out(gpu):
    """Torsions in group 3 when only group 1 is variable: the ATM force sees the
    nonbonded, bond and angle terms; group 3 is the plain torsion force at x."""
    s = SYS
    x, xd = s.positions, s.positions + s.displacements
    terms = chain_terms(x, s.num_atoms)
    system, f = rf_system(s, terms, s.params, variable=(1,), groups=(1, 1, 3))
    ctx = mm.Context(system, mm.VerletIntegrator(0.001), mm.Platform.getPlatformByName("MI355X"))
    ctx.setPositions(x)
    bonds, bp, angles, ap, tors, tp = terms

    def state(y):
        E, F = oracle.nonbonded(y, s.charges, s.sigmas, s.epsilons, s.box, s.cutoff, s.rf_dielectric,
                                exceptions=s.exceptions)[:2]
        eb, Fb = po.bonded(y, bonds, bp, angles, ap)
        return E + sum(eb.values()), F + Fb

    U1, F1 = state(x)
    U2, F2 = state(xd)
    ref, Fref = oracle_atm(s.params, U1, U2, F1, F2)
    st = ctx.getState(getEnergy=True, getForces=True, groups={2})
    assert abs(st.getPotentialEnergy() - ref["W"]) < E_TOL * abs(ref["W"])
    assert np.linalg.norm(st.getForces(asNumpy=True) - Fref) / np.linalg.norm(Fref) < F_TOL
    et, Ft = po.bonded(x, torsions=tors, torsion_params=tp)
    st3 = ctx.getState(getEnergy=True, getForces=True, groups={3})
    assert abs(st3.getPotentialEnergy() - sum(et.values())) < 1e-9 * abs(sum(et.values()))
    assert np.abs(st3.getForces(asNumpy=True) - Ft).max() < 1e-6   # fixed-point force sums


def test_other_forces_in_variable_groups_are_rejected(gpu):
    s = SYS
    terms = chain_terms(s.positions, s.num_atoms)
    # bonded forces alone (no NonbondedForce in the variable groups)
    system, f = rf_system(s, terms, s.params, variable=(3,), groups=(3, 3, 3))
    with pytest.raises(Exception, match="NonbondedForce"):
        mm.Context(system, mm.VerletIntegrator(0.001), mm.Platform.getPlatformByName("MI355X"))


def pme_system(x, q, sig, eps, L, pairs, qq, es, ee, d, terms, params, tol):
    system = mm.System()
    for _ in range(len(q)):
        system.addParticle(12.0)
    system.setDefaultPeriodicBoxVectors((L[0], 0, 0), (0, L[1], 0), (0, 0, L[2]))
    nb = mm.NonbondedForce()
    nb.setNonbondedMethod(mm.NonbondedForce.PME)
    nb.setCutoffDistance(1.0)
    nb.setEwaldErrorTolerance(tol)
    nb.setUseDispersionCorrection(True)
    for a, b, c in zip(q, sig, eps):
        nb.addParticle(float(a), float(b), float(c))
    for (i, j), a, b, c in zip(pairs, qq, es, ee):
        nb.addException(int(i), int(j), float(a), float(b), float(c))
    nb.setForceGroup(1)
    system.addForce(nb)
    add_bonded(system, terms, (1, 1, 1))
    f = atm.ATMMetaForce(params["lambda1"], params["lambda2"], params["alpha"], params["u0"], params["w0"],
                         params["umax"], params["ubcore"], params["acore"], params["direction"], [1])
    for i in range(len(q)):
        f.addParticle(i, float(d[i, 0]), float(d[i, 1]), float(d[i, 2]))
    f.setForceGroup(2)
    system.addForce(f)
    return system, f


@pytest.mark.parametrize("direction", [1.0, -1.0])
def test_pme_with_bonded_forces(gpu, direction):
    """The PME path: 400 atoms with exceptions, bonded chains of which one spans
    the 20-atom ligand (atoms 40-59) and the environment."""
    import test_pme_gpu as tp
    x, q, sig, eps, L, pairs, qq, es, ee, d = tp._small_system()
    params = dict(tp.PARAMS, direction=direction)
    # chains 36-39 | 40-43 ... 56-59 | 60-63; the molecule 58-61 spans the ligand boundary
    bonds = np.array([(i, i + 1) for i in range(34, 64)])
    bp = np.array([(0.9 * np.linalg.norm(x[j] - x[i]), 700.0) for i, j in bonds])
    angles = np.array([(i, i + 1, i + 2) for i in range(34, 63, 2)])
    ap = np.array([(1.8, 60.0)] * len(angles))
    tors = np.array([(i, i + 1, i + 2, i + 3) for i in range(34, 61, 3)])
    tpar = np.array([(2, 0.4, 3.0)] * len(tors))
    terms = (bonds, bp, angles, ap, tors, tpar)
    system, f = pme_system(x, q, sig, eps, L, pairs, qq, es, ee, d, terms, params, 5e-4)
    print(f"[vargroups] PME small system direction={direction}", flush=True)
    ctx = mm.Context(system, mm.VerletIntegrator(0.001), mm.Platform.getPlatformByName("MI355X"))
    ctx.setPositions(x)
    st = ctx.getState(getEnergy=True, getForces=True, groups={2})
    U1, F1, _ = po.nonbonded_pme(x, q, sig, eps, L, 1.0, pairs, qq, es, ee, tol=5e-4, dispersion=True)
    U2, F2, _ = po.nonbonded_pme(x + d, q, sig, eps, L, 1.0, pairs, qq, es, ee, tol=5e-4, dispersion=True)
    eb1, Fb1 = po.bonded(x, *terms)
    eb2, Fb2 = po.bonded(x + d, *terms)
    assert abs(sum(eb2.values()) - sum(eb1.values())) > 100.0
    ref, Fref = oracle_atm(params, U1 + sum(eb1.values()), U2 + sum(eb2.values()), F1 + Fb1, F2 + Fb2)
    print(st.getPotentialEnergy(), ref["W"], flush=True)
    assert abs(st.getPotentialEnergy() - ref["W"]) <= 1e-8 * max(1.0, abs(ref["W"]))
    assert abs(f.getPerturbationEnergy(ctx) - ref["usc"]) <= 1e-8 * max(1.0, abs(ref["usc"]))
    assert np.abs(st.getForces(asNumpy=True) - Fref).max() < tp.F_ABS


def test_abfe_goldens_with_bonded_forces_in_the_variable_group(gpu):
    """test_abfe.py's system with its bonded forces moved from group 0 into the
    ATM variable group 1: the ligand moves rigidly, so the bonded energy is the
    same in both states and enters W once; the goldens of test_abfe.py:147-150
    (potential energy of groups {0, ATM}, perturbation energy) hold unchanged,
    and the forces of groups {0, ATM} match those of the original layout."""
    import test_abfe_api_gpu as ab
    ctx0, force0, integ0, s, x = ab.build_abfe_context()
    F0 = ctx0.getState(getForces=True, groups={0, 2}).getForces(asNumpy=True)
    del ctx0, integ0
    print("[vargroups] temoa-g1 with bonded forces in the variable group", flush=True)
    ctx, force, integ, s, x = ab.build_abfe_context(bonded_group=1)
    st = ctx.getState(getEnergy=True, getForces=True, groups={0, 2})
    assert abs(st.getPotentialEnergy() - temoa.GOLDEN_POTENTIAL) <= temoa.GOLDEN_ATOL, st.getPotentialEnergy()
    assert abs(force.getPerturbationEnergy(ctx) - temoa.GOLDEN_PERTURBATION) <= temoa.GOLDEN_ATOL
    assert np.abs(st.getForces(asNumpy=True) - F0).max() < 1e-4


@pytest.mark.parametrize("plain", [False, True])
def test_batch_energy_offsets(gpu, plain):
    """atm_set_energy_offsets through ATMBatch: per-system offsets enter U1 and U2
    before the combine (so the weights, and with them the blended pair forces,
    change); a system without offsets is unaffected; zero offsets restore it."""
    from atmmetaforce import ATMBatch
    systems = [W.tiny_system(1500, 2.6, 0.9, 12, (0.9, 0.8, 0.7), 2, SYS.params),
               W.tiny_system(2000, 2.8, 0.9, 20, (1.1, 0.6, 0.4), 4, SYS.params)]
    offs = [(0.0, 0.0), (-250.0, 40.0)]
    with ATMBatch(systems, device=0) as b:
        for k, (e1, e2) in enumerate(offs):
            b.set_energy_offsets(k, e1, e2)
        b.compute(plain=plain)
        for k, s in enumerate(systems):
            ref, _ = oracle.atm_eval(s, forces=False)
            E1, F1 = oracle.nonbonded(s.positions, s.charges, s.sigmas, s.epsilons, s.box, s.cutoff,
                                      s.rf_dielectric)[:2]
            E2, F2 = oracle.nonbonded(s.positions + s.displacements, s.charges, s.sigmas, s.epsilons, s.box,
                                      s.cutoff, s.rf_dielectric)[:2]
            want, Fwant = oracle_atm(s.params, E1 + offs[k][0], E2 + offs[k][1], F1, F2)
            r = b.result(k)
            print(f"[vargroups] batch offsets plain={plain} system {k}: W {r['energy']} oracle {want['W']}",
                  flush=True)
            assert abs(r["state1_energy"] - want["U1"]) < E_TOL * abs(want["U1"])
            assert abs(r["state2_energy"] - want["U2"]) < E_TOL * abs(want["U2"])
            assert abs(r["energy"] - want["W"]) < E_TOL * abs(want["W"])
            assert abs(r["hybrid_weight"] - want["sp"]) < 1e-6
            assert np.linalg.norm(b.forces(k) - Fwant) / np.linalg.norm(Fwant) < F_TOL
        b.set_energy_offsets(1, 0.0, 0.0)
        b.compute(plain=plain)
        ref, _ = oracle.atm_eval(systems[1], forces=False)
        assert abs(b.result(1)["energy"] - ref["W"]) < E_TOL * abs(ref["W"])
