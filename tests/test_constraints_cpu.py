"""Distance constraints (SHAKE) in the Langevin and Verlet integrators of the
OpenMM-compatible core, on force-free rigid 3-site waters (no GPU needed):
constraint lengths hold to the tolerance, the thermostat reaches
equipartition over the 3N - constraints degrees of freedom, and the Langevin
middle integrator refuses constrained systems."""
import numpy as np
import pytest

from atmmetaforce import openmm as mm

KB = 0.00831446261815324
D_OH, D_HH = 0.09572, 0.15139


def waters(n, seed=1):
    rng = np.random.default_rng(seed)
    s = mm.System()
    x = []
    for k in range(n):
        for m in (15.999, 1.008, 1.008):
            s.addParticle(m)
        o = rng.uniform(0.5, 2.5, 3)
        h1 = o + np.array([D_OH, 0.0, 0.0])
        ang = np.deg2rad(104.52)
        h2 = o + D_OH * np.array([np.cos(ang), np.sin(ang), 0.0])
        x += [o, h1, h2]
        s.addConstraint(3 * k, 3 * k + 1, D_OH)
        s.addConstraint(3 * k, 3 * k + 2, D_OH)
        s.addConstraint(3 * k + 1, 3 * k + 2, np.linalg.norm(h1 - h2))
    s.setDefaultPeriodicBoxVectors((3, 0, 0), (0, 3, 0), (0, 0, 3))
    return s, np.array(x)


def max_violation(s, x):
    worst = 0.0
    for k in range(s.getNumConstraints()):
        a, b, d = s.getConstraintParameters(k)
        worst = max(worst, abs(np.linalg.norm(x[a] - x[b]) - d) / d)
    return worst


@pytest.mark.parametrize("make", [lambda: mm.LangevinIntegrator(300.0, 5.0, 0.002),
                                  lambda: mm.VerletIntegrator(0.002)])
def test_constraints_hold(make):
    s, x = waters(8)
    integ = make()
    ctx = mm.Context(s, integ, mm.Platform.getPlatformByName("MI355X"))
    ctx.setPositions(x)
    if isinstance(integ, mm.VerletIntegrator):
        rng = np.random.default_rng(3)
        ctx.setVelocities(rng.normal(0.0, 1.0, x.shape))
    for _ in range(20):
        integ.step(10)
        xs = ctx.getState(getPositions=True).getPositions(asNumpy=True)
        assert max_violation(s, xs) < 2e-5
    assert not np.allclose(xs, x)


def test_thermostat_equipartition_with_constraints():
    s, x = waters(20, seed=4)
    integ = mm.LangevinIntegrator(300.0, 10.0, 0.001)
    integ.setRandomNumberSeed(11)
    ctx = mm.Context(s, integ, mm.Platform.getPlatformByName("MI355X"))
    ctx.setPositions(x)
    integ.step(2000)
    ke = []
    for _ in range(800):
        integ.step(5)
        ke.append(ctx.getState(getEnergy=True).getKineticEnergy())
    dof = 3 * s.getNumParticles() - s.getNumConstraints()
    expect = 0.5 * dof * KB * 300.0
    assert abs(np.mean(ke) / expect - 1.0) < 0.06, (np.mean(ke), expect)


def test_langevin_middle_refuses_constraints():
    s, x = waters(2)
    integ = mm.LangevinMiddleIntegrator(300.0, 1.0, 0.002)
    ctx = mm.Context(s, integ, mm.Platform.getPlatformByName("MI355X"))
    ctx.setPositions(x)
    with pytest.raises(Exception, match="does not apply constraints"):
        integ.step(1)
