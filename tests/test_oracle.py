"""Pinning the FP64 oracle (oracle/atm_oracle.c) before trusting it.

Each check is independent of the oracle's own formulas: closed-form values of
the Lennard-Jones and reaction-field textbook potentials, F = -grad U by
central differences, Newton's third law, translation invariance, the cell list
against brute force, soft-core/softplus invariants and limits, the blend end
points, analytic gradients against central differences, and k_e against CODATA.
"""
import math

import numpy as np
import pytest

from atmmetaforce import workloads as W
from oracle import oracle

BIG = np.array([6.0, 6.0, 6.0])


def two_atoms(r, s1=0.3, s2=0.3, e1=0.5, e2=0.5, q1=0.0, q2=0.0, rc=1.0, erf=78.5):
    x = np.array([[1.0, 1.0, 1.0], [1.0 + r, 1.0, 1.0]])
    E, F, n, parts = oracle.nonbonded(x, [q1, q2], [s1, s2], [e1, e2], BIG, rc, erf)
    return E, F, n, parts


# ---------------------------------------------------------------- pair term

def test_lj_zero_crossing_and_force_at_sigma():
    s, e = 0.3, 0.5
    E, F, n, _ = two_atoms(s, s, s, e, e)
    assert n == 1
    assert abs(E) < 1e-12
    # F(sigma) = 4 eps (12 - 6)/sigma, pushing atom 0 towards -x
    assert F[0, 0] == pytest.approx(-24 * e / s, rel=1e-12)
    assert F[1, 0] == pytest.approx(24 * e / s, rel=1e-12)


def test_lj_minimum_and_combining_rules():
    s1, s2, e1, e2 = 0.25, 0.33, 0.2, 0.7
    sig = 0.5 * (s1 + s2)           # Lorentz
    eps = math.sqrt(e1 * e2)        # Berthelot
    rmin = 2 ** (1 / 6) * sig
    E, F, _, _ = two_atoms(rmin, s1, s2, e1, e2)
    assert E == pytest.approx(-eps, rel=1e-12)
    assert abs(F[0, 0]) < 1e-9
    E0, _, _, _ = two_atoms(sig, s1, s2, e1, e2)
    assert abs(E0) < 1e-12


def test_reaction_field_closed_forms():
    ke = 138.935456
    rc = 1.0
    # eps_rf = 1: k_rf = 0, c_rf = 1/rc -> plain Coulomb shifted to zero at rc
    for r in (0.3, 0.5, 0.9):
        E, _, _, _ = two_atoms(r, e1=0.0, e2=0.0, q1=0.7, q2=-0.4, erf=1.0)
        assert E == pytest.approx(ke * 0.7 * -0.4 * (1 / r - 1 / rc), rel=1e-12)
    # any eps_rf: E(rc) -> 0 and F(rc) = k_e q q / rc^2 * 3 / (2 eps_rf + 1)
    for erf in (2.0, 78.5, 1e12):
        r = rc - 1e-9
        E, F, _, _ = two_atoms(r, e1=0.0, e2=0.0, q1=0.5, q2=0.5, erf=erf)
        assert abs(E) < 1e-6
        assert -F[0, 0] == pytest.approx(ke * 0.25 / rc ** 2 * 3 / (2 * erf + 1), rel=1e-6, abs=1e-6)


def test_cutoff_is_strict():
    rc = 0.5  # exactly representable: the pair at r == rc is excluded
    _, _, n_at, _ = two_atoms(rc, rc=rc)
    _, _, n_in, _ = two_atoms(rc - 2.0 ** -20, rc=rc)
    assert n_at == 0 and n_in == 1


def test_coulomb_constant_against_codata():
    e = 1.602176634e-19
    na = 6.02214076e23
    eps0 = 8.8541878128e-12
    ke = e * e * na / (4 * math.pi * eps0) * 1e9 / 1e3  # kJ nm / (mol e^2)
    assert ke == pytest.approx(138.935456, rel=2e-8)


@pytest.mark.parametrize("seed", [1, 2])
def test_forces_are_minus_gradient(seed):
    s = W.tiny_system(120, 2.2, 0.9, 5, (0.5, 0.4, 0.3), seed)
    x = s.positions.copy()
    E0, F, _, _ = oracle.nonbonded(x, s.charges, s.sigmas, s.epsilons, s.box, s.cutoff, s.rf_dielectric)
    rng = np.random.default_rng(seed)
    h = 1e-6
    for a in rng.choice(len(x), 5, replace=False):
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[a, k] += h
            xm[a, k] -= h
            Ep = oracle.nonbonded(xp, s.charges, s.sigmas, s.epsilons, s.box, s.cutoff, s.rf_dielectric, forces=False)[0]
            Em = oracle.nonbonded(xm, s.charges, s.sigmas, s.epsilons, s.box, s.cutoff, s.rf_dielectric, forces=False)[0]
            assert -(Ep - Em) / (2 * h) == pytest.approx(F[a, k], rel=1e-5, abs=1e-4)


def test_newton_third_law_and_translation_invariance():
    s = W.tiny_system(400, 2.6, 0.9, 5, (0.5, 0.4, 0.3), 7)
    E, F, n, _ = oracle.nonbonded(s.positions, s.charges, s.sigmas, s.epsilons, s.box, s.cutoff, s.rf_dielectric)
    assert np.abs(F.sum(0)).max() < 1e-8 * np.abs(F).sum()
    shift = W.quantize([0.37, -1.91, 2.5])
    E2, F2, n2, _ = oracle.nonbonded(s.positions + shift, s.charges, s.sigmas, s.epsilons, s.box, s.cutoff,
                                     s.rf_dielectric)
    assert n2 == n
    assert E2 == pytest.approx(E, rel=1e-12)
    assert np.allclose(F2, F, rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("n,box,rc", [(300, 2.8, 0.9), (1500, 3.6, 1.0), (3000, 3.1, 0.9)])
def test_cell_list_equals_brute_force(n, box, rc):
    s = W.tiny_system(n, box, rc, 10, (0.5, 0.4, 0.3), n)
    pc = oracle.pairs(s.positions, s.box, rc)
    pb = oracle.pairs(s.positions, s.box, rc, brute=True)
    assert np.array_equal(pc, pb)
    Ec = oracle.nonbonded(s.positions, s.charges, s.sigmas, s.epsilons, s.box, rc, 78.5)[0]
    Eb = oracle.nonbonded(s.positions, s.charges, s.sigmas, s.epsilons, s.box, rc, 78.5, brute=True)[0]
    assert Ec == pytest.approx(Eb, rel=1e-12)


def test_pair_count_by_brute_force_python():
    """Tiny system: the in-cutoff pair set by an independent numpy brute force."""
    s = W.tiny_system(60, 2.0, 0.7, 3, (0.3, 0.2, 0.1), 3)
    x = s.positions
    L = s.box
    d = x[:, None, :] - x[None, :, :]
    d -= L * np.round(d / L)
    r2 = (d ** 2).sum(-1)
    i, j = np.nonzero(np.triu(r2 < 0.7 * 0.7, 1))
    ref = np.stack([i, j], 1)
    assert np.array_equal(oracle.pairs(x, L, 0.7), ref)


# ---------------------------------------------------------------- ATM scalars

UMAX, UB, AC = 200.0, 100.0, 0.0625


def test_softcore_invariants():
    for u in (-50.0, 0.0, 99.9, 100.0):
        v, fp = oracle.softcore(u, UMAX, AC, UB)
        assert v == u and fp == 1.0
    v, fp = oracle.softcore(UB + 1e-9, UMAX, AC, UB)
    assert v == pytest.approx(UB, abs=1e-6) and fp == pytest.approx(1.0, abs=1e-6)
    v, _ = oracle.softcore(1e30, UMAX, AC, UB)
    assert v == pytest.approx(UMAX, rel=1e-3)
    us = np.linspace(100.5, 5000, 200)
    vals = [oracle.softcore(u, UMAX, AC, UB)[0] for u in us]
    assert all(b > a for a, b in zip(vals, vals[1:]))
    for u in (101.0, 150.0, 400.0, 3e3):
        h = 1e-4 * u
        d = (oracle.softcore(u + h, UMAX, AC, UB)[0] - oracle.softcore(u - h, UMAX, AC, UB)[0]) / (2 * h)
        assert oracle.softcore(u, UMAX, AC, UB)[1] == pytest.approx(d, rel=1e-6)


def P(l1, l2, alpha=0.1, u0=110.0, w0=0.0, direction=1.0):
    return W.ATMParams(l1, l2, alpha, u0, w0, UMAX, UB, AC, direction)


def test_softplus_equal_lambdas():
    # lambda1 = lambda2 = lambda: W = e0 + lambda u_sc + w0 and bfp = lambda
    for lam, u1, u2 in ((0.3, -100.0, -40.0), (0.5, 10.0, 500.0)):
        r = oracle.combine(P(lam, lam, w0=2.5), u1, u2)
        assert r["W"] == pytest.approx(u1 + lam * r["usc"] + 2.5, rel=1e-14)
        assert r["bfp"] == pytest.approx(lam, rel=1e-14)


def test_softplus_limits_and_derivative():
    l1, l2, a, u0 = 0.1, 0.4, 0.5, 50.0
    # u_sc >> u0: W -> e0 + l2 u_sc ; u_sc << u0: W -> e0 + l1 u_sc + (l2 - l1) u0
    r = oracle.combine(P(l1, l2, a, u0), 0.0, 95.0)
    assert r["W"] == pytest.approx(l2 * r["usc"], rel=1e-8)
    r = oracle.combine(P(l1, l2, a, u0), 0.0, -80.0)
    assert r["W"] == pytest.approx(l1 * r["usc"] + (l2 - l1) * u0, rel=1e-8)
    # dW/du_sc = bfp (below the soft-core threshold u_sc = u)
    for u in (-10.0, 40.0, 60.0):
        h = 1e-5
        d = (oracle.combine(P(l1, l2, 0.1, u0), 0.0, u + h)["W"] - oracle.combine(P(l1, l2, 0.1, u0), 0.0, u - h)["W"]) / (2 * h)
        assert oracle.combine(P(l1, l2, 0.1, u0), 0.0, u)["bfp"] == pytest.approx(d, rel=1e-7)


def test_direction_symmetry_and_alpha_zero_quirk():
    a = oracle.combine(P(0.2, 0.35, direction=1.0), -30.0, 120.0)
    b = oracle.combine(P(0.2, 0.35, direction=-1.0), 120.0, -30.0)
    assert a["W"] == b["W"] and a["usc"] == b["usc"] and a["sp"] == b["sp"]
    z = oracle.combine(P(0.2, 0.35, direction=0.0), 120.0, -30.0)  # 0 counts as reverse
    assert z["W"] == b["W"]
    # alpha = 0: energy drops the softplus term, force weight uses bfp = (l1 + l2)/2
    q = oracle.combine(P(0.2, 0.35, alpha=0.0), 0.0, 40.0)
    assert q["W"] == pytest.approx(0.35 * 40.0)
    assert q["bfp"] == pytest.approx(0.275)


def test_blend_end_points():
    rng = np.random.default_rng(0)
    F1, F2 = rng.normal(size=(7, 3)), rng.normal(size=(7, 3))
    assert np.array_equal(oracle.blend(0.0, 1.0, F1, F2), F1)
    assert np.array_equal(oracle.blend(1.0, 1.0, F1, F2), F2)
    assert np.array_equal(oracle.blend(0.0, -1.0, F1, F2), F2)
    assert np.array_equal(oracle.blend(1.0, -1.0, F1, F2), F1)
    acc = oracle.blend(0.25, 1.0, F1, F2, F=np.ones((7, 3)))
    assert np.allclose(acc, 1.0 + 0.75 * F1 + 0.25 * F2)


def test_atm_eval_zero_displacement():
    s = W.tiny_system(200, 2.4, 0.8, 0, (0, 0, 0), 4)
    r, F = oracle.atm_eval(s)
    assert r["U1"] == r["U2"] and r["u"] == 0.0
    E, F1, _, _ = oracle.nonbonded(s.positions, s.charges, s.sigmas, s.epsilons, s.box, s.cutoff, s.rf_dielectric)
    assert np.allclose(F, F1)


# ---------------------------------------------------------------- gradients

def _perturbed(s, field, i, h):
    arr = getattr(s, field).copy()
    arr[i] += h
    import dataclasses
    return dataclasses.replace(s, **{field: arr})


@pytest.mark.parametrize("direction", [1.0, -1.0])
def test_analytic_gradient_against_central_differences(direction):
    # sparse box: no clashes, so u stays moderate in both directions
    p = W.ATMParams(0.15, 0.45, 0.08, 30.0, 1.0, 200.0, 100.0, 0.0625, direction)
    s = W.tiny_system(250, 3.2, 0.9, 8, (1.1, 0.7, 0.45), 5, p)
    g = oracle.atm_grad(s)
    for field, key, h in (("charges", "dq", 1e-6), ("sigmas", "dsigma", 1e-7), ("epsilons", "deps", 1e-7)):
        for i in (0, 3, 77):
            wp = oracle.atm_eval(_perturbed(s, field, i, h), forces=False)[0]["W"]
            wm = oracle.atm_eval(_perturbed(s, field, i, -h), forces=False)[0]["W"]
            assert g[key][i] == pytest.approx((wp - wm) / (2 * h), rel=1e-5, abs=1e-6), (field, i)
    import dataclasses
    for k, name in enumerate(("lambda1", "lambda2", "alpha", "u0", "w0")):
        h = 1e-6
        pp = dataclasses.replace(p, **{name: getattr(p, name) + h})
        pm = dataclasses.replace(p, **{name: getattr(p, name) - h})
        d = (oracle.atm_eval(s, pp, forces=False)[0]["W"] - oracle.atm_eval(s, pm, forces=False)[0]["W"]) / (2 * h)
        assert g["dparams"][k] == pytest.approx(d, rel=1e-6, abs=1e-7), name


def test_gradient_at_zero_sigma_and_zero_epsilon():
    """sigma_i = sigma_j = 0: E_LJ = 4 eps ((s/r)^12 - (s/r)^6) is even in s, so
    dW/dsigma is 0 there (central differences through +-h agree); the quotient
    form of the derivative would be 0/0.  eps_i = 0: the derivative of
    sqrt(eps_i eps_j) does not exist at 0 and the restatement returns 0 for that
    atom, while the other atoms' gradients stay the central differences."""
    import dataclasses
    p = W.ATMParams(0.15, 0.45, 0.08, 30.0, 1.0, 200.0, 100.0, 0.0625, 1.0)
    s = W.tiny_system(250, 3.2, 0.9, 8, (1.1, 0.7, 0.45), 5, p)
    sig = s.sigmas.copy()
    eps = s.epsilons.copy()
    sig[[2, 9]] = 0.0          # atoms 2 and 9: a sigma_ij = 0 pair if within the cutoff
    eps[[4, 30]] = 0.0
    s = dataclasses.replace(s, sigmas=sig, epsilons=eps)
    x = s.positions
    d = x[2] - x[9]
    d -= s.box * np.round(d / s.box)
    assert np.linalg.norm(d) < s.cutoff
    g = oracle.atm_grad(s)
    assert np.isfinite(g["dsigma"]).all() and np.isfinite(g["deps"]).all()
    assert g["deps"][4] == 0.0 and g["deps"][30] == 0.0
    h = 1e-7
    for field, key, i in (("sigmas", "dsigma", 2), ("sigmas", "dsigma", 9), ("sigmas", "dsigma", 3),
                          ("epsilons", "deps", 5), ("charges", "dq", 4)):
        wp = oracle.atm_eval(_perturbed(s, field, i, h), forces=False)[0]["W"]
        wm = oracle.atm_eval(_perturbed(s, field, i, -h), forces=False)[0]["W"]
        assert g[key][i] == pytest.approx((wp - wm) / (2 * h), rel=1e-5, abs=1e-6), (field, i)


# ---------------------------------------------------------------- restraints

def _num_grad(f, x, h=1e-6):
    g = np.zeros_like(x)
    for i in range(x.shape[0]):
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, k] += h
            xm[i, k] -= h
            g[i, k] = (f(xp) - f(xm)) / (2 * h)
    return g


def test_centroid_restraint_closed_form_and_gradient():
    # two single-atom groups: the centroid distance is the atom distance
    x = np.array([[0.0, 0.0, 0.0], [1.5, 0.0, 0.0]])
    E, F = oracle.centroid_restraint(x, [1], [0], [1.0], [1.0], kf=100.0, tol=0.5, offset=[0.0, 0.0, 0.0])
    assert E == pytest.approx(0.5 * 100.0 * 1.0 ** 2)
    assert F[1] == pytest.approx([-100.0, 0.0, 0.0]) and F[0] == pytest.approx([100.0, 0.0, 0.0])
    # inside the flat bottom: nothing
    E0, F0 = oracle.centroid_restraint(x, [1], [0], [1.0], [1.0], kf=100.0, tol=2.0, offset=[0, 0, 0])
    assert E0 == 0.0 and not F0.any()
    # the offset is subtracted from c1 - c2
    E1, _ = oracle.centroid_restraint(x, [1], [0], [1.0], [1.0], kf=100.0, tol=0.0, offset=[1.5, 0.0, 0.0])
    assert E1 == 0.0
    # weighted groups: F = -grad E, sum F = 0, weights matter
    rng = np.random.default_rng(4)
    y = rng.uniform(0, 3, (9, 3))
    w = rng.uniform(1, 16, 9)
    args = dict(kf=250.0, tol=0.3, offset=[0.2, -0.1, 0.4])
    f = lambda z: oracle.centroid_restraint(z, [0, 1, 2, 3], [4, 5, 6, 7, 8], w[:4], w[4:], **args)[0]
    E, F = oracle.centroid_restraint(y, [0, 1, 2, 3], [4, 5, 6, 7, 8], w[:4], w[4:], **args)
    assert E > 0
    assert np.allclose(F, -_num_grad(f, y), rtol=1e-6, atol=1e-6)
    assert np.allclose(F.sum(0), 0.0, atol=1e-9)
    Eu, _ = oracle.centroid_restraint(y, [0, 1, 2, 3], [4, 5, 6, 7, 8], np.ones(4), np.ones(5), **args)
    # an atom in both groups: each group uses its own weight for it
    Eo, Fo = oracle.centroid_restraint(y, [0, 1, 4], [4, 5], [1.0, 2.0, 3.0], [5.0, 1.0], **args)
    fo = lambda z: oracle.centroid_restraint(z, [0, 1, 4], [4, 5], [1.0, 2.0, 3.0], [5.0, 1.0], **args)[0]
    assert np.allclose(Fo, -_num_grad(fo, y), rtol=1e-6, atol=1e-6)
    assert Eu != pytest.approx(E)


def test_position_restraints_closed_form_periodic_and_gradient():
    x = np.array([[0.1, 0.2, 0.3], [3.9, 0.0, 0.0]])
    E, F = oracle.position_restraints(x, [0], [[0.1, 0.2, 1.3]], kf=10.0, tol=0.4)
    assert E == pytest.approx(0.5 * 10.0 * 0.6 ** 2)
    assert F[0] == pytest.approx([0.0, 0.0, 10.0 * 0.6]) and not F[1].any()
    # periodic: the image 0.1 nm away, not the one 3.9 nm away
    Ep, _ = oracle.position_restraints(x, [1], [[0.0, 0.0, 0.0]], kf=10.0, tol=0.05, box=[4.0, 4.0, 4.0])
    En, _ = oracle.position_restraints(x, [1], [[0.0, 0.0, 0.0]], kf=10.0, tol=0.05)
    assert Ep == pytest.approx(0.5 * 10.0 * 0.05 ** 2) and En == pytest.approx(0.5 * 10.0 * 3.85 ** 2)
    rng = np.random.default_rng(5)
    y = rng.uniform(0, 4, (6, 3))
    refs = rng.uniform(0, 4, (4, 3))
    kf = rng.uniform(10, 100, 4)
    f = lambda z: oracle.position_restraints(z, [0, 2, 3, 5], refs, kf, 0.2, box=[4.0, 4.5, 5.0])[0]
    E, F = oracle.position_restraints(y, [0, 2, 3, 5], refs, kf, 0.2, box=[4.0, 4.5, 5.0])
    assert np.allclose(F, -_num_grad(f, y), rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("threads", [1, 4])
def test_multithreaded_cpu_port_agrees_with_oracle(threads):
    """The multi-threaded CPU baseline (FP32 pair math) reproduces the FP64
    oracle to FP32 rounding on the 3,000-atom workload."""
    s = W.config1()
    ref, Fref = oracle.atm_eval(s)
    got, F = oracle.cpu_mt_atm_eval(s, threads)
    for k in ("U1", "U2", "W", "usc"):
        assert abs(got[k] - ref[k]) <= 1e-6 * abs(ref[k]), (k, got[k], ref[k])
    assert abs(got["npairs1"] - ref["npairs1"]) <= 4 and abs(got["npairs2"] - ref["npairs2"]) <= 4
    assert np.linalg.norm(F - Fref) / np.linalg.norm(Fref) < 1e-5


# ---------------------------------------------------------------- exceptions

def _ex(pairs, qq=None, sig=None, eps=None):
    pairs = np.asarray(pairs, dtype=np.int32).reshape(-1, 2)
    m = pairs.shape[0]
    z = np.zeros(m)
    return W.Exceptions(pairs, z if qq is None else np.asarray(qq, float), z if sig is None else np.asarray(sig, float),
                        z if eps is None else np.asarray(eps, float))


def _mol(n=300, box=2.8, seed=21, **kw):
    s = W.tiny_system(n, box, 0.9, 6, (0.7, 0.5, 0.3), seed)
    return W.with_molecules(s, **kw)


def test_exception_pairs_leave_the_pair_set():
    """Pair sets with exceptions: brute force = cell list = the plain pair set
    minus the listed pairs (set difference in numpy)."""
    s = _mol(size=4, crosslink_every=7)
    ex = s.exceptions
    keys = {(min(i, j), max(i, j)) for i, j in ex.pairs}
    full = oracle.pairs(s.positions, s.box, s.cutoff)
    want = np.array([p for p in map(tuple, full) if p not in keys], dtype=np.int32)
    assert len(want) < len(full)  # some listed pairs are inside the cutoff
    got = oracle.pairs(s.positions, s.box, s.cutoff, exceptions=ex)
    got_b = oracle.pairs(s.positions, s.box, s.cutoff, brute=True, exceptions=ex)
    assert np.array_equal(got, want) and np.array_equal(got_b, want)
    # the crosslinks reach beyond a 31-atom index window
    assert (np.abs(ex.pairs[:, 1] - ex.pairs[:, 0]) > 31).any()


def test_exclusions_remove_exactly_their_pair_terms():
    """E and F with exclusions only = E and F without exceptions minus, for every
    excluded pair inside the cutoff, the energy and forces of that pair alone (a
    two-atom evaluation at the same separation)."""
    s = _mol(size=3, crosslink_every=9)
    ex = s.exceptions
    excl = ex.charge_prod == 0
    exo = _ex(ex.pairs[excl])
    args = (s.positions, s.charges, s.sigmas, s.epsilons, s.box, s.cutoff, s.rf_dielectric)
    E0, F0, n0, _ = oracle.nonbonded(*args)
    E1, F1, n1, _ = oracle.nonbonded(*args, exceptions=exo)
    dE, dF, inside = 0.0, np.zeros_like(F0), 0
    for i, j in exo.pairs:
        d = s.positions[j] - s.positions[i]
        d -= s.box * np.round(d / s.box)
        if np.dot(d, d) >= s.cutoff ** 2:
            continue
        inside += 1
        x2 = np.array([[1.0, 1.0, 1.0], 1.0 + d])
        e, f, _, _ = oracle.nonbonded(x2, s.charges[[i, j]], s.sigmas[[i, j]], s.epsilons[[i, j]], s.box, s.cutoff,
                                      s.rf_dielectric)
        dE += e
        dF[i] += f[0]
        dF[j] += f[1]
    assert inside > 50
    assert n0 - n1 == inside
    assert E1 == pytest.approx(E0 - dE, rel=1e-11, abs=1e-9)
    assert np.allclose(F1, F0 - dF, rtol=1e-9, atol=1e-7)


def test_exception_term_special_values():
    """One exception pair, in and beyond the cutoff: at r = sigma' the LJ part
    vanishes (E = k_e q'/r); at r = 2^(1/6) sigma' with q' = 0 it is the minimum
    -eps' with zero force; the reaction-field pair term is gone either way."""
    ke = 138.935456
    for r in (0.35, 1.5):                          # inside and beyond rc = 1
        x = np.array([[1.0, 1.0, 1.0], [1.0 + r, 1.0, 1.0]])
        q, sg, ep = [0.8, -0.6], [0.3, 0.3], [0.5, 0.5]
        E, F, n, parts = oracle.nonbonded(x, q, sg, ep, BIG, 1.0, 78.5, exceptions=_ex([[0, 1]], [0.25], [r], [0.7]))
        assert n == 0
        assert E == pytest.approx(ke * 0.25 / r, rel=1e-13)
        assert parts[3] == pytest.approx(E, rel=1e-15)
        rmin = 2 ** (1 / 6) * 0.31
        x = np.array([[1.0, 1.0, 1.0], [1.0 + rmin, 1.0, 1.0]])
        E, F, n, _ = oracle.nonbonded(x, q, sg, ep, BIG, 1.0, 78.5, exceptions=_ex([[1, 0]], [0.0], [0.31], [0.7]))
        assert E == pytest.approx(-0.7, rel=1e-12)
        assert np.abs(F).max() < 1e-9
    # an exclusion inside the cutoff leaves nothing at all
    x = np.array([[1.0, 1.0, 1.0], [1.3, 1.0, 1.0]])
    E, F, n, _ = oracle.nonbonded(x, q, sg, ep, BIG, 1.0, 78.5, exceptions=_ex([[0, 1]]))
    assert n == 0 and E == 0.0 and not F.any()


def test_exception_term_uses_the_plain_distance():
    """Atoms 0.4 nm apart through the periodic boundary but L - 0.4 apart in the
    given coordinates: the exception term uses L - 0.4 (no periodic image)."""
    ke = 138.935456
    L = 3.0
    x = np.array([[0.2, 1.0, 1.0], [L - 0.2, 1.0, 1.0]])
    E, _, n, _ = oracle.nonbonded(x, [0.0, 0.0], [0.3, 0.3], [0.0, 0.0], np.array([L, L, L]), 1.0, 78.5,
                                  exceptions=_ex([[0, 1]], [0.5], [0.0], [0.0]))
    assert n == 0
    assert E == pytest.approx(ke * 0.5 / (L - 0.4), rel=1e-13)


def test_forces_with_exceptions_are_minus_gradient():
    s = _mol(size=4, crosslink_every=5, seed=23)
    ex = s.exceptions
    x = s.positions.copy()
    args = (s.charges, s.sigmas, s.epsilons, s.box, s.cutoff, s.rf_dielectric)
    E0, F, _, parts = oracle.nonbonded(x, *args, exceptions=ex)
    assert parts[3] != 0.0
    rng = np.random.default_rng(5)
    h = 1e-6
    scaled_atoms = np.unique(ex.pairs[ex.charge_prod != 0])
    for a in list(rng.choice(len(x), 3, replace=False)) + list(scaled_atoms[:3]):
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[a, k] += h
            xm[a, k] -= h
            Ep = oracle.nonbonded(xp, *args, forces=False, exceptions=ex)[0]
            Em = oracle.nonbonded(xm, *args, forces=False, exceptions=ex)[0]
            assert -(Ep - Em) / (2 * h) == pytest.approx(F[a, k], rel=1e-5, abs=1e-4)
    assert np.abs(F.sum(0)).max() < 1e-8 * np.abs(F).sum()


def test_atm_gradient_with_exceptions_against_central_differences():
    import dataclasses
    p = W.ATMParams(0.15, 0.45, 0.08, 30.0, 1.0, 200.0, 100.0, 0.0625, 1.0)
    s = W.with_molecules(W.tiny_system(250, 3.2, 0.9, 8, (1.1, 0.7, 0.45), 5, p), size=4, crosslink_every=6)
    r, _ = oracle.atm_eval(s, forces=False)
    r0, _ = oracle.atm_eval(dataclasses.replace(s, exceptions=None), forces=False)
    assert r["U1"] != r0["U1"] and r["npairs1"] < r0["npairs1"]
    g = oracle.atm_grad(s)
    for field, key, h in (("charges", "dq", 1e-6), ("sigmas", "dsigma", 1e-7), ("epsilons", "deps", 1e-7)):
        for i in (0, 3, 77, 8):
            wp = oracle.atm_eval(_perturbed(s, field, i, h), forces=False)[0]["W"]
            wm = oracle.atm_eval(_perturbed(s, field, i, -h), forces=False)[0]["W"]
            assert g[key][i] == pytest.approx((wp - wm) / (2 * h), rel=1e-5, abs=1e-6), (field, i)
