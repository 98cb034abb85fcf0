"""The PME variable group and the bonded terms on the MI355X
(csrc/atm_pme.hip) against the FP64 oracle (oracle/pme_oracle.py), and the
reference's end-to-end goldens on temoa-g1 (python/tests/test_abfe.py:147-150)
computed through the C ABI."""
from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest

from atmmetaforce import _capi as C, pme
from oracle import oracle, pme_oracle as po

import temoa

pytestmark = pytest.mark.gpu

E_REL = 1e-9      # FP64 sums in another order; charge grid in 2^-36 e fixed point
F_ABS = 2e-5      # forces summed in 2^-24 kJ/mol/nm fixed point (hundreds of terms per atom)


def _small_system(seed=5, n=400, L=(3.3, 3.4, 3.6)):
    rng = np.random.default_rng(seed)
    L = np.array(L)
    g = int(math.ceil(n ** (1 / 3)))
    lat = np.array([(i, j, k) for i in range(g) for j in range(g) for k in range(g)][:n], dtype=float)
    x = (lat + 0.5) * (L / g) + rng.uniform(-0.04, 0.04, (n, 3))
    x += rng.integers(-1, 2, (n, 3)) * L          # some atoms outside the box: wrapping is the kernel's job
    q = rng.normal(0, 0.5, n)
    q -= q.mean()
    sig = rng.uniform(0.25, 0.35, n)
    eps = rng.uniform(0.1, 0.6, n)
    eps[::17] = 0.0
    # chains 0-1-2-3-4 and 10-11-12 plus a triangle: exclusions and scaled 1-4 pairs
    pairs = np.array([(0, 1), (1, 2), (2, 3), (3, 4), (0, 2), (1, 3), (2, 4), (0, 3), (1, 4),
                      (10, 11), (11, 12), (10, 12), (30, 31), (31, 32), (30, 32)])
    scaled = {(0, 3), (1, 4)}
    qq = np.array([q[i] * q[j] / 1.2 if (i, j) in scaled else 0.0 for i, j in pairs])
    es = 0.5 * (sig[pairs[:, 0]] + sig[pairs[:, 1]])
    ee = np.array([math.sqrt(eps[i] * eps[j]) / 2 if (i, j) in scaled else 0.0 for i, j in pairs])
    # keep the excepted pairs bonded-close in plain coordinates
    for i, j in pairs:
        x[j] = x[i] + (x[j] - x[i]) - L * np.round((x[j] - x[i]) / L)
    d = np.zeros((n, 3))
    lig = np.arange(40, 60)
    d[lig] = np.array([1.1, 1.2, 1.3])
    return x, q, sig, eps, L, pairs, qq, es, ee, d


PARAMS = dict(lambda1=0.2, lambda2=0.35, alpha=0.05, u0=40.0, w0=1.5, umax=800.0, ubcore=400.0, acore=0.0625,
              direction=1.0)


def _oracle_atm(x, q, sig, eps, L, pairs, qq, es, ee, d, params, cutoff=1.0, **kw):
    U1, F1, p1 = po.nonbonded_pme(x, q, sig, eps, L, cutoff, pairs, qq, es, ee, **kw)
    U2, F2, p2 = po.nonbonded_pme(x + d, q, sig, eps, L, cutoff, pairs, qq, es, ee, **kw)
    r = oracle.combine(params, U1, U2)
    a = r["sp"] if params["direction"] > 0 else 1.0 - r["sp"]
    return r, (1 - a) * F1 + a * F2, p1, p2


@pytest.mark.parametrize("direction", [1.0, -1.0])
def test_pme_small_system_against_oracle(gpu, direction):
    x, q, sig, eps, L, pairs, qq, es, ee, d = _small_system()
    params = dict(PARAMS, direction=direction)
    print(f"[pme] small system n={len(q)} box={L} direction={direction}", flush=True)
    with pme.ATMPMEForce(q, sig, eps, d, L, 1.0, exceptions=(pairs, qq, es, ee), params=params) as f:
        alpha, grid = f.info()
        r, F, parts = f.evaluate(x)
    ref, Fref, p1, p2 = _oracle_atm(x, q, sig, eps, L, pairs, qq, es, ee, d, params)
    assert grid == p1["grid"] and abs(alpha - p1["alpha"]) < 1e-15
    for got, want in ((r["state1_energy"], ref["U1"]), (r["state2_energy"], ref["U2"]), (r["energy"], ref["W"]),
                      (r["perturbation_energy"], ref["usc"])):
        assert abs(got - want) <= E_REL * max(1.0, abs(want)), (got, want)
    for s, p in ((1, p1), (2, p2)):
        assert abs(parts[f"direct{s}"] - (p["direct_coulomb"] + p["direct_lj"])) < 1e-8 * abs(p["direct_coulomb"])
        assert abs(parts[f"exceptions{s}"] - (p["exceptions"] + p["exclusion_correction"])) < 1e-9 * 1e3
        assert abs(parts[f"reciprocal{s}"] - p["reciprocal"]) < 1e-8 * abs(p["reciprocal"])
        assert abs(parts[f"self{s}"] - p["self"]) < 1e-10 * abs(p["self"])
        assert abs(parts[f"dispersion{s}"] - p["dispersion"]) < 1e-10 * abs(p["dispersion"])
    assert np.abs(F - Fref).max() < F_ABS, np.abs(F - Fref).max()


def test_pme_is_bit_reproducible(gpu):
    x, q, sig, eps, L, pairs, qq, es, ee, d = _small_system(seed=9)
    with pme.ATMPMEForce(q, sig, eps, d, L, 1.0, exceptions=(pairs, qq, es, ee), params=PARAMS) as f:
        r1, F1, _ = f.evaluate(x)
        r2, F2, _ = f.evaluate(x)
    assert r1 == r2 and np.array_equal(F1, F2)


def test_pme_rejects_bad_inputs(gpu):
    x, q, sig, eps, L, pairs, qq, es, ee, d = _small_system()
    with pme.ATMPMEForce(q, sig, eps, d, L, 1.0, exceptions=(pairs, qq, es, ee), params=PARAMS) as f:
        bad = x.copy()
        bad[3, 1] = np.nan
        with pytest.raises(C.ATMError, match="non-finite"):
            f.evaluate(bad)
        with pytest.raises(C.ATMError, match="three cutoff"):
            f.evaluate(x, box=np.array([2.5, 3.4, 3.6]))
        with pytest.raises(C.ATMError, match="duplicate"):
            f.set_exceptions(np.array([(0, 1), (1, 0)]), [0.0, 0.0], [0.1, 0.1], [0.0, 0.0])
        r, _, _ = f.evaluate(x)     # still usable
        assert math.isfinite(r["energy"])


def test_bonded_against_oracle(gpu):
    s, box, x, _ = temoa.load()
    b = pme.BondedForces(s.num_atoms, s.bonds, s.bond_params, s.angles, s.angle_params, s.torsions,
                         s.torsion_params)
    E, F = b.evaluate(x)
    b.close()
    Eref, Fref = po.bonded(x, s.bonds, s.bond_params, s.angles, s.angle_params, s.torsions, s.torsion_params)
    for k in ("bond", "angle", "torsion"):
        assert abs(E[k] - Eref[k]) < 1e-10 * max(1.0, abs(Eref[k])), (k, E[k], Eref[k])
    assert np.abs(F - Fref).max() < 1e-5


def _restraints_energy(s, x, ref40, box):
    L = C.load()
    rcpt, lig, posr = temoa.restraint_setup(s)
    members = np.ascontiguousarray(np.concatenate([lig, rcpt]).astype(np.int32))
    w = np.ascontiguousarray(np.concatenate([s.masses[lig], s.masses[rcpt]]))
    cent = (C.CentroidRestraint * 1)()
    cent[0] = C.CentroidRestraint(0, len(lig), len(lig), len(rcpt), temoa.RESTRAINT_K, temoa.CM_TOL,
                                  (ctypes.c_double * 3)(0.0, 0.0, 0.0))
    pr = (C.PositionRestraint * len(posr))()
    for k, a in enumerate(posr):
        pr[k] = C.PositionRestraint(int(a), 1, (ctypes.c_double * 3)(*ref40[a]), temoa.RESTRAINT_K, temoa.POS_TOL)
    h = ctypes.c_void_p()
    C.check(L.atm_restraints_create(0, ctypes.byref(h)))
    try:
        C.check(L.atm_restraints_set(h, s.num_atoms, len(members), members.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                     w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1, cent, len(posr), pr))
        e = ctypes.c_double()
        xx = np.ascontiguousarray(x)
        bb = np.ascontiguousarray(box)
        C.check(L.atm_restraints_eval(h, xx.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      bb.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), None, ctypes.byref(e)))
    finally:
        L.atm_restraints_destroy(h)
    return e.value


def test_temoa_goldens_on_gpu(gpu):
    """test_abfe.py:147-150 on the device: potential energy of groups
    {0, ATM} = bonded + restraints + W, and the perturbation energy."""
    s, box, x, ref40 = temoa.load()
    rcpt, lig, posr = temoa.restraint_setup(s)
    d = np.zeros((s.num_atoms, 3))
    d[lig] = temoa.DISPLACEMENT
    print(f"[pme] temoa-g1 n={s.num_atoms} box={box}", flush=True)
    f, b = pme.from_amber(s, d, box, params=temoa.ATM_PARAMS, cutoff=temoa.CUTOFF, ewald_tolerance=temoa.EWALD_TOL)
    try:
        alpha, grid = f.info()
        r, F, parts = f.evaluate(x, box)
        Eb, Fb = b.evaluate(x)
    finally:
        f.close()
        b.close()
    assert grid == (35, 40, 35)
    pot = sum(Eb.values()) + _restraints_energy(s, x, ref40, box) + r["energy"]
    assert abs(pot - temoa.GOLDEN_POTENTIAL) <= temoa.GOLDEN_ATOL, pot
    assert abs(r["perturbation_energy"] - temoa.GOLDEN_PERTURBATION) <= temoa.GOLDEN_ATOL, r["perturbation_energy"]
    # and against the oracle, tightly
    ref, Fref, p1, p2 = _oracle_atm(x, s.charges, s.sigmas, s.epsilons, box, s.exception_pairs, s.exception_qq,
                                    s.exception_sigma, s.exception_eps, d, temoa.ATM_PARAMS, tol=temoa.EWALD_TOL)
    assert abs(r["state1_energy"] - ref["U1"]) <= E_REL * abs(ref["U1"])
    assert abs(r["state2_energy"] - ref["U2"]) <= E_REL * abs(ref["U2"])
    assert abs(r["perturbation_energy"] - ref["usc"]) < 1e-6
    assert np.abs(F - Fref).max() < 1e-4, np.abs(F - Fref).max()


def test_pme_parameter_and_box_changes_between_evaluations(gpu):
    """The evaluation is a captured graph: new ATM parameters and a new box
    must take effect on the next call (parameters through device memory, box
    by a new capture)."""
    x, q, sig, eps, L, pairs, qq, es, ee, d = _small_system(seed=7)
    pa = dict(PARAMS, lambda1=0.1, lambda2=0.2, direction=1.0)
    pb = dict(PARAMS, lambda1=0.4, lambda2=0.45, alpha=0.0, direction=-1.0)  # (no softplus: e^x overflows here)
    L2 = L * 1.01
    with pme.ATMPMEForce(q, sig, eps, d, L, 1.0, exceptions=(pairs, qq, es, ee), params=pa) as f:
        for params, box in ((pa, L), (pb, L), (pb, L2), (pa, L)):
            f.set_params(params)
            r, F, _ = f.evaluate(x, box=box)
            ref, Fref, _, _ = _oracle_atm(x, q, sig, eps, box, pairs, qq, es, ee, d, params,
                                          alpha=f.info()[0], grid=f.info()[1])
            assert abs(r["energy"] - ref["W"]) <= E_REL * abs(ref["W"]), (params, box, r["energy"], ref["W"])
            assert np.abs(F - Fref).max() < F_ABS


def test_pme_dense_cell_takes_the_capacity_retry(gpu):
    """Most atoms packed into one cell exceed the first per-cell capacity
    (3x the mean + 64); the evaluation detects it, widens the capacity and
    returns the right result."""
    rng = np.random.default_rng(21)
    n, L = 400, np.array([3.3, 3.3, 3.3])
    x = rng.uniform(0.0, L, (n, 3))
    g = int(math.ceil(250 ** (1 / 3)))
    lat = np.array([(i, j, k) for i in range(g) for j in range(g) for k in range(g)][:250], dtype=float)
    x[:250] = 0.1 + lat * (0.9 / g)          # 250 atoms in the first 1.1 nm cell, >= 0.14 nm apart
    q = rng.normal(0, 0.3, n)
    q -= q.mean()
    sig = np.full(n, 0.12)
    eps = np.full(n, 0.2)
    pairs = np.zeros((0, 2), dtype=np.int64)
    d = np.zeros((n, 3))
    d[300:310] = 0.5
    with pme.ATMPMEForce(q, sig, eps, d, L, 1.0, params=PARAMS) as f:
        r, F, _ = f.evaluate(x)
    ref, Fref, _, _ = _oracle_atm(x, q, sig, eps, L, pairs, [], [], [], d, PARAMS)
    assert abs(r["state1_energy"] - ref["U1"]) <= E_REL * abs(ref["U1"])
    assert np.abs(F - Fref).max() < F_ABS
