"""Pinning the oracle's triclinic boxes (oracle_pairs_tri, oracle_atm_eval_tri).

The reference hands the box vectors to its inner contexts
(platforms/reference/src/ReferenceATMMetaForceKernels.cpp:126-129), whose
NonbondedForce takes a pair's image by OpenMM's triclinic rule.  Checks that do
not retype that rule:
  * zero tilts reduce the triclinic loop to the pinned orthorhombic oracle;
  * supercells: a triclinic cell whose one tilt is half an edge tiles an
    orthorhombic box twice its size along one axis, so the orthorhombic oracle on
    the two copies gives twice the energy and the same forces;
  * for cutoffs below half the cell's smallest width, the rule's image is the
    nearest one: pairs and distances equal a brute-force minimum over 5^3 lattice
    images;
  * periodicity: moving atoms by lattice vectors changes nothing.
"""
import ctypes

import numpy as np
import pytest

from atmmetaforce import workloads as W
from oracle import oracle

GRID = W.GRID


def random_cell(n, box, tilt, seed):
    """n grid-rounded atoms spread over the triclinic cell (fractional coordinates
    uniform), with LJ + charge parameters."""
    rng = np.random.default_rng(seed)
    H = W.box_vectors(box, tilt)
    f = rng.uniform(0.0, 1.0, size=(n, 3))
    x = W.quantize(f @ H)
    q = rng.normal(0.0, 0.4, size=n)
    q -= q.mean()
    sig = rng.uniform(0.25, 0.35, size=n)
    eps = rng.uniform(0.1, 0.8, size=n)
    return x, q, sig, eps


def system_of(x, q, sig, eps, box, tilt, rc, d=None):
    n = x.shape[0]
    disp = np.zeros((n, 3)) if d is None else d
    return W.System(positions=x, charges=q, sigmas=sig, epsilons=eps, displacements=disp,
                    box=np.asarray(box, dtype=np.float64), cutoff=rc, tilt=None if tilt is None else np.asarray(tilt))


def test_zero_tilt_triclinic_loop_equals_orthorhombic_oracle():
    x, q, sig, eps = random_cell(400, [3.0, 3.2, 2.9], None, 3)
    box = np.array([3.0, 3.2, 2.9])
    L = oracle.lib()
    xp = np.ascontiguousarray(x)
    dp = ctypes.POINTER(ctypes.c_double)
    z = np.zeros(3)
    ip = ctypes.POINTER(ctypes.c_int)
    n_tri = L.oracle_pairs_tri(400, xp.ctypes.data_as(dp), box.ctypes.data_as(dp), z.ctypes.data_as(dp), 1.0,
                               0, None, None, 0)
    out = np.zeros((n_tri, 2), dtype=np.int32)
    L.oracle_pairs_tri(400, xp.ctypes.data_as(dp), box.ctypes.data_as(dp), z.ctypes.data_as(dp), 1.0, 0, None,
                       out.ctypes.data_as(ip), n_tri)
    ref = oracle.pairs(x, box, 1.0, brute=True)
    assert sorted(map(tuple, out.tolist())) == sorted(map(tuple, ref.tolist()))


@pytest.mark.parametrize("axis", ["b_x", "c_x", "c_y"])
def test_half_tilt_supercell_doubles_the_energy(axis):
    """a = (L,0,0), b = (L/2, L, 0): 2b - a = (0, 2L, 0), so the orthorhombic box
    (L, 2L, L) holding r and r + b is a supercell of two triclinic cells; likewise
    c = (L/2, 0, L) with (L, L, 2L) holding r and r + c (2c - a = (0, 0, 2L)), and
    c = (0, L/2, L) with 2c - b = (0, 0, 2L)."""
    Lb = 3.0
    box = np.array([Lb, Lb, Lb])
    tilt = {"b_x": [Lb / 2, 0.0, 0.0], "c_x": [0.0, Lb / 2, 0.0], "c_y": [0.0, 0.0, Lb / 2]}[axis]
    H = W.box_vectors(box, tilt)
    shift = H[1] if axis == "b_x" else H[2]
    sbox = box.copy()
    sbox[1 if axis == "b_x" else 2] *= 2.0
    n = 300
    x, q, sig, eps = random_cell(n, box, tilt, 11)
    rc = 0.9  # below half of every width of the cell (>= 2.68 nm / 2) and of the supercell
    n_ortho = oracle.nonbonded(x, q, sig, eps, box, rc, 78.5, forces=False)[2]  # the same atoms, no tilt
    # triclinic oracle through the ATM evaluation with d = 0 (U1 = U2 = the energy)
    r_tri, _ = oracle.atm_eval(system_of(x, q, sig, eps, box, tilt, rc), forces=False)
    x2 = np.concatenate([x, W.quantize(x + shift)])
    cat = lambda a: np.concatenate([a, a])
    E_sup, F_sup, n_sup, _ = oracle.nonbonded(x2, cat(q), cat(sig), cat(eps), sbox, rc, 78.5, brute=True)
    assert n_sup == 2 * r_tri["npairs1"]
    assert E_sup == pytest.approx(2.0 * r_tri["U1"], rel=1e-11, abs=1e-9)
    # forces: the triclinic evaluation's forces at x equal the supercell's on either copy
    sys1 = system_of(x, q, sig, eps, box, tilt, rc)
    sys1.displacements = np.zeros((n, 3))
    sys1.params = W.ATMParams(lambda1=0.0, lambda2=0.0, alpha=0.1, u0=0.0, w0=0.0, direction=1.0)
    r1, F1 = oracle.atm_eval(sys1)  # lambda = 0, d = 0: W = U1, F = F1
    assert r1["W"] == pytest.approx(r_tri["U1"], rel=1e-12)
    scale = np.abs(F_sup).max()
    assert np.abs(F_sup[:n] - F1).max() <= 1e-10 * scale
    assert np.abs(F_sup[n:] - F1).max() <= 1e-10 * scale
    assert n_ortho != r_tri["npairs1"]  # the tilt matters for this cell


def nearest_image_pairs(x, box, tilt, rc):
    H = W.box_vectors(box, tilt)
    n = x.shape[0]
    rng = np.arange(-2, 3)
    shifts = np.array([i * H[0] + j * H[1] + k * H[2] for i in rng for j in rng for k in rng])
    out = {}
    for i in range(n):
        d = x[i] - x[i + 1:]                                  # (m, 3)
        dd = d[:, None, :] - shifts[None, :, :]               # every image of j
        r2 = np.min(np.einsum("mik,mik->mi", dd, dd), axis=1)
        for j in np.nonzero(r2 < rc * rc)[0]:
            out[(i, i + 1 + j)] = r2[j]
    return out


@pytest.mark.parametrize("seed,tilt", [(5, [1.2, -1.1, 0.9]), (6, [-1.5, 1.5, -1.4]), (7, [0.3, 0.0, 1.45])])
def test_rule_picks_the_nearest_image_below_half_the_width(seed, tilt):
    box = np.array([3.0, 2.9, 3.1])
    tilt = W.quantize(tilt)
    x, q, sig, eps = random_cell(250, box, tilt, seed)
    rc = 1.0
    got = oracle.pairs(x, box, rc, tilt=tilt)
    ref = nearest_image_pairs(x, box, tilt, rc)
    assert sorted(map(tuple, got.tolist())) == sorted(ref)
    # and the energy is the one of those distances: compare the pair count through the evaluation
    r, _ = oracle.atm_eval(system_of(x, q, sig, eps, box, tilt, rc), forces=False)
    assert r["npairs1"] == len(ref)


def test_lattice_translations_change_nothing():
    # (no exceptions: their terms use the plain distance, which moving one atom of a pair changes)
    s = W.triclinic(W.tiny_system(1000, 2.6, 0.9, 12, (0.8, 0.8, 0.8), seed=4), sites=(3, -2, 4))
    H = W.box_vectors(s.box, s.tilt)
    r0, F0 = oracle.atm_eval(s)
    rng = np.random.default_rng(9)
    k = rng.integers(-2, 3, size=(s.num_atoms, 3))
    moved = W.quantize(s.positions + k @ H)
    s2 = W.System(**{**s.__dict__, "positions": moved})
    r1, F1 = oracle.atm_eval(s2)
    for key in ("U1", "U2", "W"):
        assert r1[key] == pytest.approx(r0[key], rel=1e-12)
    assert r1["npairs1"] == r0["npairs1"] and r1["npairs2"] == r0["npairs2"]
    assert np.abs(F1 - F0).max() <= 1e-9 * np.abs(F0).max()
    # a wrong image (the orthorhombic one) does see the move
    ortho = W.System(**{**s2.__dict__, "tilt": None})
    r2, _ = oracle.atm_eval(ortho)
    assert r2["U1"] != pytest.approx(r0["U1"], rel=1e-6)
