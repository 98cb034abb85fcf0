"""FP64 numpy restatement of the PME / AMBER energy path behind the
temoa-g1 binding-energy goldens (python/tests/test_abfe.py:147-150).

TEST INFRASTRUCTURE ONLY: imported by tests/ and by bench.py's cpu_baseline
leg, never by the product package.  It is the checker of the device kernels in
csrc/atm_pme.hip and csrc/atm_bonded.hip.

What it restates.  test_abfe.py:42-43 builds the system with OpenMM's
AmberPrmtopFile.createSystem(nonbondedMethod=PME, nonbondedCutoff=1 nm,
constraints=HBonds); the ATM force's variable group is that NonbondedForce
(ATMMetaForceImpl.cpp:58-63 clones it into the inner contexts), and group 0
holds the bonded terms and restraints.  OpenMM itself is not vendored (OpenMM
7.6 per example/abfe/temoa-g1-equil.xml:2); these are its documented models:

  nonbonded_pme   NonbondedForce with PME:
                  alpha = sqrt(-ln(2 tol)) / rc; grid K_k = ceil(2 alpha L_k /
                  (3 tol^(1/5))) raised to the next 2,3,5,7-smooth size (at
                  least 6); smooth PME (Essmann et al., J. Chem. Phys. 103,
                  8577 (1995)) with order-5 cardinal B-splines and the Euler
                  exponential spline moduli (a modulus below 1e-7 replaced by
                  the mean of its neighbours);
                  direct space k_e q_i q_j erfc(alpha r)/r + 4 eps_ij
                  [(s_ij/r)^12 - (s_ij/r)^6] over minimum-image pairs with
                  r < rc that are not exceptions (Lorentz-Berthelot combining);
                  every exception pair is taken out of the direct sum, adds its
                  own k_e qq'/r + 4 eps' [(s'/r)^12 - (s'/r)^6] (plain distance,
                  no periodic image) and subtracts the reciprocal-space part
                  k_e q_i q_j erf(alpha r)/r; self energy -k_e alpha/sqrt(pi)
                  sum q^2; dispersion correction 8 pi N^2 (<eps s^12>/(9 rc^9) -
                  <eps s^6>/(3 rc^3)) / V averaged over the N(N+1)/2 pairs of
                  (sigma, epsilon) classes.
  bonded          HarmonicBondForce 0.5 k (r - r0)^2, HarmonicAngleForce
                  0.5 k (theta - theta0)^2, PeriodicTorsionForce
                  k (1 + cos(n phi - phase)) (IUPAC dihedral sign).

How it is pinned (tests/test_pme_oracle.py): the two goldens of
test_abfe.py:147-150 on the shipped temoa-g1 fixtures (through the package's
AMBER reader), the Madelung constant of rock salt, the independence of the
total Ewald energy from alpha, central differences of every energy, and
Newton's third law.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.spatial import cKDTree
from scipy.special import erf, erfc

KE = 138.935456  # kJ mol^-1 nm e^-2 (OpenMM's ONE_4PI_EPS0)
ORDER = 5


def smooth_size(n: int) -> int:
    """Smallest m >= n whose prime factors are all in {2, 3, 5, 7}."""
    while True:
        m = n
        for f in (2, 3, 5, 7):
            while m % f == 0:
                m //= f
        if m == 1:
            return n
        n += 1


def pme_parameters(box, cutoff, tol=5e-4):
    """(alpha, (K1, K2, K3)) for a rectangular box."""
    alpha = math.sqrt(-math.log(2.0 * tol)) / cutoff
    grid = tuple(smooth_size(max(6, int(math.ceil(2.0 * alpha * L / (3.0 * tol ** 0.2))))) for L in box)
    return alpha, grid


def bspline(t, n=ORDER):
    """Cardinal B-spline M_n(t) (support [0, n]) by the two-term recursion."""
    t = np.asarray(t, dtype=np.float64)
    if n == 2:
        return np.where((t >= 0.0) & (t <= 2.0), 1.0 - np.abs(t - 1.0), 0.0)
    return t / (n - 1) * bspline(t, n - 1) + (n - t) / (n - 1) * bspline(t - 1.0, n - 1)


def bspline_moduli(K, n=ORDER):
    """|sum_{k=0}^{n-2} M_n(k+1) exp(2 pi i m k / K)|^2 for m = 0..K-1."""
    w = np.array([float(bspline(k + 1.0, n)) for k in range(n - 1)])
    m = np.arange(K)[:, None] * np.arange(n - 1)[None, :]
    ang = 2.0 * np.pi * m / K
    mod = (w * np.cos(ang)).sum(1) ** 2 + (w * np.sin(ang)).sum(1) ** 2
    for i in range(K):
        if mod[i] < 1e-7:
            mod[i] = 0.5 * (mod[(i - 1) % K] + mod[(i + 1) % K])
    return mod


def reciprocal(x, q, box, alpha, grid, forces=True):
    """Smooth-PME reciprocal energy and forces (rectangular box)."""
    L = np.asarray(box, dtype=np.float64)
    K = np.asarray(grid)
    u = (np.asarray(x) / L) % 1.0 * K
    g0 = np.floor(u).astype(np.int64)
    w = u - g0
    th = [np.stack([bspline(w[:, d] + j) for j in range(ORDER)], 1) for d in range(3)]
    dth = [np.stack([bspline(w[:, d] + j, ORDER - 1) - bspline(w[:, d] + j - 1.0, ORDER - 1) for j in range(ORDER)], 1)
           for d in range(3)]
    idx = [(g0[:, d][:, None] - np.arange(ORDER)[None, :]) % K[d] for d in range(3)]
    Q = np.zeros(tuple(K))
    for a in range(ORDER):
        for b in range(ORDER):
            wab = q * th[0][:, a] * th[1][:, b]
            for c in range(ORDER):
                np.add.at(Q, (idx[0][:, a], idx[1][:, b], idx[2][:, c]), wab * th[2][:, c])
    F = np.fft.fftn(Q)
    m = [np.fft.fftfreq(k, d=1.0 / k) / Lk for k, Lk in zip(K, L)]
    mx, my, mz = np.meshgrid(*m, indexing="ij")
    m2 = mx * mx + my * my + mz * mz
    mods = [bspline_moduli(k) for k in K]
    Bm = mods[0][:, None, None] * mods[1][None, :, None] * mods[2][None, None, :]
    V = float(np.prod(L))
    with np.errstate(divide="ignore", invalid="ignore"):
        wm = np.exp(-np.pi ** 2 * m2 / alpha ** 2) / (m2 * Bm)
    wm[0, 0, 0] = 0.0
    E = KE / (2.0 * np.pi * V) * float(np.sum(wm * np.abs(F) ** 2))
    if not forces:
        return E, None
    # dE/dQ(k) = phi(k) = k_e/(pi V) Re sum_m w(m) F(m) exp(+2 pi i m.k/K)
    phi = KE / (np.pi * V) * np.real(np.fft.ifftn(wm * F)) * float(np.prod(K))
    Fx = np.zeros((len(q), 3))
    for a in range(ORDER):
        for b in range(ORDER):
            for c in range(ORDER):
                p = phi[idx[0][:, a], idx[1][:, b], idx[2][:, c]]
                Fx[:, 0] -= K[0] / L[0] * dth[0][:, a] * th[1][:, b] * th[2][:, c] * p
                Fx[:, 1] -= K[1] / L[1] * th[0][:, a] * dth[1][:, b] * th[2][:, c] * p
                Fx[:, 2] -= K[2] / L[2] * th[0][:, a] * th[1][:, b] * dth[2][:, c] * p
    return E, Fx * q[:, None]


def _pair_codes(i, j, n):
    lo, hi = np.minimum(i, j), np.maximum(i, j)
    return lo.astype(np.int64) * n + hi


def direct(x, q, sig, eps, box, cutoff, alpha, ex_pairs, forces=True):
    """Direct-space erfc Coulomb + LJ over non-exception minimum-image pairs with r < rc."""
    L = np.asarray(box, dtype=np.float64)
    n = len(q)
    xw = np.asarray(x) % L
    pr = cKDTree(xw, boxsize=L).query_pairs(cutoff, output_type="ndarray")
    if len(ex_pairs):
        keep = ~np.isin(_pair_codes(pr[:, 0], pr[:, 1], n), _pair_codes(ex_pairs[:, 0], ex_pairs[:, 1], n))
        pr = pr[keep]
    i, j = pr[:, 0], pr[:, 1]
    d = xw[i] - xw[j]
    d -= L * np.round(d / L)
    r2 = (d * d).sum(1)
    sel = r2 < cutoff * cutoff
    i, j, d, r2 = i[sel], j[sel], d[sel], r2[sel]
    r = np.sqrt(r2)
    qq = KE * q[i] * q[j]
    s = 0.5 * (sig[i] + sig[j])
    e = np.sqrt(eps[i] * eps[j])
    sr6 = (s * s / r2) ** 3
    ec = qq * erfc(alpha * r) / r
    elj = 4.0 * e * (sr6 * sr6 - sr6)
    E_c, E_lj = float(ec.sum()), float(elj.sum())
    Fx = None
    if forces:
        # -dE/dr / r
        fr = (ec + qq * 2.0 * alpha / math.sqrt(math.pi) * np.exp(-(alpha * r) ** 2)) / r2 \
            + 4.0 * e * (12.0 * sr6 * sr6 - 6.0 * sr6) / r2
        f = fr[:, None] * d
        Fx = np.zeros((n, 3))
        np.add.at(Fx, i, f)
        np.add.at(Fx, j, -f)
    return E_c, E_lj, len(r), Fx


def exceptions(x, q, alpha, ex_pairs, ex_qq, ex_sig, ex_eps, forces=True):
    """Exception pairs: own Coulomb + LJ (plain distance) minus k_e q_i q_j erf(alpha r)/r."""
    x = np.asarray(x)
    n = len(q)
    if len(ex_pairs) == 0:
        return 0.0, 0.0, (np.zeros((n, 3)) if forces else None)
    i, j = ex_pairs[:, 0], ex_pairs[:, 1]
    d = x[i] - x[j]
    r2 = (d * d).sum(1)
    r = np.sqrt(r2)
    qq = KE * ex_qq
    sr6 = (ex_sig * ex_sig / r2) ** 3
    e_own = qq / r + 4.0 * ex_eps * (sr6 * sr6 - sr6)
    qiqj = KE * q[i] * q[j]
    e_corr = -qiqj * erf(alpha * r) / r
    Fx = None
    if forces:
        fr = qq / (r2 * r) + 4.0 * ex_eps * (12.0 * sr6 * sr6 - 6.0 * sr6) / r2 \
            + (e_corr + qiqj * 2.0 * alpha / math.sqrt(math.pi) * np.exp(-(alpha * r) ** 2)) / r2
        f = fr[:, None] * d
        Fx = np.zeros((n, 3))
        np.add.at(Fx, i, f)
        np.add.at(Fx, j, -f)
    return float(e_own.sum()), float(e_corr.sum()), Fx


def dispersion_correction(sig, eps, box, cutoff):
    classes = {}
    for s, e in zip(np.asarray(sig, dtype=np.float64), np.asarray(eps, dtype=np.float64)):
        classes[(s, e)] = classes.get((s, e), 0) + 1
    keys = list(classes)
    s1 = s2 = 0.0
    for a in range(len(keys)):
        for b in range(a, len(keys)):
            (sa, ea), (sb, eb) = keys[a], keys[b]
            na, nb = classes[keys[a]], classes[keys[b]]
            cnt = na * (na + 1) / 2.0 if a == b else float(na * nb)
            s = 0.5 * (sa + sb)
            e = math.sqrt(ea * eb)
            s6 = s ** 6
            s1 += cnt * e * s6 * s6
            s2 += cnt * e * s6
    n = float(len(sig))
    pairs = n * (n + 1) / 2.0
    s1 /= pairs
    s2 /= pairs
    return 8.0 * n * n * math.pi * (s1 / (9.0 * cutoff ** 9) - s2 / (3.0 * cutoff ** 3)) / float(np.prod(box))


def nonbonded_pme(x, q, sig, eps, box, cutoff, ex_pairs, ex_qq, ex_sig, ex_eps, tol=5e-4, alpha=None, grid=None,
                  dispersion=True, forces=True):
    """Total NonbondedForce (PME) energy and forces; parts in a dict."""
    q = np.asarray(q, dtype=np.float64)
    sig = np.asarray(sig, dtype=np.float64)
    eps = np.asarray(eps, dtype=np.float64)
    ex_pairs = np.asarray(ex_pairs, dtype=np.int64).reshape(-1, 2)
    a0, g0 = pme_parameters(box, cutoff, tol)
    alpha = a0 if alpha is None else alpha
    grid = g0 if grid is None else tuple(grid)
    Ec, Elj, npairs, Fd = direct(x, q, sig, eps, box, cutoff, alpha, ex_pairs, forces)
    Eown, Ecorr, Fe = exceptions(x, q, alpha, ex_pairs, np.asarray(ex_qq, dtype=np.float64),
                                 np.asarray(ex_sig, dtype=np.float64), np.asarray(ex_eps, dtype=np.float64), forces)
    Erec, Fr = reciprocal(x, q, box, alpha, grid, forces)
    Eself = -KE * alpha / math.sqrt(math.pi) * float(np.sum(q * q))
    Edisp = dispersion_correction(sig, eps, box, cutoff) if dispersion else 0.0
    parts = {"direct_coulomb": Ec, "direct_lj": Elj, "exceptions": Eown, "exclusion_correction": Ecorr,
             "reciprocal": Erec, "self": Eself, "dispersion": Edisp, "pairs": npairs, "alpha": alpha, "grid": grid}
    E = Ec + Elj + Eown + Ecorr + Erec + Eself + Edisp
    F = (Fd + Fe + Fr) if forces else None
    return E, F, parts


def bonded(x, bonds=None, bond_params=None, angles=None, angle_params=None, torsions=None, torsion_params=None):
    """Energies (bond, angle, torsion) and total forces.  bond_params (r0, k);
    angle_params (theta0, k); torsion_params (periodicity, phase, k)."""
    x = np.asarray(x, dtype=np.float64)
    F = np.zeros_like(x)
    Eb = Ea = Et = 0.0
    if bonds is not None and len(bonds):
        b = np.asarray(bonds)
        r0, k = np.asarray(bond_params).T
        d = x[b[:, 0]] - x[b[:, 1]]
        r = np.linalg.norm(d, axis=1)
        Eb = float(np.sum(0.5 * k * (r - r0) ** 2))
        f = (-k * (r - r0) / r)[:, None] * d
        np.add.at(F, b[:, 0], f)
        np.add.at(F, b[:, 1], -f)
    if angles is not None and len(angles):
        a = np.asarray(angles)
        t0, k = np.asarray(angle_params).T
        v1 = x[a[:, 0]] - x[a[:, 1]]
        v2 = x[a[:, 2]] - x[a[:, 1]]
        r1, r2 = np.linalg.norm(v1, axis=1), np.linalg.norm(v2, axis=1)
        c = np.clip((v1 * v2).sum(1) / (r1 * r2), -1.0, 1.0)
        th = np.arccos(c)
        Ea = float(np.sum(0.5 * k * (th - t0) ** 2))
        dEdth = k * (th - t0)
        s = np.sqrt(np.maximum(1.0 - c * c, 1e-300))
        # dtheta/dv1 = -(v2/(r1 r2) - c v1/r1^2) / sin
        g1 = -(v2 / (r1 * r2)[:, None] - (c / r1 ** 2)[:, None] * v1) / s[:, None]
        g2 = -(v1 / (r1 * r2)[:, None] - (c / r2 ** 2)[:, None] * v2) / s[:, None]
        f1 = -dEdth[:, None] * g1
        f3 = -dEdth[:, None] * g2
        np.add.at(F, a[:, 0], f1)
        np.add.at(F, a[:, 2], f3)
        np.add.at(F, a[:, 1], -f1 - f3)
    if torsions is not None and len(torsions):
        t = np.asarray(torsions)
        per, ph, k = np.asarray(torsion_params).T
        xi, xj, xk, xl = (x[t[:, m]] for m in range(4))
        rij, rkj, rkl = xi - xj, xk - xj, xk - xl
        mv = np.cross(rij, rkj)
        nv = np.cross(rkj, rkl)
        m2, n2 = (mv * mv).sum(1), (nv * nv).sum(1)
        nrkj = np.linalg.norm(rkj, axis=1)
        cosp = np.clip((mv * nv).sum(1) / np.sqrt(m2 * n2), -1.0, 1.0)
        phi = np.arccos(cosp) * np.sign((rij * nv).sum(1) + 0.0) + 0.0
        phi = np.where((rij * nv).sum(1) == 0.0, np.arccos(cosp), phi)
        Et = float(np.sum(k * (1.0 + np.cos(per * phi - ph))))
        dVdphi = -k * per * np.sin(per * phi - ph)
        fi = (-dVdphi * nrkj / m2)[:, None] * mv
        fl = (dVdphi * nrkj / n2)[:, None] * nv
        p = (rij * rkj).sum(1) / nrkj ** 2
        qf = (rkl * rkj).sum(1) / nrkj ** 2
        sv = p[:, None] * fi - qf[:, None] * fl
        fj = fi - sv
        fk = fl + sv
        np.add.at(F, t[:, 0], fi)
        np.add.at(F, t[:, 1], -fj)
        np.add.at(F, t[:, 2], -fk)
        np.add.at(F, t[:, 3], fl)
    return {"bond": Eb, "angle": Ea, "torsion": Et}, F
