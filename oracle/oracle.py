"""ctypes front end of the FP64 oracle (oracle/atm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  Each function
names the reference code it restates in atm_oracle.c's header.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libatm_oracle.so")
_lib = None


class Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("lambda1", "lambda2", "alpha", "u0", "w0", "umax", "ubcore", "acore", "direction")]


class Result(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("W", "usc", "U1", "U2", "u", "fp", "bfp", "sp")] + \
               [("npairs1", ctypes.c_longlong), ("npairs2", ctypes.c_longlong),
                ("abs1", ctypes.c_double), ("abs2", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def build():
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH) or \
                os.path.getmtime(_LIB_PATH) < os.path.getmtime(os.path.join(_HERE, "atm_oracle.c")):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_softcore.restype = ctypes.c_double
        L.oracle_softcore.argtypes = [ctypes.c_double] * 4 + [dp]
        L.oracle_combine.argtypes = [ctypes.POINTER(Params), ctypes.c_double, ctypes.c_double, ctypes.POINTER(Result)]
        L.oracle_blend.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, dp, dp, dp]
        L.oracle_pairs.restype = ctypes.c_longlong
        L.oracle_pairs.argtypes = [ctypes.c_int, dp, dp, ctypes.c_double, ctypes.c_int, ip, ctypes.c_longlong]
        L.oracle_nonbonded.restype = ctypes.c_double
        L.oracle_nonbonded.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int, dp, ctypes.POINTER(ctypes.c_longlong), dp]
        L.oracle_atm_eval.restype = ctypes.c_int
        L.oracle_atm_eval.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp, dp, ctypes.c_double, ctypes.c_double,
                                      ctypes.POINTER(Params), ctypes.c_int, dp, ctypes.POINTER(Result)]
        L.oracle_atm_grad.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp, dp, ctypes.c_double, ctypes.c_double,
                                      ctypes.POINTER(Params), ctypes.POINTER(Result), dp, dp, dp, dp]
        L.oracle_pairs_ex.restype = ctypes.c_longlong
        L.oracle_pairs_ex.argtypes = [ctypes.c_int, dp, dp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ip, ip,
                                      ctypes.c_longlong]
        L.oracle_nonbonded_ex.restype = ctypes.c_double
        L.oracle_nonbonded_ex.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_int, ctypes.c_int, ip, dp, dp, dp, dp,
                                          ctypes.POINTER(ctypes.c_longlong), dp]
        L.oracle_atm_eval_ex.restype = ctypes.c_int
        L.oracle_atm_eval_ex.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp, dp, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_int, ip, dp, dp, dp, ctypes.POINTER(Params), ctypes.c_int, dp,
                                         ctypes.POINTER(Result)]
        L.oracle_pairs_tri.restype = ctypes.c_longlong
        L.oracle_pairs_tri.argtypes = [ctypes.c_int, dp, dp, dp, ctypes.c_double, ctypes.c_int, ip, ip,
                                       ctypes.c_longlong]
        L.oracle_atm_eval_tri.restype = ctypes.c_int
        L.oracle_atm_eval_tri.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp, dp, dp, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_int, ip, dp, dp, dp, ctypes.POINTER(Params), dp,
                                          ctypes.POINTER(Result)]
        L.oracle_atm_grad_ex.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp, dp, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_int, ip, ctypes.POINTER(Params), ctypes.POINTER(Result), dp, dp, dp,
                                         dp]
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def params_of(p) -> Params:
    if isinstance(p, Params):
        return p
    if hasattr(p, "as_tuple"):
        return Params(*p.as_tuple())
    if isinstance(p, dict):
        return Params(**p)
    return Params(*p)


def softcore(u, umax, acore, ubcore):
    fp = ctypes.c_double()
    v = lib().oracle_softcore(u, umax, acore, ubcore, ctypes.byref(fp))
    return v, fp.value


def combine(params, U1, U2) -> dict:
    r = Result()
    p = params_of(params)
    lib().oracle_combine(ctypes.byref(p), U1, U2, ctypes.byref(r))
    return r.as_dict()


def blend(sp, direction, F1, F2, F=None):
    F1, p1 = _d(F1)
    F2, p2 = _d(F2)
    out = np.zeros_like(F1) if F is None else np.array(F, dtype=np.float64, copy=True)
    out = np.ascontiguousarray(out)
    lib().oracle_blend(F1.shape[0], sp, direction, p1, p2, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def _exc(ex):
    """(count, pairs, qq, sigma, epsilon) ctypes arguments of a workloads.Exceptions (or None)."""
    ip = ctypes.POINTER(ctypes.c_int)
    if ex is None or ex.count == 0:
        return 0, None, None, None, None, ()
    pr = np.ascontiguousarray(ex.pairs, dtype=np.int32)
    qq, pq = _d(ex.charge_prod)
    sg, ps = _d(ex.sigma)
    ep, pe = _d(ex.epsilon)
    return ex.count, pr.ctypes.data_as(ip), pq, ps, pe, (pr, qq, sg, ep)


def _tilt(tilt):
    """(array, pointer) of a triclinic box's (b_x, c_x, c_y), or None for an orthorhombic box."""
    if tilt is None or not np.any(np.asarray(tilt, dtype=np.float64) != 0.0):
        return None
    return _d(tilt)


def pairs(x, box, rc, brute=False, exceptions=None, tilt=None) -> np.ndarray:
    """Sorted (i, j) in-cutoff pairs, i < j, that are not exceptions, shape (P, 2).  tilt:
    (b_x, c_x, c_y) of a triclinic box with edges box = (a_x, b_y, c_z)."""
    x, px = _d(x)
    box, pb = _d(box)
    n = x.shape[0]
    ne, pex, _, _, _, keep = _exc(exceptions)
    t = _tilt(tilt)
    ip = ctypes.POINTER(ctypes.c_int)
    if t is not None:
        count = lib().oracle_pairs_tri(n, px, pb, t[1], rc, ne, pex, None, 0)
        out = np.zeros((max(count, 1), 2), dtype=np.int32)
        lib().oracle_pairs_tri(n, px, pb, t[1], rc, ne, pex, out.ctypes.data_as(ip), count)
    else:
        count = lib().oracle_pairs_ex(n, px, pb, rc, int(brute), ne, pex, None, 0)
        out = np.zeros((max(count, 1), 2), dtype=np.int32)
        lib().oracle_pairs_ex(n, px, pb, rc, int(brute), ne, pex, out.ctypes.data_as(ip), count)
    out = out[:count]
    order = np.lexsort((out[:, 1], out[:, 0]))
    return out[order]


def nonbonded(x, q, sig, eps, box, rc, eps_rf, brute=False, forces=True, exceptions=None):
    """(energy, forces (N,3) or None, npairs, (E_LJ, E_C, sum |e_ij|, E_exceptions))."""
    x, px = _d(x)
    q, pq = _d(q)
    sig, ps = _d(sig)
    eps, pe = _d(eps)
    box, pb = _d(box)
    n = x.shape[0]
    F = np.zeros((n, 3)) if forces else None
    np_ = ctypes.c_longlong()
    parts = np.zeros(4)
    ne, pex, pqq, psg, pep, keep = _exc(exceptions)
    E = lib().oracle_nonbonded_ex(n, px, pq, ps, pe, pb, rc, eps_rf, int(brute), ne, pex, pqq, psg, pep,
                                  F.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if forces else None,
                                  ctypes.byref(np_), parts.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return E, F, np_.value, (parts[0], parts[1], parts[2], parts[3])


def atm_eval(system, params=None, brute=False, forces=True):
    """Oracle ATM evaluation of a workloads.System: (result dict, ATM force (N,3) or None)."""
    p = params_of(params if params is not None else system.params)
    x, px = _d(system.positions)
    d, pd = _d(system.displacements)
    q, pq = _d(system.charges)
    s, ps = _d(system.sigmas)
    e, pe = _d(system.epsilons)
    b, pb = _d(system.box)
    n = x.shape[0]
    F = np.zeros((n, 3)) if forces else None
    r = Result()
    ne, pex, pqq, psg, pep, keep = _exc(getattr(system, "exceptions", None))
    pF = F.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if forces else None
    t = _tilt(getattr(system, "tilt", None))
    if t is not None:
        # triclinic box (system.tilt = (b_x, c_x, c_y)): O(N^2) pair loop
        rc = lib().oracle_atm_eval_tri(n, px, pd, pq, ps, pe, pb, t[1], system.cutoff, system.rf_dielectric, ne,
                                       pex, pqq, psg, pep, ctypes.byref(p), pF, ctypes.byref(r))
    else:
        rc = lib().oracle_atm_eval_ex(n, px, pd, pq, ps, pe, pb, system.cutoff, system.rf_dielectric, ne, pex, pqq,
                                      psg, pep, ctypes.byref(p), int(brute), pF, ctypes.byref(r))
    if rc != 0:
        raise MemoryError("oracle_atm_eval failed")
    return r.as_dict(), F


def atm_grad(system, params=None):
    """Analytic FP64 gradient: dict with dq, dsigma, deps (N,) and dparams (5,)."""
    p = params_of(params if params is not None else system.params)
    res, _ = atm_eval(system, p, forces=False)
    r = Result(**{k: res[k] for k in ("W", "usc", "U1", "U2", "u", "fp", "bfp", "sp")})
    x, px = _d(system.positions)
    d, pd = _d(system.displacements)
    q, pq = _d(system.charges)
    s, ps = _d(system.sigmas)
    e, pe = _d(system.epsilons)
    b, pb = _d(system.box)
    n = x.shape[0]
    dq, ds, de, dp = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(5)
    c = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    ne, pex, _, _, _, keep = _exc(getattr(system, "exceptions", None))
    lib().oracle_atm_grad_ex(n, px, pd, pq, ps, pe, pb, system.cutoff, system.rf_dielectric, ne, pex, ctypes.byref(p),
                             ctypes.byref(r), c(dq), c(ds), c(de), c(dp))
    return {"dq": dq, "dsigma": ds, "deps": de, "dparams": dp}


# ---------------------------------------------------------------- CPU-equivalent
# Multi-threaded FP32-pair / FP64-sum CPU evaluation (atm_cpu_mt.c): the
# speed baseline timed beside the device library, not a checker.

_MT_PATH = os.path.join(_HERE, "libatm_cpu_mt.so")
_mt = None


def mt_lib():
    global _mt
    if _mt is None:
        lib()
        if not os.path.exists(_MT_PATH) or \
                os.path.getmtime(_MT_PATH) < os.path.getmtime(os.path.join(_HERE, "atm_cpu_mt.c")):
            build()
        L = ctypes.CDLL(_MT_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.cpu_mt_atm_eval.restype = ctypes.c_int
        L.cpu_mt_atm_eval.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp, dp, ctypes.c_double, ctypes.c_double,
                                      ctypes.POINTER(Params), ctypes.c_int, dp, ctypes.POINTER(Result)]
        L.cpu_mt_max_threads.restype = ctypes.c_int
        _mt = L
    return _mt


def cpu_mt_atm_eval(system, threads=0, params=None, forces=True):
    """ATM evaluation of a workloads.System on `threads` host threads (0: all
    OpenMP threads): (result dict, ATM force (N,3) or None)."""
    p = params_of(params if params is not None else system.params)
    x, px = _d(system.positions)
    d, pd = _d(system.displacements)
    q, pq = _d(system.charges)
    s, ps = _d(system.sigmas)
    e, pe = _d(system.epsilons)
    b, pb = _d(system.box)
    n = x.shape[0]
    F = np.zeros((n, 3)) if forces else None
    r = Result()
    rc = mt_lib().cpu_mt_atm_eval(n, px, pd, pq, ps, pe, pb, system.cutoff, system.rf_dielectric, ctypes.byref(p),
                                  int(threads), F.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if forces else None,
                                  ctypes.byref(r))
    if rc != 0:
        raise ValueError(f"cpu_mt_atm_eval failed ({rc}): the box needs at least 3 cells of rc per edge")
    return r.as_dict(), F


def cpu_mt_threads() -> int:
    return mt_lib().cpu_mt_max_threads()


# ---------------------------------------------------------------- restraints
# FP64 numpy restatement of the flat-bottom restraints that
# python/ATMMetaForceUtils.py builds as OpenMM custom forces (test
# infrastructure: the checker of the device kernel atm_restraints.hip).

def centroid_restraint(x, group1, group2, weights1, weights2, kf, tol, offset):
    """Energy and forces of one centroid-centroid restraint,
    python/ATMMetaForceUtils.py:113-114 (CustomCentroidBondForce expression
    "(kfcm/2)*step(d12-tolcm)*(d12-tolcm)^2; d12 = |c1 - off - c2|"; member
    weights per group, aligned with the member lists; positions not wrapped)."""
    x = np.asarray(x, dtype=np.float64)
    g1, g2 = np.asarray(group1), np.asarray(group2)
    w1, w2 = np.asarray(weights1, dtype=np.float64), np.asarray(weights2, dtype=np.float64)
    c1 = (w1[:, None] * x[g1]).sum(0) / w1.sum()
    c2 = (w2[:, None] * x[g2]).sum(0) / w2.sum()
    d = c1 - np.asarray(offset, dtype=np.float64) - c2
    r = float(np.sqrt(d @ d))
    F = np.zeros_like(x)
    if r <= tol:
        return 0.0, F
    s = r - tol
    g = kf * s * d / r                                # dE/dc1
    np.add.at(F, g1, -g[None, :] * (w1 / w1.sum())[:, None])
    np.add.at(F, g2, g[None, :] * (w2 / w2.sum())[:, None])
    return 0.5 * kf * s * s, F


def position_restraints(x, atoms, refs, kf, tol, box=None):
    """Energy and forces of flat-bottom position restraints,
    python/ATMMetaForceUtils.py:733-736 ("0.5*fc*select(step(dist-tol),
    (dist-tol)^2, 0)", dist = periodicdistance(x, x0) when `box` is given,
    else |x - x0|).  kf, tol: scalars or one per restrained atom."""
    x = np.asarray(x, dtype=np.float64)
    atoms = np.asarray(atoms)
    d = x[atoms] - np.asarray(refs, dtype=np.float64)
    if box is not None:
        L = np.asarray(box, dtype=np.float64)
        d = d - L * np.round(d / L)
    r = np.sqrt((d * d).sum(1))
    kf = np.broadcast_to(np.asarray(kf, dtype=np.float64), r.shape)
    tol = np.broadcast_to(np.asarray(tol, dtype=np.float64), r.shape)
    s = np.where(r > tol, r - tol, 0.0)
    E = float((0.5 * kf * s * s).sum())
    F = np.zeros_like(x)
    k = np.where(r > tol, -kf * s / np.where(r > 0, r, 1.0), 0.0)
    np.add.at(F, atoms, k[:, None] * d)
    return E, F
