"""Batched ATM evaluation on one MI355X through the C ABI (include/atm_mi355x.h).

A batch holds independent systems (lambda replicas of one system, or systems of
uneven size).  One evaluation computes, for every system, U1/F1 at x, U2/F2 at
x + d, the soft-core/softplus combine and the hybrid force, in one captured HIP
graph; the reference-structured path (one system and one coordinate set per
launch sequence, energies read back to the host) is available with plain=True.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from . import _capi as C
from .workloads import box_vectors


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _params(p) -> C.Params:
    if isinstance(p, C.Params):
        return p
    if hasattr(p, "as_tuple"):
        return C.Params(*p.as_tuple())
    if isinstance(p, dict):
        return C.Params(**p)
    return C.Params(*p)


class ATMBatch:
    """A context of `len(systems)` ATM systems on GPU `device`.

    `systems` are objects with positions (N,3), charges, sigmas, epsilons (N,),
    displacements (N,3), box (3,), cutoff, rf_dielectric and params (nine ATM
    parameters) -- e.g. atmmetaforce.workloads.System.
    """

    def __init__(self, systems: Sequence, device: int = 0):
        self._lib = C.load()
        self.systems = list(systems)
        n = len(self.systems)
        descs = (C.SystemDesc * n)()
        for k, s in enumerate(self.systems):
            descs[k].num_atoms = int(s.positions.shape[0])
            for a in range(3):
                descs[k].box[a] = float(s.box[a])
            descs[k].cutoff = float(s.cutoff)
            descs[k].rf_dielectric = float(s.rf_dielectric)
        h = ctypes.c_void_p()
        C.check(self._lib.atm_create(descs, n, device, ctypes.byref(h)))
        self._h = h
        self._eval_id = 0
        self.num_atoms = [int(s.positions.shape[0]) for s in self.systems]
        # what the device holds, for ATMEnergyFunction's re-evaluation: positions as last set
        # from the host (None once set on the device), boxes, and a version bumped on changes
        self._pos_version = 0
        self._host_positions = [None] * n
        self._boxes = [np.array(s.box, dtype=np.float64).reshape(3) for s in self.systems]
        for k, s in enumerate(self.systems):
            tilt = getattr(s, "tilt", None)
            if tilt is not None and np.any(np.asarray(tilt) != 0.0):
                self.set_box_vectors(k, box_vectors(s.box, tilt))
            self.set_particles(k, s.charges, s.sigmas, s.epsilons, s.displacements)
            self.set_positions(k, s.positions)
            self.set_params(k, s.params)
            ex = getattr(s, "exceptions", None)
            if ex is not None and ex.count > 0:
                self.set_exceptions(k, ex.pairs, ex.charge_prod, ex.sigma, ex.epsilon)

    # ------------------------------------------------------------ lifecycle
    def close(self):
        if getattr(self, "_h", None):
            C.check(self._lib.atm_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    # ------------------------------------------------------------ inputs
    def set_particles(self, k, charges, sigmas, epsilons, displacements=None):
        q = np.ascontiguousarray(charges, dtype=np.float64)
        s = np.ascontiguousarray(sigmas, dtype=np.float64)
        e = np.ascontiguousarray(epsilons, dtype=np.float64)
        n = self.num_atoms[k]
        if q.shape != (n,) or s.shape != (n,) or e.shape != (n,):
            raise ValueError(f"system {k}: particle arrays must have shape ({n},)")
        d = None
        if displacements is not None:
            d = np.ascontiguousarray(displacements, dtype=np.float64).reshape(n, 3)
        C.check(self._lib.atm_set_particles(self._h, k, _dptr(q), _dptr(s), _dptr(e),
                                            _dptr(d) if d is not None else None))

    def set_exceptions(self, k, pairs, charge_prod, sigma, epsilon):
        """NonbondedForce exceptions of system k (atm_set_exceptions): pairs (m, 2),
        charge products (e^2), sigmas (nm), epsilons (kJ/mol); m = 0 removes them."""
        pr = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        m = pr.shape[0]
        qq = np.ascontiguousarray(charge_prod, dtype=np.float64).reshape(m)
        sg = np.ascontiguousarray(sigma, dtype=np.float64).reshape(m)
        ep = np.ascontiguousarray(epsilon, dtype=np.float64).reshape(m)
        C.check(self._lib.atm_set_exceptions(self._h, k, m, pr.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                             _dptr(qq), _dptr(sg), _dptr(ep)))

    def set_energy_offsets(self, k, state1, state2):
        """Energies (kJ/mol) of the other variable-group forces of system k at x and at
        x + d (atm_set_energy_offsets): added to U1 and U2 before the soft-core combine."""
        C.check(self._lib.atm_set_energy_offsets(self._h, k, float(state1), float(state2)))

    def set_displacements(self, k, displacements):
        d = np.ascontiguousarray(displacements, dtype=np.float64).reshape(self.num_atoms[k], 3)
        C.check(self._lib.atm_set_displacements(self._h, k, _dptr(d)))

    def set_positions(self, k, positions):
        x = np.ascontiguousarray(positions, dtype=np.float64).reshape(self.num_atoms[k], 3)
        C.check(self._lib.atm_set_positions(self._h, k, _dptr(x)))
        self._host_positions[k] = x.copy()
        self._pos_version += 1

    def set_positions_device(self, ptr: int):
        C.check(self._lib.atm_set_positions_device(self._h, ctypes.c_void_p(ptr)))
        self.positions_changed_on_device()

    def positions_changed_on_device(self):
        """Record that device code (a device position upload, MD steps) moved the atoms."""
        self._host_positions = [None] * len(self.systems)
        self._pos_version += 1

    def set_box(self, k, box):
        """Box of system k: edges (3,) of an orthorhombic box, or vectors (3, 3) / (9,)
        (rows a, b, c in OpenMM's reduced form; triclinic boxes included)."""
        b = np.ascontiguousarray(box, dtype=np.float64).reshape(-1)
        if b.size == 9:
            self.set_box_vectors(k, b)
            return
        b = b.reshape(3)
        C.check(self._lib.atm_set_box(self._h, k, _dptr(b)))
        self._boxes[k] = b.copy()
        self._pos_version += 1

    def set_box_vectors(self, k, vectors):
        """Periodic box vectors of system k (atm_set_box_vectors): rows a, b, c of a (3, 3)
        array in OpenMM's reduced form, nm."""
        v = np.ascontiguousarray(vectors, dtype=np.float64).reshape(9)
        C.check(self._lib.atm_set_box_vectors(self._h, k, _dptr(v)))
        self._boxes[k] = v.copy()
        self._pos_version += 1

    def position_snapshot(self):
        """(version, host positions or None per system, boxes) of the current inputs."""
        return (self._pos_version, list(self._host_positions), [b.copy() for b in self._boxes])

    def restore_positions(self, snap):
        """Upload the positions and boxes of a snapshot again (no-op if nothing changed)."""
        version, xs, boxes = snap
        if version == self._pos_version:
            return
        if any(x is None for x in xs):
            raise RuntimeError("the positions of this evaluation were set on the device and have changed "
                               "since; they cannot be restored")
        for k, (x, b) in enumerate(zip(xs, boxes)):
            if not np.array_equal(b, self._boxes[k]):
                self.set_box(k, b)
            self.set_positions(k, x)
        self._pos_version = version

    def set_params(self, k, params):
        p = _params(params)
        C.check(self._lib.atm_set_params(self._h, k, ctypes.byref(p)))

    def set_stream(self, stream_ptr: int | None):
        C.check(self._lib.atm_set_stream(self._h, ctypes.c_void_p(stream_ptr) if stream_ptr else None))

    # ------------------------------------------------------------ evaluation
    def evaluate(self, forces: bool = True, clear: bool = True, plain: bool = False, graph: bool = True,
                 fp32_pairs: bool = False, param_grads: bool = False):
        """Enqueue one evaluation of every system (asynchronous).  param_grads: energies only,
        with the per-atom parameter derivatives for a following backward in the same pass."""
        flags = C.ATM_EVAL_FP32_PAIRS if fp32_pairs else 0
        if param_grads:
            flags |= C.ATM_EVAL_PARAM_GRADS
        if forces:
            flags |= C.ATM_EVAL_FORCES
        if clear:
            flags |= C.ATM_EVAL_CLEAR_FORCES
        if plain:
            flags |= C.ATM_EVAL_PLAIN
        if not graph:
            flags |= C.ATM_EVAL_NO_GRAPH
        self._eval_id += 1
        C.check(self._lib.atm_eval(self._h, flags))

    def synchronize(self):
        C.check(self._lib.atm_synchronize(self._h))

    def compute(self, forces: bool = True, plain: bool = False, graph: bool = True, retries: int = 3,
                fp32_pairs: bool = False, param_grads: bool = False):
        """Evaluate (clearing forces) and wait; re-runs if a neighbour list overflowed."""
        for attempt in range(retries + 1):
            self.evaluate(forces=forces, clear=True, plain=plain, graph=graph, fp32_pairs=fp32_pairs,
                          param_grads=param_grads)
            r = C.Result()
            rc = self._lib.atm_get_results(self._h, 0, ctypes.byref(r))
            if rc == C.ATM_ERR_CAPACITY and attempt < retries:
                continue
            C.check(rc)
            return

    def result(self, k) -> dict:
        r = C.Result()
        C.check(self._lib.atm_get_results(self._h, k, ctypes.byref(r)))
        return r.as_dict()

    def results(self):
        return [self.result(k) for k in range(len(self.systems))]

    def forces(self, k) -> np.ndarray:
        out = np.zeros((self.num_atoms[k], 3))
        C.check(self._lib.atm_get_forces(self._h, k, _dptr(out)))
        return out

    def forces_fixed(self, k) -> np.ndarray:
        out = np.zeros((self.num_atoms[k], 3), dtype=np.int64)
        C.check(self._lib.atm_get_forces_fixed(self._h, k, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return out

    def clear_forces(self):
        C.check(self._lib.atm_clear_forces(self._h))

    def pairs(self, k, state) -> np.ndarray:
        """Sorted (i, j) in-cutoff pairs (i < j) of system k, state 0 (x) or 1 (x + d)."""
        cnt = ctypes.c_int64()
        C.check(self._lib.atm_get_pairs(self._h, k, state, None, 0, ctypes.byref(cnt)))
        out = np.zeros((max(cnt.value, 1), 2), dtype=np.int32)
        C.check(self._lib.atm_get_pairs(self._h, k, state, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                        cnt.value, ctypes.byref(cnt)))
        return out[:cnt.value]

    def stats(self) -> dict:
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        C.check(self._lib.atm_get_stats(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"pairs_in_cutoff": a.value, "pairs_evaluated": b.value, "list_entries": c.value}

    def list_stats(self) -> dict:
        """Per coordinate set (index 2 k + state) of the last evaluation: in-cutoff pairs counted
        by the pair kernel, the longest group list and the list capacity per group."""
        n = 2 * len(self.systems)
        pairs = np.zeros(n, dtype=np.int64)
        mx = np.zeros(n, dtype=np.int32)
        cap = np.zeros(n, dtype=np.int32)
        C.check(self._lib.atm_get_list_stats(self._h, pairs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                             mx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                             cap.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return {"pairs": pairs, "max_entries": mx, "capacity": cap}

    def set_particles_device(self, charges_ptr: int, sigmas_ptr: int, epsilons_ptr: int):
        """Charges, sigmas, epsilons of every system from device float64 arrays (pointers)."""
        C.check(self._lib.atm_set_particles_device(self._h, ctypes.c_void_p(charges_ptr), ctypes.c_void_p(sigmas_ptr),
                                                   ctypes.c_void_p(epsilons_ptr)))

    def backward_device(self, grad_energy, dq_ptr: int, dsigma_ptr: int, deps_ptr: int) -> np.ndarray:
        """atm_backward with the per-atom gradients written to device float64 arrays
        (pointers, 0 = skip); returns dW/d(lambda1, lambda2, alpha, u0, w0) per system."""
        g = np.ascontiguousarray(grad_energy, dtype=np.float64).reshape(len(self.systems))
        dp = np.zeros(5 * len(self.systems))
        vp = lambda p: ctypes.c_void_p(p) if p else None
        C.check(self._lib.atm_backward_device(self._h, _dptr(g), vp(dq_ptr), vp(dsigma_ptr), vp(deps_ptr), _dptr(dp)))
        return dp.reshape(-1, 5)

    def backward(self, grad_energy) -> dict:
        g = np.ascontiguousarray(grad_energy, dtype=np.float64).reshape(len(self.systems))
        tot = sum(self.num_atoms)
        dq, ds, de = np.zeros(tot), np.zeros(tot), np.zeros(tot)
        dp = np.zeros(5 * len(self.systems))
        C.check(self._lib.atm_backward(self._h, _dptr(g), _dptr(dq), _dptr(ds), _dptr(de), _dptr(dp)))
        return {"dq": dq, "dsigma": ds, "deps": de, "dparams": dp.reshape(-1, 5)}

    def profile(self, reps: int = 10, forces: bool = True, fp32_pairs: bool = False) -> dict:
        """Per-stage device times (ms) of the batched evaluation, HIP events between stages."""
        t = C.StageTimes()
        flags = C.ATM_EVAL_CLEAR_FORCES | (C.ATM_EVAL_FORCES if forces else 0) | \
            (C.ATM_EVAL_FP32_PAIRS if fp32_pairs else 0)
        C.check(self._lib.atm_profile(self._h, flags, reps, ctypes.byref(t)))
        return t.as_dict()

    # ------------------------------------------------------------ tuning
    def autotune(self, repeats: int = 3) -> dict:
        t = C.Tuning()
        C.check(self._lib.atm_autotune(self._h, repeats, ctypes.byref(t)))
        return _tuning_dict(t)

    def tuning(self) -> dict:
        t = C.Tuning()
        C.check(self._lib.atm_get_tuning(self._h, ctypes.byref(t)))
        return _tuning_dict(t)

    def set_tuning(self, column_scale=1.3, pair_waves_per_block=2, zbins_per_column=4, pair_kernel=0,
                   list_capacity=0, pair_split=1, full_state2=False, pair_order=2):
        """pair_kernel 0: 8x8 cluster-pair tiles; 1-4: compacted in-cutoff pairs with (list entries
        per batch, copies of the j-force sums, drain after every test step) = (4, 2, yes), (4, 2, no),
        (4, 4, yes), (8, 2, yes); 5: atom-level j lists per i-cluster, 8 i x 8 j lanes per step;
        6: filtered atom queues.  list_capacity: list entries per group to start from (0: sized from
        the density).  pair_split: waves sharing one group's list in the tile kernel (1, 2, 4, 8).
        full_state2: bin and pair x + d in full instead of as the difference from x at the
        displaced atoms (used when every system displaces at most a quarter of its atoms).
        pair_order: the tile kernel's group order (0 launch order, 1 longest lists first over all
        sets, 2 the same over windows of 8 sets, 3 each set longest first, one set after another)."""
        t = C.Tuning(column_scale, pair_waves_per_block, zbins_per_column, 0.0, pair_kernel, list_capacity,
                     pair_split, 1 if full_state2 else 0, pair_order, 0)
        C.check(self._lib.atm_set_tuning(self._h, ctypes.byref(t)))


def _tuning_dict(t) -> dict:
    return {"column_scale": t.column_scale, "pair_waves_per_block": t.pair_waves_per_block,
            "zbins_per_column": t.zbins_per_column, "pair_kernel": t.pair_kernel, "best_ms": t.best_ms,
            "list_capacity": t.list_capacity, "pair_split": t.pair_split, "full_state2": bool(t.full_state2),
            "pair_order": t.pair_order}
