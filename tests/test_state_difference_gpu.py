"""x + d evaluated as the difference from x.

When every system of a batch displaces at most a quarter of its atoms, the
batched evaluation bins and pairs only x; the displaced-atom kernel evaluates
every pair with a displaced atom at x and at x + d from the x cells (the other
atoms do not move), and the exception terms whose atoms are displaced
differently, so U2 = U1 - E_d(x) + E_d(x + d) and F2 = F1 + dF.  Checked here
against the same batch evaluated with x + d binned and paired in full
(full_state2) and against the FP64 oracle: energies, forces, the per-set pair
counts the kernels report and the x + d pair set element by element (the
state-0 list traversal without the displaced kernel's x pairs, with its
x + d pairs).
"""
import dataclasses

import numpy as np
import pytest

from atmmetaforce import ATMBatch, workloads as W
from oracle import oracle

from test_gpu_parity import check_system

pytestmark = pytest.mark.gpu

P0 = W.ATMParams(0.3, 0.45, 0.1, 110.0, 1.5, 200.0, 100.0, 0.0625, 1.0)


def two_ligands(n=2500, box=2.9, seed=7):
    """An RBFE-like layout: atoms 0-19 displaced by +v and 20-39 by -v, so the
    pairs between the two ligands change with the state.  v moves a lattice site
    to the centre of a lattice cube (no overlaps)."""
    s = W.tiny_system(n, box, 0.9, 40, (0.0, 0.0, 0.0), seed, P0)
    a = s.box[0] / 14   # 14^3 >= 2500 lattice sites per edge
    v = np.array([2.5 * a, 0.5 * a, 0.5 * a])
    d = np.zeros((n, 3))
    d[:20] = v
    d[20:40] = -v
    return dataclasses.replace(s, displacements=W.quantize(d))


def systems():
    return [
        W.with_molecules(W.tiny_system(3000, 3.1, 0.9, 30, (1.03, 0.77, 0.62), 1, P0), size=4, crosslink_every=9),
        two_ligands(),
        W.tiny_system(2500, 2.9, 0.9, 25, (0.21, 0.17, 0.13), 3, P0),   # displaced ligand clashes (side sums)
        W.tiny_system(64, 1.7, 0.8, 4, (0.6, 0.4, 0.3), 11, P0),         # box < 2 (rc + column): no image shifts
    ]


def run(batch, full):
    batch.set_tuning(1.3, 1, 4, 0, pair_split=2, full_state2=full)
    assert batch.tuning()["full_state2"] == full
    batch.compute()
    n = len(batch.systems)
    return {
        "res": [batch.result(k) for k in range(n)],
        "F": [batch.forces(k) for k in range(n)],
        "counts": batch.list_stats()["pairs"].copy(),
        "pairs1": [batch.pairs(k, 1) for k in range(n)],
    }


def test_state_difference_matches_full_and_oracle(gpu):
    sy = systems()
    # the exception batch and the others in separate batches (a batch with exceptions runs
    # the tile kernel with exclusion tests on every system)
    for group in (sy[:1], sy[1:]):
        with ATMBatch(group) as b:
            full = run(b, True)
            diff = run(b, False)
            for k, s in enumerate(group):
                rd, rf = diff["res"][k], full["res"][k]
                tag = f"state difference sys {k} N={s.num_atoms}"
                print(f"[{tag}] U2 {rd['state2_energy']:.12g} full {rf['state2_energy']:.12g}", flush=True)
                assert rd["state1_energy"] == rf["state1_energy"]   # the same x evaluation
                assert abs(rd["state2_energy"] - rf["state2_energy"]) <= 1e-11 * max(1.0, abs(rf["state2_energy"]))
                assert abs(rd["energy"] - rf["energy"]) <= 1e-11 * max(1.0, abs(rf["energy"]))
                dF = np.linalg.norm(diff["F"][k] - full["F"][k]) / np.linalg.norm(full["F"][k])
                assert dF < 1e-6, dF
                assert diff["counts"][2 * k + 1] == full["counts"][2 * k + 1]
                assert np.array_equal(diff["pairs1"][k], full["pairs1"][k])
                po = oracle.pairs(s.positions + s.displacements, s.box, s.cutoff,
                                  exceptions=getattr(s, "exceptions", None))
                assert np.array_equal(diff["pairs1"][k], po)
            b.set_tuning(1.3, 1, 4, 0, pair_split=2)
            b.compute()
            counts = b.list_stats()["pairs"]
            for k, s in enumerate(group):
                check_system(b, k, s, pairs=True, tag="state difference", counts=counts)


def test_many_displaced_atoms_take_the_full_path(gpu):
    """More than a quarter of the atoms displaced: x + d is binned in full, and the
    results still match the oracle."""
    s = W.tiny_system(800, 2.2, 0.8, 300, (0.4, 0.3, 0.2), 5, P0)
    with ATMBatch([s]) as b:
        b.compute()
        check_system(b, 0, s, pairs=True, tag="many displaced", counts=b.list_stats()["pairs"])


def test_displacement_changes_between_evaluations(gpu):
    """New displacements (another ligand size) re-plan the displaced-atom kernel;
    the captured evaluation follows."""
    s = W.tiny_system(2000, 2.8, 0.9, 20, (1.1, 0.6, 0.4), 4, P0)
    with ATMBatch([s]) as b:
        b.compute()
        check_system(b, 0, s, pairs=True, tag="ligand 20", counts=b.list_stats()["pairs"])
        d = np.zeros_like(s.displacements)
        d[:60] = (0.9, -0.5, 0.3)
        d = W.quantize(d)
        s2 = dataclasses.replace(s, displacements=d)
        b.set_displacements(0, d)
        b.compute()
        check_system(b, 0, s2, pairs=True, tag="ligand 60", counts=b.list_stats()["pairs"])
        b.set_displacements(0, np.zeros_like(d))   # nothing displaced: U2 = U1
        b.compute()
        r = b.result(0)
        assert r["state2_energy"] == r["state1_energy"]


def test_backward_after_a_state_difference_evaluation(gpu):
    """batch.backward after a plain (state-difference) evaluation: the gradients of
    both states come from a full evaluation with parameter gradients."""
    s = W.tiny_system(1500, 2.6, 0.9, 12, (0.9, 0.8, 0.7), 2, P0)
    with ATMBatch([s]) as b:
        b.compute(forces=False)
        g = b.backward(np.array([1.0]))
    ref = oracle.atm_grad(s)
    for key in ("dq", "dsigma", "deps"):
        err = np.linalg.norm(g[key] - ref[key]) / np.linalg.norm(ref[key])
        print(f"[backward] {key} rel {err:.2e}", flush=True)
        assert err < 2e-5
