"""NonbondedForce exceptions on the reaction-field path (atm_set_exceptions):
excluded pairs leave the cutoff sum, scaled pairs add their own Coulomb + LJ
term at the plain distance.  Checked against the FP64 oracle with the bars of
the plain parity tests (pair sets element by element, pair counts exact, U1/U2
within 1e-10 of sum |e_ij|, forces within 1e-5 relative RMS), for exceptions
inside the 31-atom index window of the near masks and for crosslinks far
outside it, in the batched and the reference-structured paths, in FP32 pair
mode, with a pair-kernel variant that has no exclusion support asked for, and
through the backward pass."""
import dataclasses

import numpy as np
import pytest
import torch

from atmmetaforce import ATMBatch, ATMEnergyFunction, workloads as W
from oracle import oracle
from test_gpu_parity import check_system

pytestmark = pytest.mark.gpu

P = W.ATMParams(0.3, 0.45, 0.1, 110.0, 1.5, 200.0, 100.0, 0.0625, 1.0)


def systems():
    return [
        W.with_molecules(W.tiny_system(3000, 3.1, 0.9, 30, (1.03, 0.77, 0.62), 1, P), size=3, crosslink_every=11),
        W.with_molecules(W.tiny_system(2500, 2.9, 0.9, 24, (0.41, 0.37, 0.23), 3, P), size=4, crosslink_every=5),
        W.with_molecules(W.tiny_system(64, 1.7, 0.8, 4, (0.6, 0.4, 0.3), 11, P), size=4, crosslink_every=3),
        W.tiny_system(1200, 2.6, 0.9, 10, (0.5, 0.5, 0.5), 9, P),          # no exceptions in the same batch
    ]


def test_exception_workloads_cover_both_paths():
    for s in systems()[:2]:
        d = np.abs(s.exceptions.pairs[:, 1] - s.exceptions.pairs[:, 0])
        assert (d <= 31).any() and (d > 31).any()
        assert (s.exceptions.charge_prod != 0).any() and (s.exceptions.charge_prod == 0).any()


@pytest.mark.parametrize("plain", [False, True])
def test_exceptions_batch(gpu, plain):
    sy = systems()
    with ATMBatch(sy) as b:
        b.compute(plain=plain)
        counts = b.list_stats()["pairs"] if not plain else None
        for k, s in enumerate(sy):
            check_system(b, k, s, pairs=not plain, tag=f"exceptions plain={plain}", strict=False,
                         counts=counts if not plain else None)


def test_exceptions_fp32_pairs_and_variant_fallback(gpu):
    sy = systems()
    with ATMBatch(sy) as b:
        b.compute(fp32_pairs=True)
        for k, s in enumerate(sy):
            check_system(b, k, s, pairs=False, tag="exceptions fp32", strict=False, fp32=True)
        for kernel in (1, 5, 6):  # no exclusion support: the tile kernel runs instead
            b.set_tuning(1.0, 2, 4, kernel)
            b.compute()
            counts = b.list_stats()["pairs"]
            for k, s in enumerate(sy):
                check_system(b, k, s, pairs=(k == 0), tag=f"exceptions kernel {kernel}", strict=False, counts=counts)


def test_exceptions_replaced_and_removed(gpu):
    """Setting new exceptions relays out the batch; count 0 removes them."""
    sy = systems()[:2]
    with ATMBatch(sy) as b:
        b.compute()
        b.set_exceptions(0, np.zeros((0, 2), np.int32), [], [], [])
        b.compute()
        check_system(b, 0, dataclasses.replace(sy[0], exceptions=None), pairs=True, tag="removed", strict=False)
        check_system(b, 1, sy[1], pairs=False, tag="kept", strict=False)
        other = W.with_molecules(sy[0], size=2, crosslink_every=0)
        ex = other.exceptions
        b.set_exceptions(0, ex.pairs, ex.charge_prod, ex.sigma, ex.epsilon)
        b.compute()
        check_system(b, 0, other, pairs=True, tag="replaced", strict=False)


def test_exceptions_rejected_inputs(gpu):
    s = systems()[2]
    with ATMBatch([s]) as b:
        for pairs in ([[0, 0]], [[0, 64]], [[-1, 3]], [[1, 2], [2, 1]]):
            m = len(pairs)
            with pytest.raises(Exception):
                b.set_exceptions(0, pairs, [0.0] * m, [0.0] * m, [0.0] * m)
        with pytest.raises(Exception):
            b.set_exceptions(0, [[1, 2]], [0.0], [-0.1], [0.0])
        # the batch still evaluates with the original exceptions
        b.compute()
        check_system(b, 0, s, pairs=True, tag="after rejects", strict=False)


@pytest.mark.parametrize("device", ["cpu", "cuda"])
def test_exceptions_backward(gpu, device):
    sy = [W.with_molecules(W.tiny_system(1200, 3.0, 0.9, 10, (0.9, 0.8, 0.7), 40 + k,
                                         W.ATMParams(0.1 + 0.1 * k, 0.3 + 0.05 * k, 0.1, 110.0, 0.5, 200.0, 100.0,
                                                     0.0625, 1.0)), size=3 + k, crosslink_every=7)
          for k in range(2)]
    with ATMBatch(sy) as b:
        q = torch.tensor(np.concatenate([s.charges for s in sy]), device=device, requires_grad=True)
        sg = torch.tensor(np.concatenate([s.sigmas for s in sy]), device=device, requires_grad=True)
        ep = torch.tensor(np.concatenate([s.epsilons for s in sy]), device=device, requires_grad=True)
        sc = [torch.tensor([getattr(s.params, n) for s in sy], device=device, dtype=torch.float64, requires_grad=True)
              for n in ("lambda1", "lambda2", "alpha", "u0", "w0")]
        Wt = ATMEnergyFunction.apply(b, q, sg, ep, *sc)
        Wt.sum().backward()
        off = 0
        for k, s in enumerate(sy):
            assert Wt[k].item() == pytest.approx(oracle.atm_eval(s, forces=False)[0]["W"], rel=1e-9)
            ref = oracle.atm_grad(s)
            n = s.num_atoms
            for name, g in (("dq", q), ("dsigma", sg), ("deps", ep)):
                got = g.grad[off:off + n].cpu().numpy()
                err = np.abs(got - ref[name]).max() / np.abs(ref[name]).max()
                print(f"sys {k} {name}: max err / max |grad| = {err:.2e}", flush=True)
                assert err < 2e-5
            for j in range(5):
                assert sc[j].grad[k].item() == pytest.approx(ref["dparams"][j], rel=1e-6, abs=1e-9)
            off += n
