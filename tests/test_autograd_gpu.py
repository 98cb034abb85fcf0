"""Differentiation: dW/d(charges, sigma, epsilon, lambda1, lambda2, alpha, u0, w0)
from the backward kernel behind ATMEnergyFunction, checked against the FP64
analytic oracle gradient and against FP64-scalar central differences of the
GPU energy (step 1e-4 in each parameter's unit)."""
import numpy as np
import pytest
import torch

from atmmetaforce import ATMBatch, ATMEnergyFunction, workloads as W
from oracle import oracle

pytestmark = pytest.mark.gpu


def tensors(systems, device="cpu"):
    q = torch.tensor(np.concatenate([s.charges for s in systems]), device=device, requires_grad=True)
    sg = torch.tensor(np.concatenate([s.sigmas for s in systems]), device=device, requires_grad=True)
    ep = torch.tensor(np.concatenate([s.epsilons for s in systems]), device=device, requires_grad=True)
    sc = [torch.tensor([getattr(s.params, n) for s in systems], device=device, dtype=torch.float64, requires_grad=True)
          for n in ("lambda1", "lambda2", "alpha", "u0", "w0")]
    return q, sg, ep, sc


def check_against_oracle(systems, grads, rtol):
    q, sg, ep, sc = grads
    off = 0
    for k, s in enumerate(systems):
        n = s.num_atoms
        ref = oracle.atm_grad(s)
        for name, g, r in (("dq", q, ref["dq"]), ("dsigma", sg, ref["dsigma"]), ("deps", ep, ref["deps"])):
            got = g[off:off + n]
            err = np.abs(got - r).max() / np.abs(r).max()
            print(f"sys {k} {name}: max err / max |grad| = {err:.2e}", flush=True)
            assert err < rtol, (k, name, err)
        for j in range(5):
            assert sc[j][k] == pytest.approx(ref["dparams"][j], rel=1e-6, abs=1e-9), (k, j)
        off += n


def test_small_batch_gradients(gpu):
    systems = [W.tiny_system(1200, 3.0, 0.9, 10, (0.9, 0.8, 0.7), 40 + k,
                             W.ATMParams(0.1 + 0.1 * k, 0.3 + 0.05 * k, 0.1, 110.0, 0.5, 200.0, 100.0, 0.0625,
                                         1.0 if k % 2 == 0 else -1.0))
               for k in range(3)]
    # reverse-direction members use a sparse box so the reference's softplus stays finite
    systems[1] = W.tiny_system(300, 3.4, 0.9, 8, (1.1, 0.7, 0.45), 41, systems[1].params)
    with ATMBatch(systems) as b:
        q, sg, ep, sc = tensors(systems)
        Wt = ATMEnergyFunction.apply(b, q, sg, ep, *sc)
        gw = torch.tensor([1.0, -0.5, 2.0], dtype=torch.float64)
        (Wt * gw).sum().backward()
        for k, s in enumerate(systems):
            assert Wt[k].item() == pytest.approx(oracle.atm_eval(s, forces=False)[0]["W"], rel=1e-9)
        grads = (q.grad.numpy() / np.repeat(gw.numpy(), [s.num_atoms for s in systems]),
                 sg.grad.numpy() / np.repeat(gw.numpy(), [s.num_atoms for s in systems]),
                 ep.grad.numpy() / np.repeat(gw.numpy(), [s.num_atoms for s in systems]),
                 [t.grad.numpy() / gw.numpy() for t in sc])
        check_against_oracle(systems, grads, rtol=2e-5)


def test_config5_replica_batch_gradients(gpu):
    """22 replicas of the 23k-atom system: oracle gradient on three replicas,
    central differences of the GPU energy on chosen atoms and all scalars."""
    reps = W.replica_batch(22)
    with ATMBatch(reps) as b:
        q, sg, ep, sc = tensors(reps)
        Wt = ATMEnergyFunction.apply(b, q, sg, ep, *sc)
        Wt.sum().backward()
        sel = [0, 7, 21]
        off = [sum(s.num_atoms for s in reps[:k]) for k in range(23)]
        for k in sel:
            s = reps[k]
            ref = oracle.atm_grad(s)
            for name, g, r in (("dq", q.grad, ref["dq"]), ("dsigma", sg.grad, ref["dsigma"]),
                               ("deps", ep.grad, ref["deps"])):
                got = g[off[k]:off[k + 1]].numpy()
                err = np.abs(got - r).max() / np.abs(r).max()
                print(f"replica {k} {name}: {err:.2e}", flush=True)
                assert err < 2e-5
            for j in range(5):
                assert sc[j].grad[k].item() == pytest.approx(ref["dparams"][j], rel=1e-6, abs=1e-9)
        # FP64 central differences (h = 1e-4) of the GPU W on replica 0, ligand atom 3 and solvent atom 5000
        h = 1e-4
        base_q = q.detach().clone()
        for idx in (3, 5000):
            for which, tens in (("q", q), ("sigma", sg), ("eps", ep)):
                vals = []
                for sgn in (1, -1):
                    args = [q.detach().clone(), sg.detach().clone(), ep.detach().clone()]
                    pos = {"q": 0, "sigma": 1, "eps": 2}[which]
                    args[pos][idx] += sgn * h
                    vals.append(ATMEnergyFunction.apply(b, *args, *[t.detach() for t in sc])[0].item())
                fd = (vals[0] - vals[1]) / (2 * h)
                got = tens.grad[idx].item()
                print(f"atom {idx} d{which}: backward {got:.8g} central {fd:.8g}", flush=True)
                assert got == pytest.approx(fd, rel=2e-4, abs=1e-4)
        del base_q
