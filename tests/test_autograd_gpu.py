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


@pytest.mark.parametrize("device", ["cpu", "cuda"])
def test_small_batch_gradients(gpu, device):
    systems = [W.tiny_system(1200, 3.0, 0.9, 10, (0.9, 0.8, 0.7), 40 + k,
                             W.ATMParams(0.1 + 0.1 * k, 0.3 + 0.05 * k, 0.1, 110.0, 0.5, 200.0, 100.0, 0.0625,
                                         1.0 if k % 2 == 0 else -1.0))
               for k in range(3)]
    # reverse-direction members use a sparse box so the reference's softplus stays finite
    systems[1] = W.tiny_system(300, 3.4, 0.9, 8, (1.1, 0.7, 0.45), 41, systems[1].params)
    with ATMBatch(systems) as b:
        q, sg, ep, sc = tensors(systems, device)
        Wt = ATMEnergyFunction.apply(b, q, sg, ep, *sc)
        assert Wt.device.type == device
        gw = torch.tensor([1.0, -0.5, 2.0], dtype=torch.float64, device=device)
        (Wt * gw).sum().backward()
        for k, s in enumerate(systems):
            assert Wt[k].item() == pytest.approx(oracle.atm_eval(s, forces=False)[0]["W"], rel=1e-9)
        gwn = gw.cpu().numpy()
        rep = np.repeat(gwn, [s.num_atoms for s in systems])
        grads = (q.grad.cpu().numpy() / rep, sg.grad.cpu().numpy() / rep, ep.grad.cpu().numpy() / rep,
                 [t.grad.cpu().numpy() / gwn for t in sc])
        check_against_oracle(systems, grads, rtol=2e-5)


def test_config5_replica_batch_gradients(gpu):
    """22 replicas of the 23k-atom system, parameters as device tensors: oracle
    gradient on three replicas, central differences of the GPU energy on chosen
    atoms and all scalars."""
    reps = W.replica_batch(22)
    with ATMBatch(reps) as b:
        q, sg, ep, sc = tensors(reps, "cuda")
        Wt = ATMEnergyFunction.apply(b, q, sg, ep, *sc)
        Wt.sum().backward()
        sel = [0, 7, 21]
        off = [sum(s.num_atoms for s in reps[:k]) for k in range(23)]
        for k in sel:
            s = reps[k]
            ref = oracle.atm_grad(s)
            for name, g, r in (("dq", q.grad, ref["dq"]), ("dsigma", sg.grad, ref["dsigma"]),
                               ("deps", ep.grad, ref["deps"])):
                got = g[off[k]:off[k + 1]].cpu().numpy()
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


@pytest.mark.parametrize("device", ["cpu", "cuda"])
def test_zero_epsilon_and_sigma_sites(gpu, device):
    """Force fields have sites with eps = 0 (TIP3P hydrogens) and sigma = 0.
    dW/dsigma at sigma_i + sigma_j = 0 is 0 (the limit of 4 eps (12 s^11/r^12 -
    6 s^5/r^6)); dW/deps of an eps = 0 atom does not exist and is returned as 0,
    the oracle's convention.  Every gradient is finite and matches the oracle."""
    s = W.tiny_system(1500, 3.0, 0.9, 12, (0.9, 0.8, 0.7), 61,
                      W.ATMParams(0.2, 0.35, 0.1, 110.0, 0.5, 200.0, 100.0, 0.0625, 1.0))
    eps = s.epsilons.copy()
    sig = s.sigmas.copy()
    eps[::5] = 0.0
    sig[::7] = 0.0        # every 7th atom: pairs of two such atoms have sigma_ij = 0
    eps[::7] = 0.3
    import dataclasses
    s = dataclasses.replace(s, epsilons=eps, sigmas=sig)
    with ATMBatch([s]) as b:
        q, sg, ep, sc = tensors([s], device)
        Wt = ATMEnergyFunction.apply(b, q, sg, ep, *sc)
        Wt.sum().backward()
        for g in (q.grad, sg.grad, ep.grad):
            assert torch.isfinite(g).all()
        assert (ep.grad.cpu().numpy()[eps == 0.0] == 0.0).all()
        check_against_oracle([s], (q.grad.cpu().numpy(), sg.grad.cpu().numpy(), ep.grad.cpu().numpy(),
                                   [t.grad.cpu().numpy() for t in sc]), rtol=2e-5)


@pytest.mark.parametrize("device", ["cpu", "cuda"])
def test_backward_after_re_evaluation_at_other_positions(gpu, device):
    """forward at x1, then the batch is evaluated at x2, then the backward of
    the first forward: the gradient is the one at x1 (the backward restores
    the forward's positions before it re-runs the evaluation)."""
    s1 = W.tiny_system(1200, 3.0, 0.9, 10, (0.9, 0.8, 0.7), 63,
                       W.ATMParams(0.2, 0.35, 0.1, 110.0, 0.5, 200.0, 100.0, 0.0625, 1.0))
    x2 = W.quantize(s1.positions + np.random.default_rng(5).normal(0.0, 0.05, s1.positions.shape))
    with ATMBatch([s1]) as b:
        q, sg, ep, sc = tensors([s1], device)
        Wt = ATMEnergyFunction.apply(b, q, sg, ep, *sc)
        b.set_positions(0, x2)
        b.compute()
        assert b.result(0)["energy"] != pytest.approx(Wt[0].item(), rel=1e-6)
        Wt.sum().backward()
        check_against_oracle([s1], (q.grad.cpu().numpy(), sg.grad.cpu().numpy(), ep.grad.cpu().numpy(),
                                    [t.grad.cpu().numpy() for t in sc]), rtol=2e-5)


def test_fused_parameter_gradients_match_the_separate_pass(gpu):
    """ATM_EVAL_PARAM_GRADS: the energies and pair counts of the evaluation equal
    an energy-only evaluation's, and atm_backward after it (weights and gather
    only) equals atm_backward after an energy-only evaluation (a second, weighted
    pass over the pairs), to the FP64 rounding of the per-slot sums."""
    systems = [W.tiny_system(1500, 3.0, 0.9, 12, (0.9, 0.8, 0.7), 60 + k,
                             W.ATMParams(0.2 + 0.1 * k, 0.4, 0.1, 110.0, 0.5, 200.0, 100.0, 0.0625, 1.0))
               for k in range(3)]
    systems.append(W.with_molecules(W.tiny_system(1200, 3.0, 0.9, 10, (0.9, 0.8, 0.7), 63, systems[0].params),
                                    size=3, crosslink_every=7))
    g = np.array([1.0, -0.7, 2.5, 0.3])
    with ATMBatch(systems) as b:
        b.compute(forces=False)
        e0 = [b.result(k) for k in range(len(systems))]
        n0 = b.list_stats()["pairs"].copy()
        r0 = b.backward(g)
        b.compute(forces=False, param_grads=True)
        n1 = b.list_stats()["pairs"]
        assert np.array_equal(n0, n1)
        for k in range(len(systems)):
            r = b.result(k)
            for key in ("energy", "state1_energy", "state2_energy"):
                assert r[key] == pytest.approx(e0[k][key], rel=1e-12)
        r1 = b.backward(g)
        for key in ("dq", "dsigma", "deps"):
            err = np.abs(r1[key] - r0[key]).max() / np.abs(r0[key]).max()
            print(f"{key}: fused vs separate {err:.2e}", flush=True)
            assert err < 1e-11
        assert np.allclose(r1["dparams"], r0["dparams"], rtol=1e-12, atol=1e-14)
        with pytest.raises(Exception):
            b.evaluate(forces=True, param_grads=True)


def test_partial_gradient_outputs_on_device(gpu):
    """Only the charges require a gradient: the device backward writes that array
    alone (the others are not allocated) and it matches the oracle."""
    s = W.tiny_system(1200, 3.0, 0.9, 10, (0.9, 0.8, 0.7), 71,
                      W.ATMParams(0.2, 0.4, 0.1, 110.0, 0.5, 200.0, 100.0, 0.0625, 1.0))
    with ATMBatch([s]) as b:
        q, sg, ep, sc = tensors([s], "cuda")
        sg.requires_grad_(False)
        ep.requires_grad_(False)
        Wt = ATMEnergyFunction.apply(b, q, sg, ep, *sc)
        Wt.sum().backward()
        ref = oracle.atm_grad(s)
        err = np.abs(q.grad.cpu().numpy() - ref["dq"]).max() / np.abs(ref["dq"]).max()
        assert err < 2e-5
        assert sg.grad is None and ep.grad is None
