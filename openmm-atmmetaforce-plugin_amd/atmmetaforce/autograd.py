"""torch.autograd entry point of the batched ATM evaluation.

ATMEnergyFunction.apply(batch, charges, sigmas, epsilons, lambda1, lambda2,
alpha, u0, w0) -> W (num_systems,) float64.

  charges, sigmas, epsilons: per-atom tensors over all atoms of the batch
      (systems concatenated in order), any float dtype, any device;
  lambda1, lambda2, alpha, u0, w0: per-system tensors (num_systems,).

The forward sets the parameters and runs one evaluation of every system on
the GPU.  Per-atom tensors on a GPU go to the library as device arrays
(atm_set_particles_device) and their gradients come back as device arrays
(atm_backward_device): no per-atom data crosses to the host.  When a per-atom
tensor requires a gradient, the forward evaluates with ATM_EVAL_PARAM_GRADS:
the pair kernel's backward mode computes the energies and, in the same pass
over the pairs, dU1/dp and dU2/dp per atom (p = charge, sigma, epsilon); the
backward then only weights them by dW/dU1 and dW/dU2 and adds the closed-form
derivatives of the softplus in the five scalar parameters.  Otherwise the
forward is an energy-only evaluation and the backward makes one weighted pass
over the forward's lists.  Pair terms are FP64 from the exact r^2, their
per-lane sums FP32, their per-atom sums FP64.

If the batch was evaluated again between a forward and its backward, the
backward re-runs that forward's evaluation first: at the forward's parameters
and at its positions and boxes (restored from the forward's snapshot; a
backward whose positions were since changed on the device raises).
"""
from __future__ import annotations

import numpy as np
import torch

from .batch import ATMBatch


def _np(t) -> np.ndarray:
    return t.detach().to("cpu", torch.float64).numpy()


def _dev64(t) -> torch.Tensor:
    return t.detach().to(torch.float64).contiguous()


class ATMEnergyFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, batch: ATMBatch, charges, sigmas, epsilons, lambda1, lambda2, alpha, u0, w0):
        S = len(batch.systems)
        total = sum(batch.num_atoms)
        for name, t in (("charges", charges), ("sigmas", sigmas), ("epsilons", epsilons)):
            if t.shape != (total,):
                raise ValueError(f"{name} must have shape ({total},), got {tuple(t.shape)}")
        scal = [lambda1, lambda2, alpha, u0, w0]
        for t in scal:
            if t.shape != (S,):
                raise ValueError(f"per-system parameters must have shape ({S},)")
        on_gpu = charges.is_cuda and sigmas.is_cuda and epsilons.is_cuda
        if on_gpu:
            atoms = (_dev64(charges), _dev64(sigmas), _dev64(epsilons))
        else:
            atoms = (_np(charges), _np(sigmas), _np(epsilons))
        grads = any(t.requires_grad for t in (charges, sigmas, epsilons))
        state = (on_gpu, atoms, [_np(t) for t in scal], grads)
        _upload(batch, state)
        W = np.array([batch.result(k)["energy"] for k in range(S)])
        ctx.batch = batch
        ctx.state = state
        ctx.positions = batch.position_snapshot()
        ctx.eval_id = batch._eval_id
        ctx.meta = [(t.dtype, t.device) for t in (charges, sigmas, epsilons, *scal)]
        return torch.as_tensor(W, dtype=torch.float64, device=charges.device)

    @staticmethod
    def backward(ctx, gW):
        batch = ctx.batch
        if batch._eval_id != ctx.eval_id:
            batch.restore_positions(ctx.positions)
            _upload(batch, ctx.state)
        g = _np(gW)
        on_gpu = ctx.state[0]
        need = ctx.needs_input_grad[1:]
        if on_gpu:
            total = sum(batch.num_atoms)
            dev = ctx.meta[0][1]
            outs_atoms = [torch.empty(total, dtype=torch.float64, device=dev) if need[k] else None for k in range(3)]
            dparams = batch.backward_device(g, *[o.data_ptr() if o is not None else 0 for o in outs_atoms])
        else:
            r = batch.backward(g)
            outs_atoms = [torch.as_tensor(r["dq"]), torch.as_tensor(r["dsigma"]), torch.as_tensor(r["deps"])]
            dparams = r["dparams"]
        # the five per-system parameter gradients cross to their device in one copy
        pdev = ctx.meta[3][1]
        dp = torch.as_tensor(dparams).to(pdev) if any(need[3:]) else None
        outs = outs_atoms + [dp[:, k] if dp is not None else None for k in range(5)]
        grads = [o.to(device=dev, dtype=dt) if (o is not None and n) else None
                 for o, (dt, dev), n in zip(outs, ctx.meta, need)]
        return (None,) + tuple(grads)


def _upload(batch: ATMBatch, state):
    on_gpu, atoms, scal, grads = state
    if on_gpu:
        q, s, e = atoms
        torch.cuda.current_stream(q.device).synchronize()  # the library reads on its own stream
        batch.set_particles_device(q.data_ptr(), s.data_ptr(), e.data_ptr())
    else:
        q, s, e = atoms
        off = 0
        for k, (sysk, n) in enumerate(zip(batch.systems, batch.num_atoms)):
            batch.set_particles(k, q[off:off + n], s[off:off + n], e[off:off + n], sysk.displacements)
            off += n
    for k, sysk in enumerate(batch.systems):
        p = sysk.params
        batch.set_params(k, (scal[0][k], scal[1][k], scal[2][k], scal[3][k], scal[4][k], p.umax, p.ubcore, p.acore,
                             p.direction))
    batch.compute(forces=False, param_grads=grads)


def atm_energy(batch: ATMBatch, charges, sigmas, epsilons, lambda1, lambda2, alpha, u0, w0):
    """Functional form of ATMEnergyFunction.apply."""
    return ATMEnergyFunction.apply(batch, charges, sigmas, epsilons, lambda1, lambda2, alpha, u0, w0)
