"""torch.autograd entry point of the batched ATM evaluation.

ATMEnergyFunction.apply(batch, charges, sigmas, epsilons, lambda1, lambda2,
alpha, u0, w0) -> W (num_systems,) float64.

  charges, sigmas, epsilons: per-atom tensors over all atoms of the batch
      (systems concatenated in order), any float dtype, any device;
  lambda1, lambda2, alpha, u0, w0: per-system tensors (num_systems,).

The forward uploads the parameters and runs one evaluation of every system on
the GPU.  The backward runs the hand-written backward kernel (atm_backward in
include/atm_mi355x.h), which walks the forward's cluster-pair lists again and
recomputes every in-cutoff pair term at x and at x + d, weighted by
dW/dU1 and dW/dU2 -- nothing per pair is stored between the passes -- and the
closed-form derivatives of the softplus in the five scalar parameters.
Pair terms of the backward are FP32, their sums FP64.

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


class ATMEnergyFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, batch: ATMBatch, charges, sigmas, epsilons, lambda1, lambda2, alpha, u0, w0):
        S = len(batch.systems)
        sizes = batch.num_atoms
        total = sum(sizes)
        for name, t in (("charges", charges), ("sigmas", sigmas), ("epsilons", epsilons)):
            if t.shape != (total,):
                raise ValueError(f"{name} must have shape ({total},), got {tuple(t.shape)}")
        scal = [lambda1, lambda2, alpha, u0, w0]
        for t in scal:
            if t.shape != (S,):
                raise ValueError(f"per-system parameters must have shape ({S},)")
        state = _upload(batch, _np(charges), _np(sigmas), _np(epsilons), [_np(t) for t in scal])
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
            _upload(batch, *ctx.state)
        g = batch.backward(_np(gW))
        outs = [g["dq"], g["dsigma"], g["deps"]] + [g["dparams"][:, k] for k in range(5)]
        grads = [torch.as_tensor(o).to(device=dev, dtype=dt) for o, (dt, dev) in zip(outs, ctx.meta)]
        need = ctx.needs_input_grad[1:]
        return (None,) + tuple(gr if n else None for gr, n in zip(grads, need))


def _upload(batch: ATMBatch, q, s, e, scal):
    off = 0
    for k, (sysk, n) in enumerate(zip(batch.systems, batch.num_atoms)):
        batch.set_particles(k, q[off:off + n], s[off:off + n], e[off:off + n], sysk.displacements)
        p = sysk.params
        batch.set_params(k, (scal[0][k], scal[1][k], scal[2][k], scal[3][k], scal[4][k], p.umax, p.ubcore, p.acore,
                             p.direction))
        off += n
    batch.compute(forces=False)
    return (q, s, e, scal)


def atm_energy(batch: ATMBatch, charges, sigmas, epsilons, lambda1, lambda2, alpha, u0, w0):
    """Functional form of ATMEnergyFunction.apply."""
    return ATMEnergyFunction.apply(batch, charges, sigmas, epsilons, lambda1, lambda2, alpha, u0, w0)
