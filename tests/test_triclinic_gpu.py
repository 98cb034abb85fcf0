"""Triclinic boxes on the MI355X path against the FP64 oracle (through the C ABI).

The reference copies the full box vectors into both inner contexts
(platforms/reference/src/ReferenceATMMetaForceKernels.cpp:126-129); a triclinic
box in OpenMM's reduced form reaches the library through atm_set_box_vectors.
Same bar as the orthorhombic parity tests: pair sets identical to the oracle's
(oracle_pairs_tri, pinned in tests/test_oracle_triclinic.py), U1, U2, W within
1e-6 relative and the FP64 bound of the default path, forces within 1e-5
relative RMS.  Positions and tilts sit on the 2^-20 nm grid, where every step of
the triclinic wrap and image is exact in FP32, so the pair sets are integer
objects both paths must reproduce.
"""
import numpy as np
import pytest

from atmmetaforce import ATMBatch, workloads as W
from atmmetaforce import _capi as C
from oracle import oracle

from test_gpu_parity import check_system

pytestmark = pytest.mark.gpu


def tri_config1():
    return W.triclinic(W.config1(), sites=(5, -4, 7))


@pytest.mark.parametrize("plain", [False, True])
def test_triclinic_3000_atoms(gpu, plain):
    s = tri_config1()
    with ATMBatch([s]) as b:
        b.compute(plain=plain)
        check_system(b, 0, s, tag=f"tri-3000 plain={plain}")


def test_triclinic_with_exceptions_and_fp32_pairs(gpu):
    s = W.triclinic(W.with_molecules(W.config1(seed=3), size=4, crosslink_every=7), sites=(-6, 3, -5))
    with ATMBatch([s]) as b:
        b.compute()
        check_system(b, 0, s, tag="tri-exc")
        b.compute(fp32_pairs=True)
        check_system(b, 0, s, pairs=False, strict=False, fp32=True, tag="tri-exc fp32")


def test_triclinic_23k_in_a_mixed_batch(gpu):
    """A 23k-atom triclinic system between orthorhombic ones, autotuned: the batch
    runs the tile kernel and bins x + d in full; every system still matches."""
    big = W.triclinic(W.config2(), sites=(9, -7, 12))
    ortho = W.tiny_system(3000, 3.1, 0.9, 24, (1.5, 1.5, 1.5), 8)
    small = W.triclinic(W.tiny_system(1500, 2.9, 0.9, 16, (0.9, 0.8, 0.7), 9), sites=(-2, 5, -3))
    systems = [ortho, big, small]
    with ATMBatch(systems) as b:
        b.autotune(repeats=1)
        b.compute()
        counts = b.list_stats()["pairs"]
        for k, s in enumerate(systems):
            check_system(b, k, s, tag="tri-mixed", counts=counts)


def test_lattice_translations_give_identical_results(gpu):
    """Moving atoms by lattice vectors of the triclinic box (a, b, c and sums) leaves
    the wrapped positions, and so every result, bit for bit the same."""
    s = tri_config1()
    H = W.box_vectors(s.box, s.tilt)
    k = np.random.default_rng(21).integers(-1, 2, size=(s.num_atoms, 3))
    moved = W.System(**{**s.__dict__, "positions": W.quantize(s.positions + k @ H)})
    out = []
    for sys_ in (s, moved):
        with ATMBatch([sys_]) as b:
            b.compute()
            out.append((b.result(0), b.forces_fixed(0)))
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1])


def test_box_vector_switches(gpu):
    """set_box_vectors on a context built orthorhombic, back to orthorhombic with
    set_box, and the reduced-form checks."""
    s = tri_config1()
    flat = W.System(**{**s.__dict__, "tilt": None})
    with ATMBatch([flat]) as b:
        b.set_box_vectors(0, W.box_vectors(s.box, s.tilt))
        b.compute()
        check_system(b, 0, s, tag="tri-switch")
        b.set_box(0, flat.box)
        b.compute()
        check_system(b, 0, flat, tag="ortho-again")
        bad = W.box_vectors(s.box, s.tilt)
        bad[1, 0] = 0.6 * s.box[0]  # |b_x| > a_x / 2
        with pytest.raises(C.ATMError):
            b.set_box_vectors(0, bad)
        bad = W.box_vectors(s.box, s.tilt)
        bad[0, 1] = 0.1            # a not along x
        with pytest.raises(C.ATMError):
            b.set_box_vectors(0, bad)


def test_openmm_api_triclinic_context(gpu):
    """ATMMetaForce over a CutoffPeriodic NonbondedForce (exceptions included) in a
    System with triclinic default box vectors, then Context.setPeriodicBoxVectors
    to another triclinic box, then an orthorhombic one."""
    from test_boundary_gpu import build, frms, rel
    import atmmetaforce as atm
    from atmmetaforce import openmm as mm

    s = W.triclinic(W.with_molecules(W.config1(seed=5), size=3), sites=(4, 6, -3))
    system, f = build(s)
    H = W.box_vectors(s.box, s.tilt)
    system.setDefaultPeriodicBoxVectors(tuple(H[0]), tuple(H[1]), tuple(H[2]))
    ctx = mm.Context(system, mm.VerletIntegrator(0.001), mm.Platform.getPlatformByName("MI355X"), {})
    ctx.setPositions(s.positions)
    for case in ("default", "other tilt", "orthorhombic"):
        if case == "other tilt":
            s = W.triclinic(W.System(**{**s.__dict__, "tilt": None}), sites=(-7, 2, 5))
        elif case == "orthorhombic":
            s = W.System(**{**s.__dict__, "tilt": None})
        H = W.box_vectors(s.box, s.tilt)
        if case != "default":
            ctx.setPeriodicBoxVectors(tuple(H[0]), tuple(H[1]), tuple(H[2]))
        ref, Fref = oracle.atm_eval(s)
        st = ctx.getState(getEnergy=True, getForces=True, groups={2})
        print(case, st.getPotentialEnergy(), ref["W"], flush=True)
        assert rel(st.getPotentialEnergy(), ref["W"]) < 1e-6
        assert frms(st.getForces(asNumpy=True), Fref) < 1e-5
        assert rel(f.getPerturbationEnergy(ctx), ref["usc"]) < 1e-6
    with pytest.raises(Exception, match="reduced form"):
        ctx.setPeriodicBoxVectors((3.1, 0, 0), (1.6, 3.1, 0), (0, 0, 3.1))


def test_triclinic_state_difference_and_full_state2(gpu):
    """x + d as the difference from x at the displaced atoms (the default: the
    displaced-atom kernel sweeps the x cells over the triclinic images) and x + d
    binned and paired in full both match the oracle, pair sets included."""
    s = W.triclinic(W.with_molecules(W.config1(seed=6), size=3, crosslink_every=11), sites=(-5, -6, 4))
    with ATMBatch([s]) as b:
        b.compute()
        ls = b.list_stats()
        assert ls["max_entries"][1] == 0, "x + d should be the difference from x"
        r_d, _ = check_system(b, 0, s, tag="tri difference")
        t = b.tuning()
        b.set_tuning(t["column_scale"], t["pair_waves_per_block"], t["zbins_per_column"], 0, t["list_capacity"],
                     t["pair_split"], full_state2=True, pair_order=t["pair_order"])
        b.compute()
        assert b.list_stats()["max_entries"][1] > 0
        r_f, _ = check_system(b, 0, s, tag="tri full")
        for key in ("state1_energy", "state2_energy", "energy"):
            assert abs(r_d[key] - r_f[key]) <= 1e-11 * abs(r_f[key])


def test_triclinic_parameter_gradients_against_oracle_differences(gpu):
    """dW/d(q, sigma, epsilon, lambda1) from the backward pass in a triclinic box
    against central differences of the FP64 oracle's W (oracle_atm_eval_tri), on
    ligand atoms and on atoms at the faces of the box, whose pairs take lattice
    images with tilts."""
    import torch
    from atmmetaforce import ATMEnergyFunction

    p = W.ATMParams(0.2, 0.45, 0.1, 110.0, 0.5, 200.0, 100.0, 0.0625, 1.0)
    s = W.triclinic(W.tiny_system(800, 2.9, 0.9, 10, (0.9, 0.8, 0.7), 31, p), sites=(3, -2, 4))
    dev = torch.device("cuda:0")
    q = torch.tensor(s.charges, device=dev, requires_grad=True)
    sg = torch.tensor(s.sigmas, device=dev, requires_grad=True)
    ep = torch.tensor(s.epsilons, device=dev, requires_grad=True)
    sc = [torch.tensor([getattr(p, n)], device=dev, dtype=torch.float64, requires_grad=True)
          for n in ("lambda1", "lambda2", "alpha", "u0", "w0")]
    with ATMBatch([s]) as b:
        Wt = ATMEnergyFunction.apply(b, q, sg, ep, *sc)
        Wt.sum().backward()
    W0 = oracle.atm_eval(s, forces=False)[0]["W"]
    assert Wt.item() == pytest.approx(W0, rel=1e-9)
    z, x = s.positions[:, 2], s.positions[:, 0]
    atoms = [0, 4, 9] + list(np.argsort(z)[-2:]) + list(np.argsort(x)[:2])

    def w_at(**kw):
        return oracle.atm_eval(W.System(**{**s.__dict__, **kw}), forces=False)[0]["W"]

    for name, arr, grad, h in (("charges", s.charges, q.grad, 1e-5), ("sigmas", s.sigmas, sg.grad, 1e-6),
                               ("epsilons", s.epsilons, ep.grad, 1e-5)):
        g = grad.cpu().numpy()
        scale = np.abs(g).max()
        for a in atoms:
            up, dn = arr.copy(), arr.copy()
            up[a] += h
            dn[a] -= h
            fd = (w_at(**{name: up}) - w_at(**{name: dn})) / (2 * h)
            print(f"{name}[{a}]: backward {g[a]:.8g} vs oracle differences {fd:.8g}", flush=True)
            assert abs(g[a] - fd) <= 2e-5 * scale, (name, a)
    h = 1e-6
    import dataclasses
    fd = (w_at(params=dataclasses.replace(p, lambda1=p.lambda1 + h))
          - w_at(params=dataclasses.replace(p, lambda1=p.lambda1 - h))) / (2 * h)
    assert sc[0].grad.item() == pytest.approx(fd, rel=1e-5, abs=1e-6)


@pytest.mark.parametrize("variant,split", [(1, 1), (5, 1), (6, 1), (0, 4)])
def test_triclinic_runs_the_tile_kernel_for_every_variant(gpu, variant, split):
    """A tuning that asks for another pair-kernel variant (compacted pairs, atom
    lists, filtered queues: orthorhombic images only) runs the tile kernel in a
    batch with a triclinic system; list splits apply as usual."""
    s = tri_config1()
    with ATMBatch([s]) as b:
        b.set_tuning(1.3, 2, 4, variant, pair_split=split)
        b.compute()
        check_system(b, 0, s, tag=f"tri variant {variant} split {split}")
