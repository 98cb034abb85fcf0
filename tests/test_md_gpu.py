"""Batched Langevin dynamics (atm_md_*, atm_md.hip) and the lambda-replica
exchange driver (atmmetaforce.replica) on the MI355X."""
import numpy as np
import pytest

from atmmetaforce import replica, workloads as W
from oracle import oracle

pytestmark = pytest.mark.gpu

KB = replica.KB


def _system(n=600, box=3.2, seed=41, lig=6):
    return W.tiny_system(n, box, 0.9, lig, (1.05, 0.85, 0.7), seed, W.ATMParams(0.0, 0.0, 0.0, 0.0, 0.0))


def test_leapfrog_matches_fp64_reference(gpu):
    """friction 0: the splitting reduces to kick-drift (v += dt F/m; x += dt v).
    Three steps against the same update driven by the FP64 oracle's forces."""
    s = _system()
    m = np.full(s.num_atoms, 12.0)
    with replica.ATMReplicaExchange(s, [s.params], m, temperature=300.0, friction=0.0, dt=0.001, seed=3) as re:
        x, v, _ = re.state(0)
        assert np.array_equal(x, s.positions)
        re.step(3)
        xg, vg, _ = re.state(0)
    sys_k = s
    for _ in range(3):
        _, F = oracle.atm_eval(sys_k)
        v = v + 0.001 * F / 12.0
        x = x + 0.001 * v
        sys_k = W.System(**{**sys_k.__dict__, "positions": x})
    assert np.abs(xg - x).max() < 2e-6
    assert np.abs(vg - v).max() / np.abs(v).max() < 2e-4


def test_energy_tracks_fp64_reference_without_friction(gpu):
    """200 friction-free steps: the energy series (W + kinetic energy of the
    half-step velocities) follows the same leapfrog driven by the FP64
    oracle, and stays bounded (no heating from the integrator)."""
    s = _system(500, 3.4, 43)
    m = np.full(s.num_atoms, 16.0)
    dt = 0.0005
    with replica.ATMReplicaExchange(s, [s.params], m, temperature=200.0, friction=0.0, dt=dt, seed=5) as re:
        x, v, _ = re.state(0)
        Eg = []
        for _ in range(10):
            re.step(20)
            _, _, ke = re.state(0)
            Eg.append(ke + re.batch.result(0)["energy"])
    Er = []
    sys_k = s
    for step in range(1, 201):
        _, F = oracle.atm_eval(sys_k)
        v = v + dt * F / 16.0
        x = x + dt * v
        sys_k = W.System(**{**sys_k.__dict__, "positions": x})
        if step % 20 == 0:
            ref, _ = oracle.atm_eval(sys_k, forces=False)
            Er.append(0.5 * 16.0 * (v * v).sum() + ref["W"])
    ke_scale = 1.5 * s.num_atoms * KB * 200.0
    assert np.abs(np.array(Eg) - np.array(Er)).max() < 0.005 * ke_scale, (Eg, Er)
    assert np.ptp(Eg) < 0.05 * ke_scale


def test_reproducible_and_thermostatted(gpu):
    s = _system(400, 3.6, 45)
    m = np.full(s.num_atoms, 12.0)
    T = 300.0
    finals = []
    for rep in range(2):
        with replica.ATMReplicaExchange(s, [s.params], m, temperature=T, friction=20.0, dt=0.002, seed=11) as re:
            re.step(300)
            kes = []
            for _ in range(30):
                re.step(20)
                kes.append(re.state(0)[2])
            finals.append(re.state(0)[0])
    assert np.array_equal(finals[0], finals[1])           # same seed: the same trajectory bit for bit
    assert np.mean(kes) == pytest.approx(1.5 * s.num_atoms * KB * T, rel=0.06)


def test_replica_exchange_run(gpu):
    s = _system(700, 3.0, 47, lig=8)
    states = [W.ATMParams(l1, l2, 0.1, 30.0, 0.0) for l1, l2 in ((0.0, 0.0), (0.0, 0.1), (0.1, 0.2), (0.2, 0.3))]
    m = np.full(s.num_atoms, 12.0)
    import io
    buf = io.StringIO()
    with replica.ATMReplicaExchange(s, states, m, temperature=300.0, friction=5.0, dt=0.001, seed=9) as re:
        samples = re.run(6, 25, exchange=True, sweeps=2, out=buf)
        assert sorted(re.state_of) == [0, 1, 2, 3]
        # results and forces follow the states after the exchanges
        for a in range(4):
            p = states[re.state_of[a]]
            x, _, _ = re.state(a)
            ref, _ = oracle.atm_eval(W.System(**{**s.__dict__, "positions": x, "params": p}))
            assert re.batch.result(a)["energy"] == pytest.approx(ref["W"], rel=1e-5, abs=1e-3)
        acc = re.acceptance()
        assert ((0.0 <= acc) & (acc <= 1.0)).all() and re.attempts.sum() == 6 * 2 * 3
    lines = buf.getvalue().strip().split("\n")
    assert len(lines) == 6 * 4 and len(samples) == 24
    assert all(len(l.split()) == 9 for l in lines)
    assert float(lines[0].split()[0]) == 300.0


def _md_pair(s, m, steps, capacity_from_init=False):
    """Positions after `steps` Langevin steps; optionally with the list capacity cut to
    the longest group list of the initial positions (so a later step overflows)."""
    with replica.ATMReplicaExchange(s, [s.params], m, temperature=600.0, friction=2.0, dt=0.002, seed=17) as re:
        cap0 = None
        if capacity_from_init:
            cap0 = int(re.batch.list_stats()["max_entries"].max())
            t = re.batch.tuning()
            re.batch.set_tuning(t["column_scale"], t["pair_waves_per_block"], t["zbins_per_column"],
                                t["pair_kernel"], list_capacity=cap0, pair_split=t["pair_split"])
        re.step(steps)
        x = re.state(0)[0]
        cap1 = int(re.batch.list_stats()["capacity"].max())
        print(f"list capacity {cap0} -> {cap1}", flush=True)
    return x, cap0, cap1


def test_list_overflow_inside_a_chunk_is_replayed(gpu):
    """atm_md_step checks list overflow once per chunk of steps.  The overflow
    flag is sticky over the chunk, so an overflow on any step (here: the list
    capacity is cut to the longest list at the initial positions, and the
    atoms then move at 600 K) is seen, the capacity grows and the steps since
    the chunk's snapshot are replayed: the trajectory equals one run with
    ample capacity bit for bit."""
    s = _system(1500, 3.4, 49, lig=10)
    m = np.full(s.num_atoms, 4.0)
    x_ref, _, _ = _md_pair(s, m, 40)
    x, cap0, cap1 = _md_pair(s, m, 40, capacity_from_init=True)
    assert cap1 > cap0  # an overflow happened and was replayed
    assert np.array_equal(x, x_ref)


def test_set_particles_during_md_keeps_the_md_positions(gpu):
    """New particle parameters during dynamics leave the integrator's positions
    on the device: the next evaluation is at the atm_md_get positions."""
    s = _system(600, 3.2, 53)
    m = np.full(s.num_atoms, 12.0)
    with replica.ATMReplicaExchange(s, [s.params], m, temperature=300.0, friction=5.0, dt=0.002, seed=19) as re:
        re.step(60)
        x, _, _ = re.state(0)
        q2 = s.charges * 0.5
        re.batch.set_particles(0, q2, s.sigmas, s.epsilons, s.displacements)
        re.batch.compute()
        got = re.batch.result(0)
    ref, _ = oracle.atm_eval(W.System(**{**s.__dict__, "positions": x, "charges": q2}), forces=False)
    # device positions are the FP32 roundings of the FP64 MD positions
    assert got["state1_energy"] == pytest.approx(ref["U1"], rel=1e-5, abs=1e-2)
    assert got["energy"] == pytest.approx(ref["W"], rel=1e-5, abs=1e-2)
