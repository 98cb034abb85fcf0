"""Replica exchange of lambda-states (atmmetaforce.replica) on the host side:
the bias B(u) against the oracle's combine, and the Metropolis exchange leaving
the joint distribution of state assignments invariant."""
import math
import types

import numpy as np
import pytest

from atmmetaforce import replica, workloads as W
from oracle import oracle


@pytest.mark.parametrize("p", [W.ATMParams(0.2, 0.45, 0.1, 110.0, 1.5), W.ATMParams(0.3, 0.3, 0.0, 0.0, 0.0),
                               W.ATMParams(0.0, 0.5, 0.05, 40.0, -3.0, 200.0, 100.0, 0.0625, 1.0)])
def test_bias_is_W_minus_e0(p):
    for U1, U2 in ((-1000.0, -990.0), (50.0, 400.0), (10.0, 10.0), (-5.0, 30000.0)):
        r = oracle.combine(p, U1, U2)
        assert replica.softplus_bias(p, r["usc"]) == pytest.approx(r["W"] - U1, rel=1e-12, abs=1e-9)


def test_exchange_ratio_properties():
    b = 1 / (replica.KB * 300.0)
    pi, pj = W.ATMParams(0.1, 0.2, 0.1, 100.0, 0.0), W.ATMParams(0.3, 0.4, 0.1, 100.0, 2.0)
    assert replica.exchange_log_ratio(b, pi, pi, 10.0, 50.0) == 0.0              # same state: always accepted
    assert replica.exchange_log_ratio(b, pi, pj, 30.0, 30.0) == pytest.approx(0.0, abs=1e-12)  # same u
    # reverse move has the opposite log ratio (detailed balance)
    assert replica.exchange_log_ratio(b, pi, pj, 10.0, 80.0) == pytest.approx(
        -replica.exchange_log_ratio(b, pi, pj, 80.0, 10.0))


def test_metropolis_exchanges_sample_the_assignment_distribution():
    """Fixed perturbation energies of 3 replicas; repeated exchange sweeps must
    visit the 6 assignments of states with probability prop. to
    exp(-beta sum_a B_{s(a)}(u_a))."""
    T = 300.0
    beta = 1 / (replica.KB * T)
    states = [W.ATMParams(0.0, 0.0, 0.1, 50.0, 0.0), W.ATMParams(0.2, 0.3, 0.1, 50.0, 0.0),
              W.ATMParams(0.4, 0.5, 0.1, 50.0, 0.0)]
    u = [2.0, 5.0, 9.0]

    class FakeBatch:
        def result(self, a):
            return {"perturbation_energy": u[a]}

        def set_params(self, a, p):
            pass

        def compute(self, forces=True):
            pass

    re = replica.ATMReplicaExchange.__new__(replica.ATMReplicaExchange)
    re.states, re.beta, re.batch = states, beta, FakeBatch()
    re.state_of = [0, 1, 2]
    re.rng = np.random.default_rng(7)
    re.attempts = np.zeros(2, dtype=np.int64)
    re.accepts = np.zeros(2, dtype=np.int64)
    counts = {}
    for _ in range(60000):
        re.exchange(1)
        key = tuple(re.state_of)
        counts[key] = counts.get(key, 0) + 1
    import itertools
    w = {perm: math.exp(-beta * sum(replica.softplus_bias(states[s], u[a]) for a, s in enumerate(perm)))
         for perm in itertools.permutations(range(3))}
    Z = sum(w.values())
    total = sum(counts.values())
    for perm, wt in w.items():
        assert counts.get(perm, 0) / total == pytest.approx(wt / Z, abs=0.01)
    assert 0 < re.acceptance().min() <= re.acceptance().max() <= 1


def test_context_replica_exchange_applies_states_and_samples():
    """ContextReplicaExchange on stand-in contexts (the API it drives:
    setParameter, getState, the force's parameter names and
    getPerturbationEnergy): each exchange leaves every context holding the
    parameters of its assigned state, and the assignments sample the
    Boltzmann weights of the softplus biases."""
    import itertools
    T = 300.0
    beta = 1 / (replica.KB * T)
    states = [W.ATMParams(0.0, 0.0, 0.1, 50.0, 0.0), W.ATMParams(0.2, 0.3, 0.1, 50.0, 0.0),
              W.ATMParams(0.4, 0.5, 0.1, 50.0, 0.0)]
    u = [2.0, 5.0, 9.0]
    names = replica.ContextReplicaExchange._NAMES

    class Force:
        def getForceGroup(self):
            return 2

        def getPerturbationEnergy(self, ctx):
            return ctx.u

    for n in names:
        setattr(Force, n, staticmethod(lambda n=n: "ATM" + n))

    class Ctx:
        def __init__(self, u):
            self.u, self.p = u, {}

        def setParameter(self, k, v):
            self.p[k] = v

        def getState(self, getEnergy=False, groups=None):
            return types.SimpleNamespace(getPotentialEnergy=lambda: 0.0, getKineticEnergy=lambda: 0.0)

    class Integ:
        def step(self, n):
            pass

    ctxs = [Ctx(x) for x in u]
    re = replica.ContextReplicaExchange([(c, Integ(), Force()) for c in ctxs], states, temperature=T, seed=3)
    counts = {}
    for _ in range(40000):
        re.exchange(1)
        for a, c in enumerate(ctxs):
            p = states[re.state_of[a]].as_tuple()
            assert all(c.p["ATM" + n] == v for n, v in zip(names, p))
        key = tuple(re.state_of)
        counts[key] = counts.get(key, 0) + 1
    w = {perm: math.exp(-beta * sum(replica.softplus_bias(states[s], u[a]) for a, s in enumerate(perm)))
         for perm in itertools.permutations(range(3))}
    Z = sum(w.values())
    for perm, wt in w.items():
        assert counts.get(perm, 0) / 40000 == pytest.approx(wt / Z, abs=0.012)
    lines = re.run(1, 10, exchange=False)
    assert len(lines) == 3 and all(len(s.out_line().split()) == 9 for s in lines)


def test_states_must_share_softcore_parameters():
    """The exchange ratio evaluates each state's bias at the other replica's u_sc,
    which that replica computed with its own soft-core map: a schedule whose
    umax, ubcore or acore differ between states is rejected."""
    from atmmetaforce.replica import _check_states
    base = [W.ATMParams(0.0, 0.1), W.ATMParams(0.1, 0.2)]
    _check_states(base)
    for field, value in (("umax", 250.0), ("ubcore", 80.0), ("acore", 0.125)):
        import dataclasses
        bad = [base[0], dataclasses.replace(base[1], **{field: value})]
        with pytest.raises(ValueError, match="soft-core"):
            _check_states(bad)
    with pytest.raises(ValueError, match="direction"):
        _check_states([base[0], W.ATMParams(0.1, 0.2, direction=-1.0)])
