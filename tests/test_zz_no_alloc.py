"""The no-allocation entry point atm_eval_noalloc against atm_eval.

atm_eval_noalloc runs atm_eval's kernels with the same default launch
configuration on one system, with every intermediate buffer in a static device
array instead of context allocations, so its results must equal a one-system
context's bit for bit (the same cluster composition, list order and 64-bit
fixed-point force sums) when that context also bins and pairs x + d in full
(full_state2); the context's default mode (x + d as the difference from x)
agrees to rounding.  Checked on two calls in a row, the second on a
smaller system (the scratch keeps the first call's contents, which the second
must not see), on an accumulating call, on a call captured in a HIP graph and
replayed, and on an invalid input.
"""
import ctypes

import numpy as np
import pytest

from atmmetaforce import ATMBatch, workloads as W
from atmmetaforce import _capi as C

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

RESULT_KEYS = ("energy", "perturbation_energy", "state1_energy", "state2_energy", "raw_perturbation",
               "softcore_derivative", "softplus_derivative", "hybrid_weight")


def _system(n, seed, lam=(0.2, 0.4)):
    L = (n / 33.75) ** (1.0 / 3.0)
    p = W.ATMParams(lam[0], lam[1], 0.1, 110.0, 0.0, 200.0, 100.0, 0.0625, 1.0)
    return W.lattice_system(n, L, 1.0, 20, (L / 4, L / 4, L / 4), seed, p)


class _Inputs:
    """Device copies of one system's inputs and the output buffers."""

    def __init__(self, s, forces=None):
        dev = torch.device("cuda:0")
        f64 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
        self.s = s
        self.x = f64(s.positions)
        self.q, self.sig, self.eps = f64(s.charges), f64(s.sigmas), f64(s.epsilons)
        self.d = f64(s.displacements)
        self.F = f64(forces if forces is not None else np.zeros((s.num_atoms, 3)))
        self.res = torch.full((8,), float("nan"), dtype=torch.float64, device=dev)
        self.status = torch.full((1,), 12345, dtype=torch.int32, device=dev)
        self.desc = C.SystemDesc(s.num_atoms, (ctypes.c_double * 3)(*[float(v) for v in s.box]), float(s.cutoff),
                                 float(s.rf_dielectric))
        self.params = C.Params(*s.params.as_tuple())

    def call(self, flags, stream=None):
        lib = C.load()
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        C.check(lib.atm_eval_noalloc(ctypes.byref(self.desc), ctypes.byref(self.params), self.x.data_ptr(),
                                     self.q.data_ptr(), self.sig.data_ptr(), self.eps.data_ptr(),
                                     self.d.data_ptr(), flags, self.F.data_ptr(), self.res.data_ptr(),
                                     self.status.data_ptr(), ctypes.c_void_p(st)))

    def read(self):
        torch.cuda.synchronize()
        return (dict(zip(RESULT_KEYS, self.res.cpu().numpy().tolist())), self.F.cpu().numpy().copy(),
                int(self.status.item()))


def _context(s, full_state2=True):
    """atm_eval's results for s.  The no-allocation entry point bins and pairs x + d in
    full, so the bit-for-bit comparison is with that mode (full_state2); the default
    mode (x + d as the difference from x) agrees to rounding."""
    with ATMBatch([s]) as b:
        b.set_tuning(full_state2=full_state2)
        b.compute()
        return b.result(0), b.forces(0)


def test_two_calls_match_atm_eval(gpu):
    """A 9000-atom call, then a 3500-atom call: each equals atm_eval exactly."""
    F_ALL = C.ATM_EVAL_FORCES | C.ATM_EVAL_CLEAR_FORCES
    for n, seed in ((9000, 71), (3500, 72)):
        s = _system(n, seed)
        r_ref, F_ref = _context(s)
        inp = _Inputs(s, forces=np.full((n, 3), 7.0))  # cleared by ATM_EVAL_CLEAR_FORCES
        inp.call(F_ALL)
        r, F, status = inp.read()
        print(f"n={n}: W={r['energy']!r} (atm_eval {r_ref['energy']!r}), status {status}", flush=True)
        assert status == C.ATM_OK
        for k in RESULT_KEYS:
            assert r[k] == r_ref[k], (n, k, r[k], r_ref[k])
        assert np.array_equal(F, F_ref), (n, float(np.abs(F - F_ref).max()))
        r_d, F_d = _context(s, full_state2=False)
        for k in ("energy", "state1_energy", "state2_energy"):
            assert abs(r[k] - r_d[k]) <= 1e-11 * max(1.0, abs(r_d[k])), (n, k, r[k], r_d[k])
        assert np.linalg.norm(F - F_d) / np.linalg.norm(F_d) < 1e-6


def test_accumulates_and_energy_only(gpu):
    """Without ATM_EVAL_CLEAR_FORCES the force is added to the caller's array;
    without ATM_EVAL_FORCES the array is untouched and the energies agree."""
    s = _system(5000, 73)
    r_ref, F_ref = _context(s)
    base = np.arange(3 * s.num_atoms, dtype=np.float64).reshape(-1, 3) * 0.25
    inp = _Inputs(s, forces=base)
    inp.call(C.ATM_EVAL_FORCES)
    r, F, status = inp.read()
    assert status == C.ATM_OK
    assert np.array_equal(F, base + F_ref)
    inp2 = _Inputs(s, forces=base)
    inp2.call(0)
    r2, F2, status2 = inp2.read()
    assert status2 == C.ATM_OK
    assert np.array_equal(F2, base)
    for k in ("energy", "state1_energy", "state2_energy", "perturbation_energy"):
        assert r2[k] == r_ref[k], k


def test_graph_capture_replays(gpu):
    """Captured once in a HIP graph (no allocation, no memset nodes), replayed with
    new positions copied into the same input array between replays: every
    replay equals the eager call on those positions."""
    s = _system(4000, 74)
    inp = _Inputs(s)
    flags = C.ATM_EVAL_FORCES | C.ATM_EVAL_CLEAR_FORCES
    inp.call(flags)  # first call on the device: looks the scratch address up
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    cap.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(cap):
        g.capture_begin()
        inp.call(flags, stream=cap.cuda_stream)
        g.capture_end()
    torch.cuda.current_stream().wait_stream(cap)
    rng = np.random.default_rng(5)
    for rep in range(3):
        x = W.quantize(np.mod(s.positions + rng.normal(0.0, 0.01, s.positions.shape), s.box))
        inp.x.copy_(torch.as_tensor(x, device=inp.x.device))
        inp.res.fill_(float("nan"))
        g.replay()
        r, F, status = inp.read()
        eager = _Inputs(W.System(**{**s.__dict__, "positions": x}))
        eager.call(flags)
        re, Fe, se = eager.read()
        print(f"replay {rep}: W={r['energy']!r} eager {re['energy']!r}", flush=True)
        assert status == C.ATM_OK and se == C.ATM_OK
        for k in RESULT_KEYS:
            assert r[k] == re[k], (rep, k)
        assert np.array_equal(F, Fe), rep


def test_invalid_input_status(gpu):
    """A negative sigma gives ATM_ERR_INVALID_ARGUMENT in the status word, NaN
    results and untouched forces; the next valid call is unaffected."""
    s = _system(2000, 75)
    bad = W.System(**{**s.__dict__, "sigmas": s.sigmas.copy()})
    bad.sigmas[17] = -0.3
    inp = _Inputs(bad, forces=np.ones((s.num_atoms, 3)))
    inp.call(C.ATM_EVAL_FORCES)
    r, F, status = inp.read()
    assert status == C.ATM_ERR_INVALID_ARGUMENT
    assert all(np.isnan(v) for v in r.values())
    assert np.array_equal(F, np.ones((s.num_atoms, 3)))
    r_ref, F_ref = _context(s)
    ok = _Inputs(s)
    ok.call(C.ATM_EVAL_FORCES | C.ATM_EVAL_CLEAR_FORCES)
    r2, F2, status2 = ok.read()
    assert status2 == C.ATM_OK and r2["energy"] == r_ref["energy"] and np.array_equal(F2, F_ref)


def test_limits_rejected_on_host(gpu):
    """Systems above the atom limit are refused before any launch."""
    n = 32769
    L = (n / 33.75) ** (1.0 / 3.0)
    desc = C.SystemDesc(n, (ctypes.c_double * 3)(L, L, L), 1.0, 78.5)
    p = C.Params(0.5, 0.5, 0.1, 110.0, 0.0, 200.0, 100.0, 0.0625, 1.0)
    dummy = torch.zeros(8, dtype=torch.float64, device="cuda:0")
    st = torch.zeros(1, dtype=torch.int32, device="cuda:0")
    rc = C.load().atm_eval_noalloc(ctypes.byref(desc), ctypes.byref(p), dummy.data_ptr(), dummy.data_ptr(),
                                   dummy.data_ptr(), dummy.data_ptr(), dummy.data_ptr(), C.ATM_EVAL_FORCES,
                                   dummy.data_ptr(), dummy.data_ptr(), st.data_ptr(), None)
    assert rc == C.ATM_ERR_UNSUPPORTED
    assert "32768" in C.load().atm_last_error().decode()
