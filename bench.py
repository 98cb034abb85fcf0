#!/usr/bin/env python3
"""Benchmark: ATM energy+force evaluations per second on one MI355X.

Workload (the headline configuration): 22 lambda replicas of a 23,000-atom
jittered-lattice LJ + reaction-field system (8.8 nm box, 1.0 nm cutoff,
eps_rf 78.5, 40-atom ligand displaced by (2.2, 2.2, 2.2) nm, each replica with
N(0, 0.01 nm) coordinate noise, lambda1/lambda2 on a 22-point schedule 0..0.5).
One step = one evaluation of all 22 replicas: binning, cluster lists, the pair
kernel on x and on x + d, the soft-core/softplus combine and the hybrid force,
launched as one captured HIP graph.  Inputs are resident in HBM before timing.

Printed beside it: the plain one-system-per-launch path on one 23,000-atom
system (reference-structured: two inner evaluations with host energy
read-backs, host soft-core/softplus, hybrid-force launch), the pair kernel's
roofline from HIP events around it, and the CPU baselines timed on the host (an
OpenMP FP32-pair port on every host thread; the FP64 oracle on one thread).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "openmm-atmmetaforce-plugin_amd"))

METRIC = "ATM energy+force evals/s, 23k atoms: 22-replica batch vs plain 1-GPU"
FLOP_PER_PAIR = 57            # FP32 FLOP per in-cutoff pair, FMA = 2
PEAK_FP32_TFLOPS = 157.3      # MI355X FP32 vector peak (packed); 78.6 without packed FP32
PEAK_HBM_GBS = 8000.0
# k_blend's algorithmic bytes per atom with F cleared (no read of the old F): the
# atom's slot index (int32), F1 (int64 x3) and F2 or dF (int64 x3) read, F
# (FP64 x3) written; the full mode reads a second slot index for F2
BLEND_BYTES_PER_ATOM = {"difference": 76, "full": 80}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(systems, gpu_result0, gpu_forces0, seconds=12.0):
    """CPU evaluations of replica 0 (23,000 atoms, both states), for ~`seconds`:
    the multi-threaded FP32-pair / FP64-sum cell-list port on every OpenMP
    thread (the reported baseline: what a multi-core CPU MD code does), then
    the FP64 cell-list oracle on one thread (the Reference-platform equivalent;
    its result also checks the GPU result)."""
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import oracle
    s = systems[0]
    threads = oracle.cpu_mt_threads()

    def timed(fn, budget):
        t0 = time.perf_counter()
        n, out = 0, None
        while True:
            out = fn()
            n += 1
            if time.perf_counter() - t0 > budget and n >= 2:
                break
        return n, time.perf_counter() - t0, out

    oracle.cpu_mt_atm_eval(s, threads)  # warm-up (thread pool, first touch)
    n_mt, dt_mt, _ = timed(lambda: oracle.cpu_mt_atm_eval(s, threads), 0.6 * seconds)
    n_ref, dt_ref, (ref, F) = timed(lambda: oracle.atm_eval(s), 0.4 * seconds)
    err = {
        "U1_rel": abs(gpu_result0["state1_energy"] - ref["U1"]) / abs(ref["U1"]),
        "U2_rel": abs(gpu_result0["state2_energy"] - ref["U2"]) / abs(ref["U2"]),
        "W_rel": abs(gpu_result0["energy"] - ref["W"]) / abs(ref["W"]),
        "force_rel_rms": float(np.linalg.norm(gpu_forces0 - F) / np.linalg.norm(F)),
        "oracle_pairs_in_cutoff": int(ref["npairs1"] + ref["npairs2"]),
    }
    cpu = {"value": round(n_mt / dt_mt, 3), "unit": "evals/s", "cores": threads, "kind": "port",
           "sample": f"{n_mt} energy+force evaluations of replica 0 (23,000 atoms, both states, cell list rebuilt "
                     f"each evaluation) by the OpenMP FP32-pair / FP64-sum port on {threads} host threads "
                     f"({dt_mt:.1f} s; host has {os.cpu_count()} logical CPUs)",
           "fp64_oracle_one_thread": {"value": round(n_ref / dt_ref, 3), "unit": "evals/s", "cores": 1,
                                      "sample": f"{n_ref} evaluations of the same system by the FP64 cell-list "
                                                f"oracle ({dt_ref:.1f} s)"}}
    return cpu, err


def config_report(steps=50, cpu_seconds=5.0):
    """The other configurations, as parity and throughput figures beside the
    headline (bench.py --configs prints this instead of the headline line):
      config1: the 3,000-atom CPU-runnable case -- FP64 oracle on one thread,
               the OpenMP port on every host thread, the GPU plain path;
      config2: the 23,000-atom system on the GPU plain path against the oracle;
      config4: 64 systems of 5k-60k atoms in one batched launch, each checked
               against the FP64 oracle;
      config5: forward + backward (dW/d charges, sigma, epsilon, 5 scalars) of
               the 22-replica 23k-atom batch through ATMEnergyFunction."""
    import numpy as np
    import torch
    from atmmetaforce import ATMBatch, ATMEnergyFunction, workloads as W
    sys.path.insert(0, ROOT)
    from oracle import oracle

    def timed(fn, budget, least=3):
        t0 = time.perf_counter()
        n, out = 0, None
        while n < least or time.perf_counter() - t0 < budget:
            out = fn()
            n += 1
        return n / (time.perf_counter() - t0), out

    def errors(r, F, ref, Fref):
        return {"U1_rel": abs(r["state1_energy"] - ref["U1"]) / abs(ref["U1"]),
                "U2_rel": abs(r["state2_energy"] - ref["U2"]) / abs(ref["U2"]),
                "W_rel": abs(r["energy"] - ref["W"]) / max(abs(ref["W"]), 1e-300),
                "force_rel_rms": float(np.linalg.norm(F - Fref) / np.linalg.norm(Fref))}

    out = {"metric": "configuration report", "n_gpus": 1}
    log("[bench] config1: 3,000 atoms")
    s1 = W.config1()
    threads = oracle.cpu_mt_threads()
    ref_eps, (ref, Fref) = timed(lambda: oracle.atm_eval(s1), cpu_seconds)
    oracle.cpu_mt_atm_eval(s1, threads)
    mt_eps, (rmt, Fmt) = timed(lambda: oracle.cpu_mt_atm_eval(s1, threads), cpu_seconds)
    err_mt = errors({"state1_energy": rmt["U1"], "state2_energy": rmt["U2"], "energy": rmt["W"]}, Fmt, ref, Fref)
    with ATMBatch([s1]) as b:
        for _ in range(3):
            b.compute(plain=True)
        b.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps * 4):
            b.evaluate(forces=True, clear=True, plain=True)
        b.synchronize()
        gpu_eps = steps * 4 / (time.perf_counter() - t0)
        b.compute(plain=True)
        err1 = errors(b.result(0), b.forces(0), ref, Fref)
    out["config1"] = {"atoms": s1.num_atoms, "fp64_oracle_evals_per_s": round(ref_eps, 2), "oracle_cores": 1,
                      "cpu_port_evals_per_s": round(mt_eps, 2), "cpu_port_cores": threads,
                      "cpu_port_error_vs_fp64_oracle": err_mt,
                      "gpu_plain_evals_per_s": round(gpu_eps, 1), "error_vs_fp64_oracle": err1,
                      "pairs_in_cutoff_both_states": int(ref["npairs1"] + ref["npairs2"])}

    log("[bench] config2: 23,000 atoms, plain path")
    s2 = W.config2()
    ref2, Fref2 = oracle.atm_eval(s2)
    with ATMBatch([s2]) as b:
        for _ in range(3):
            b.compute(plain=True)
        b.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.evaluate(forces=True, clear=True, plain=True)
        b.synchronize()
        plain_eps = steps / (time.perf_counter() - t0)
        b.compute(plain=True)
        err2 = errors(b.result(0), b.forces(0), ref2, Fref2)
    out["config2"] = {"atoms": s2.num_atoms, "gpu_plain_evals_per_s": round(plain_eps, 1),
                      "error_vs_fp64_oracle": err2, "pairs_in_cutoff_both_states": int(ref2["npairs1"] + ref2["npairs2"])}

    log("[bench] config4: 64 uneven systems")
    systems = W.uneven_batch(64)
    with ATMBatch(systems) as b:
        b.compute()
        b.autotune(repeats=3)
        b.compute()
        tuning4 = b.tuning()
        for _ in range(5):
            b.evaluate(forces=True, clear=True)
        b.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.evaluate(forces=True, clear=True)
        b.synchronize()
        dt = time.perf_counter() - t0
        ls = b.list_stats()
        pairs = int(ls["pairs"][ls["max_entries"] > 0].sum())  # the sets the pair kernel paired
        b.compute()
        worst = {"U1_rel": 0.0, "U2_rel": 0.0, "W_rel": 0.0, "force_rel_rms": 0.0}
        t_or = time.perf_counter()
        for k, s in enumerate(systems):
            rk, Fk = oracle.atm_eval(s)
            e = errors(b.result(k), b.forces(k), rk, Fk)
            for key in worst:
                worst[key] = max(worst[key], e[key])
        log(f"[bench] config4 oracle pass {time.perf_counter() - t_or:.1f} s")
    out["config4"] = {"systems": len(systems), "atoms": sum(s.num_atoms for s in systems),
                      "atoms_min_max": [min(s.num_atoms for s in systems), max(s.num_atoms for s in systems)],
                      "system_evals_per_s": round(len(systems) * steps / dt, 1), "ms_per_batch": round(dt * 1e3 / steps, 4),
                      "pair_interactions_per_s": round(pairs * steps / dt, 1), "tuning": tuning4,
                      "max_error_vs_fp64_oracle_over_systems": worst}

    log("[bench] config5: forward + backward, 22 replicas")
    reps = W.replica_batch(22)

    def tensors():
        dev = torch.device("cuda:0")
        q = torch.tensor(np.concatenate([s.charges for s in reps]), device=dev, requires_grad=True)
        sg = torch.tensor(np.concatenate([s.sigmas for s in reps]), device=dev, requires_grad=True)
        ep = torch.tensor(np.concatenate([s.epsilons for s in reps]), device=dev, requires_grad=True)
        sc = [torch.tensor([getattr(s.params, n) for s in reps], dtype=torch.float64, device=dev, requires_grad=True)
              for n in ("lambda1", "lambda2", "alpha", "u0", "w0")]
        return q, sg, ep, sc

    with ATMBatch(reps) as b:
        b.compute()
        b.autotune(repeats=3)
        q, sg, ep, sc = tensors()

        def fwd_bwd():
            Wt = ATMEnergyFunction.apply(b, q, sg, ep, *sc)
            Wt.sum().backward()
            for t in (q, sg, ep, *sc):
                t.grad = None
            return Wt

        for _ in range(3):
            fwd_bwd()
        torch.cuda.synchronize()
        n = 20
        t0 = time.perf_counter()
        for _ in range(n):
            fwd_bwd()
        torch.cuda.synchronize()
        fb_ms = (time.perf_counter() - t0) * 1e3 / n
    out["config5"] = {"replicas": 22, "ms_per_forward_backward": round(fb_ms, 3),
                      "note": "device tensors in and device gradients out; gradients checked against the FP64 "
                              "oracle and central differences in tests/test_autograd_gpu.py"}
    return out


def backward_timing(batch, reps, eval_ms, n=20):
    """dW/d(charges, sigmas, epsilons, lambda1, lambda2, alpha, u0, w0) of the
    whole replica batch through ATMEnergyFunction with the per-atom tensors on
    the GPU (parameters in and gradients out as device arrays).  With gradients
    required, the forward (parameter upload, one evaluation of every replica, W
    read back) computes the energies and the per-atom parameter derivatives of
    U1 and U2 in one pass over the pairs; the backward weights and gathers them
    and adds the softplus derivatives.  The energy-only forward (inputs that
    need no gradient) is timed for comparison."""
    import numpy as np
    import torch
    from atmmetaforce import ATMEnergyFunction
    dev = torch.device("cuda:0")
    q = torch.tensor(np.concatenate([s.charges for s in reps]), device=dev, requires_grad=True)
    sg = torch.tensor(np.concatenate([s.sigmas for s in reps]), device=dev, requires_grad=True)
    ep = torch.tensor(np.concatenate([s.epsilons for s in reps]), device=dev, requires_grad=True)
    sc = [torch.tensor([getattr(s.params, k) for s in reps], dtype=torch.float64, device=dev, requires_grad=True)
          for k in ("lambda1", "lambda2", "alpha", "u0", "w0")]

    def fwd():
        return ATMEnergyFunction.apply(batch, q, sg, ep, *sc)

    def fwd_bwd():
        Wt = fwd()
        Wt.sum().backward()
        for t in (q, sg, ep, *sc):
            t.grad = None
        return Wt

    def fwd_plain():
        return ATMEnergyFunction.apply(batch, q.detach(), sg.detach(), ep.detach(), *[t.detach() for t in sc])

    for _ in range(3):
        fwd_bwd()
        fwd_plain()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n):
        fwd_plain()
    torch.cuda.synchronize()
    e_ms = (time.perf_counter() - t0) * 1e3 / n
    t0 = time.perf_counter()
    for _ in range(n):
        fwd()
    torch.cuda.synchronize()
    f_ms = (time.perf_counter() - t0) * 1e3 / n
    t0 = time.perf_counter()
    for _ in range(n):
        fwd_bwd()
    torch.cuda.synchronize()
    fb_ms = (time.perf_counter() - t0) * 1e3 / n
    return {"replicas": len(reps), "ms_per_forward_energy_only": round(e_ms, 4),
            "ms_per_forward": round(f_ms, 4), "ms_per_forward_backward": round(fb_ms, 4),
            "ms_per_backward": round(fb_ms - f_ms, 4),
            "forward_backward_over_energy_only_forward": round(fb_ms / e_ms, 3),
            "forward_backward_over_batched_eval": round(fb_ms / eval_ms, 3),
            "note": "device tensors in, device gradients out; with gradients required the forward computes the "
                    "energies and the per-atom parameter derivatives of U1, U2 in one pass over the pairs; "
                    "gradients are checked against the FP64 oracle and central differences in "
                    "tests/test_autograd_gpu.py"}


def summarize_profiles(root: str, tuning_label: str, out_json: str) -> str:
    """Markdown tables of the rocprofv3 runs of bench.py under `root` (kernel
    trace in trace/, FETCH_SIZE in pmc_fetch/, WRITE_SIZE in pmc_write/, SQ
    counters in pmc_valu/ and pmc_lds/) and the per-launch figures bench.py
    reads for roofline.traffic, written to out_json.  The HBM scale factors
    are the 256 MiB device copies of --calibrate-copy-mib 256: known bytes over
    what each counter reads for them."""
    import collections
    import csv
    import glob

    def rows(sub, name):
        f = glob.glob(os.path.join(root, sub, "**", name), recursive=True)
        return list(csv.DictReader(open(f[0]))) if f else []

    def kname(r):
        return r.get("Kernel_Name", r.get("Name", "")).split("(")[0].replace("void ", "")

    def per_kernel(rs):
        d = collections.defaultdict(list)
        for r in rs:
            d[(kname(r), r["Counter_Name"])].append(float(r["Counter_Value"]))
        return d

    out = [f"Tuning of every run: {tuning_label}", "",
           "| kernel | calls | avg us | min us | max us | % of GPU time |", "|---|---|---|---|---|---|"]
    for r in sorted(rows("trace", "*kernel_stats.csv"), key=lambda r: -float(r["TotalDurationNs"]))[:16]:
        out.append(f"| `{kname(r)}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | {float(r['MinNs']) / 1e3:.2f} | "
                   f"{float(r['MaxNs']) / 1e3:.2f} | {float(r['Percentage']):.2f} |")
    fr, wr = rows("pmc_fetch", "*counter_collection.csv"), rows("pmc_write", "*counter_collection.csv")
    copy_kib = 256 * 1024
    big = lambda rs: [float(r["Counter_Value"]) for r in rs
                      if kname(r).startswith("__amd_rocclr_copyBuffer") and float(r["Counter_Value"]) > copy_kib / 4]
    cf, cw = big(fr), big(wr)
    fs, ws = copy_kib / (sum(cf) / len(cf)), copy_kib / (sum(cw) / len(cw))
    f, w = per_kernel(fr), per_kernel(wr)
    out += ["", f"Calibration: {len(cf)} copies of 256 MiB read FETCH_SIZE {sum(cf) / len(cf):.1f} KiB (scale {fs:.4f}), "
                f"WRITE_SIZE {sum(cw) / len(cw):.1f} KiB (scale {ws:.4f}).", "",
            "| kernel | launches | FETCH_SIZE KiB/launch | WRITE_SIZE KiB/launch | HBM read MB (scaled) | "
            "HBM write MB (scaled) |", "|---|---|---|---|---|---|"]
    traffic = {}
    for (k, c), v in sorted(f.items()):
        wv = w.get((k, "WRITE_SIZE"))
        if c != "FETCH_SIZE" or not k.startswith("atm::") or not wv:
            continue
        rd, wrb = sum(v) / len(v) * 1024 * fs, sum(wv) / len(wv) * 1024 * ws
        traffic[k] = (rd, wrb, len(v))
        out.append(f"| `{k.replace('atm::', '')}` | {len(v)} | {sum(v) / len(v):.1f} | {sum(wv) / len(wv):.1f} | "
                   f"{rd / 1e6:.2f} | {wrb / 1e6:.2f} |")
    # the pair kernel of the timed loop: the forward k_pairs instantiation with the most launches
    kf = max((k for k in traffic if k.startswith("atm::k_pairs<")), key=lambda k: traffic[k][2])
    sq = per_kernel(rows("pmc_valu", "*counter_collection.csv"))
    lds = per_kernel(rows("pmc_lds", "*counter_collection.csv"))
    mean = lambda d, c: sum(d[(kf, c)]) / len(d[(kf, c)]) if d.get((kf, c)) else None
    out += ["", f"| counter (`{kf.replace('atm::', '')}`, mean per launch) | value |", "|---|---|"]
    for d in (sq, lds):
        for (k, c), v in sorted(d.items()):
            if k == kf:
                out.append(f"| {c} | {sum(v) / len(v):,.0f} |")
    valu, grbm = mean(sq, "SQ_INSTS_VALU"), mean(sq, "GRBM_GUI_ACTIVE")
    # GRBM_GUI_ACTIVE sums the 8 XCDs' busy cycles; 256 CUs x 4 SIMDs issue at most
    # one wave64 VALU instruction per 4 cycles each
    issue = valu / (grbm / 8.0 * 256.0) if valu and grbm else None
    rec = {"k_pairs": {"kernel": kf.replace("atm::", ""), "tuning": tuning_label,
                       "hbm_bytes_per_launch": round(traffic[kf][0] + traffic[kf][1]),
                       "hbm_read_bytes_per_launch": round(traffic[kf][0]),
                       "hbm_write_bytes_per_launch": round(traffic[kf][1]),
                       "fetch_scale": fs, "write_scale": ws,
                       "valu_insts_per_launch": round(valu) if valu else None,
                       "busy_cycles_per_xcd": round(grbm / 8.0) if grbm else None,
                       "valu_issue_frac": round(issue, 3) if issue else None,
                       "workload": "22 replicas x 23,000 atoms; the sets the pair kernel pairs per launch "
                                   "(the 22 x sets, x + d as the difference at the displaced atoms)"},
           "k_blend": {"hbm_bytes_per_launch": round(sum(traffic.get("atm::k_blend", (0, 0, 0))[:2])),
                       "hbm_read_bytes_per_launch": round(traffic.get("atm::k_blend", (0, 0, 0))[0]),
                       "hbm_write_bytes_per_launch": round(traffic.get("atm::k_blend", (0, 0, 0))[1]),
                       "trace_avg_us": next((round(float(r["AverageNs"]) / 1e3, 3)
                                             for r in rows("trace", "*kernel_stats.csv")
                                             if kname(r) == "atm::k_blend"), None)}}
    with open(out_json, "w") as fh:
        json.dump(rec, fh, indent=1)
    return "\n".join(out)


def restraint_timing(system, evals=200):
    """Per-evaluation time of the flat-bottom restraints an ABFE script keeps in
    force group 0 beside the ATM force (python/tests/test_abfe.py:58-79): one
    centroid restraint (the 40 ligand atoms against 500 receptor atoms) and 200
    position restraints, evaluated through the C ABI from host positions of the
    23,000-atom system (gather, upload, one kernel, download, scatter)."""
    import ctypes
    import numpy as np
    from atmmetaforce import _capi as C
    L = C.load()
    n = system.num_atoms
    x = np.ascontiguousarray(system.positions)
    box = np.ascontiguousarray(system.box, dtype=np.float64)
    lig, rcpt = np.arange(40), np.arange(1000, 1500)
    members = np.ascontiguousarray(np.concatenate([lig, rcpt]).astype(np.int32))
    weights = np.ascontiguousarray(np.full(len(members), 12.0))
    cent = (C.CentroidRestraint * 1)(C.CentroidRestraint(0, 40, 40, 500, 104600.0, 0.05, (ctypes.c_double * 3)(0, 0, 0)))
    pos = (C.PositionRestraint * 200)()
    for k in range(200):
        a = 2000 + 7 * k
        pos[k] = C.PositionRestraint(a, 1, (ctypes.c_double * 3)(*(x[a] + 0.03)), 10460.0, 0.05)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    h = ctypes.c_void_p()
    C.check(L.atm_restraints_create(0, ctypes.byref(h)))
    try:
        C.check(L.atm_restraints_set(h, n, len(members), members.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                     dp(weights), 1, cent, 200, pos))
        F = np.zeros((n, 3))
        e = ctypes.c_double()
        for _ in range(10):
            C.check(L.atm_restraints_eval(h, dp(x), dp(box), dp(F), ctypes.byref(e)))
        t0 = time.perf_counter()
        for _ in range(evals):
            C.check(L.atm_restraints_eval(h, dp(x), dp(box), dp(F), ctypes.byref(e)))
        ms = (time.perf_counter() - t0) * 1e3 / evals
    finally:
        L.atm_restraints_destroy(h)
    return {"ms_per_eval": round(ms, 4), "bound": "latency (host round trip of 740 atoms; one kernel launch)",
            "restraints": "1 centroid (40 vs 500 atoms) + 200 position restraints on the 23,000-atom system"}


def md_timing(reps, steps=100):
    """Batched Langevin MD of the whole replica batch (atm_md_step): every step
    integrates all replicas and evaluates their ATM forces in one graph launch.
    2 fs steps, friction 1/ps, 300 K, 12 amu atoms."""
    import numpy as np
    from atmmetaforce import replica
    base = reps[0]
    re = replica.ATMReplicaExchange(base, [r.params for r in reps], np.full(base.num_atoms, 12.0), temperature=300.0,
                                    friction=1.0, dt=0.002, seed=1, positions=[r.positions for r in reps])
    try:
        re.step(10)
        re.batch.synchronize()
        t0 = time.perf_counter()
        re.step(steps)
        re.batch.synchronize()
        dt = time.perf_counter() - t0
    finally:
        re.close()
    sps = steps / dt
    return {"replicas": len(reps), "steps_per_s": round(sps, 1),
            "ns_per_day_per_replica": round(sps * 0.002 * 86400 / 1000, 2),
            "ms_per_step": round(1e3 / sps, 4),
            "note": "each step: Langevin update of all replicas + one batched ATM evaluation (both states)"}


def pme_timing(reps=30, cpu=True):
    """The reference's end-to-end system (python/tests/test_abfe.py): temoa-g1
    host-guest in water (8,479 atoms, AMBER, PME with a 1 nm cutoff, ligand
    displaced by 2.2 nm), ATM variable group = the PME NonbondedForce.  One
    evaluation = U1, U2, combine and hybrid force on the device, positions in
    and forces out through host memory; the bonded terms and restraints of
    group 0 beside it.  Checked against the reference's goldens (potential
    energy -116071.0 and perturbation energy 58.2 kJ/mol, within 0.1)."""
    import ctypes
    import numpy as np
    from atmmetaforce import _capi as C, amber, pme
    s, arr = amber.system_from_npz(os.path.join(ROOT, "tests", "golden", "temoa-g1.npz"))
    box, x, ref40 = arr["state/box"], arr["state/positions"], arr["inpcrd/positions40"]
    kcal = 4.184
    rcpt, lig = s.residue_atoms(1), s.residue_atoms(2)
    d = np.zeros((s.num_atoms, 3))
    d[lig] = 2.2
    params = dict(lambda1=0.5, lambda2=0.5, alpha=0.0, u0=0.0, w0=0.0, umax=200.0 * kcal, ubcore=100.0 * kcal,
                  acore=0.0625, direction=0.0625)
    f, b = pme.from_amber(s, d, box, params=params)
    try:
        alpha, grid = f.info()
        for _ in range(3):
            r, F, parts = f.evaluate(x, box)
        t0 = time.perf_counter()
        for _ in range(reps):
            r, F, parts = f.evaluate(x, box)
        dt = (time.perf_counter() - t0) / reps
        Eb, _ = b.evaluate(x)
        t0 = time.perf_counter()
        for _ in range(reps):
            b.evaluate(x)
        dtb = (time.perf_counter() - t0) / reps
    finally:
        f.close()
        b.close()
    # group-0 restraints of test_abfe.py:58-79 through the C ABI
    L = C.load()
    kf = 25.0 * kcal * 100.0
    posr = np.array([a for a in rcpt if s.atom_names[a].startswith("C") and a < 40])
    members = np.ascontiguousarray(np.concatenate([lig, rcpt]).astype(np.int32))
    w = np.ascontiguousarray(np.concatenate([s.masses[lig], s.masses[rcpt]]))
    cent = (C.CentroidRestraint * 1)()
    cent[0] = C.CentroidRestraint(0, len(lig), len(lig), len(rcpt), kf, 0.5, (ctypes.c_double * 3)(0.0, 0.0, 0.0))
    pr = (C.PositionRestraint * len(posr))()
    for k, a in enumerate(posr):
        pr[k] = C.PositionRestraint(int(a), 1, (ctypes.c_double * 3)(*ref40[a]), kf, 0.05)
    h = ctypes.c_void_p()
    C.check(L.atm_restraints_create(0, ctypes.byref(h)))
    er = ctypes.c_double()
    try:
        dp = lambda a: np.ascontiguousarray(a).ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        xx, bb = np.ascontiguousarray(x), np.ascontiguousarray(box)
        C.check(L.atm_restraints_set(h, s.num_atoms, len(members),
                                     members.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), dp(w), 1, cent,
                                     len(posr), pr))
        C.check(L.atm_restraints_eval(h, dp(xx), dp(bb), None, ctypes.byref(er)))
    finally:
        L.atm_restraints_destroy(h)
    pot = sum(Eb.values()) + er.value + r["energy"]
    out = {"workload": "temoa-g1 host-guest in water, 8,479 atoms, AMBER, PME (alpha %.4f/nm, grid %s), rc 1 nm; "
                       "ATM variable group = PME NonbondedForce, ligand displaced by 2.2 nm" % (alpha, "x".join(map(str, grid))),
           "ms_per_eval": round(dt * 1e3, 4), "evals_per_s": round(1.0 / dt, 2),
           "bonded_ms_per_eval": round(dtb * 1e3, 4),
           "potential_energy": round(pot, 4), "perturbation_energy": round(r["perturbation_energy"], 4),
           "golden": {"potential_energy": -116071.0, "perturbation_energy": 58.2, "atol": 0.1,
                      "ok": bool(abs(pot + 116071.0) <= 0.1 and abs(r["perturbation_energy"] - 58.2) <= 0.1)},
           "note": "host positions in, host forces out (PCIe included); FP64 throughout"}
    if cpu:
        sys.path.insert(0, ROOT)
        from oracle import pme_oracle as po
        t0 = time.perf_counter()
        U1, _, _ = po.nonbonded_pme(x, s.charges, s.sigmas, s.epsilons, box, 1.0, s.exception_pairs, s.exception_qq,
                                    s.exception_sigma, s.exception_eps)
        x2 = x.copy()
        x2[lig] += 2.2
        U2, _, _ = po.nonbonded_pme(x2, s.charges, s.sigmas, s.epsilons, box, 1.0, s.exception_pairs,
                                    s.exception_qq, s.exception_sigma, s.exception_eps)
        tc = time.perf_counter() - t0
        out["cpu_oracle"] = {"s_per_eval": round(tc, 3), "cores": 1, "kind": "port",
                             "sample": "one evaluation (both states, forces) by the FP64 numpy oracle",
                             "U1_rel_err": abs(r["state1_energy"] - U1) / abs(U1),
                             "U2_rel_err": abs(r["state2_energy"] - U2) / abs(U2)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--replicas", type=int, default=22)
    ap.add_argument("--plain-steps", type=int, default=40)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--tuning", default=None,
                    help="column_scale,pair_waves_per_block,zbins_per_column[,pair_kernel[,pair_split[,pair_order]]] "
                         "(skips the autotuner)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--configs", action="store_true",
                    help="print the report of the other configurations (3k-atom CPU case, uneven batch, backward) "
                         "instead of the headline line")
    ap.add_argument("--no-backward", action="store_true", help="skip the forward + backward timing")
    ap.add_argument("--no-legs", action="store_true",
                    help="skip the restraint, replica-MD and PME legs (profiles of the batched evaluation alone)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first run a device-to-device copy of this many MiB five times (a kernel of known "
                         "traffic for scaling the FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--summarize-profiles", default=None, metavar="DIR",
                    help="print markdown tables of the rocprofv3 runs under DIR and write DIR/traffic.json "
                         "(copied to profiles/traffic.json, which the roofline reads); no GPU work")
    ap.add_argument("--tuning-label", default="", help="the tuning the profiled runs used (with --summarize-profiles)")
    args = ap.parse_args()
    if args.summarize_profiles:
        print(summarize_profiles(args.summarize_profiles, args.tuning_label,
                                 os.path.join(args.summarize_profiles, "traffic.json")))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(replicas and systems would shard across GPUs with no exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from atmmetaforce import ATMBatch, workloads as W

    if args.configs:
        print(json.dumps(config_report()), flush=True)
        return

    if args.calibrate_copy_mib > 0:
        n = args.calibrate_copy_mib * 1024 * 1024 // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy: {args.calibrate_copy_mib} MiB read + {args.calibrate_copy_mib} MiB written, x5")

    log("[bench] generating workload")
    base = W.config2()
    reps = W.replica_batch(args.replicas, base=base)
    batch = ATMBatch(reps)
    batch.compute()
    tuning = batch.tuning()
    if args.tuning:
        f = args.tuning.split(",")
        batch.set_tuning(float(f[0]), int(f[1]), int(f[2]), int(f[3]) if len(f) > 3 else 0,
                         pair_split=int(f[4]) if len(f) > 4 else 1, pair_order=int(f[5]) if len(f) > 5 else 2)
        batch.compute()
        tuning = batch.tuning()
    elif not args.no_autotune:
        log("[bench] autotuning")
        tuning = batch.autotune(repeats=5)
        batch.compute()
    # in-cutoff pairs counted by the pair kernel itself (lane masks of its exact r < rc tests), over
    # the coordinate sets it paired: with x + d evaluated as the difference from x (the default when
    # few atoms are displaced) only the x sets have lists (max_entries > 0)
    ls = batch.list_stats()
    paired = ls["max_entries"] > 0
    pairs = int(ls["pairs"][paired].sum())
    nsets_paired = int(paired.sum())
    state2_mode = "difference" if nsets_paired < len(ls["pairs"]) else "full"

    log(f"[bench] warmup {args.warmup}")
    for _ in range(args.warmup):
        batch.evaluate(forces=True, clear=True)
    batch.synchronize()
    torch.cuda.synchronize()
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.evaluate(forces=True, clear=True)
    batch.synchronize()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt * 1e3 / args.steps
    evals_per_s = args.replicas * args.steps / dt

    fp32_eps = None
    if not args.no_fp32:
        log("[bench] FP32 scalar pair-function mode")
        for _ in range(args.warmup):
            batch.evaluate(forces=True, clear=True, fp32_pairs=True)
        batch.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            batch.evaluate(forces=True, clear=True, fp32_pairs=True)
        batch.synchronize()
        fp32_eps = args.replicas * args.steps / (time.perf_counter() - t1)

    full_eps = None
    if state2_mode == "difference" and not args.no_fp32:
        log("[bench] x + d binned and paired in full (full_state2)")
        t = batch.tuning()
        batch.set_tuning(t["column_scale"], t["pair_waves_per_block"], t["zbins_per_column"], t["pair_kernel"],
                         t["list_capacity"], t["pair_split"], full_state2=True, pair_order=t["pair_order"])
        for _ in range(args.warmup):
            batch.evaluate(forces=True, clear=True)
        batch.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            batch.evaluate(forces=True, clear=True)
        batch.synchronize()
        full_eps = args.replicas * args.steps / (time.perf_counter() - t1)
        batch.set_tuning(t["column_scale"], t["pair_waves_per_block"], t["zbins_per_column"], t["pair_kernel"],
                         t["list_capacity"], t["pair_split"], pair_order=t["pair_order"])
        batch.compute()

    tri_eps = tri_stage = None
    if not args.no_legs:
        # the same replicas in a triclinic box of the same edges (tilts (9, -7, 12) lattice
        # spacings): tile kernel, every column a list candidate
        log("[bench] triclinic boxes")
        t = batch.tuning()
        tri = ATMBatch([W.triclinic(r, sites=(9, -7, 12)) for r in reps])
        tri.set_tuning(t["column_scale"], t["pair_waves_per_block"], t["zbins_per_column"], 0,
                       pair_split=t["pair_split"], pair_order=t["pair_order"])
        tri.compute()
        for _ in range(args.warmup):
            tri.evaluate(forces=True, clear=True)
        tri.synchronize()
        nt = max(10, args.steps // 4)
        t1 = time.perf_counter()
        for _ in range(nt):
            tri.evaluate(forces=True, clear=True)
        tri.synchronize()
        tri_eps = args.replicas * nt / (time.perf_counter() - t1)
        tri_stage = tri.profile(reps=10)
        tri.close()

    log("[bench] per-stage HIP-event timing")
    prof = batch.profile(reps=max(10, args.steps // 10))
    r0 = batch.result(0)
    batch.compute()
    r0 = batch.result(0)
    F0 = batch.forces(0)

    pair_tflops = FLOP_PER_PAIR * pairs / (prof["pairs_ms"] * 1e-3) / 1e12
    traffic = issue = valu_launch = None
    tsource = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):  # HBM bytes and VALU issue rate per launch from the committed rocprofv3 --pmc summaries
        with open(tpath) as fh:
            kp = json.load(fh).get("k_pairs", {})
        traffic, issue = kp.get("hbm_bytes_per_launch"), kp.get("valu_issue_frac")
        valu_launch = kp.get("valu_insts_per_launch")
        tsource = {"file": "profiles/traffic.json", "measured_with": "rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE, "
                   "SQ_INSTS_VALU, GRBM_GUI_ACTIVE (separate runs of bench.py, see profiles/k_pairs_22rep_23k.md)",
                   "kernel": kp.get("kernel"), "tuning": kp.get("tuning")}
    roofline = {"bound": "valu", "achieved": round(pair_tflops, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(pair_tflops / PEAK_FP32_TFLOPS, 4), "traffic": traffic,
                "kernel": f"k_pairs (LJ + reaction-field pair kernel, {nsets_paired} coordinate sets per launch: "
                          + ("the x set of every replica; x + d is the difference at the displaced atoms, "
                             "stage_ms.displaced_ms)" if state2_mode == "difference" else "x and x + d of every replica)"),
                "per_launch_flop": FLOP_PER_PAIR * pairs, "pairs_per_launch": pairs,
                "avg_launch_ms": round(prof["pairs_ms"], 4),
                "valu_issue_frac": issue,
                "valu_insts_per_pair": round(valu_launch / pairs, 3) if valu_launch else None,
                "traffic_and_issue_source": tsource,
                "note": "achieved counts 57 FLOP per in-cutoff pair only; the kernel issues VALU instructions "
                        "at valu_issue_frac of the chip's rate (SQ_INSTS_VALU over GRBM_GUI_ACTIVE, "
                        "profiles/k_pairs_22rep_23k.md): 8x8 cluster tiles hold 21 % in-cutoff pairs"}
    total_atoms = sum(s.num_atoms for s in reps)
    blend_bytes = BLEND_BYTES_PER_ATOM[state2_mode] * total_atoms
    blend_gbs = blend_bytes / (prof["blend_ms"] * 1e-3) / 1e9
    blend_traffic = blend_trace_us = None
    if os.path.exists(tpath):
        with open(tpath) as fh:
            kb = json.load(fh).get("k_blend", {})
        blend_traffic, blend_trace_us = kb.get("hbm_bytes_per_launch"), kb.get("trace_avg_us")
    backward = None
    if not args.no_backward:
        log("[bench] forward + backward through ATMEnergyFunction (device tensors)")
        backward = backward_timing(batch, reps, ms_per_step)
    batch.close()

    plain_eps = None
    if not args.no_plain:
        log("[bench] plain one-system path")
        plain = ATMBatch([base])
        for _ in range(3):
            plain.compute(plain=True)
        plain.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            plain.evaluate(forces=True, clear=True, plain=True)
        plain.synchronize()
        plain_dt = time.perf_counter() - t0
        plain_eps = args.plain_steps / plain_dt
        plain.close()

    restraints = md = pme_v = None
    if not args.no_legs:
        log("[bench] companion restraints")
        restraints = restraint_timing(base)
        log("[bench] batched lambda-replica MD")
        md = md_timing(reps)
        log("[bench] PME variable group (temoa-g1)")
        pme_v = pme_timing(cpu=not args.no_cpu)

    cpu = None
    err = None
    if not args.no_cpu:
        log("[bench] CPU baselines (multi-threaded port, FP64 oracle)")
        cpu, err = cpu_baseline(reps, r0, F0, args.cpu_seconds)

    out = {
        "metric": METRIC, "value": round(evals_per_s, 2), "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32+fp64", "data": "synthetic",
        "config": {"workload": f"{args.replicas}-replica batch of 23,000-atom LJ + reaction-field systems "
                               "(8.8 nm box, rc 1.0 nm, 40-atom ligand, d=(2.2,2.2,2.2) nm), "
                               "one HIP-graph launch per evaluation of all replicas",
                   "replicas": args.replicas, "atoms_per_replica": base.num_atoms, "cutoff_nm": 1.0,
                   "parallelism": "single GPU, replicas batched in one launch", "tuning": tuning},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "pair_interactions_per_s": round(pairs * args.steps / dt, 1),
        "pair_interactions_note": "in-cutoff pairs the pair kernel evaluates per second (the x sets when "
                                  "x + d is the difference from x)",
        "state2_mode": state2_mode,
        "full_state2_evals_per_s": round(full_eps, 2) if full_eps else None,
        "triclinic_evals_per_s": round(tri_eps, 2) if tri_eps else None,
        "triclinic_stage_ms": {k: round(v, 4) for k, v in tri_stage.items() if k.endswith("_ms")} if tri_eps else None,
        "triclinic_note": ("the same 22 replicas in triclinic boxes of the same edges (tilts b_x, c_x, c_y = "
                           "9, -7, 12 lattice spacings): tile kernel, every column a list candidate, x + d as "
                           "the difference from x at the displaced atoms") if tri_eps else None,
        "plain_evals_per_s": round(plain_eps, 2) if plain_eps else None,
        "speedup_vs_plain": round(evals_per_s / plain_eps, 2) if plain_eps else None,
        "fp32_pair_mode_evals_per_s": round(fp32_eps, 2) if fp32_eps else None,
        "stage_ms": {k: round(v, 4) for k, v in prof.items() if k.endswith("_ms")},
        "blend_kernel": {"bound": "hbm", "achieved": round(blend_gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": round(blend_gbs / PEAK_HBM_GBS, 4), "algorithmic_bytes": blend_bytes,
                         "bytes_per_atom": BLEND_BYTES_PER_ATOM[state2_mode], "traffic": blend_traffic,
                         "traffic_over_algorithmic": round(blend_traffic / blend_bytes, 3) if blend_traffic else None,
                         "kernel_trace_us": blend_trace_us,
                         "kernel_trace_frac": round(blend_bytes / (blend_trace_us * 1e-6) / 1e9 / PEAK_HBM_GBS, 4)
                         if blend_trace_us else None,
                         "note": "achieved: HIP events around the blend stage (stage_ms.blend_ms, launch gap "
                                 "included); kernel_trace_us and traffic (FETCH_SIZE + WRITE_SIZE): rocprofv3 runs "
                                 "of this bench, profiles/traffic.json"},
        "backward": backward,
        "companion_restraints": restraints,
        "replica_md": md,
        "pme_variable_group": pme_v,
        "error_vs_fp64_oracle_replica0": err,
        "precision": "coordinates, displacement, cutoff geometry and force vectors FP32; scalar pair function "
                     "(r^2, 1/r, (s/r)^6, energies, force magnitude) FP64; U1/U2/u/W FP64; forces 64-bit "
                     "fixed point (2^-24 kJ/mol/nm)",
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
