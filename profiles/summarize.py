"""Summarise a rocprofv3 kernel trace and FETCH_SIZE / WRITE_SIZE / SQ counter
runs of bench.py into markdown tables and traffic.json (read by bench.py for
roofline.traffic).  Usage, after the runs listed in k_pairs_22rep_23k.md:

    python profiles/summarize.py gpurun_out > table.md

The HBM scale factors come from the 256 MiB device copies that
`bench.py --calibrate-copy-mib 256` runs before the workload: known bytes over
what each counter reads for them.
"""
import collections
import csv
import json
import os
import sys

COPY_KIB = 256 * 1024


def counters(path):
    d = collections.defaultdict(list)
    rows = list(csv.DictReader(open(path)))
    for r in rows:
        d[(r["Kernel_Name"].split("(")[0].replace("void ", ""), r["Counter_Name"])].append(float(r["Counter_Value"]))
    return d, rows


def trace_table(path, top=20):
    """Markdown table of a rocprofv3 kernel_stats.csv, slowest total first."""
    rows = sorted(csv.DictReader(open(path)), key=lambda r: -float(r["TotalDurationNs"]))[:top]
    out = ["| kernel | calls | avg us | min us | max us | % of GPU time |", "|---|---|---|---|---|---|"]
    for r in rows:
        out.append(f"| `{r['Name'].split('(')[0].replace('void ', '')}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                   f"{float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | {float(r['Percentage']):.2f} |")
    return "\n".join(out)


def main(root):
    out = []
    stats = list(csv.DictReader(open(os.path.join(root, "trace", "run_kernel_stats.csv"))))
    out.append("| kernel | calls | avg ns | min ns | max ns | % of GPU time |")
    out.append("|---|---|---|---|---|---|")
    for r in stats:
        out.append(f"| `{r['Name'].split('(')[0].replace('void ', '')}` | {r['Calls']} | {float(r['AverageNs']):.0f} | "
                   f"{float(r['MinNs']):.0f} | {float(r['MaxNs']):.0f} | {float(r['Percentage']):.2f} |")
    f, frows = counters(os.path.join(root, "pmc_fetch", "run_counter_collection.csv"))
    w, wrows = counters(os.path.join(root, "pmc_write", "run_counter_collection.csv"))
    big = lambda rows: [float(r["Counter_Value"]) for r in rows
                        if r["Kernel_Name"].startswith("__amd_rocclr_copyBuffer") and float(r["Counter_Value"]) > COPY_KIB / 4]
    cf, cw = big(frows), big(wrows)
    fs, ws = COPY_KIB / (sum(cf) / len(cf)), COPY_KIB / (sum(cw) / len(cw))
    out.append("")
    out.append(f"Calibration: {len(cf)} copies of 256 MiB read FETCH_SIZE {sum(cf) / len(cf):.1f} KiB (scale {fs:.4f}), "
               f"WRITE_SIZE {sum(cw) / len(cw):.1f} KiB (scale {ws:.4f}).")
    out.append("")
    out.append("| kernel | launches | FETCH_SIZE KiB/launch | WRITE_SIZE KiB/launch | HBM read MB (scaled) | HBM write MB (scaled) |")
    out.append("|---|---|---|---|---|---|")
    traffic = {}
    for (k, c), v in sorted(f.items()):
        if c != "FETCH_SIZE" or not k.startswith("atm::"):
            continue
        wv = w.get((k, "WRITE_SIZE"))
        if not wv:
            continue
        rd, wr = sum(v) / len(v) * 1024 * fs, sum(wv) / len(wv) * 1024 * ws
        traffic[k] = (rd, wr, len(v))
        out.append(f"| `{k.replace('atm::', '')}` | {len(v)} | {sum(v) / len(v):.1f} | {sum(wv) / len(wv):.1f} | "
                   f"{rd / 1e6:.2f} | {wr / 1e6:.2f} |")
    s, _ = counters(os.path.join(root, "pmc_valu", "run_counter_collection.csv"))
    pk = max((k for (k, c) in s if k.startswith("atm::k_pairs<") or k.startswith("atm::k_pairs_q<")),
             key=lambda k: len(s[(k, "SQ_INSTS_VALU")]))
    out.append("")
    out.append(f"| counter (`{pk.replace('atm::', '')}`, mean per launch) | value |")
    out.append("|---|---|")
    for (k, c), v in sorted(s.items()):
        if k == pk:
            out.append(f"| {c} | {sum(v) / len(v):,.0f} |")
    print("\n".join(out))
    # the pair kernel of the timed loop: the k_pairs instantiation with the most launches
    kf = max((k for k in traffic if k.startswith("atm::k_pairs") and "dump" not in k), key=lambda k: traffic[k][2])
    # VALU issue rate: GRBM_GUI_ACTIVE sums the 8 XCDs' busy cycles; 256 CUs x 4
    # SIMDs issue at most one wave64 VALU instruction per 4 cycles each
    mean = lambda c: sum(s[(pk, c)]) / len(s[(pk, c)])
    valu, grbm = mean("SQ_INSTS_VALU"), mean("GRBM_GUI_ACTIVE")
    issue = valu / (grbm / 8.0 * 256.0)
    json.dump({"k_pairs": {"kernel": kf.replace("atm::", ""),
                           "hbm_bytes_per_launch": round(traffic[kf][0] + traffic[kf][1]),
                           "hbm_read_bytes_per_launch": round(traffic[kf][0]),
                           "hbm_write_bytes_per_launch": round(traffic[kf][1]),
                           "fetch_scale": fs, "write_scale": ws,
                           "valu_insts_per_launch": round(valu), "busy_cycles_per_xcd": round(grbm / 8.0),
                           "valu_issue_frac": round(issue, 3),
                           "workload": "22 replicas x 23,000 atoms, both states (44 sets) per launch"},
               "k_blend": {"hbm_bytes_per_launch": round(sum(traffic.get("atm::k_blend", (0, 0, 0))[:2]))}},
              open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "traffic.json"), "w"), indent=1)


if __name__ == "__main__" and len(sys.argv) > 2 and sys.argv[1] == "--trace":
    print(trace_table(sys.argv[2]))
    sys.exit(0)

if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "gpurun_out")
