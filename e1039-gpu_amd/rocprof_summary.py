"""Summaries of the chain kernel's rocprofv3 runs, for profiles/.

    python e1039-gpu_amd/rocprof_summary.py --dir RUNS --workload c3 --events 8000 --mode plain|pruned

RUNS holds one rocprofv3 output directory per run (each with run_results.db),
made by the commands the summary lists (bench.py on one workload, one counter
block per run): kt_<mode> (kernel trace), fetch_<mode>, write_<mode>
(FETCH_SIZE, WRITE_SIZE), copy_fetch, copy_write (the same counters on a 1 GiB
device copy, the calibration), and sq1, sq2 (instruction and LDS counters,
plain chain only).  The chain launch of the mode is the first dispatch of the
dominant chain kernel (largest duration) in a plain run, the second in a
pruned run (bench.py first runs the plain chain once as the recall reference).

Writes profiles/chain_kernel_<workload>_<mode>.md and
profiles/chain_kernel_<workload>_<mode>_counters.json; bench.py reads the
json when its kernel-source hash, batch and settings match the timed run.
"""
import argparse
import hashlib
import json
import csv
import os
import re
import sqlite3

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
KERNEL_SOURCES = ["or_chain.hip", "or_device.h", "or_prune.h"]
HBM_PEAK_GBS = 8000.0
CLOCK_GHZ = 2.4   # MI355X peak engine clock
SIMDS = 256 * 4


def kernel_source_sha():
    """sha256 (16 hex digits) over the chain kernel's sources."""
    h = hashlib.sha256()
    for n in KERNEL_SOURCES:
        with open(os.path.join(HERE, "csrc", n), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def dispatches(db):
    c = sqlite3.connect(db)
    cols = [r[1] for r in c.execute("pragma table_info(kernels)")]
    return [dict(zip(cols, r)) for r in c.execute("select * from kernels order by start")]


def counters(db):
    """{dispatch_id: {"name": kernel, counter: value summed over its instances}}"""
    c = sqlite3.connect(db)
    cols = [r[1] for r in c.execute("pragma table_info(counters_collection)")]
    out = {}
    for r in c.execute("select * from counters_collection"):
        d = dict(zip(cols, r))
        e = out.setdefault(d["dispatch_id"], {"name": d["kernel_name"]})
        e[d["counter_name"]] = e.get(d["counter_name"], 0.0) + float(d["value"])
    return out


def pick(entries, name, index):
    """index-th dispatch (in dispatch order) whose kernel name is name"""
    ids = sorted(k for k, v in entries.items() if v["name"] == name)
    return entries[ids[index]]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--dir", required=True)
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--events", type=int, required=True)
    ap.add_argument("--mode", choices=["plain", "pruned"], required=True)
    ap.add_argument("--waves-per-event", type=int, default=8, help="the session's setting (at most; 4 on the 2,048-hit tile)")
    ap.add_argument("--yz-batch", type=int, default=1024)
    ap.add_argument("--algorithmic-bytes", type=float, required=True, help="bytes per launch (bench.py's figure)")
    ap.add_argument("--commands", default="", help="file holding the commands that made the runs")
    ap.add_argument("--out", default=os.path.join(ROOT, "profiles"))
    a = ap.parse_args()
    idx = 0 if a.mode == "plain" else 1

    kt = [d for d in dispatches(os.path.join(a.dir, f"kt_{a.mode}", "run_results.db")) if "chain_kernel" in d["name"]]
    tot = {}
    for d in kt:
        tot[d["name"]] = tot.get(d["name"], 0) + d["duration"]
    dom = max(tot, key=tot.get)
    runs = [d for d in kt if d["name"] == dom]
    launch = runs[idx]
    ms = launch["duration"] / 1e6

    def one(run, counter, name=dom, index=idx):
        return pick(counters(os.path.join(a.dir, run, "run_results.db")), name, index)[counter]

    copy_name = "__amd_rocclr_copyBuffer"
    cf = counters(os.path.join(a.dir, "copy_fetch", "run_results.db"))
    cw = counters(os.path.join(a.dir, "copy_write", "run_results.db"))
    cfk = max(v["FETCH_SIZE"] for v in cf.values() if v["name"] == copy_name)
    cwk = max(v["WRITE_SIZE"] for v in cw.values() if v["name"] == copy_name)
    copy_kib = 1 << 20   # 1 GiB read and written
    f_scale, w_scale = copy_kib / cfk, copy_kib / cwk
    fetch_kib = one(f"fetch_{a.mode}", "FETCH_SIZE")
    write_kib = one(f"write_{a.mode}", "WRITE_SIZE")
    fetch_b = fetch_kib * 1024 * f_scale
    write_b = write_kib * 1024 * w_scale
    hbm = fetch_b + write_b

    sq = {}
    if a.mode == "plain" and os.path.exists(os.path.join(a.dir, "sq1")):
        for run in ("sq1", "sq2"):
            e = pick(counters(os.path.join(a.dir, run, "run_results.db")), dom, 0)
            sq.update({k: v for k, v in e.items() if k != "name"})
    derived = {}
    if sq:
        # GRBM_GUI_ACTIVE is summed over the 8 XCDs; per-XCD active cycles
        cyc = sq["GRBM_GUI_ACTIVE"] / 8.0
        derived = {
            "valu_busy_frac": sq["SQ_ACTIVE_INST_VALU"] * 4.0 / SIMDS / cyc,
            "valu_insts_per_event": sq["SQ_INSTS_VALU"] / a.events,
            "salu_insts_per_event": sq["SQ_INSTS_SALU"] / a.events,
            "lds_insts_per_event": sq["SQ_INSTS_LDS"] / a.events,
            "vmem_rd_insts_per_event": sq["SQ_INSTS_VMEM_RD"] / a.events,
            "vmem_wr_insts_per_event": sq["SQ_INSTS_VMEM_WR"] / a.events,
            "lds_util_frac": sq["SQ_LDS_IDX_ACTIVE"] / (cyc * 256),
            "lds_bank_conflict_frac": sq["SQ_LDS_BANK_CONFLICT"] / max(1.0, sq["SQ_LDS_IDX_ACTIVE"]),
            "wait_any_over_wave_cycles": sq["SQ_WAIT_ANY"] / sq["SQ_WAVE_CYCLES"],
        }

    res = {
        "kernel": dom, "workload": a.workload, "events": a.events, "mode": a.mode,
        "waves_per_event": a.waves_per_event, "yz_batch": a.yz_batch,
        "kernel_source_sha256_16": kernel_source_sha(),
        "launch_ms": ms, "all_launch_ms": [d["duration"] / 1e6 for d in runs],
        "workgroups": launch["grid_x"] // launch["workgroup_x"], "workgroup_size": launch["workgroup_x"],
        "lds_bytes": launch["lds_size"], "scratch_bytes_per_lane": launch["scratch_size"],
        "fetch_kib": fetch_kib, "write_kib": write_kib, "fetch_scale": f_scale, "write_scale": w_scale,
        "hbm_bytes_per_launch": hbm, "hbm_gbs": hbm / (ms * 1e-3) / 1e9,
        "algorithmic_bytes_per_launch": a.algorithmic_bytes,
        "traffic_over_algorithmic": hbm / a.algorithmic_bytes,
        "counters": sq, "derived": derived,
    }
    os.makedirs(a.out, exist_ok=True)
    base = os.path.join(a.out, f"chain_kernel_{a.workload}_{a.mode}")
    with open(base + "_counters.json", "w") as f:
        json.dump(res, f, indent=1)
    cmds = open(a.commands).read().strip() if a.commands else ""
    kname = re.search(r"chain_kernel<[^>]*>", dom).group(0)
    lines = [f"# {kname}, {a.workload}, {a.mode} chain, {a.events} events",
             "", f"Settings: waves per event at most {a.waves_per_event} ({min(a.waves_per_event, 4)} on the 1,024- and "
             "2,048-hit tiles), YZ survivor batch {a.yz_batch}; kernel sources "
             f"sha256 {res['kernel_source_sha256_16']} ({', '.join(KERNEL_SOURCES)}).", "",
             "Commands (repository root on the GPU box as `$REPO`; one counter block per run):", "", "```", cmds,
             "```", "", "## Kernel trace", "",
             f"Launch of the {a.mode} chain: **{ms:.1f} ms**, {res['workgroups']} workgroups of "
             f"{res['workgroup_size']} lanes, {launch['lds_size']} B LDS, {launch['scratch_size']} B/lane scratch. "
             f"All launches of this kernel in the run (ms): {', '.join(f'{x:.1f}' for x in res['all_launch_ms'])}.",
             "", "rocprofv3 --stats of the same run (kernel-trace run's run_kernel_stats.csv), chain kernels:", "",
             "| kernel | calls | total ms | average ms |", "|---|---|---|---|"]
    stats = os.path.join(a.dir, f"kt_{a.mode}", "run_kernel_stats.csv")
    if os.path.exists(stats):
        with open(stats) as f:
            for r in csv.DictReader(f):
                if "chain_kernel" in r["Name"]:
                    lines.append(f"| {re.search(r'chain_kernel<[^>]*>', r['Name']).group(0)} | {r['Calls']} | "
                                 f"{float(r['TotalDurationNs']) / 1e6:.1f} | {float(r['AverageNs']) / 1e6:.1f} |")
    lines += ["", "## HBM bytes", "",
             f"FETCH_SIZE {fetch_kib:,.0f} KiB, WRITE_SIZE {write_kib:,.0f} KiB per launch. Calibration: the 1 GiB "
             f"device copy reads {cfk:,.1f} KiB and writes {cwk:,.1f} KiB by the same counters, so FETCH is scaled by "
             f"{f_scale:.4f} and WRITE by {w_scale:.4f}.", "",
             f"Traffic per launch: {fetch_b / 1e9:.1f} GB read + {write_b / 1e9:.1f} GB written = "
             f"**{hbm / 1e9:.1f} GB** ({hbm / a.events / 1e6:.2f} MB per event, {res['hbm_gbs']:.0f} GB/s over the "
             f"launch), against {a.algorithmic_bytes / 1e9:.3f} GB of algorithmic bytes "
             f"({res['traffic_over_algorithmic']:.0f}x)."]
    if sq:
        lines += ["", "## Instruction and LDS counters (plain chain)", "", "| counter | value |", "|---|---|"]
        lines += [f"| {k} | {v:.4g} |" for k, v in sorted(sq.items())]
        lines += ["", "| derived | value |", "|---|---|"]
        lines += [f"| {k} | {v:.4g} |" for k, v in derived.items()]
        lines += ["", "valu_busy_frac = SQ_ACTIVE_INST_VALU x 4 / (1024 SIMDs x per-XCD GRBM_GUI_ACTIVE); "
                  "lds_util_frac = SQ_LDS_IDX_ACTIVE / (256 CUs x per-XCD GRBM_GUI_ACTIVE); "
                  "lds_bank_conflict_frac = SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE."]
    with open(base + ".md", "w") as f:
        f.write("\n".join(lines) + "\n")
    print(json.dumps({k: v for k, v in res.items() if k not in ("counters",)}, indent=1))


if __name__ == "__main__":
    main()
