"""Summaries of rocprofv3 CSV output for profiles/.

Kernel times and HBM traffic:

    python -m ns3d_cpp_amd.prof --stats DIR/ks_kernel_stats.csv \\
        --fetch DIR/fetch_counter_collection.csv --write DIR/write_counter_collection.csv \\
        --n 256 --ng 4 --alg sweep_x=40,... --stage-bytes 140,160,200 \\
        --out profiles/NAME.md --traffic-json profiles/traffic_PREC_hit256.json

FETCH_SIZE / WRITE_SIZE (KiB per dispatch) are scaled by what they read
for the streaming copy kernel k_copy, whose bytes are known (12 cell arrays
read and written).  Kernel kinds are mapped to the names bench.py reports;
the stage kernel's three instantiations (RK3 stages 1, 2, 3) are compared
with their own algorithmic bytes (--stage-bytes).

Counter summary (one or more --pmc runs of the same command):

    python -m ns3d_cpp_amd.prof --counters DIR/pmc1_counter_collection.csv,... \\
        --n 256 --out profiles/NAME.md

Per kernel: duration in GPU cycles (GRBM_GUI_ACTIVE over the 8 XCDs), VALU
and vector-memory instructions per cell, the fraction of those cycles the
texture addresser (TA_TA_BUSY over the 256 CUs) and the vector L1 waiting on
L2 (TCP_PENDING_STALL_CYCLES) were busy, L2 hit rate and the L1-to-L2 read
requests per cell.
"""
import argparse
import collections
import csv
import json
import re


def kind_of(name):
    """bench.py kernel kind of a HIP kernel symbol."""
    n = name.replace("ns3d::", "")
    m = re.search(r"ka_sweep(?:_line|c)<(\d)", n)
    if m:
        return ["adj_sweep_x", "adj_sweep_y", "adj_sweep_z"][int(m.group(1))]
    for key, kind in (("ka_vflux", "adj_vflux"), ("ka_grad", "adj_grad"), ("ka_fill", "adj_fill"),
                      ("k_dissip", "dissipation"), ("k_sum_partials", "adj_sum")):
        if key in n:
            return kind
    if re.search(r"k_sweep_xrow|km_sweep_xrow", n):
        return "sweep_x"
    m = re.search(r"km?_sweep<(\d)", n)
    if m:
        return ["sweep_x", "sweep_y", "sweep_z"][int(m.group(1))]
    if "gradflux" in n:
        return "gradflux"
    if re.search(r"km_vdiv_zs<", n):
        return "vdiv_zs_stage"
    m = re.search(r"km?_vdiv(?:_z|_lds)?<\d+, (\d)(?:, (?:\d|true|false))?>", n)
    if m:
        return "vdiv_rhs" if m.group(1) == "0" else "vdiv_stage"
    if "fill_periodic" in n or "fill_bc" in n:
        return "fill_ghosts"
    if "k_cfl" in n or "km_cfl" in n:
        return "cfl"
    if "k_copy_nt" in n:
        return "copy_nt"
    if "k_copy" in n:
        return "copy"
    return None


def stage_of(name):
    """RK3 stage (1..3) of a stage-kernel instantiation, else None."""
    m = re.search(r"km?_vdiv(?:_z|_lds|_zs)?<\d+, (\d)(?:, (?:\d|true|false))*>", name.replace("ns3d::", ""))
    return int(m.group(1)) if m and m.group(1) != "0" else None


def short(name):
    return name.replace("ns3d::", "").replace("void ", "").split("(")[0]


def load_counter(path, counter):
    d = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r.get("Counter_Name") == counter:
            d[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in d.items()}


def load_all_counters(paths):
    """{kernel name: {counter: mean value per dispatch}} over several CSVs."""
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    for path in paths:
        for r in csv.DictReader(open(path)):
            acc[r["Kernel_Name"]][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in acc.items()}


N_XCD, N_CU = 8, 256


def counter_rows(counters, n3):
    """Derived per-kernel figures from the raw counters (None where a
    counter was not collected)."""
    rows = []
    for k, c in counters.items():
        kind = kind_of(k)
        if kind is None:
            continue
        waves_cells = n3 / 64.0

        def g(key):
            return c.get(key)
        cyc = g("GRBM_GUI_ACTIVE") / N_XCD if g("GRBM_GUI_ACTIVE") else None
        row = {"kernel": short(k), "kind": kind, "cycles": cyc}
        row["valu_per_cell"] = g("SQ_INSTS_VALU") / waves_cells if g("SQ_INSTS_VALU") else None
        vm = (g("SQ_INSTS_VMEM_RD") or 0) + (g("SQ_INSTS_VMEM_WR") or 0)
        row["vmem_per_cell"] = vm / waves_cells if vm else None
        row["ta_busy"] = g("TA_TA_BUSY_sum") / N_CU / cyc if cyc and g("TA_TA_BUSY_sum") else None
        row["tcp_pending"] = (g("TCP_PENDING_STALL_CYCLES_sum") / N_CU / cyc
                              if cyc and g("TCP_PENDING_STALL_CYCLES_sum") else None)
        hit, miss = g("TCC_HIT_sum"), g("TCC_MISS_sum")
        row["l2_hit"] = hit / (hit + miss) if hit is not None and miss is not None and hit + miss > 0 else None
        row["l1_l2_read_B_per_cell"] = 64.0 * g("TCP_TCC_READ_REQ_sum") / n3 if g("TCP_TCC_READ_REQ_sum") else None
        row["l2_latency"] = (g("TCP_TCC_READ_REQ_LATENCY_sum") / g("TCP_TCC_READ_REQ_sum")
                             if g("TCP_TCC_READ_REQ_LATENCY_sum") and g("TCP_TCC_READ_REQ_sum") else None)
        row["valu_activity"] = (g("SQ_ACTIVE_INST_VALU") / g("SQ_BUSY_CYCLES")
                                if g("SQ_ACTIVE_INST_VALU") and g("SQ_BUSY_CYCLES") else None)
        rows.append(row)
    order = ["copy", "copy_nt", "sweep_x", "sweep_y", "sweep_z", "gradflux", "vdiv_stage", "vdiv_zs_stage",
             "fill_ghosts", "cfl"]
    rows.sort(key=lambda r: (order.index(r["kind"]) if r["kind"] in order else 99, r["kernel"]))
    return rows


def fmt(v, spec):
    return "" if v is None else format(v, spec)


def counters_md(rows):
    lines = ["| kernel | kind | cycles | VALU / cell | VMEM / cell | TA busy | L1 waiting on L2 | L2 hit | "
             "L1->L2 read B / cell | L2 read latency (cycles) | VALU activity / busy cycle |",
             "|---|---|---|---|---|---|---|---|---|---|---|"]
    for r in rows:
        lines.append(f"| `{r['kernel']}` | {r['kind']} | {fmt(r['cycles'], '.3g')} | {fmt(r['valu_per_cell'], '.0f')} | "
                     f"{fmt(r['vmem_per_cell'], '.0f')} | {fmt(r['ta_busy'], '.0%')} | {fmt(r['tcp_pending'], '.0%')} | "
                     f"{fmt(r['l2_hit'], '.0%')} | {fmt(r['l1_l2_read_B_per_cell'], '.0f')} | "
                     f"{fmt(r['l2_latency'], '.0f')} | {fmt(r['valu_activity'], '.1f')} |")
    return lines


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--stats")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--counters", help="comma-separated counter_collection.csv files")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--ng", type=int, default=4)
    ap.add_argument("--title", default="")
    ap.add_argument("--command", default="")
    ap.add_argument("--alg", default="", help="kind=bytes_per_cell,... algorithmic bytes")
    ap.add_argument("--stage-bytes", default="", help="algorithmic bytes of the stage kernel's stages 1,2,3")
    ap.add_argument("--zs-stage-bytes", default="",
                    help="the same for the stage kernel with the z sweep folded in (vdiv_zs_stage)")
    ap.add_argument("--scale", default="",
                    help="FETCH,WRITE scale factors for runs without the copy kernel (from a run with it)")
    ap.add_argument("--out", required=True)
    ap.add_argument("--traffic-json")
    a = ap.parse_args()
    n3 = a.n ** 3
    ncell = (a.n + 2 * a.ng) ** 3
    alg = {}
    for kv in filter(None, a.alg.split(",")):
        k, v = kv.split("=")
        alg[k] = float(v)
    stage_bytes = [float(x) for x in a.stage_bytes.split(",")] if a.stage_bytes else None
    zs_bytes = [float(x) for x in a.zs_stage_bytes.split(",")] if a.zs_stage_bytes else None
    lines = [f"# {a.title}", ""]
    if a.command:
        lines += ["Commands (run from the repository root on the GPU box):", "", "```", a.command, "```", ""]
    if a.counters:
        rows = counter_rows(load_all_counters(a.counters.split(",")), n3)
        lines += counters_md(rows)
        open(a.out, "w").write("\n".join(lines) + "\n")
        return
    lines += ["| kernel | kind | calls | avg us | share % |", "|---|---|---|---|---|"]
    for r in csv.DictReader(open(a.stats)):
        lines.append(f"| `{short(r['Name'])}` | {kind_of(r['Name']) or ''} | {r['Calls']} | "
                     f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['Percentage']):.2f} |")
    traffic = {}
    if a.fetch and a.write:
        f = load_counter(a.fetch, "FETCH_SIZE")
        w = load_counter(a.write, "WRITE_SIZE")
        copy = [k for k in f if kind_of(k) == "copy"]
        copy_kib = 12 * ncell * 8 / 1024.0
        if copy:
            fs, ws = copy_kib / f[copy[0]], copy_kib / w[copy[0]]
            cal = (f"Calibration: the copy kernel `k_copy` moves {12 * ncell * 8 / 1e9:.3f} GB each way per launch; "
                   f"FETCH_SIZE reads {1 / fs:.3f} of that (scale x{fs:.3f}), WRITE_SIZE {1 / ws:.3f} "
                   f"(scale x{ws:.3f}).")
        else:
            fs, ws = [float(x) for x in a.scale.split(",")] if a.scale else (1.0, 1.0)
            cal = (f"Calibration: this run has no copy kernel; FETCH_SIZE is scaled x{fs:.3f} and WRITE_SIZE "
                   f"x{ws:.3f}, what the copy kernel of the same counters gave in the runs that have it.")
        lines += ["", "HBM traffic per dispatch from FETCH_SIZE and WRITE_SIZE (separate runs). " + cal +
                  " Each kernel is compared with its own algorithmic bytes (the stage kernel stage by stage).", "",
                  "| kernel | kind | FETCH KiB raw | WRITE KiB raw | traffic GB | algorithmic GB | ratio |",
                  "|---|---|---|---|---|---|---|"]
        per_kind = collections.defaultdict(list)
        for k in f:
            kind = kind_of(k)
            if kind is None or kind in ("copy", "copy_nt"):
                continue
            tr = (f[k] * fs + w.get(k, 0.0) * ws) * 1024.0
            al = alg.get(kind)
            st = stage_of(k)
            sb = zs_bytes if kind == "vdiv_zs_stage" else stage_bytes
            if st is not None and sb:
                al = sb[st - 1]
                per_kind[f"{kind}_{st}"].append(tr)
            per_kind[kind].append(tr)
            lines.append(f"| `{short(k)}` | {kind}{'' if st is None else f' (stage {st})'} | {f[k]:.0f} | "
                         f"{w.get(k, 0.0):.0f} | {tr / 1e9:.3f} | "
                         f"{'' if al is None else f'{al * n3 / 1e9:.3f}'} | "
                         f"{'' if al is None else f'{tr / (al * n3):.2f}'} |")
        traffic = {k: sum(v) / len(v) for k, v in per_kind.items()}
    open(a.out, "w").write("\n".join(lines) + "\n")
    if a.traffic_json and traffic:
        json.dump({k: round(v) for k, v in traffic.items()}, open(a.traffic_json, "w"), indent=1)


if __name__ == "__main__":
    main()
