"""Summaries of rocprofv3 CSV output for profiles/.

    python -m ns3d_cpp_amd.prof --stats DIR/ks_kernel_stats.csv \
        --fetch DIR/fetch_counter_collection.csv --write DIR/write_counter_collection.csv \
        --n 256 --ng 4 --out profiles/NAME.md --traffic-json profiles/traffic_PREC_hit256.json

FETCH_SIZE / WRITE_SIZE (KiB per dispatch) are scaled by what they read
for the streaming copy kernel k_copy, whose bytes are known (12 cell arrays
read and written).  Kernel kinds are mapped to the names bench.py reports.
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
    m = re.search(r"km?_vdiv(_z)?<\d+, (\d)(?:, \d)?>", n)
    if m:
        return "vdiv_rhs" if m.group(2) == "0" else "vdiv_stage"
    if "fill_periodic" in n or "fill_bc" in n:
        return "fill_ghosts"
    if "k_cfl" in n or "km_cfl" in n:
        return "cfl"
    if "k_copy" in n:
        return "copy"
    return None


def short(name):
    return name.replace("ns3d::", "").replace("void ", "").split("(")[0]


def load_counter(path, counter):
    d = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r.get("Counter_Name") == counter:
            d[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in d.items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--stats", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--ng", type=int, default=4)
    ap.add_argument("--title", default="")
    ap.add_argument("--command", default="")
    ap.add_argument("--alg", default="", help="kind=bytes_per_cell,... algorithmic bytes")
    ap.add_argument("--out", required=True)
    ap.add_argument("--traffic-json")
    a = ap.parse_args()
    n3 = a.n ** 3
    ncell = (a.n + 2 * a.ng) ** 3
    alg = {}
    for kv in filter(None, a.alg.split(",")):
        k, v = kv.split("=")
        alg[k] = float(v)
    lines = [f"# {a.title}", ""]
    if a.command:
        lines += ["Commands (run from the repository root on the GPU box):", "", "```", a.command, "```", ""]
    lines += ["| kernel | kind | calls | avg us | share % |", "|---|---|---|---|---|"]
    for r in csv.DictReader(open(a.stats)):
        lines.append(f"| `{short(r['Name'])}` | {kind_of(r['Name']) or ''} | {r['Calls']} | "
                     f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['Percentage']):.2f} |")
    traffic = {}
    if a.fetch and a.write:
        f = load_counter(a.fetch, "FETCH_SIZE")
        w = load_counter(a.write, "WRITE_SIZE")
        copy = [k for k in f if "k_copy" in k]
        copy_kib = 12 * ncell * 8 / 1024.0
        fs = copy_kib / f[copy[0]] if copy else 1.0
        ws = copy_kib / w[copy[0]] if copy else 1.0
        lines += ["", f"HBM traffic per dispatch from FETCH_SIZE and WRITE_SIZE (separate runs). Calibration: the copy "
                      f"kernel `k_copy` moves {12 * ncell * 8 / 1e9:.3f} GB each way per launch; FETCH_SIZE reads "
                      f"{1 / fs:.3f} of that (scale x{fs:.3f}), WRITE_SIZE {1 / ws:.3f} (scale x{ws:.3f}).", "",
                  "| kernel | kind | FETCH KiB raw | WRITE KiB raw | traffic GB | algorithmic GB | ratio |",
                  "|---|---|---|---|---|---|---|"]
        per_kind = collections.defaultdict(list)
        for k in f:
            kind = kind_of(k)
            if kind is None or kind == "copy":
                continue
            tr = (f[k] * fs + w.get(k, 0.0) * ws) * 1024.0
            al = alg.get(kind)
            per_kind[kind].append(tr)
            lines.append(f"| `{short(k)}` | {kind} | {f[k]:.0f} | {w.get(k, 0.0):.0f} | {tr / 1e9:.3f} | "
                         f"{'' if al is None else f'{al * n3 / 1e9:.3f}'} | "
                         f"{'' if al is None else f'{tr / (al * n3):.2f}'} |")
        traffic = {k: sum(v) / len(v) for k, v in per_kind.items()}
    open(a.out, "w").write("\n".join(lines) + "\n")
    if a.traffic_json and traffic:
        json.dump({k: round(v) for k, v in traffic.items()}, open(a.traffic_json, "w"), indent=1)


if __name__ == "__main__":
    main()
