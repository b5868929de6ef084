"""Summarise rocprofv3 CSV output of the MPC-MD library as markdown.

    python -m starpolymersmd_amd.profsum --stats <dir>/<name>_kernel_stats.csv \
        [--fetch <fetch counter_collection.csv>] [--write <write counter_collection.csv>] \
        [--copy-fetch <csv> --copy-write <csv> --copy-bytes N] [--particles N]
    python -m starpolymersmd_amd.profsum closure --trace <dir>/<name>_kernel_trace.csv [--stars 256]

Kernel statistics come from `rocprofv3 --kernel-trace --stats --output-format
csv`; byte counters from `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE`
(KiB per dispatch).  A plain device copy of known size, profiled the same
way, gives the factor between the counters and real bytes.
"""
from __future__ import annotations

import argparse
import csv
from collections import defaultdict

PIPELINES = {
    # sort path: bin, scan, permutation, collide
    "sort": ("k_bin", "k_scan", "k_perm", "k_collide"),
    # row-tile path: advect, the cells' Gamma draws, the redo kernel, row-window collide
    "rows": ("k_advect", "k_cell_gamma", "k_rows_redo", "k_collide_rows"),
}


def short(name: str) -> str:
    n = name.split("(")[0]
    for k in ("k_md", "k_prologue", "k_bin", "k_sort_redo", "k_scan_redo", "k_scan", "k_scatter", "k_perm",
              "k_collide_rows", "k_collide", "k_advect", "k_cell_gamma", "k_rows_redo", "k_fluid_init",
              "k_fluid_mean", "k_temperature"):
        if k in n:
            return k
    return n[:48]


def load_stats(path):
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append((short(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"]), float(r["AverageNs"]),
                         float(r["Percentage"])))
    return rows


def load_counter(path, counter):
    """mean counter value per kernel (KiB per dispatch), dispatch count"""
    per = defaultdict(list)
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            per[short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return {k: (sum(v) / len(v), len(v)) for k, v in per.items()}


def calibrate(fetch_csv, write_csv, kernel, groups, read_bytes, write_bytes, skip=0):
    """Counter calibration on a probe of known traffic: the dispatches of
    `kernel` (substring) in dispatch order, split into `groups` equal runs
    (one per probe variant); per run the mean FETCH_SIZE and WRITE_SIZE
    (KiB) against the bytes the probe must move."""
    def per_dispatch(path, counter):
        vals = defaultdict(float)
        with open(path) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] == counter and kernel in r["Kernel_Name"]:
                    vals[int(r["Dispatch_Id"])] += float(r["Counter_Value"])
        return [vals[k] for k in sorted(vals)][skip:]
    fv, wv = per_dispatch(fetch_csv, "FETCH_SIZE"), per_dispatch(write_csv, "WRITE_SIZE")
    out = []
    for g in range(groups):
        fg = fv[g * len(fv) // groups:(g + 1) * len(fv) // groups]
        wg = wv[g * len(wv) // groups:(g + 1) * len(wv) // groups]
        f, w = sum(fg) / len(fg) * 1024, sum(wg) / len(wg) * 1024
        out.append((f, w, read_bytes / f, write_bytes / w))
    return out


def closure_summary(trace_csv, stars=256, top=24):
    """Markdown of a kernel trace of the closure benchmark: the whole run's
    kernels, then the closure kernels (MD at `stars` workgroups, features,
    forward pass, kick; by grid size) launched between the first and the
    last `stars`-workgroup k_md launch, the closure intervals of `stars`
    stars."""
    rows = []
    with open(trace_csv) as f:
        for r in csv.DictReader(f):
            wg = int(r.get("Workgroup_Size_X") or 1)
            gx = int(r.get("Grid_Size_X") or 0)
            rows.append((r["Kernel_Name"].split("(")[0].replace("void ", ""), int(r["Start_Timestamp"]),
                         int(r["End_Timestamp"]), gx // max(wg, 1)))
    rows.sort(key=lambda t: t[1])
    out = []
    tot = defaultdict(lambda: [0, 0])
    for n, a, b, _ in rows:
        tot[n][0] += 1
        tot[n][1] += b - a
    out.append("| kernel | calls | total ms | avg us |\n|---|---:|---:|---:|")
    for n, (c, t) in sorted(tot.items(), key=lambda kv: -kv[1][1])[:top]:
        out.append(f"| {n} | {c} | {t / 1e6:.2f} | {t / c / 1e3:.2f} |")
    md = [i for i, r in enumerate(rows) if "k_md" in r[0] and r[3] == stars]
    if md:
        win = defaultdict(list)
        for n, a, b, g in rows[md[0]:md[-1] + 1]:
            if ("k_md" in n and g == stars) or any(k in n for k in ("k_closure_", "k_forward_fused")):
                win[(n, g)].append(b - a)
        out.append(f"\nLaunches of the {stars}-star intervals (first to last {stars}-workgroup k_md):\n")
        out.append("| kernel | workgroups | launches | avg us |\n|---|---:|---:|---:|")
        for (n, g), v in sorted(win.items(), key=lambda kv: -sum(kv[1]) / len(kv[1])):
            out.append(f"| {n} | {g} | {len(v)} | {sum(v) / len(v) / 1e3:.2f} |")
    return "\n".join(out)


def sq_summary(csvs, kernels):
    """Markdown table of SQ counters per kernel (dispatch means, merged over
    the passes in `csvs`): resident waves per SIMD = SQ_WAVE_CYCLES /
    SQ_BUSY_CU_CYCLES; waiting, issuing and VALU active over
    SQ_WAVE_CYCLES; VALU instructions per wave = SQ_INSTS_VALU / SQ_WAVES;
    conflict cycles per LDS access = SQ_LDS_BANK_CONFLICT /
    (SQ_LDS_IDX_ACTIVE - SQ_LDS_BANK_CONFLICT); conflict share of the
    kernel's time = SQ_LDS_BANK_CONFLICT / (GRBM_GUI_ACTIVE / 8) / 256."""
    per = defaultdict(lambda: defaultdict(list))
    names = {}
    for path in csvs:
        with open(path) as f:
            for r in csv.DictReader(f):
                full = r["Kernel_Name"].split("(")[0].replace("void ", "").replace("spmd::", "")
                k = short(r["Kernel_Name"])
                names.setdefault(k, full)
                per[k][r["Counter_Name"]].append(float(r["Counter_Value"]))
    out = ["| kernel | waves/SIMD resident | waiting | issuing | VALU active | VALU instructions per wave "
           "| LDS conflict cycles per access | LDS conflict share of kernel time |",
           "|---|---:|---:|---:|---:|---:|---:|---:|"]
    for k in kernels:
        if k not in per:
            continue
        m = {c: sum(v) / len(v) for c, v in per[k].items()}
        wc = m.get("SQ_WAVE_CYCLES", 0.0) or 1.0
        lds = m.get("SQ_LDS_IDX_ACTIVE", 0.0) - m.get("SQ_LDS_BANK_CONFLICT", 0.0)
        share = m.get("SQ_LDS_BANK_CONFLICT", 0.0) / max(m.get("GRBM_GUI_ACTIVE", 0.0) / 8.0, 1.0) / 256.0
        out.append(f"| {names[k]} | {wc / max(m.get('SQ_BUSY_CU_CYCLES', 1.0), 1.0):.1f} "
                   f"| {m.get('SQ_WAIT_ANY', 0.0) / wc:.0%} | {m.get('SQ_ACTIVE_INST_ANY', 0.0) / wc:.0%} "
                   f"| {m.get('SQ_ACTIVE_INST_VALU', 0.0) / wc:.0%} "
                   f"| {m.get('SQ_INSTS_VALU', 0.0) / max(m.get('SQ_WAVES', 1.0), 1.0):.0f} "
                   f"| {m.get('SQ_LDS_BANK_CONFLICT', 0.0) / max(lds, 1.0):.2f} | {share:.1%} |")
    return "\n".join(out)


def timeline(trace_csv, last="k_collide_rows", intervals=(150, 151, 152)):
    """Kernel start/end (us) of whole intervals, each from the end of the
    previous `last` kernel to the end of its own, with the hardware queue
    each ran on: shows the gaps the dependencies between queues cost."""
    rows = []
    with open(trace_csv) as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]),
                         r["Kernel_Name"].split("(")[0].replace("void ", "").replace("spmd::", ""),
                         r.get("Queue_Id") or r.get("Stream_Id") or "?"))
    rows.sort()
    ends = [i for i, r in enumerate(rows) if last in r[2]]
    out = []
    for c in intervals:
        if c < 1 or c >= len(ends):
            continue
        t0 = rows[ends[c - 1]][1]
        out.append(f"\ninterval ending with {last} #{c} (us after the previous one ended):\n")
        out.append("| kernel | queue | start | end | duration |\n|---|---:|---:|---:|---:|")
        for a, b, n, q in rows[ends[c - 1] + 1:ends[c] + 1]:
            out.append(f"| {n} | {q} | {(a - t0) / 1e3:.2f} | {(b - t0) / 1e3:.2f} | {(b - a) / 1e3:.2f} |")
    return "\n".join(out)


def main():
    import sys
    if len(sys.argv) > 1 and sys.argv[1] == "timeline":
        ap = argparse.ArgumentParser(prog="profsum timeline")
        ap.add_argument("--trace", required=True, help="<dir>/<name>_kernel_trace.csv")
        ap.add_argument("--last", default="k_collide_rows")
        ap.add_argument("--intervals", default="150,151,152")
        a = ap.parse_args(sys.argv[2:])
        print(timeline(a.trace, a.last, [int(x) for x in a.intervals.split(",")]))
        return
    if len(sys.argv) > 1 and sys.argv[1] == "sq":
        ap = argparse.ArgumentParser(prog="profsum sq")
        ap.add_argument("csvs", nargs="+", help="counter_collection.csv of each --pmc pass")
        ap.add_argument("--kernels", default="k_advect,k_cell_gamma,k_rows_redo,k_collide_rows,k_md")
        a = ap.parse_args(sys.argv[2:])
        print(sq_summary(a.csvs, a.kernels.split(",")))
        return
    if len(sys.argv) > 1 and sys.argv[1] == "closure":
        ap = argparse.ArgumentParser(prog="profsum closure")
        ap.add_argument("--trace", required=True, help="<dir>/<name>_kernel_trace.csv")
        ap.add_argument("--stars", type=int, default=256)
        a = ap.parse_args(sys.argv[2:])
        print(closure_summary(a.trace, a.stars))
        return
    if len(sys.argv) > 1 and sys.argv[1] == "calib":
        ap = argparse.ArgumentParser(prog="profsum calib")
        ap.add_argument("--fetch", required=True)
        ap.add_argument("--write", required=True)
        ap.add_argument("--kernel", required=True)
        ap.add_argument("--groups", type=int, default=1)
        ap.add_argument("--skip", type=int, default=0, help="leading dispatches of the kernel that are not probes")
        ap.add_argument("--names", default="")
        ap.add_argument("--read-bytes", type=float, required=True)
        ap.add_argument("--write-bytes", type=float, required=True)
        a = ap.parse_args(sys.argv[2:])
        names = a.names.split(",") if a.names else [str(g) for g in range(a.groups)]
        print("| probe | bytes read | FETCH_SIZE bytes | read scale | bytes written | WRITE_SIZE bytes | write scale |")
        print("|---|---:|---:|---:|---:|---:|---:|")
        for nm, (f, w, fs, ws) in zip(names, calibrate(a.fetch, a.write, a.kernel, a.groups, a.read_bytes,
                                                         a.write_bytes, a.skip)):
            print(f"| {nm} | {a.read_bytes:.0f} | {f:.0f} | {fs:.3f} | {a.write_bytes:.0f} | {w:.0f} | {ws:.3f} |")
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--stats", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-fetch")
    ap.add_argument("--copy-write")
    ap.add_argument("--copy-bytes", type=float, default=0.0)
    ap.add_argument("--particles", type=int, default=0)
    ap.add_argument("--bytes-per-particle", type=float, default=48.0)
    ap.add_argument("--pipeline", choices=sorted(PIPELINES), default="sort",
                    help="which collision path's kernels make up the pipeline")
    ap.add_argument("--read-scale", type=float, default=0.0,
                    help="FETCH_SIZE scale from a probe of the kernels' access shape (overrides the copy's)")
    ap.add_argument("--write-scale", type=float, default=0.0)
    ap.add_argument("--kernel-read-scale", action="append", default=[],
                    help="kernel=scale: FETCH_SIZE scale for one kernel (a gather's scattered lines are counted "
                         "whole, a coalesced stream's at half)")
    a = ap.parse_args()
    kscale = {kv.split("=")[0]: float(kv.split("=")[1]) for kv in a.kernel_read_scale}
    pipeline = PIPELINES[a.pipeline]
    coll = pipeline[-1]  # the collide runs once per collision

    stats = load_stats(a.stats)
    merged = defaultdict(lambda: [0, 0.0])
    for k, n, tot, _, _ in stats:
        merged[k][0] += n
        merged[k][1] += tot
    print("| kernel | calls | total ms | avg us |")
    print("|---|---:|---:|---:|")
    for k, (n, tot) in sorted(merged.items(), key=lambda kv: -kv[1][1]):
        print(f"| {k} | {n} | {tot / 1e6:.3f} | {tot / n / 1e3:.2f} |")
    # per collision: every kernel's total is divided by the number of
    # collisions of the pipeline's path
    ncoll = merged[coll][0] if coll in merged else 1
    print(f"\ncollisions profiled: {ncoll}")
    print("\n| kernel | us per collision |")
    print("|---|---:|")
    for k in ("k_prologue", "k_md") + pipeline + ("k_sort_redo",):
        if k in merged:
            print(f"| {k} | {merged[k][1] / ncoll / 1e3:.2f} |")
    pipe_us = sum(merged[k][1] / ncoll / 1e3 for k in pipeline if k in merged)
    print(f"\ncollision pipeline ({' + '.join(pipeline)}): {pipe_us:.2f} us per collision")
    if a.particles:
        alg = a.particles * a.bytes_per_particle
        print(f"algorithmic bytes per collision: {alg:.0f} ({a.bytes_per_particle:g} B x {a.particles} particles)")
        print(f"achieved (algorithmic): {alg / (pipe_us * 1e-6) / 1e9:.1f} GB/s")

    if a.fetch and a.write:
        fs, ws = 1.0, 1.0
        if a.copy_fetch and a.copy_write and a.copy_bytes:
            cf = max(load_counter(a.copy_fetch, "FETCH_SIZE").values(), key=lambda t: t[0])[0] * 1024
            cw = max(load_counter(a.copy_write, "WRITE_SIZE").values(), key=lambda t: t[0])[0] * 1024
            fs, ws = a.copy_bytes / cf, a.copy_bytes / cw
            print(f"\ncopy calibration ({a.copy_bytes:.0f} B copied): FETCH_SIZE reads {cf:.0f} B -> scale {fs:.3f}; "
                  f"WRITE_SIZE reads {cw:.0f} B -> scale {ws:.3f}")
        if a.read_scale:
            fs = a.read_scale
        if a.write_scale:
            ws = a.write_scale
        print(f"\nscales applied: FETCH_SIZE x {fs:.3f}"
              + "".join(f" ({k}: x {v:.3f})" for k, v in kscale.items()) + f", WRITE_SIZE x {ws:.3f}")
        f = load_counter(a.fetch, "FETCH_SIZE")
        w = load_counter(a.write, "WRITE_SIZE")
        nf = f.get(coll, (0.0, 1))[1]
        print("\nper collision (dispatch means x dispatches / collisions):")
        print("\n| kernel | dispatches | FETCH KiB | WRITE KiB | scaled read MB | scaled write MB |")
        print("|---|---:|---:|---:|---:|---:|")
        tot = 0.0
        for k in sorted(set(f) | set(w)):
            fk, nk = f.get(k, (0.0, 0))
            wk, _ = w.get(k, (0.0, 0))
            per = nk / nf if k in pipeline else 1.0
            rb, wb = fk * 1024 * kscale.get(k, fs) * per, wk * 1024 * ws * per
            if k in pipeline:
                tot += rb + wb
            print(f"| {k} | {nk} | {fk * per:.0f} | {wk * per:.0f} | {rb / 1e6:.1f} | {wb / 1e6:.1f} |")
        print(f"\nHBM traffic of the collision pipeline: {tot:.0f} B per collision")
        if pipe_us:
            print(f"measured bandwidth of the pipeline: {tot / (pipe_us * 1e-6) / 1e9:.1f} GB/s")
        if a.particles:
            print(f"amplification over algorithmic bytes: {tot / (a.particles * a.bytes_per_particle):.2f}")
        print(f"traffic_bytes_per_collision: {tot:.0f}")


if __name__ == "__main__":
    main()
