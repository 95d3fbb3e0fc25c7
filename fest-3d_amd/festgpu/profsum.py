"""Summaries of rocprofv3 CSV output for the engine's kernels.

  python -m festgpu.profsum stats DIR
      kernel table from the *_kernel_stats.csv under DIR (--kernel-trace
      --stats --output-format csv): calls, average and total duration.
  python -m festgpu.profsum sq DIR [DIR ...]
      per-kernel VALU and wait figures from --pmc runs with SQ_WAVES,
      SQ_INSTS_VALU, SQ_WAVE_CYCLES, SQ_WAIT_INST_ANY, SQ_INSTS_VMEM_RD,
      GRBM_GUI_ACTIVE (and SQC_ICACHE_REQ/HITS).
  python -m festgpu.profsum traffic --fetch DIR --write DIR --case NAME
      HBM bytes per launch of each kernel from two counter runs of the
      engine (--pmc FETCH_SIZE, --pmc WRITE_SIZE; the counters are KiB per
      dispatch), scaled by what the same counters read for the engine's
      k_step_begin, whose traffic is known exactly (the state read, Temp
      and the RK base written).  The table rows are
      `| group | kernel | fetch B | write B | traffic B |`.

Run from fest-3d_amd/ (or with it on sys.path).
"""
import argparse
import csv
import glob
import os
import sys
from collections import defaultdict

from . import cases as C
from . import roofline as R


GROUPS = [  # (substring of the demangled kernel name, kernel group)
    ("k_flux<0", "flux_i"), ("k_flux<1", "flux_j"), ("k_flux<2", "flux_k"),
    ("k_grad<", "grad"), ("k_grad_bc", "grad_bc"), ("k_mut_ghost_tp", "grad_bc"), ("k_update", "update"),
    ("k_bc_face", "bc_face"), ("k_bc_edges", "bc_edges"), ("k_step_begin", "step_begin"),
    ("k_dt_min", "time_step"), ("k_dt_fill", "time_step"), ("k_dt", "time_step"),
    ("k_norm", "resnorm"), ("k_node_dist", "setup"), ("k_cell_dist", "setup"), ("k_cells", "setup"),
    ("k_face_metrics", "setup"), ("k_ghost_nodes", "setup"), ("k_fill", "setup"),
    ("k_lusgs_sweep", "lusgs_sweep"), ("k_lusgs_prep", "lusgs_diag"),
    ("k_to_skew", "lusgs_diag"), ("k_lusgs_apply", "lusgs_apply"), ("k_exchange", "exchange"),
    ("k_kkl_source", "kkl_source"), ("k_to_f32", "f32_copy"), ("k_g32_widen", "f32_copy"),
    ("k_vjp_", "vjp"),
    # batched group steps: one launch per phase over all blocks
    ("km_flux<0", "flux_i"), ("km_flux<1", "flux_j"), ("km_flux<2", "flux_k"), ("km_grad<", "grad"),
    ("km_grad_bc", "grad_bc"), ("km_update", "update"), ("km_bc_face", "bc_face"), ("km_bc_edges", "bc_edges"),
    ("km_step_begin", "step_begin"), ("km_dt", "time_step"), ("km_to_f32", "f32_copy"),
]


def group_of(name):
    for key, grp in GROUPS:
        if key in name:
            return grp
    return "-"


def short(name):
    """fest::k_flux<0>(fest::Blk, int) -> k_flux<0>"""
    n = name.split("(")[0]
    for pre in ("void ", "fest::"):
        n = n.replace(pre, "")
    return n.strip()


def _csvs(d, suffix):
    out = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))
    if not out:
        sys.exit(f"no *{suffix} under {d}")
    return out


def stats(d):
    rows = []
    for path in _csvs(d, "kernel_stats.csv"):
        with open(path) as f:
            for r in csv.DictReader(f):
                rows.append((short(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"]),
                             float(r["AverageNs"]), float(r["Percentage"])))
    rows.sort(key=lambda r: -r[2])
    lines = ["| kernel | group | calls | avg us | total ms | % |", "|---|---|---|---|---|---|"]
    for n, c, tot, avg, pct in rows:
        lines.append(f"| {n} | {group_of(n)} | {c} | {avg / 1e3:.1f} | {tot / 1e6:.3f} | {pct:.1f} |")
    return "\n".join(lines)


def counters(d, counter):
    """Average counter value per dispatch, by kernel."""
    acc = defaultdict(list)
    for path in _csvs(d, "counter_collection.csv"):
        with open(path) as f:
            for r in csv.DictReader(f):
                if r.get("Counter_Name") != counter:
                    continue
                acc[short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def counter_max(d, counter, kernel):
    """Largest counter value of one dispatch of `kernel`."""
    best = None
    for path in _csvs(d, "counter_collection.csv"):
        with open(path) as f:
            for r in csv.DictReader(f):
                if r.get("Counter_Name") == counter and short(r["Kernel_Name"]) == kernel:
                    v = float(r["Counter_Value"])
                    best = v if best is None else max(best, v)
    return best


def traffic(fetch_dir, write_dir, case_name):
    """HBM bytes per launch of every kernel.  The counters' scale comes from
    the engine's own k_step_begin at the start of an RK4 step, whose traffic
    is exact: it streams the n_var state planes in and writes Temp and the
    n_var RK-base planes (8 B per value, nc padded cells per plane).  Its
    largest dispatch is that one (the state upload launches it without the
    RK-base copy)."""
    fe, wr = counters(fetch_dir, "FETCH_SIZE"), counters(write_dir, "WRITE_SIZE")
    workload, size = case_name.rsplit("_", 1)
    case = C.WORKLOADS[workload](*(int(v) for v in size.split("x")))[0]
    nc = (case.imx + 5) * (case.jmx + 5) * (case.kmx + 5)
    known_f, known_w = 8.0 * nc * case.n_var, 8.0 * nc * (case.n_var + 1)
    pf = counter_max(fetch_dir, "FETCH_SIZE", "k_step_begin")
    pw = counter_max(write_dir, "WRITE_SIZE", "k_step_begin")
    kf = known_f / (1024.0 * pf)
    kw = known_w / (1024.0 * pw)
    lines = [f"k_step_begin (RK-base copy): FETCH_SIZE {pf:.0f} KiB and WRITE_SIZE {pw:.0f} KiB per launch for "
             f"{known_f:.0f} B read and {known_w:.0f} B written: scale factors fetch {kf:.4f}, write {kw:.4f}",
             "",
             "| group | kernel | fetch B | write B | traffic B |", "|---|---|---|---|---|"]
    extra = ["", "| group | kernel | algorithmic B | traffic / algorithmic |", "|---|---|---|---|"]
    for k in sorted(fe, key=lambda k: -fe[k]):
        grp = group_of(k)
        if grp == "-":
            continue
        f_b = fe[k] * 1024.0 * kf
        w_b = wr.get(k, 0.0) * 1024.0 * kw
        lines.append(f"| {grp} | {k} | {f_b:.4g} | {w_b:.4g} | {f_b + w_b:.6g} |")
        alg = R.bytes_per_launch(case, grp) if case is not None else None
        if alg:
            extra.append(f"| {grp} | {k} | {alg:.6g} | {(f_b + w_b) / alg:.3f} |")
    return "\n".join(lines + extra)


def sq(d):
    """Per kernel: waves, VALU instructions per wave, the fraction of SIMD
    cycles issuing VALU (SQ_INSTS_VALU x 4 cycles over 1024 SIMDs x the
    kernel's cycles; GRBM_GUI_ACTIVE sums the 8 XCDs), vector-memory reads
    per wave, the fraction of wave cycles waiting on dependencies, and the
    instruction-cache hit rate when those counters were collected."""
    acc = defaultdict(lambda: defaultdict(list))
    for path in _csvs(d, "counter_collection.csv"):
        with open(path) as f:
            for r in csv.DictReader(f):
                acc[short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    lines = ["| kernel | group | waves | VALU/wave | VALU busy | VMEM rd/wave | dep. wait | I-cache hit |",
             "|---|---|---|---|---|---|---|---|"]
    for k in sorted(acc):
        grp = group_of(k)
        if grp in ("setup", "-", "resnorm"):
            continue
        m = {c: sum(v) / len(v) for c, v in acc[k].items()}
        waves = m.get("SQ_WAVES", 0.0)

        def fmt(x, f):
            return f.format(x) if x is not None else "-"
        valu = m["SQ_INSTS_VALU"] / waves if waves and "SQ_INSTS_VALU" in m else None
        busy = (m["SQ_INSTS_VALU"] * 4.0 / (1024.0 * m["GRBM_GUI_ACTIVE"] / 8.0)
                if "SQ_INSTS_VALU" in m and m.get("GRBM_GUI_ACTIVE") else None)
        vmem = m["SQ_INSTS_VMEM_RD"] / waves if waves and "SQ_INSTS_VMEM_RD" in m else None
        wait = (m["SQ_WAIT_INST_ANY"] / m["SQ_WAVE_CYCLES"]
                if "SQ_WAIT_INST_ANY" in m and m.get("SQ_WAVE_CYCLES") else None)
        ic = (m["SQC_ICACHE_HITS"] / m["SQC_ICACHE_REQ"] if m.get("SQC_ICACHE_REQ") else None)
        lines.append(f"| {k} | {grp} | {waves:.0f} | {fmt(valu, '{:.0f}')} | {fmt(busy, '{:.2f}')} | "
                     f"{fmt(vmem, '{:.0f}')} | {fmt(wait, '{:.2f}')} | {fmt(ic, '{:.3f}')} |")
    return "\n".join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    q = sub.add_parser("sq")
    q.add_argument("dir", nargs="+")
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--case", required=True)
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        print(stats(a.dir))
    elif a.cmd == "sq":
        print("\n\n".join(sq(d) for d in a.dir))
    else:
        print(traffic(a.fetch, a.write, a.case))


if __name__ == "__main__":
    main()
