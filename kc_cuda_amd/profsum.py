"""Summaries of rocprofv3 CSV output, small enough to keep with the sources.

    python -m kc_cuda_amd.profsum stats <kernel_stats.csv> --solves S --title T --command C [--note N] [--rows R]
        per-kernel time table (markdown) from `rocprofv3 --kernel-trace --stats --output-format csv`
    python -m kc_cuda_amd.profsum pmc <counter_collection.csv> [<counter_collection.csv> ...] --out F.json
        per-kernel sums of every counter (launches, totals) from `rocprofv3 --pmc ... --output-format csv`
    python -m kc_cuda_amd.profsum traffic <fetch.json> <write.json> --kernel K --workload W --copy-bytes B --out F.json
        HBM bytes per launch of kernel K (several kernels joined by '+'), or per solve with --kernel solve:
        FETCH_SIZE and WRITE_SIZE (KiB) scaled by what the same counters read for the largest
        device-to-device copy in the run (B bytes)
    python -m kc_cuda_amd.profsum lds <pmc.json> --title T --command C [--note N] [--rows R]
        LDS bank-conflict table: SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE per kernel
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import re
import sys

COPY_KERNEL = "__amd_rocclr_copyBuffer"


def short_name(name: str) -> str:
    """Kernel name without the argument list, `void ` and anonymous namespaces."""
    n = re.sub(r"\(anonymous namespace\)::", "", name)
    n = re.sub(r"^void ", "", n)
    depth, out = 0, []
    for ch in n:   # drop the argument list, keep template arguments
        if ch == "(" and depth == 0:
            break
        depth += ch == "<"
        depth -= ch == ">"
        out.append(ch)
    return "".join(out).strip()


def base_name(name: str) -> str:
    """Short name without namespace and template arguments (kt::k_x<8> -> k_x)."""
    return short_name(name).split("<")[0].split("::")[-1]


def stats_markdown(path: str, solves: float, title: str, command: str, note: str = "", rows: int = 25) -> str:
    recs = list(csv.DictReader(open(path)))
    tot = sum(float(r["TotalDurationNs"]) for r in recs)
    per = f" ({tot / 1e6 / solves:.1f} ms per solve)" if solves else ""
    out = [f"# {title}", "", "Command (from the repository root):", "", f"    {command}", ""]
    if note:
        out += [note, ""]
    ps_head, ps_rule = (" per solve ms |", "---:|") if solves else ("", "")
    out += [f"Total kernel time {tot / 1e6:.1f} ms{per}.", "",
            f"| kernel | calls | total ms |{ps_head} avg us | max us | share |",
            f"|---|---:|---:|{ps_rule}---:|---:|---:|"]
    recs.sort(key=lambda r: -float(r["TotalDurationNs"]))
    for r in recs[:rows]:
        t = float(r["TotalDurationNs"])
        ps = f" {t / 1e6 / solves:.3f} |" if solves else ""
        out.append(f"| `{short_name(r['Name'])[:70]}` | {r['Calls']} | {t / 1e6:.3f} |{ps} "
                   f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | {100 * t / tot:.1f}% |")
    if len(recs) > rows:
        rest = sum(float(r["TotalDurationNs"]) for r in recs[rows:])
        blank = " |" if solves else ""
        out.append(f"| ({len(recs) - rows} more kernels) | | {rest / 1e6:.3f} |{blank} | | {100 * rest / tot:.1f}% |")
    return "\n".join(out) + "\n"


def pmc_sums(paths: list[str]) -> dict:
    """Per kernel (short name): launches and the sum of every counter; the
    device-to-device copies are also kept one by one (for calibration)."""
    agg: dict = collections.defaultdict(lambda: collections.defaultdict(float))
    launches: dict = collections.defaultdict(set)
    copies: dict = collections.defaultdict(lambda: collections.defaultdict(float))
    for path in paths:
        for r in csv.DictReader(open(path)):
            k = short_name(r["Kernel_Name"])
            did = (path, r["Dispatch_Id"])
            agg[k][r["Counter_Name"]] += float(r["Counter_Value"])
            launches[k].add(did)
            if COPY_KERNEL in k:
                copies[did[1]][r["Counter_Name"]] += float(r["Counter_Value"])
    kernels = {k: {"launches": len(launches[k]), **{c: v for c, v in agg[k].items()}} for k in agg}
    biggest = max(copies.values(), key=lambda c: sum(c.values()), default={})
    return {"kernels": kernels, "largest_copy": dict(biggest)}


def traffic(fetch: dict, write: dict, kernel: str, workload: str, copy_bytes: int, what: str,
            commands: list[str]) -> dict:
    copy_kib = copy_bytes / 1024.0
    cf = fetch["largest_copy"].get("FETCH_SIZE", 0.0)
    cw = write["largest_copy"].get("WRITE_SIZE", 0.0)
    if cf <= 0 or cw <= 0:
        raise SystemExit("no device-to-device copy with FETCH_SIZE / WRITE_SIZE in the run")
    ff, wf = copy_kib / cf, copy_kib / cw

    def select(d: dict) -> list[str]:
        if kernel == "solve":  # every kernel but the calibration copy
            return list(d["kernels"])
        names = kernel.split("+")
        return [k for k in d["kernels"] if base_name(k) in names]

    fk, wk = select(fetch), select(write)
    if not fk:
        raise SystemExit(f"kernel {kernel} not in the run")
    fsum = sum(fetch["kernels"][k].get("FETCH_SIZE", 0.0) for k in fk)
    wsum = sum(write["kernels"][k].get("WRITE_SIZE", 0.0) for k in wk)
    if kernel == "solve":
        fsum -= cf
        wsum -= cw
    launches = sum(fetch["kernels"][k]["launches"] for k in fk)
    hbm = fsum * ff * 1024 + wsum * wf * 1024
    rec = {"kernel": kernel, "workload": workload, "what": what,
           "kernels": sorted(fk), "launches": launches,
           "fetch_size_kib_total": fsum, "write_size_kib_total": wsum,
           "calibration": {"copy_bytes": copy_bytes, "copy_fetch_kib": cf, "copy_write_kib": cw,
                           "fetch_factor": ff, "write_factor": wf},
           "hbm_fetch_bytes": fsum * ff * 1024, "hbm_write_bytes": wsum * wf * 1024,
           "commands": commands}
    if kernel == "solve":
        rec["hbm_bytes_per_solve"] = hbm
    else:
        rec["hbm_bytes_per_launch"] = hbm / max(launches, 1)
    return rec


def lds_markdown(pmc: dict, title: str, command: str, note: str = "", rows: int = 20) -> str:
    ks = pmc["kernels"]
    recs = [(k, v) for k, v in ks.items() if v.get("SQ_LDS_IDX_ACTIVE", 0) > 0]
    recs.sort(key=lambda kv: -kv[1]["SQ_LDS_IDX_ACTIVE"])
    tot_a = sum(v["SQ_LDS_IDX_ACTIVE"] for _, v in recs)
    tot_c = sum(v.get("SQ_LDS_BANK_CONFLICT", 0) for _, v in recs)
    out = [f"# {title}", "", "Command (from the repository root):", "", f"    {command}", ""]
    if note:
        out += [note, ""]
    out += [f"All kernels: {tot_c:.4g} bank-conflict cycles over {tot_a:.4g} LDS-active cycles "
            f"({100 * tot_c / max(tot_a, 1):.1f}%).", "",
            "| kernel | launches | LDS-active cycles | bank-conflict cycles | conflict / active | LDS instructions |",
            "|---|---:|---:|---:|---:|---:|"]
    for k, v in recs[:rows]:
        a, c = v["SQ_LDS_IDX_ACTIVE"], v.get("SQ_LDS_BANK_CONFLICT", 0)
        ins = v.get("SQ_INSTS_LDS")
        out.append(f"| `{k[:70]}` | {v['launches']} | {a:.4g} | {c:.4g} | {100 * c / a:.1f}% | "
                   f"{'-' if ins is None else f'{ins:.4g}'} |")
    return "\n".join(out) + "\n"


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(prog="python -m kc_cuda_amd.profsum")
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--solves", type=float, default=0.0)
    s.add_argument("--title", required=True)
    s.add_argument("--command", required=True)
    s.add_argument("--note", default="")
    s.add_argument("--rows", type=int, default=25)
    s.add_argument("--out")
    p = sub.add_parser("pmc")
    p.add_argument("csv", nargs="+")
    p.add_argument("--out", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("fetch")
    t.add_argument("write")
    t.add_argument("--kernel", required=True)
    t.add_argument("--workload", required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("--what", default="")
    t.add_argument("--command", action="append", default=[])
    t.add_argument("--out", required=True)
    l_ = sub.add_parser("lds")
    l_.add_argument("pmc")
    l_.add_argument("--title", required=True)
    l_.add_argument("--command", required=True)
    l_.add_argument("--note", default="")
    l_.add_argument("--rows", type=int, default=20)
    l_.add_argument("--out")
    a = ap.parse_args(argv)

    def emit(text: str, out: str | None) -> None:
        if out:
            with open(out, "w") as f:
                f.write(text)
        else:
            sys.stdout.write(text)

    if a.cmd == "stats":
        emit(stats_markdown(a.csv, a.solves, a.title, a.command, a.note, a.rows), a.out)
    elif a.cmd == "pmc":
        emit(json.dumps(pmc_sums(a.csv), indent=1) + "\n", a.out)
    elif a.cmd == "traffic":
        rec = traffic(json.load(open(a.fetch)), json.load(open(a.write)), a.kernel, a.workload,
                      a.copy_bytes, a.what, a.command)
        emit(json.dumps(rec, indent=1) + "\n", a.out)
    else:
        emit(lds_markdown(json.load(open(a.pmc)), a.title, a.command, a.note, a.rows), a.out)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
