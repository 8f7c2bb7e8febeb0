"""Short text summaries of rocprofv3 CSV output (kernel statistics, SQ
counters, HBM traffic) for the committed profile records.

    python -m stater_amd.profsum stats    OUT_DIR/PREFIX_kernel_stats.csv  --cmd "..."
    python -m stater_amd.profsum counters OUT_DIR/PREFIX_counter_collection.csv --cmd "..."
    python -m stater_amd.profsum traffic  FETCH.csv WRITE.csv CAL_FETCH.csv CAL_WRITE.csv
                                          --cal-bytes N --cmd "..." [--json-dir DIR]

``traffic`` scales each kernel's FETCH_SIZE and WRITE_SIZE (KiB per dispatch)
by the factor that makes the same counter read the known size of a plain
device-to-device copy (``python -m stater_amd.copycal``), and can write
``traffic_<kernel>.json`` files ({"bytes_per_launch": ...}) that the bench
attaches to its roofline entry.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import re
import sys
from pathlib import Path

N_SIMD = 256 * 4        # MI355X: 256 CUs x 4 SIMDs
CLOCK_GHZ = 2.4
CYCLES_PER_VALU = 4     # a wave64 FP32 VALU instruction on a 16-lane SIMD

# kernel symbol -> the library timer whose launches it belongs to (a timer
# can cover several kernels: the variance filter is a pack pass (with the
# prepass) and a filter pass per set group, the mode-2 Welch filter a
# decision pass and a weighting pass)
TIMER_NAMES = [("k_welch", "welch_filter"), ("k_welch_decide", "welch_filter"),
               ("k_welch_accum2", "welch_filter"), ("k_welch_accum", "welch_filter"),
               ("k_var_filter", "var_filter"), ("k_var_pack", "var_filter"),
               ("k_accumulate_merge", "accumulate_merge64"), ("k_accumulate", "accumulate"),
               ("k_accumulate_px", "accumulate"), ("k_film_merge", "accumulate"), ("k_film_resolve", "film_resolve")]


def short(name: str) -> str:
    name = re.sub(r"\(.*$", "", name.replace("void ", ""))
    return name.replace("stater::", "")


def timer_name(kernel: str) -> str | None:
    base = re.sub(r"<.*$", "", short(kernel))
    for sym, t in TIMER_NAMES:
        if base == sym:
            return t
    return None


def _rows(path: str) -> list[dict]:
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def stats(path: str, cmd: str) -> str:
    rows = _rows(path)
    out = [f"Command: `{cmd}`", "", "| kernel | calls | avg ms | min ms | max ms | share |", "|---|---|---|---|---|---|"]
    for r in rows:
        out.append(f"| `{short(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e6:.4f} | "
                   f"{float(r['MinNs']) / 1e6:.4f} | {float(r['MaxNs']) / 1e6:.4f} | {float(r['Percentage']):.1f}% |")
    return "\n".join(out) + "\n"


def _per_dispatch(path: str):
    """{kernel: {counter: [values per dispatch]}} and {kernel: [durations ns]}."""
    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    dur = collections.defaultdict(dict)
    for r in _rows(path):
        k = short(r["Kernel_Name"])
        vals[k][r["Counter_Name"]].append(float(r["Counter_Value"]))
        dur[k][r["Dispatch_Id"]] = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    return vals, {k: list(v.values()) for k, v in dur.items()}


def counters(path: str, cmd: str) -> str:
    vals, dur = _per_dispatch(path)
    names = sorted({c for v in vals.values() for c in v})
    out = [f"Command: `{cmd}`", "", "Averages per dispatch (whole chip).", "",
           "| kernel | " + " | ".join(names) + " | VALU issue util. |",
           "|---|" + "---|" * (len(names) + 1)]
    for k, cs in vals.items():
        if k.startswith("__amd") or k in ("k_zero", "k_features"):
            continue
        cells = [f"{sum(cs[n]) / len(cs[n]):.4g}" if n in cs else "" for n in names]
        util = ""
        if "SQ_INSTS_VALU" in cs and dur.get(k):
            d = sum(dur[k]) / len(dur[k])
            v = sum(cs["SQ_INSTS_VALU"]) / len(cs["SQ_INSTS_VALU"])
            util = f"{v * CYCLES_PER_VALU / (N_SIMD * d * CLOCK_GHZ):.2f}"
        out.append(f"| `{k}` | " + " | ".join(cells) + f" | {util} |")
    out += ["", "VALU issue utilisation = SQ_INSTS_VALU x 4 cycles / (1024 SIMDs x dispatch time x 2.4 GHz), "
            "counting every VALU instruction as a full-rate wave64 op on a 16-lane SIMD; dispatch times are those "
            "of the counter run.  Values near or above 1 mean the kernel issues VALU work every cycle it can "
            "(instructions that complete in fewer than 4 cycles push the figure past 1)."]
    return "\n".join(out) + "\n"


def _kib(path: str, counter: str) -> dict[str, list[float]]:
    vals, _ = _per_dispatch(path)
    return {k: v[counter] for k, v in vals.items() if counter in v}


def traffic(fetch: str, write: str, cal_f: str, cal_w: str, cal_bytes: int, cmd: str, json_dir: str | None,
            calls: int = 1) -> str:
    """Per kernel: mean HBM bytes per dispatch.  Per timer (json): all its
    kernels' bytes over the profiled run divided by `calls` (the number of
    times the program called each timed entry point)."""
    def cal(path, counter):
        for k, v in _kib(path, counter).items():
            if "copyBuffer" in k:
                return cal_bytes / (sum(v) / len(v) * 1024.0)
        raise SystemExit(f"no copy kernel in {path}")

    ff, fw = cal(cal_f, "FETCH_SIZE"), cal(cal_w, "WRITE_SIZE")
    f, w = _kib(fetch, "FETCH_SIZE"), _kib(write, "WRITE_SIZE")
    out = [f"Command: `{cmd}`", "",
           f"Calibration (device-to-device copy of {cal_bytes} bytes): FETCH_SIZE x {ff:.3f}, WRITE_SIZE x {fw:.3f}.",
           "", "| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM read MB | HBM write MB | HBM MB/dispatch |",
           "|---|---|---|---|---|---|---|"]
    per_timer = collections.defaultdict(lambda: [0.0, 0.0, []])
    for k in f:
        if k.startswith("__amd") or k not in w:
            continue
        fk, wk = sum(f[k]) / len(f[k]), sum(w[k]) / len(w[k])
        rd, wr = fk * 1024 * ff, wk * 1024 * fw
        out.append(f"| `{k}` | {len(f[k])} | {fk:.0f} | {wk:.0f} | {rd / 1e6:.1f} | {wr / 1e6:.1f} | {(rd + wr) / 1e6:.1f} |")
        t = timer_name(k)
        if t:
            acc = per_timer[t]
            acc[0] += sum(f[k]) * 1024 * ff
            acc[1] += sum(w[k]) * 1024 * fw
            acc[2].append(k)
    if per_timer:
        out += ["", f"Per timed call ({calls} call(s) of each entry point profiled):", "",
                "| timer | kernels | HBM read MB | HBM write MB | HBM MB/call |", "|---|---|---|---|---|"]
    for t, (rd, wr, ks) in per_timer.items():
        rd, wr = rd / calls, wr / calls
        out.append(f"| {t} | {', '.join('`%s`' % k for k in ks)} | {rd / 1e6:.1f} | {wr / 1e6:.1f} | {(rd + wr) / 1e6:.1f} |")
        if json_dir:
            Path(json_dir, f"traffic_{t}.json").write_text(json.dumps(
                {"kernels": ks, "bytes_per_launch": round(rd + wr), "read_bytes": round(rd), "write_bytes": round(wr),
                 "command": cmd}) + "\n")
    return "\n".join(out) + "\n"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["stats", "counters", "traffic"])
    ap.add_argument("files", nargs="+")
    ap.add_argument("--cmd", required=True)
    ap.add_argument("--cal-bytes", type=int, default=2**30)
    ap.add_argument("--json-dir")
    ap.add_argument("--calls", type=int, default=1, help="traffic: calls of each timed entry point in the run")
    a = ap.parse_args()
    if a.mode == "stats":
        sys.stdout.write(stats(a.files[0], a.cmd))
    elif a.mode == "counters":
        sys.stdout.write(counters(a.files[0], a.cmd))
    else:
        if len(a.files) != 4:
            raise SystemExit("traffic needs FETCH WRITE CAL_FETCH CAL_WRITE csv files")
        sys.stdout.write(traffic(*a.files, a.cal_bytes, a.cmd, a.json_dir, a.calls))
    return 0


if __name__ == "__main__":
    sys.exit(main())
