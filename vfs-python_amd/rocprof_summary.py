"""Summaries of rocprofv3 CSV output for the measurement records.

    python vfs-python_amd/rocprof_summary.py stats DIR TITLE COMMAND > summary.md
        per-kernel calls / total / average / min / max duration of a
        ``rocprofv3 --kernel-trace --stats --output-format csv -d DIR`` run, and the
        blocked and fused kernels per launch shape;

    python vfs-python_amd/rocprof_summary.py kernels TRACE_DIR FETCH_DIR WRITE_DIR SQ_DIR SCALES_JSON TITLE COMMAND
        per kernel of one command run five times (``--kernel-trace``; ``--pmc FETCH_SIZE``;
        ``--pmc WRITE_SIZE``; ``--pmc`` with each half of SQ_COUNTERS, SQ_DIR = the two
        output folders joined by a comma): average
        duration, HBM bytes per call (scaled as in pmc_traffic.json, SCALES_JSON) and
        GB/s, LDS bytes per call and GB/s, LDS bank conflicts per access, and VALU
        instructions per wave;

    python vfs-python_amd/rocprof_summary.py pmc FETCH_DIR WRITE_DIR PROBE_LOG N_ITER > pmc_traffic.json
        HBM bytes per k_pd_block launch per pyramid level from two ``rocprofv3 --pmc``
        runs (FETCH_SIZE, then WRITE_SIZE) of ``bench.py --probe``: each counter is
        scaled by its reading for the probe's copies of 2^26 floats (256 MiB read and
        256 MiB written per copy), and each level's second warp is the one counted.
"""

import collections
import csv
import glob
import json
import re
import sys

COPY_FLOATS = 1 << 26
COPY_BYTES = 4 * COPY_FLOATS


def short_name(n):
    n = n.replace("(anonymous namespace)::", "")
    n = re.sub(r"^void ", "", n)
    i = n.find("(")
    return n[:i] if i >= 0 else n


def _one(d, pattern):
    files = glob.glob(f"{d}/**/{pattern}", recursive=True)
    if not files:
        raise FileNotFoundError(f"no {pattern} under {d}")
    return files[0]


def stats(d, title, cmd, top=18):
    rows = list(csv.DictReader(open(_one(d, "*kernel_trace.csv"))))
    agg = collections.OrderedDict()
    shape = collections.OrderedDict()
    for r in rows:
        n = short_name(r["Kernel_Name"])
        us = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1000.0
        a = agg.setdefault(n, [0, 0.0, float("inf"), 0.0])
        a[0] += 1
        a[1] += us
        a[2] = min(a[2], us)
        a[3] = max(a[3], us)
        if n.startswith(("k_pd_block", "k_adj_fused", "k_adj_block")):
            g = shape.setdefault((n, int(r.get("Grid_Size", r.get("Grid_Size_X", 0)))), [0, 0.0])
            g[0] += 1
            g[1] += us
    tot = sum(a[1] for a in agg.values())
    out = [f"# {title}", "", "Command (one MI355X; `cd /tmp && export TMPDIR=/tmp` first, then from the repository root):",
           "", "```", cmd, "```", "", "| kernel | calls | total us | avg us | min us | max us | % |",
           "|---|---|---|---|---|---|---|"]
    for n, a in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
        out.append(f"| {n[:70]} | {a[0]} | {a[1]:.0f} | {a[1] / a[0]:.3f} | {a[2]:.3f} | {a[3]:.3f} | "
                   f"{100 * a[1] / tot:.2f} |")
    if shape:
        out += ["", "Per launch shape (grid = threads):", "", "| kernel | grid | calls | avg us |", "|---|---|---|---|"]
        for (n, g), v in sorted(shape.items(), key=lambda kv: (kv[0][0], -kv[0][1])):
            out.append(f"| {n} | {g} | {v[0]} | {v[1] / v[0]:.3f} |")
        pd = [v for (n, _), v in shape.items() if n.startswith("k_pd_block")]
        if pd:
            c, t = sum(v[0] for v in pd), sum(v[1] for v in pd)
            out += ["", f"All k_pd_block launches: {c} calls, average {t / c:.3f} us."]
    return "\n".join(out)


SQ_COUNTERS = ["SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_LDS_LOAD_BANDWIDTH", "SQ_INSTS_LDS_STORE_BANDWIDTH",
               "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_LDS_BANK_CONFLICT", "SQ_LDS_IDX_ACTIVE"]
HBM_PEAK_GBS = 8000.0
# 128 B per clock per CU, 256 CUs, 2.4 GHz
LDS_PEAK_GBS = 128 * 256 * 2.4


def _per_kernel(d, names):
    """{kernel: {counter: mean per dispatch}} over the dispatches of a --pmc run."""
    per = collections.OrderedDict()
    for r in csv.DictReader(open(_one(d, "*counter_collection.csv"))):
        if r["Counter_Name"] not in names:
            continue
        e = per.setdefault(int(r["Dispatch_Id"]), [short_name(r["Kernel_Name"]), {}])
        e[1][r["Counter_Name"]] = e[1].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    agg = collections.OrderedDict()
    for n, c in per.values():
        a = agg.setdefault(n, [0, collections.Counter()])
        a[0] += 1
        a[1].update(c)
    return {n: {k: v / a[0] for k, v in a[1].items()} for n, a in agg.items()}


def kernels(trace_dir, fetch_dir, write_dir, sq_dir, scales_json, title, cmd, top=14):
    rows = list(csv.DictReader(open(_one(trace_dir, "*kernel_trace.csv"))))
    dur = collections.OrderedDict()
    for r in rows:
        a = dur.setdefault(short_name(r["Kernel_Name"]), [0, 0.0])
        a[0] += 1
        a[1] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1000.0
    ck = json.load(open(scales_json))["copy_kernel"]
    fe = _per_kernel(fetch_dir, {"FETCH_SIZE"})
    wr = _per_kernel(write_dir, {"WRITE_SIZE"})
    sq = {}
    for d in sq_dir.split(","):  # the SQ counters may be split over several runs
        for n, c in _per_kernel(d, set(SQ_COUNTERS)).items():
            sq.setdefault(n, {}).update(c)
    tot = sum(a[1] for a in dur.values())
    out = [f"# {title}", "", "Commands (one MI355X; `cd /tmp && export TMPDIR=/tmp` first, then from the repository root):",
           "", "```"]
    out += [f"rocprofv3 --kernel-trace --stats --output-format csv -d gpurun_out/k_trace -o t -- {cmd}",
            f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d gpurun_out/k_fetch -o f -- {cmd}",
            f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d gpurun_out/k_write -o w -- {cmd}",
            f"rocprofv3 --pmc {' '.join(SQ_COUNTERS[:4])} --output-format csv -d gpurun_out/k_sq1 -o s -- {cmd}",
            f"rocprofv3 --pmc {' '.join(SQ_COUNTERS[4:])} --output-format csv -d gpurun_out/k_sq2 -o s -- {cmd}",
            "python3 vfs-python_amd/rocprof_summary.py kernels gpurun_out/k_trace gpurun_out/k_fetch gpurun_out/k_write "
            "gpurun_out/k_sq1,gpurun_out/k_sq2 profiles/pmc_traffic.json TITLE COMMAND", "```", "",
            f"HBM bytes: FETCH_SIZE x {ck['fetch_scale']} and WRITE_SIZE x {ck['write_scale']} KiB (the scales measured "
            "on a copy of known size, pmc_traffic.json).  LDS bytes: (SQ_INSTS_LDS_LOAD_BANDWIDTH + "
            "SQ_INSTS_LDS_STORE_BANDWIDTH) x 64.  Rates are per call over the trace's average duration; peaks "
            f"{HBM_PEAK_GBS:.0f} GB/s HBM and {LDS_PEAK_GBS:.0f} GB/s LDS (128 B/clk/CU x 256 CUs x 2.4 GHz).  "
            "Bank conflicts: SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE.  VALU / wave: SQ_INSTS_VALU / SQ_WAVES.", "",
            "| kernel | calls | % time | avg us | HBM MB / call | HBM GB/s | % HBM peak | LDS MB / call | LDS GB/s | "
            "% LDS peak | bank confl. / access | VALU / wave |",
            "|---|---|---|---|---|---|---|---|---|---|---|---|"]
    for n, (c, t) in sorted(dur.items(), key=lambda kv: -kv[1][1])[:top]:
        us = t / c
        hb = None
        if n in fe and n in wr:
            hb = (fe[n]["FETCH_SIZE"] * ck["fetch_scale"] + wr[n]["WRITE_SIZE"] * ck["write_scale"]) * 1024
        q = sq.get(n, {})
        lb = (q.get("SQ_INSTS_LDS_LOAD_BANDWIDTH", 0.0) + q.get("SQ_INSTS_LDS_STORE_BANDWIDTH", 0.0)) * 64 if q else None
        conf = (q["SQ_LDS_BANK_CONFLICT"] / q["SQ_LDS_IDX_ACTIVE"]) if q.get("SQ_LDS_IDX_ACTIVE") else None
        vw = (q["SQ_INSTS_VALU"] / q["SQ_WAVES"]) if q.get("SQ_WAVES") else None
        f = lambda v, fmt: "-" if v is None else format(v, fmt)  # noqa: E731
        hgbs = None if hb is None else hb / (us * 1e3)
        lgbs = None if lb is None else lb / (us * 1e3)
        out.append(f"| {n[:60]} | {c} | {100 * t / tot:.1f} | {us:.2f} | {f(hb and hb / 1e6, '.2f')} | "
                   f"{f(hgbs, '.0f')} | {f(hgbs and 100 * hgbs / HBM_PEAK_GBS, '.1f')} | "
                   f"{f(lb and lb / 1e6, '.2f')} | {f(lgbs, '.0f')} | {f(lgbs and 100 * lgbs / LDS_PEAK_GBS, '.1f')} | "
                   f"{f(conf, '.3f')} | {f(vw, '.0f')} |")
    return "\n".join(out)


def _counter(d, name):
    per = collections.OrderedDict()
    for r in csv.DictReader(open(_one(d, "*counter_collection.csv"))):
        if r["Counter_Name"] != name:
            continue
        e = per.setdefault(int(r["Dispatch_Id"]), [short_name(r["Kernel_Name"]), int(r["Grid_Size"]), 0.0])
        e[2] += float(r["Counter_Value"])
    return [v for _, v in sorted(per.items())]


def pmc(fetch_dir, write_dir, probe_log, n_iter):
    fe, wr = _counter(fetch_dir, "FETCH_SIZE"), _counter(write_dir, "WRITE_SIZE")
    levels = []
    for line in open(probe_log):
        m = re.match(r"\[probe\] (\d+)x(\d+) tile \((\d+), (\d+), (\d+), (\d+)\)", line)
        if m:
            h, w, tx, ty, k, nt = map(int, m.groups())
            levels.append(((h, w), (tx, ty, k, nt)))

    def copies(p):
        return sorted(v[2] for v in p if v[0] in ("k_copy", "k_copy4") and v[1] == COPY_FLOATS // 4)

    cf, cw = copies(fe), copies(wr)
    if not cf or not cw:
        raise ValueError("no copy-kernel readings to scale by")
    sf = COPY_BYTES / (cf[len(cf) // 2] * 1024)
    sw = COPY_BYTES / (cw[len(cw) // 2] * 1024)
    pf = [v for v in fe if v[0].startswith("k_pd_block")]
    pw = [v for v in wr if v[0].startswith("k_pd_block")]
    if len(pf) != len(pw):
        raise ValueError("the two runs dispatched different numbers of k_pd_block launches")
    out, i = [], 0
    for (h, w), (tx, ty, k, nt) in levels:
        per_warp = -(-n_iter // k)
        cfe, cwr = pf[i + per_warp:i + 2 * per_warp], pw[i + per_warp:i + 2 * per_warp]
        i += 2 * per_warp
        want = f"k_pd_block<{tx}, {ty}, {k}, {nt}>"
        if not all(c[0] == want for c in cfe):
            raise ValueError(f"expected {want}, found {cfe[0][0]}")
        out.append(dict(size=[h, w], tile=[tx, ty, k, nt], launches_per_warp=per_warp,
                        fetch_bytes=round(sum(c[2] for c in cfe) * 1024 * sf / per_warp),
                        write_bytes=round(sum(c[2] for c in cwr) * 1024 * sw / per_warp)))
    if i != len(pf):
        raise ValueError(f"{len(pf) - i} k_pd_block launches not assigned to a level")
    return dict(kernel="k_pd_block", counters=["FETCH_SIZE", "WRITE_SIZE"], n_iter=n_iter,
                copy_kernel=dict(bytes_read=COPY_BYTES, bytes_written=COPY_BYTES, fetch_kib=cf, write_kib=cw,
                                 fetch_scale=round(sf, 5), write_scale=round(sw, 5)),
                levels=out)


def main(argv):
    if len(argv) >= 4 and argv[0] == "stats":
        print(stats(argv[1], argv[2], argv[3]))
    elif len(argv) == 8 and argv[0] == "kernels":
        print(kernels(*argv[1:8]))
    elif len(argv) == 5 and argv[0] == "pmc":
        print(json.dumps(pmc(argv[1], argv[2], argv[3], int(argv[4])), indent=1))
    else:
        sys.exit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
