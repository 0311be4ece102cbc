"""Summaries of rocprofv3 CSV output for the step kernels.

Turns `--kernel-trace --stats` and `--pmc FETCH_SIZE` / `--pmc WRITE_SIZE`
results into short text tables, and derives the HBM bytes each step launch
moves. The memory counters are calibrated against the plain device copy
(`lbfx_copy`) that bench.py runs first: a copy of B bytes reads B and writes B,
so factor = B / counter (over the B-byte copies only, not bench.py's shorter
warm-up copies), applied to every other kernel of the same run.

    python -m lbflow_amd.profiling trace  <kernel_stats.csv> --cmd "..."
    python -m lbflow_amd.profiling traffic <fetch_counter_collection.csv> \
        <write_counter_collection.csv> --copy-bytes 2147483648 --cmd-fetch "..." --cmd-write "..."
    python -m lbflow_amd.profiling run-geometries --steps 10      # the program to trace
    python -m lbflow_amd.profiling counters <counter_collection.csv>... --cmd "..." [--cmd ...]

`run-geometries` steps the 256^3 D3Q19 TRT porous channel and then the
all-fluid periodic box, in FP16-shifted and then FP32 storage, through the
in-place kernel; `counters` assigns each step launch to its geometry by
dispatch order and tabulates the memory-pipeline counters per wavefront
memory instruction.
"""
import argparse
import collections
import csv
import re

# storage tag in the kernel's template arguments -> bench.py variant name
_STORAGE = {"StF32": "f32", "StF16S": "f16s", "StBF16S": "bf16s", "StF32S": "f32s", "StF64": "f64"}


def variant_of(kernel_name):
    """bench.py variant a step kernel belongs to, or None."""
    m = re.search(r"momentum_step_(aa|plain|tb2)<lbf::(\w+), lbf::(\w+),", kernel_name)
    if not m:
        return None
    scheme, _lat, storage = m.groups()
    v = _STORAGE.get(storage)
    if v is None:
        return None
    if scheme == "plain":
        return "f64" if v == "f64" else v + "_two_array"
    if scheme == "tb2":
        # the FP32 two-step kernel is the headline variant "f32"
        return "f32" if v == "f32" else v + "_two_step"
    if v == "f64":
        return "f64_aa"
    return "f32_aa" if v == "f32" else v


def short_name(kernel_name):
    n = kernel_name.replace("lbf::", "").replace("(KArgs)", "")
    n = re.sub(r"HIP_vector_type<unsigned int, 4u>", "uint4", n)
    return n[:88]


def trace_summary(stats_csv, command):
    rows = list(csv.DictReader(open(stats_csv)))
    out = [f"command: {command}", "",
           f"{'kernel':88s} {'calls':>6s} {'avg_us':>9s} {'min_us':>9s} {'max_us':>9s} {'%time':>6s}"]
    for r in rows:
        out.append(f"{short_name(r['Name']):88s} {int(r['Calls']):6d} "
                   f"{float(r['AverageNs']) / 1e3:9.1f} {float(r['MinNs']) / 1e3:9.1f} "
                   f"{float(r['MaxNs']) / 1e3:9.1f} {float(r['Percentage']):6.2f}")
    return "\n".join(out) + "\n"


def _per_kernel(path, counter):
    vals = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] == counter:
            vals[r["Kernel_Name"]].append(float(r["Counter_Value"]) * 1024.0)  # KiB -> bytes
    return vals


def traffic(fetch_csv, write_csv, copy_bytes):
    """(factors, {kernel: (read_bytes, write_bytes, dispatches)}) with calibrated bytes."""
    f = _per_kernel(fetch_csv, "FETCH_SIZE")
    w = _per_kernel(write_csv, "WRITE_SIZE")
    copy = [k for k in f if "copy_kernel" in k]
    if not copy:
        raise ValueError("no copy_kernel dispatch to calibrate against")

    def largest(vals):
        # the copy_bytes copies of bench.py's bandwidth probe; its shorter
        # warm-up copies before each leg read and write less
        top = max(vals)
        big = [v for v in vals if v >= 0.9 * top]
        return sum(big) / len(big)

    ff = copy_bytes / largest(f[copy[0]])
    fw = copy_bytes / largest(w[copy[0]])
    per = {}
    for k in f:
        if k in w:
            per[k] = (ff * sum(f[k]) / len(f[k]), fw * sum(w[k]) / len(w[k]), len(f[k]))
    return (ff, fw), per


def traffic_summary(fetch_csv, write_csv, copy_bytes, cmd_fetch, cmd_write, nodes, q=19):
    (ff, fw), per = traffic(fetch_csv, write_csv, copy_bytes)
    out = [f"command (reads):  {cmd_fetch}", f"command (writes): {cmd_write}", "",
           f"calibration on lbfx_copy of {copy_bytes} B: FETCH_SIZE x {ff:.3f}, WRITE_SIZE x {fw:.3f}",
           "", f"{'kernel':88s} {'n':>3s} {'read_GB':>8s} {'write_GB':>8s} {'alg_GB':>7s} {'ratio':>6s}"]
    by_variant = collections.defaultdict(list)
    for k, (rb, wb, cnt) in per.items():
        v = variant_of(k)
        s = {"f32": 4, "f32_aa": 4, "f16s": 2, "bf16s": 2, "f32s": 4, "f64": 8, "f64_aa": 8,
             "f32_two_array": 4, "f64_two_step": 8, "f16s_two_step": 2, "bf16s_two_step": 2,
             "f32s_two_step": 4}.get(v)
        # a two-step launch (momentum_step_tb2) advances two time steps
        steps = 2 if "momentum_step_tb2" in k else 1
        alg = 2.0 * q * nodes * s * steps if s else None
        ratio = f"{(rb + wb) / alg:6.2f}" if alg else "     -"
        out.append(f"{short_name(k):88s} {cnt:3d} {rb / 1e9:8.3f} {wb / 1e9:8.3f} "
                   f"{(alg or 0) / 1e9:7.3f} {ratio}")
        # streaming steps (no moment store) are the launches a run is made of
        mom = re.search(r"momentum_step_aa<[^>]*, (true|false), (true|false)>", k)
        mom2 = re.search(r"momentum_step_tb2<[^>]*, (true|false), \d+>", k)
        if v and not (mom and mom.group(2) == "true") and not (mom2 and mom2.group(1) == "true"):
            by_variant[v].append(rb + wb)
    out.append("")
    out.append("HBM bytes per step launch (mean of the streaming kernels: in place, the even "
               "and odd steps; two-step, one launch per two steps):")
    for v, xs in sorted(by_variant.items()):
        out.append(f"traffic_bytes_per_launch {v} {sum(xs) / len(xs):.0f}")
    return "\n".join(out) + "\n"


def read_traffic(path):
    """{variant: bytes per launch} from a traffic summary, {} if absent."""
    res = {}
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch "):
                _, v, b = line.split()
                res[v] = float(b)
    except OSError:
        pass
    return res


GEOMETRY_STORAGES = ("f16s", "f32")
GEOMETRIES = ("porous", "all_fluid")


def run_geometries(steps, n=256):
    """Steps each geometry in each storage (order: storage, then porous before
    all-fluid), one case at a time, on the in-place kernel."""
    from . import abi, configs
    from .engine import Case
    prec = {"f16s": abi.F16_SHIFTED, "f32": abi.F32}
    for st in GEOMETRY_STORAGES:
        for geo in GEOMETRIES:
            if geo == "porous":
                sp = configs.porous_channel(n=n, precision=prec[st], streaming=abi.STREAM_AA)
            else:
                sp = configs.periodic_box(n=n, precision=prec[st], streaming=abi.STREAM_AA).replace(
                    initial_velocity=(0.05, 0.0, 0.0))
            c = Case(sp)
            c.initialize_nodes()
            c.main_steps(steps)
            c.synchronize()
            c.close()
            print(f"[run-geometries] {st} {geo}: {steps} steps", flush=True)


# memory-pipeline counters per wavefront memory instruction; the _sum counters
# add up every instance (one TCP per CU, one TCC per L2 channel),
# GRBM_GUI_ACTIVE counts the GPU's busy cycles over the dispatch
def counters_summary(paths, commands):
    per = collections.defaultdict(lambda: collections.defaultdict(list))
    for path in paths:
        disp = collections.defaultdict(dict)
        names = {}
        for r in csv.DictReader(open(path)):
            d = int(r["Dispatch_Id"])
            disp[d][r["Counter_Name"]] = disp[d].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
            names[d] = r["Kernel_Name"]
        by_kernel = collections.defaultdict(list)
        for d in sorted(disp):
            m = re.search(r"momentum_step_aa<lbf::D3Q19, lbf::(\w+), 1, (true|false), (true|false)>",
                          names[d])
            if m and m.group(3) == "false":  # streaming launches (no moment store)
                by_kernel[(_STORAGE.get(m.group(1)), m.group(2) == "true")].append(d)
        for (st, odd), ds in by_kernel.items():
            half = len(ds) // 2
            for gi, geo in enumerate(GEOMETRIES):
                part = ds[gi * half:(gi + 1) * half]
                if not odd:
                    part = part[1:]  # the first even launch reads natural slots
                for d in part:
                    for cn, v in disp[d].items():
                        per[(st, geo, "odd" if odd else "even")][cn].append(v)
    mean = {k: {cn: sum(v) / len(v) for cn, v in c.items()} for k, c in per.items()}
    out = [f"command: {c}" for c in commands]
    out += ["", "Per launch of the in-place step kernel (D3Q19 TRT, 256^3), means over the "
            "streaming launches of each parity:", ""]
    cols = [("VMEM rd/wave", lambda c: c["SQ_INSTS_VMEM_RD"] / c["SQ_WAVES"]),
            ("VMEM wr/wave", lambda c: c["SQ_INSTS_VMEM_WR"] / c["SQ_WAVES"]),
            ("TCP acc/instr", lambda c: c["TCP_TOTAL_CACHE_ACCESSES_sum"] /
             (c["SQ_INSTS_VMEM_RD"] + c["SQ_INSTS_VMEM_WR"])),
            ("L2 rd req/ld", lambda c: c["TCP_TCC_READ_REQ_sum"] / c["SQ_INSTS_VMEM_RD"]),
            ("L2 wr req/st", lambda c: c["TCP_TCC_WRITE_REQ_sum"] / c["SQ_INSTS_VMEM_WR"]),
            ("L2 hit %", lambda c: 100 * c["TCC_HIT_sum"] / (c["TCC_HIT_sum"] + c["TCC_MISS_sum"])),
            ("busy Mcyc", lambda c: c["GRBM_GUI_ACTIVE"] / 1e6)]
    out.append(f"{'storage':8s} {'geometry':10s} {'step':5s} " +
               " ".join(f"{name:>13s}" for name, _ in cols))
    for st in GEOMETRY_STORAGES:
        for geo in GEOMETRIES:
            for par in ("even", "odd"):
                c = mean.get((st, geo, par))
                if not c:
                    continue
                vals = []
                for _, f in cols:
                    try:
                        vals.append(f"{f(c):13.3f}")
                    except (KeyError, ZeroDivisionError):
                        vals.append(f"{'-':>13s}")
                out.append(f"{st:8s} {geo:10s} {par:5s} " + " ".join(vals))
    return "\n".join(out) + "\n"


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    t = sub.add_parser("trace")
    t.add_argument("stats_csv")
    t.add_argument("--cmd", required=True)
    m = sub.add_parser("traffic")
    m.add_argument("fetch_csv")
    m.add_argument("write_csv")
    m.add_argument("--copy-bytes", type=int, default=2 << 30)
    m.add_argument("--nodes", type=int, default=256 ** 3)
    m.add_argument("--cmd-fetch", required=True)
    m.add_argument("--cmd-write", required=True)
    g = sub.add_parser("run-geometries")
    g.add_argument("--steps", type=int, default=10)
    c = sub.add_parser("counters")
    c.add_argument("csv", nargs="+")
    c.add_argument("--cmd", action="append", default=[])
    a = ap.parse_args()
    if a.what == "trace":
        print(trace_summary(a.stats_csv, a.cmd), end="")
    elif a.what == "run-geometries":
        run_geometries(a.steps)
    elif a.what == "counters":
        print(counters_summary(a.csv, a.cmd), end="")
    else:
        print(traffic_summary(a.fetch_csv, a.write_csv, a.copy_bytes, a.cmd_fetch, a.cmd_write,
                              a.nodes), end="")


if __name__ == "__main__":
    main()
