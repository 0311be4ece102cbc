"""Benchmark of the lattice step on one MI355X.

Headline workload: D3Q19 TRT body-force channel through a Bernoulli(0.3)
porous solid, 256^3 nodes, periodic in x, no-slip y/z walls, force 1e-6
along x (magic parameter 3/16, tau 0.8), populations in FP32. One "step" is
one lattice time step of the whole 256^3 lattice. MLUPS = N * steps / t / 1e6.

Prints one JSON line. The headline (FP32) runs two time steps per launch
(LBF_STREAM_TWO_STEP). Besides it, the same workload is timed under "extra"
in FP32 in place (f32_aa, the same bits), FP64 (f64_aa: the default FP64
path, in place, with the reference's bits; f64: the two-array scheme;
f64_two_step), FP32 on two arrays and the shifted FP32/FP16/BF16 storage
tiers. The roofline block uses algorithmic
bytes 2*Q*s per node update; the CPU baseline times the oracle (the
line-by-line f64 restatement of the reference solver) on the host cores for
a bounded sample of the same workload. The same kernels on an all-fluid, fully
periodic 256^3 box (uniform initial velocity 0.05 along x, no bounce-back
links) are reported as extra.all_fluid_periodic: the clean roofline point.
Every achieved_gbs / hbm_frac counts those algorithmic bytes, so a two-step
variant, which moves fewer real bytes than two single steps, can pass 1.0;
roofline.traffic holds the HBM bytes the headline kernel really moves.

drop_in_loop: the headline workload through lbf_run, the reference's time loop
with residual norms and the stop rule after every step (src/momentum.rs:273-292).
extra.other_workloads: GPU and oracle (CPU, one thread per core) rates on the
128^2 lid-driven cavity and the 1024x512 coupled discs.
Each timed leg starts with about 50 ms of plain copies on the device (the
case's state untouched), then its warm-up steps, then exactly `steps` timed
steps: from idle the device lowers its core clock for about the first 20 ms
of a load, which a 20-step window would otherwise measure (extra.timing).
serving: the 2,000-case serving mix as specified, on a seeded subsample run at
full step counts (--serve-sample 2000 runs the whole mix), against sequential
plain FP64 runs and the oracle on the host.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per step launch measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE),
# calibrated on the plain copy kernel; written by lbflow_amd.profiling.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_d3q19_trt_porous_256.txt")
METRIC = "MLUPS & % HBM peak, D3Q19 256³; served cases/s + p99 latency vs plain f64"


class Hip:
    """Minimal HIP event timing through libamdhip64 (stream of the case)."""

    def __init__(self):
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                self.lib = C.CDLL(name)
                break
            except OSError:
                continue
        L = self.lib
        L.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        L.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        L.hipEventSynchronize.argtypes = [C.c_void_p]
        L.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        L.hipEventDestroy.argtypes = [C.c_void_p]
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        L.hipDeviceSynchronize.argtypes = []

    def event(self):
        e = C.c_void_p()
        assert self.lib.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.lib.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.lib.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.lib.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value


def copy_bandwidth(hip, mib=2048, reps=10):
    """Device-to-device copy rate (GB/s counted as bytes read + written)."""
    from lbflow_amd.engine import load_library
    lib = load_library()
    lib.lbfx_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
    nbytes = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.lib.hipMalloc(C.byref(a), nbytes) == 0
    assert hip.lib.hipMalloc(C.byref(b), nbytes) == 0
    for _ in range(2):
        lib.lbfx_copy(b, a, nbytes, None)
    e0, e1 = hip.event(), hip.event()
    hip.record(e0, None)
    for _ in range(reps):
        lib.lbfx_copy(b, a, nbytes, None)
    hip.record(e1, None)
    ms = hip.elapsed_ms(e0, e1)
    hip.lib.hipFree(a)
    hip.lib.hipFree(b)
    return 2 * nbytes * reps / (ms / 1e3) / 1e9


def sweep_bandwidth(hip, n_nodes, es, q=19, reps=10):
    """Rate of the in-place kernel's access pattern alone (lbfx_sweep: every
    lane reads its element of each of q planes and writes it to the paired
    plane, no arithmetic, no bounce-back) over n_nodes elements of es bytes,
    GB/s counted as bytes read + written: the attainable bound for that
    pattern on this device."""
    from lbflow_amd.engine import load_library
    lib = load_library()
    ld = (n_nodes + 63) // 64 * 64
    nbytes = q * ld * es
    a = C.c_void_p()
    assert hip.lib.hipMalloc(C.byref(a), nbytes) == 0
    assert hip.lib.hipMemset(a, 0, nbytes) == 0
    for _ in range(2):
        assert lib.lbfx_sweep(a, q, ld, es, n_nodes, None) == 0
    e0, e1 = hip.event(), hip.event()
    hip.record(e0, None)
    for _ in range(reps):
        assert lib.lbfx_sweep(a, q, ld, es, n_nodes, None) == 0
    hip.record(e1, None)
    ms = hip.elapsed_ms(e0, e1)
    hip.lib.hipFree(a)
    return 2 * q * n_nodes * es * reps / (ms / 1e3) / 1e9


PREROLL_MS = 50.0
_preroll_bufs = []


def preroll(hip, stream, ms=PREROLL_MS):
    """Keeps the device busy with plain copies (256 MiB, on the case's stream,
    about `ms` of device time) right before a leg's warm-up steps. From idle,
    the device's power management lowers the core clock for roughly the
    first 20 ms of a sustained load and then restores it; a short timed
    window right after the warm-up would otherwise measure that transient
    instead of the kernel. The case's state is not touched."""
    from lbflow_amd.engine import load_library
    lib = load_library()
    lib.lbfx_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
    nbytes = 256 << 20
    if not _preroll_bufs:
        a, b = C.c_void_p(), C.c_void_p()
        assert hip.lib.hipMalloc(C.byref(a), nbytes) == 0
        assert hip.lib.hipMalloc(C.byref(b), nbytes) == 0
        assert hip.lib.hipMemset(a, 0, nbytes) == 0
        _preroll_bufs.extend([a, b])
    a, b = _preroll_bufs
    # 2 x 256 MiB per copy at about 6 TB/s: about 0.09 ms each
    for _ in range(max(1, int(ms / 0.09))):
        assert lib.lbfx_copy(b, a, nbytes, stream) == 0


def time_case(spec, steps, warmup, hip, loop=False):
    """Times `steps` steps after `warmup` on the case's own stream. loop: the
    drop-in time loop instead (lbf_run: residuals and the stop rule after
    every step, src/momentum.rs:273-292), run for the same number of steps."""
    from lbflow_amd.engine import Case, load_library
    lib = load_library()
    lib.lbfx_case_stream.restype = C.c_void_p
    lib.lbfx_case_stream.argtypes = [C.c_void_p]
    case = Case(spec)
    stream = C.c_void_p(lib.lbfx_case_stream(case.h))
    case.initialize_nodes()
    preroll(hip, stream)
    case.main_steps(warmup)
    case.synchronize()
    e0, e1 = hip.event(), hip.event()
    t0 = time.perf_counter()
    hip.record(e0, stream)
    if loop:
        # the loop stops after step index max_iterations + 1 (at most max + 2 steps)
        done = case.run(warmup + steps - 2)
        assert done == warmup + steps, done
    else:
        case.main_steps(steps)
    hip.record(e1, stream)
    case.synchronize()
    wall = time.perf_counter() - t0
    ms = hip.elapsed_ms(e0, e1)
    from lbflow_amd import abi
    nonfinite = case.reduce(abi.REDUCE_NONFINITE)
    mass = case.reduce(abi.REDUCE_TOTAL_MASS)
    case.close()
    return ms, wall, nonfinite, mass


def pin_threads():
    """One OpenMP thread per host core, bound (the reference's --pin,
    src/cli.rs:159-190); takes effect when the oracle library loads."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")


def other_cpu_baselines(hip, budget_s=8.0):
    """The oracle and the GPU on the other single-case workloads: the 128^2
    D2Q9 BGK lid-driven cavity (FP64, the reference's loop) and the 1024x512
    coupled discs (FP32 populations on the GPU as the workload states, FP64 in
    the oracle). CPU: steps grow until a run takes about budget_s."""
    from lbflow_amd import abi, configs
    from oracle import oracle
    lib = oracle.load()
    cores = lib.lbo_max_threads()
    out = {}
    for name, make, gpu_precs in (
            ("cavity_128", lambda p: configs.cavity(n=128, tau=0.6, lid=0.1, precision=p),
             (abi.F64,)),
            ("discs_1024x512", lambda p: configs.coupled_discs(precision=p), (abi.F32, abi.F64))):
        sp = make(abi.F64)
        o = oracle.OracleCase(sp)
        o.initialize_nodes()
        o.main_steps(1)
        steps, dt = 8, 0.0
        while dt < budget_s / 4:
            t0 = time.perf_counter()
            o.main_steps(steps)
            dt = time.perf_counter() - t0
            if dt < budget_s / 4:
                steps *= 4
        o.close()
        entry = {"cpu_mlups": round(sp.num_nodes * steps / dt / 1e6, 2), "cpu_cores": int(cores),
                 "cpu_sample": f"oracle, {steps} steps after 1 warm-up ({dt:.1f} s)"}
        for p in gpu_precs:
            ms, _, _, _ = time_case(make(p), 2000, 20, hip)
            entry[f"gpu_{abi.PRECISION_NAMES[p]}_mlups"] = round(sp.num_nodes * 2000 / ms * 1e3 / 1e6, 1)
        print(f"[bench] {name}: " + ", ".join(f"{k} {v}" for k, v in entry.items() if "mlups" in k),
              file=sys.stderr, flush=True)
        out[name] = entry
    return out


def cpu_baseline(n=128, steps=450, threads=None):
    """Times the oracle on a bounded sample (n^3 porous channel, same physics)."""
    from lbflow_amd import configs
    from oracle import oracle
    oracle.load()
    lib = oracle.load()
    if threads:
        lib.lbo_set_threads(int(threads))
    cores = lib.lbo_max_threads()
    sp = configs.porous_channel(n=n, precision=0)
    o = oracle.OracleCase(sp)
    o.initialize_nodes()
    o.main_steps(1)
    t0 = time.perf_counter()
    o.main_steps(steps)
    dt = time.perf_counter() - t0
    mlups = sp.num_nodes * steps / dt / 1e6
    o.close()
    return {"value": round(mlups, 3), "unit": "MLUPS", "cores": int(cores), "kind": "port",
            "sample": f"oracle (C f64 restatement of the reference step, OpenMP) on the same "
                      f"D3Q19 TRT porous channel at {n}^3, {steps} timed steps after 1 warm-up "
                      f"({dt:.1f} s)"}


def serving_cpu_baseline(params, spec_of, nodes_steps, budget_s=15.0, seed=7, case_s=1.5):
    """The oracle (C f64 restatement of the reference step, OpenMP on all host
    cores) on the serving mix: cases taken in a seeded random order until
    budget_s has passed. Each case is created, initialised, read back, and
    stepped for as many of its steps as fit about case_s (at least 4); its
    full run is estimated as set-up + steps x the measured time per step. The
    estimates are scaled to the whole mix by node updates."""
    import numpy as np
    from lbflow_amd import abi
    from oracle import oracle
    lib = oracle.load()
    cores = lib.lbo_max_threads()
    order = np.random.default_rng(seed).permutation(len(params))
    t_sum = est_sum = 0.0
    est_updates = ran_updates = 0.0
    ran = 0
    for i in order:
        p = params[i]
        sp = spec_of(p, abi.F64)
        t0 = time.perf_counter()
        o = oracle.OracleCase(sp)
        o.initialize_nodes()
        t1 = time.perf_counter()
        k = 4
        o.main_steps(k)
        t2 = time.perf_counter()
        per_step = (t2 - t1) / k
        more = int(min(p["steps"] - k, max(0.0, case_s - (t2 - t0)) / max(per_step, 1e-9)))
        if more > 0:
            o.main_steps(more)
            k += more
            per_step = (time.perf_counter() - t1) / k
        t3 = time.perf_counter()
        o.fields()
        o.close()
        setup = (t1 - t0) + (time.perf_counter() - t3)
        t_sum += time.perf_counter() - t0
        est_sum += setup + per_step * p["steps"]
        est_updates += sp.num_nodes * p["steps"]
        ran_updates += sp.num_nodes * k
        ran += 1
        if t_sum >= budget_s and ran >= 3:
            break
    total_s = est_sum * nodes_steps / est_updates
    return {"cases_per_s": round(len(params) / total_s, 4),
            "mlups": round(est_updates / est_sum / 1e6, 2), "cores": int(cores), "kind": "port",
            "sample": f"oracle (C f64 restatement of the reference step, OpenMP) on {ran} cases of "
                      f"the mix in seeded random order, each stepped for up to {case_s} s and "
                      f"extrapolated to its full step count ({t_sum:.1f} s measured, "
                      f"{ran_updates:.3g} node updates run), scaled to the whole mix by node "
                      f"updates"}


def serving_leg(sample=48, plain_sample=12, two_array_sample=6, seed=5, cpu=True):
    """The serving mix (configs.serving_mix: 2,000 cases, D2Q9 / D3Q19 50/50,
    sides log-uniform in [32, 256], tau ~ U[0.55, 1.2], steps ~ U{1000, 20000},
    porosity ~ U[0.6, 0.9], FP32 or FP16-shifted 50/50, periodic faces, BGK)
    as written, on a seeded uniform subsample of `sample` of its cases run to
    their full step counts (sample >= 2000: the whole mix). Cases/s for the
    whole mix are the sample's times scaled by node updates (sum of nodes x
    steps over the mix / over the sample).
      - served: the sample submitted at once, each result fetched as its case
        completes;
      - plain: `plain_sample` of those cases run one after another on the
        default FP64 path (in place, the reference's bits: create, steps,
        fields), and `two_array_sample` on the two-array FP64 kernel;
      - latency: the plain-sample cases arriving as a Poisson process at 1.2x
        the default FP64 path's capacity; served time to solution is measured
        live, the plain FIFO queue's from the measured service times
        (finish_i = max(arrival_i, finish_i-1) + service_i);
      - CPU: the oracle on a time-bounded sample of the mix (cpu=True).
    """
    import numpy as np
    from lbflow_amd import abi, configs
    from lbflow_amd.engine import Case
    from lbflow_amd.serve import Server

    count = 2000
    mix = configs.serving_mix(count, seed=seed)

    def updates(p):
        return float(p["side"] ** (2 if p["velocity_set"] == abi.D2Q9 else 3) * p["steps"])

    def spec_of(p, prec=None, streaming=abi.STREAM_AUTO):
        return configs.serving_case(p["velocity_set"], p["side"], p["tau"], p["porosity"],
                                    p["seed"], p["precision"] if prec is None else prec).replace(
                                        streaming=streaming)

    rng = np.random.default_rng(seed + 1)
    idx = np.arange(count) if sample >= count else np.sort(rng.choice(count, sample, replace=False))
    params = [mix[i] for i in idx]
    mix_updates = sum(updates(p) for p in mix)
    sample_updates = sum(updates(p) for p in params)
    specs = [(spec_of(p), p["steps"]) for p in params]

    def plain_service(ids, streaming):
        times = []
        for i in ids:
            sp = spec_of(params[i], abi.F64, streaming)
            t0 = time.perf_counter()
            c = Case(sp)
            c.initialize_nodes()
            c.main_steps(params[i]["steps"])
            c.fields()
            c.close()
            times.append(time.perf_counter() - t0)
        return np.array(times)

    # served throughput on the sample
    print(f"[serve] served: {len(params)} of the {count} cases at once "
          f"({sample_updates:.3g} of {mix_updates:.3g} node updates)", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    with Server() as srv:
        tickets = [srv.submit(sp, n) for sp, n in specs]
        # each result is fetched as its case completes (finished results would
        # otherwise wait in host memory), with a progress line every minute
        pending, last = set(tickets), t0
        while pending:
            for t in list(pending):
                if srv.info(t).state in (abi.REQ_DONE, abi.REQ_FAILED):
                    srv.result(t)
                    pending.discard(t)
            if pending:
                time.sleep(0.02)
            if time.perf_counter() - last > 60:
                last = time.perf_counter()
                print(f"[serve] {len(tickets) - len(pending)} of {len(tickets)} cases done after "
                      f"{last - t0:.0f} s", file=sys.stderr, flush=True)
            if time.perf_counter() - t0 > 3600:
                raise RuntimeError("serving sample did not finish within an hour")
        st = srv.stats()
    wall = time.perf_counter() - t0
    served_mix_s = wall * mix_updates / sample_updates

    # plain FP64 baselines on sub-samples of the sample, scaled by node updates
    plain_ids = rng.choice(len(params), size=min(plain_sample, len(params)), replace=False)
    print(f"[serve] plain FP64 (default path) on {len(plain_ids)} cases", file=sys.stderr, flush=True)
    svc = plain_service(plain_ids, abi.STREAM_AUTO)
    plain_mix_s = svc.sum() / sum(updates(params[i]) for i in plain_ids) * mix_updates
    two_ids = plain_ids[: min(two_array_sample, len(plain_ids))]
    print(f"[serve] plain FP64 two-array on {len(two_ids)} cases", file=sys.stderr, flush=True)
    svc2 = plain_service(two_ids, abi.STREAM_TWO_ARRAY)
    two_mix_s = svc2.sum() / sum(updates(params[i]) for i in two_ids) * mix_updates

    # latency: the plain-sample cases under Poisson arrivals at 1.2x plain capacity
    rate = 1.2 * len(plain_ids) / svc.sum()
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(plain_ids)))
    fin = np.zeros(len(plain_ids))
    prev = 0.0
    for k in range(len(plain_ids)):
        prev = max(arrivals[k], prev) + svc[k]
        fin[k] = prev
    plain_lat = fin - arrivals
    print(f"[serve] Poisson arrivals: {len(plain_ids)} cases at {rate:.2f}/s", file=sys.stderr,
          flush=True)
    with Server() as srv:
        start = time.perf_counter()
        tickets = []
        for k, i in enumerate(plain_ids):
            delay = arrivals[k] - (time.perf_counter() - start)
            if delay > 0:
                time.sleep(delay)
            tickets.append(srv.submit(specs[i][0], specs[i][1]))
        srv.drain(3600)
        lat = []
        for t in tickets:
            inf = srv.info(t)
            lat.append(inf.t_done - inf.t_submit)
            srv.result(t)
        st_lat = srv.stats()
    lat = np.array(lat)
    cpu_base = None
    if cpu:
        print("[serve] CPU baseline: the oracle on a time-bounded sample of the mix",
              file=sys.stderr, flush=True)
        cpu_base = serving_cpu_baseline(mix, spec_of, mix_updates)
    served_cps = count / served_mix_s
    plain_cps = count / plain_mix_s
    # algorithmic bytes of the served sample (2 Q s per node update) over the
    # time the server's graph replays took (HIP events)
    served_bytes = sum(updates(p) * 2 * (9 if p["velocity_set"] == abi.D2Q9 else 19) *
                       (4 if p["precision"] == abi.F32 else 2) for p in params)
    return {
        "metric": "served cases/s and time to solution, serving mix",
        "cases": count,
        "mix": "2,000 cases: D2Q9/D3Q19 50/50, sides log-uniform [32,256], tau U[0.55,1.2], "
               "steps U{1000,20000}, porosity U[0.6,0.9], FP32/FP16-shifted 50/50, periodic, "
               "BGK, body force 1e-5",
        "sample": (f"all {count} cases" if len(params) == count else
                   f"seeded uniform subsample of {len(params)} cases at their full step counts; "
                   f"whole-mix times = sample times x {mix_updates / sample_updates:.2f} "
                   f"(node updates of the mix / of the sample)"),
        "node_updates": mix_updates,
        "sample_node_updates": sample_updates,
        "served_cases_per_s": round(served_cps, 3),
        "served_mlups": round(sample_updates / wall / 1e6, 1),
        "served_sample_wall_s": round(wall, 2),
        "plain_f64_cases_per_s": round(plain_cps, 3),
        "plain_f64_sample": int(len(plain_ids)),
        "throughput_ratio": round(served_cps / plain_cps, 3),
        "plain_f64_two_array_cases_per_s": round(count / two_mix_s, 3),
        "plain_f64_two_array_sample": int(len(two_ids)),
        "throughput_ratio_vs_two_array": round(served_cps * two_mix_s / count, 3),
        "gpu_busy_frac": round(st["gpu_seconds"] / wall, 3),
        "served_kernel_gbs": round(served_bytes / st["gpu_seconds"] / 1e9, 1),
        "hbm_peak_bytes": int(st["hbm_peak"]),
        "tuned_grid": int(st["tuned_grid"]),
        "tuned_tile": int(st["tuned_tile"]),
        "poisson_rate_per_s": round(rate, 3),
        "latency_cases": int(len(plain_ids)),
        "served_p50_s": round(float(np.percentile(lat, 50)), 4),
        "served_p99_s": round(float(np.percentile(lat, 99)), 4),
        "plain_p50_s": round(float(np.percentile(plain_lat, 50)), 4),
        "plain_p99_s": round(float(np.percentile(plain_lat, 99)), 4),
        "cpu_baseline": cpu_base,
        "throughput_ratio_vs_cpu": (round(served_cps / cpu_base["cases_per_s"], 1)
                                    if cpu_base else None),
        "graphs_built_throughput": int(st["graphs_built"]),
        "graphs_built_latency": int(st_lat["graphs_built"]),
    }


def precision_leg(scale=1.0, only=None):
    """Precision tier per configuration: relative L2 velocity error and global
    mass drift of each storage against the FP64 parity path after the
    configuration's full step count (x scale), plus wall time.

      cavity   D2Q9 BGK lid-driven cavity 128^2, tau 0.6, lid 0.1, 20,000 steps
      porous   D3Q19 TRT porous channel 256^3, 5,000 steps
      sphere   D3Q27 MRT (this build's basis) flow past a sphere, 384x128x128,
               radius 16, 3-D Zou-He velocity inlet u = 0.05 and pressure
               outlet rho = 1, periodic y/z, 10,000 steps
      discs    D2Q9 BGK + D2Q9 scalar with adsorption, 1024x512, 200 discs,
               Zou-He inlet/outlet, 50,000 steps (FP32 two-array vs FP64)
    """
    import numpy as np
    from lbflow_amd import abi, configs
    from lbflow_amd.case import Collision
    from lbflow_amd.engine import Case

    def run(spec, steps, scalar=False):
        t0 = time.perf_counter()
        c = Case(spec)
        c.initialize_nodes()
        done = 0
        while done < steps:  # batches keep the host loop short
            k = min(1000, steps - done)
            c.main_steps(k)
            done += k
        c.synchronize()
        wall = time.perf_counter() - t0
        rho, u = c.fields()
        mass = c.reduce(abi.REDUCE_TOTAL_MASS)
        out = dict(u=u, rho=rho, mass=mass, wall=wall)
        if scalar:
            cv, qa = c.scalar(0)
            out.update(c=cv, ads=c.reduce(abi.REDUCE_ADSORBED),
                       ctot=c.reduce(abi.REDUCE_SCALAR_TOTAL))
        c.close()
        return out

    def l2rel(a, b, fluid):
        a, b = a[fluid], b[fluid]
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    tau = 0.6
    mrt27 = Collision.mrt([1.0 / tau] * 10 + [1.2] * 17)
    cases = [
        ("cavity", lambda p: configs.cavity(n=128, tau=0.6, lid=0.1, precision=p), 20000,
         [abi.F32, abi.F32_SHIFTED, abi.F16_SHIFTED, abi.BF16_SHIFTED]),
        ("porous", lambda p: configs.porous_channel(n=256, precision=p), 5000,
         [abi.F32, abi.F32_SHIFTED, abi.F16_SHIFTED, abi.BF16_SHIFTED]),
        ("sphere", lambda p: configs.sphere_channel(precision=p, op=mrt27), 10000,
         [abi.F32, abi.F32_SHIFTED, abi.F16_SHIFTED, abi.BF16_SHIFTED]),
        ("discs", lambda p: configs.coupled_discs(precision=p), 50000, [abi.F32]),
    ]
    report = {}
    for name, make, steps, precs in cases:
        if only and name not in only:
            continue
        steps = max(1, int(steps * scale))
        print(f"[precision] {name}: f64 reference, {steps} steps", file=sys.stderr, flush=True)
        ref_spec = make(abi.F64)
        fluid = ~ref_spec.node_type.astype(bool).reshape(-1) if ref_spec.node_type is not None \
            else np.ones(ref_spec.num_nodes, bool)
        scalar = bool(ref_spec.scalars)
        ref = run(ref_spec, steps, scalar)
        entry = {"steps": steps, "nodes": ref_spec.num_nodes, "f64_wall_s": round(ref["wall"], 3)}
        for p in precs:
            name_p = abi.PRECISION_NAMES[p]
            print(f"[precision] {name}: {name_p}", file=sys.stderr, flush=True)
            r = run(make(p), steps, scalar)
            e = {"l2_u_rel": l2rel(r["u"], ref["u"], fluid),
                 "mass_drift_rel": abs(r["mass"] - ref["mass"]) / abs(ref["mass"]),
                 "wall_s": round(r["wall"], 3),
                 "nonfinite": bool(not np.isfinite(r["u"][fluid]).all())}
            if scalar:
                e["l2_c_rel"] = l2rel(r["c"], ref["c"], fluid)
                e["adsorbed_rel"] = abs(r["ads"] - ref["ads"]) / max(abs(ref["ads"]), 1e-300)
            entry[name_p] = e
        report[name] = entry
    return report


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", default="", help="comma list of variants (f32,f64,f16s,...)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-mix leg")
    ap.add_argument("--no-all-fluid", action="store_true",
                    help="skip the all-fluid periodic box (the kernel without bounce-back)")
    ap.add_argument("--serve-sample", type=int, default=48,
                    help="cases of the 2,000-case serving mix run in full (>= 2000: all)")
    ap.add_argument("--serve-plain", type=int, default=12,
                    help="cases of that sample run on the plain FP64 path")
    ap.add_argument("--no-other", action="store_true",
                    help="skip the cavity and coupled-discs GPU / CPU rates")
    ap.add_argument("--precision", action="store_true",
                    help="only the precision-tier report (errors vs FP64 per configuration)")
    ap.add_argument("--precision-scale", type=float, default=1.0)
    ap.add_argument("--precision-cases", default="",
                    help="comma-separated subset of cavity,porous,sphere,discs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build "
              f"(every case fits one MI355X's HBM; multi-GPU sharding is not built)", file=sys.stderr)
        sys.exit(2)

    if args.precision:
        print(json.dumps({"precision": precision_leg(args.precision_scale,
                                                  [c for c in args.precision_cases.split(",") if c])}))
        return
    from lbflow_amd import abi, configs
    pin_threads()
    hip = Hip()
    n = args.n
    N = n ** 3
    variants = [("f32", abi.F32, abi.STREAM_TWO_STEP, 4), ("f32_aa", abi.F32, abi.STREAM_AA, 4),
                ("f64", abi.F64, abi.STREAM_TWO_ARRAY, 8),
                ("f16s", abi.F16_SHIFTED, abi.STREAM_AA, 2),
                ("bf16s", abi.BF16_SHIFTED, abi.STREAM_AA, 2),
                ("f32s", abi.F32_SHIFTED, abi.STREAM_AA, 4),
                ("f32_two_array", abi.F32, abi.STREAM_TWO_ARRAY, 4),
                ("f64_aa", abi.F64, abi.STREAM_AA, 8),
                ("f64_two_step", abi.F64, abi.STREAM_TWO_STEP, 8),
                ("f16s_two_step", abi.F16_SHIFTED, abi.STREAM_TWO_STEP, 2),
                ("bf16s_two_step", abi.BF16_SHIFTED, abi.STREAM_TWO_STEP, 2),
                ("f32s_two_step", abi.F32_SHIFTED, abi.STREAM_TWO_STEP, 4)]
    if args.only:
        by_name = {v[0]: v for v in variants}
        variants = [by_name[k] for k in args.only.split(",")]
    results = {}
    copy_gbs = copy_bandwidth(hip)
    print(f"[bench] device copy: {copy_gbs:.0f} GB/s", file=sys.stderr, flush=True)
    sweep_gbs = {es: sweep_bandwidth(hip, n ** 3, es) for es in (2, 4, 8)}
    print("[bench] in-place 19-plane sweep: " +
          ", ".join(f"{8 * es}-bit {g:.0f} GB/s" for es, g in sweep_gbs.items()),
          file=sys.stderr, flush=True)
    for name, prec, streaming, sbytes in variants:
        sp = configs.porous_channel(n=n, precision=prec, streaming=streaming)
        print(f"[bench] {name}: D3Q19 TRT porous {n}^3, warmup {args.warmup}, steps {args.steps}",
              file=sys.stderr, flush=True)
        ms, wall, nonfinite, mass = time_case(sp, args.steps, args.warmup, hip)
        mlups = N * args.steps / (ms / 1e3) / 1e6
        gbs = mlups * 1e6 * 2 * 19 * sbytes / 1e9
        results[name] = dict(mlups=round(mlups, 1), ms_per_step=round(ms / args.steps, 4),
                             achieved_gbs=round(gbs, 1), hbm_frac=round(gbs / HBM_PEAK_GBS, 4),
                             wall_s=round(wall, 3), nonfinite=nonfinite, mass=mass,
                             bytes_per_site=2 * 19 * sbytes, storage_bytes=sbytes,
                             frac_of_sweep=round(gbs / sweep_gbs[sbytes], 4),
                             streaming={abi.STREAM_AA: "aa", abi.STREAM_TWO_ARRAY: "two-array",
                                        abi.STREAM_TWO_STEP: "two-step"}[streaming])
        print(f"[bench] {name}: {mlups:.0f} MLUPS, {gbs:.0f} GB/s algorithmic "
              f"({100 * gbs / HBM_PEAK_GBS:.1f}% of HBM peak)", file=sys.stderr, flush=True)
    # the same kernel on an all-fluid, fully periodic box: no bounce-back links,
    # so this isolates the streaming pattern from the porous geometry
    all_fluid = {}
    if not args.no_all_fluid:
        for name, prec, sbytes, streaming in (
                ("f32", abi.F32, 4, abi.STREAM_TWO_STEP), ("f32_aa", abi.F32, 4, abi.STREAM_AA),
                ("f32s", abi.F32_SHIFTED, 4, abi.STREAM_AA),
                ("f16s", abi.F16_SHIFTED, 2, abi.STREAM_AA)):
            sp = configs.periodic_box(n=n, precision=prec, streaming=streaming).replace(
                initial_velocity=(0.05, 0.0, 0.0))
            ms, wall, nonfinite, mass = time_case(sp, args.steps, args.warmup, hip)
            mlups = N * args.steps / (ms / 1e3) / 1e6
            gbs = mlups * 1e6 * 2 * 19 * sbytes / 1e9
            all_fluid[name] = dict(mlups=round(mlups, 1), ms_per_step=round(ms / args.steps, 4),
                                   achieved_gbs=round(gbs, 1),
                                   hbm_frac=round(gbs / HBM_PEAK_GBS, 4),
                                   frac_of_sweep=round(gbs / sweep_gbs[sbytes], 4),
                                   nonfinite=nonfinite,
                                   streaming="two-step" if streaming == abi.STREAM_TWO_STEP
                                   else "aa")
            print(f"[bench] all-fluid periodic {n}^3 {name}: {mlups:.0f} MLUPS, {gbs:.0f} GB/s "
                  f"algorithmic ({100 * gbs / HBM_PEAK_GBS:.1f}% of HBM peak)", file=sys.stderr,
                  flush=True)
    head = results.get("f32") or next(iter(results.values()))
    head_name = "f32" if "f32" in results else next(iter(results))
    # the same workload through the drop-in time loop (lbf_run): residual
    # norms and the stop rule after every step, as the reference's loop does
    drop_in = None
    # the loop computes residuals every step, which the two-step passes leave
    # to single steps: it runs in place (the FP32 schemes give the same bits)
    loop_name = "f32_aa" if head_name == "f32" and "f32_aa" in results else head_name
    if loop_name in ("f32_aa", "f64_aa", "f16s", "f32s", "bf16s", "f64"):
        name, prec, streaming, sbytes = next(v for v in variants if v[0] == loop_name)
        sp = configs.porous_channel(n=n, precision=prec, streaming=streaming)
        ms, _, _, _ = time_case(sp, args.steps, args.warmup, hip, loop=True)
        mlups = N * args.steps / (ms / 1e3) / 1e6
        drop_in = dict(mlups=round(mlups, 1), ms_per_step=round(ms / args.steps, 4),
                       frac_of_value=round(mlups / head["mlups"], 4), variant=loop_name,
                       loop="lbf_run: residuals and stop rule every step (src/momentum.rs:273-292)")
        print(f"[bench] {loop_name} drop-in loop (lbf_run): {mlups:.0f} MLUPS", file=sys.stderr,
              flush=True)
    other = None if args.no_other or args.no_cpu else other_cpu_baselines(hip)
    serving = None if args.no_serve else serving_leg(sample=args.serve_sample,
                                                     plain_sample=args.serve_plain,
                                                     cpu=not args.no_cpu)
    cpu = None if args.no_cpu else cpu_baseline()
    # a two-step launch advances the lattice by two time steps
    steps_per_launch = 2 if head["streaming"] == "two-step" else 1
    bytes_per_launch = N * head["bytes_per_site"] * steps_per_launch
    from lbflow_amd.profiling import read_traffic
    traffic = read_traffic(TRAFFIC_FILE)
    line = {
        "metric": METRIC,
        "value": head["mlups"],
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": head_name,
        "data": "synthetic (seeded Bernoulli(0.3) solid mask, rest-state start)",
        "config": {"workload": f"D3Q19 TRT porous channel {n}^3, Bernoulli(0.3) solids seed 2, "
                               f"periodic x, no-slip y/z, force 1e-6 x, magic 3/16, tau 0.8, "
                               f"{head_name} populations, "
                               + ("two time steps per launch (two arrays, first step held in LDS)"
                                  if head["streaming"] == "two-step"
                                  else "in-place AA-pattern fused kernel"),
                   "lattice": "D3Q19", "nodes": N, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": head["achieved_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": head["hbm_frac"],
                     "traffic": traffic.get(head_name) if n == 256 else None,
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "steps_per_launch": steps_per_launch,
                     "attainable_copy_gbs": round(copy_gbs, 1),
                     "frac_of_copy": round(head["achieved_gbs"] / copy_gbs, 4),
                     "attainable_sweep_gbs": round(sweep_gbs[head["storage_bytes"]], 1),
                     "frac_of_sweep": round(head["achieved_gbs"] / sweep_gbs[head["storage_bytes"]], 4)},
        "cpu_baseline": cpu,
        "serving": serving,
        "drop_in_loop": drop_in,
        "extra": dict(results, copy_gbs=round(copy_gbs, 1),
                      other_workloads=other,
                      sweep_gbs={f"{8 * es}-bit": round(g, 1) for es, g in sweep_gbs.items()},
                      all_fluid_periodic=all_fluid or None,
                      timing={"device_preroll_ms": PREROLL_MS,
                              "preroll": "plain 256 MiB copies on the case's stream before "
                                         "the warm-up steps; timed region = exactly `steps` steps"}),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
