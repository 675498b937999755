#!/usr/bin/env python3
"""Benchmark: batched ML-KEM NTT + inverse NTT round trip on one MI355X.

A step is one forward NTT pass and one inverse NTT pass over the whole batch
(2^20 polys of the edge-value mix, seed 3: 70% uniform in [-3328, 3328], 10%
zero, 10% all 3328, 5% all -3328, 5% unit impulse), inputs resident in HBM:
forward out-of-place (in -> buf), inverse in place (buf).  The throughput is
npoly / (t_fwd + t_inv), the round-trip rate the reference reports
(ntt_bench.cu:480-488), for the INT8-MFMA method; the plain butterfly kernel
is timed on the same inputs in the same run, and the reference C code on the
host cores (``cpu_baseline``).

Prints one JSON line.  ``--gpus N`` with N > 1 exits: this build measures one
GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_POLY_PASS = 1024  # 512 B read + 512 B written per poly per transform


EXTRAS = ("c1", "c2", "basemul", "ragged", "matvec", "packed", "kpke", "sched", "host")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--npoly", type=int, default=1 << 20)
    ap.add_argument("--method", default="mfma16")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-check", action="store_true", help="skip the parity check")
    ap.add_argument("--only", default=None, help="time only this method (profiling runs)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the ragged-plan and fused-matvec sections")
    ap.add_argument("--extras", default=",".join(EXTRAS),
                    help="comma list of extra sections to run: " + ", ".join(EXTRAS))
    ap.add_argument("--extras-only", action="store_true",
                    help="profiling run: only the selected extra sections, printed as JSON")
    ap.add_argument("--profile-launches", action="store_true",
                    help="profiling run: --steps launches of each transform kernel and of a "
                         "copy of the batch, nothing timed")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="write profiles/traffic.json from two rocprofv3 --pmc runs of "
                         "--profile-launches (FETCH_SIZE, WRITE_SIZE); no GPU needed")
    ap.add_argument("--kernels-only", action="store_true",
                    help="profiling run: launch each kernel `steps` times plus a same-size "
                         "device copy, print per-kernel times, no CPU leg or parity check")
    return ap.parse_args()


def copy_bandwidth(torch, src, dst, steps):
    """A plain device copy of the batch: achievable-bandwidth and counter reference."""
    for _ in range(2):
        dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    return {"ms": ms, "GBs": 2 * src.numel() * 2 / (ms / 1e3) / 1e9}


def warm(torch, fn, min_runs=2, min_s=0.5):
    """Runs `fn` at least `min_runs` times and until `min_s` seconds of GPU
    work have passed, so the clocks have left their idle state."""
    t0, n = time.perf_counter(), 0
    while n < min_runs or time.perf_counter() - t0 < min_s:
        fn()
        n += 1
        if n % 8 == 0:
            torch.cuda.synchronize()
    torch.cuda.synchronize()


def time_method(tk, torch, method, src, buf, steps, warmup, npoly):
    """Times the round trip and each kernel on the current stream.

    * ``rt_ms``: `steps` forward+inverse pairs back to back between two events
      (the whole-job step time, launch gaps included);
    * ``fwd_ms`` / ``inv_ms``: `steps` launches of one kernel back to back
      between two events, divided by `steps` (the per-launch average that the
      rocprofv3 kernel-trace average is compared with);
    * ``rt_steps_ms``: per-step times with an event between steps, for spread.
    """
    stream = torch.cuda.current_stream()

    def ev():
        return torch.cuda.Event(enable_timing=True)

    # Warm-up: `warmup` steps, then more until 0.5 s of work has run, so the
    # GPU clocks have left their idle state before anything is timed.
    def step():
        tk.ntt(src, out=buf, method=method)
        tk.invntt(buf, method=method)
    warm(torch, step, warmup)
    e0, e1 = ev(), ev()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        tk.ntt(src, out=buf, method=method)
        tk.invntt(buf, method=method)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    rt_ms = e0.elapsed_time(e1) / steps

    def per_kernel(fn):
        a, b = ev(), ev()
        a.record(stream)
        for _ in range(steps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / steps

    fwd_ms = per_kernel(lambda: tk.ntt(src, out=buf, method=method))
    inv_ms = per_kernel(lambda: tk.invntt(buf, method=method))

    evs = [ev() for _ in range(steps + 1)]
    evs[0].record(stream)
    for s in range(steps):
        tk.ntt(src, out=buf, method=method)
        tk.invntt(buf, method=method)
        evs[s + 1].record(stream)
    torch.cuda.synchronize()
    per = np.array([evs[s].elapsed_time(evs[s + 1]) for s in range(steps)])
    rt = npoly / (per / 1e3)
    return {
        "rt_ms": rt_ms, "fwd_ms": fwd_ms, "inv_ms": inv_ms,
        "rt_poly_s": npoly / (rt_ms / 1e3),
        "rt_poly_s_median": float(np.median(rt)),
        "rt_poly_s_sd": float(rt.std(ddof=1)) if steps > 1 else 0.0,
        "wall_ms_per_step": wall,
    }


def cpu_baseline(x):
    """Reference C ntt()+invntt() round trip on the host cores (bounded sample):
    all host threads, and one thread on a smaller sample."""
    from oracle.oracle import Oracle, Reference, reference_available
    kind = "reference" if reference_available() else "port"
    impl = Reference() if kind == "reference" else Oracle()
    cores = impl.max_threads()

    def rate(npoly, nthreads, budget_s):
        sample = np.ascontiguousarray(x[:npoly])
        reps, t = 0, 0.0
        while t < budget_s and reps < 20:
            a = sample.copy()
            t0 = time.perf_counter()
            impl.roundtrip_inplace(a, nthreads)
            t += time.perf_counter() - t0
            reps += 1
        return reps * npoly / t, reps, npoly, t

    v, reps, n, t = rate(min(x.shape[0], 1 << 18), cores, 3.0)
    v1, reps1, n1, t1 = rate(min(x.shape[0], 1 << 14), 1, 2.0)
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            model = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    return {"value": v, "unit": "poly/s", "cores": cores, "kind": kind,
            "one_core_poly_s": v1, "cpu_model": model, "nproc": os.cpu_count(),
            "sample": f"{reps} x {n} polys of the same workload, ntt() then invntt() per "
                      f"poly, OpenMP over polys on {cores} host threads, {t:.2f} s wall "
                      f"({t * cores:.1f} core-s); one thread: {reps1} x {n1} polys, "
                      f"{t1:.2f} s"}


def hip_graph_runtime():
    """hipGraphInstantiate / Launch / AddChildGraphNode of the HIP runtime this
    process already loaded (torch's, shared with libtkyber), via ctypes."""
    import ctypes
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
    lib = ctypes.CDLL(path)
    vp, sz, pvp = ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)
    for name, args in (("hipGraphCreate", [pvp, ctypes.c_uint]), ("hipGraphDestroy", [vp]),
                       ("hipGraphAddChildGraphNode", [pvp, vp, vp, sz, vp]),
                       ("hipGraphInstantiate", [pvp, vp, vp, vp, sz]),
                       ("hipGraphLaunch", [vp, vp]), ("hipGraphExecDestroy", [vp])):
        getattr(lib, name).argtypes, getattr(lib, name).restype = args, ctypes.c_int
    return lib


def run_stats(ms, npoly):
    """The reference's statistics of timed runs (ntt_bench.cu:428-488): mean,
    median and sample standard deviation of the per-run throughput, and the
    median latency."""
    ms = np.asarray(ms, dtype=np.float64)
    tp = npoly / (ms / 1e3)
    return {"mean_poly_s": float(tp.mean()), "median_poly_s": float(np.median(tp)),
            "sd_poly_s": float(tp.std(ddof=1)) if tp.size > 1 else 0.0,
            "median_latency_ms": float(np.median(ms))}


def bench_config_rows(tk, torch, config, runs, methods=("mfma16", "mfma32", "butterfly")):
    """Configurations 1 and 2 as the reference runs them (ntt_bench.cu:632-663,
    548-587): per run, the batch is re-staged from a pristine device copy, then
    the forward NTT (in place) and the inverse NTT (in place) are each timed
    alone between two events with a synchronisation after each; `runs` timed
    runs after 2 warm-ups.  Per direction: mean / median / sd poly/s and median
    latency; round trip npoly / (t_fwd + t_inv) per run.  Bit-exact: the last
    run's forward and inverse outputs against the oracle on every poly.  The
    reference C code (ntt() then invntt() per poly) is timed on the host on all
    threads and on one.

    The passes are called through the C ABI (ntt_batch_ex / invntt_batch_ex),
    as the reference's benchmark launches its kernels from C++: after a
    synchronisation the host's time to reach the launch falls inside the
    timed region, and the Python wrapper adds about 2 us of it."""
    from oracle.oracle import Oracle, Reference, reference_available
    x = tk.workload(config)
    npoly = x.shape[0]
    src = torch.from_numpy(x).to("cuda:0")
    buf = torch.empty_like(src)
    stream = torch.cuda.current_stream()
    o = Oracle()
    want_f = np.mod(o.ntt(x).astype(np.int32), 3329).astype(np.int16)
    inv = np.mod(o.invntt(want_f).astype(np.int32), 3329)
    want_i = np.where(inv > 1664, inv - 3329, inv).astype(np.int16)
    out = {"workload": f"configuration {config}: {npoly} polys uniform [0,q), seed {config}",
           "npoly": npoly, "runs": runs}
    L, ptr, sp = tk.lib(), buf.data_ptr(), stream.cuda_stream
    for m in methods:
        mc = tk.METHODS[m]

        def fwd_call():
            tk._check(L.ntt_batch_ex(ptr, ptr, npoly, mc, tk.FORM_DEFAULT, sp), "ntt_batch_ex")

        def inv_call():
            tk._check(L.invntt_batch_ex(ptr, ptr, npoly, mc, tk.FORM_DEFAULT, sp),
                      "invntt_batch_ex")

        def one():
            buf.copy_(src)
            fwd_call()
            inv_call()
        warm(torch, one, 2)
        ms_f, ms_i = [], []
        e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        for _ in range(runs):
            buf.copy_(src)
            e[0].record(stream)
            fwd_call()
            e[1].record(stream)
            torch.cuda.synchronize()
            ms_f.append(e[0].elapsed_time(e[1]))
            fwd = buf.cpu().numpy() if _ == runs - 1 else None
            e[2].record(stream)
            inv_call()
            e[3].record(stream)
            torch.cuda.synchronize()
            ms_i.append(e[2].elapsed_time(e[3]))
        rt = npoly / ((np.asarray(ms_f) + np.asarray(ms_i)) / 1e3)
        out[m] = {"ntt": run_stats(ms_f, npoly), "invntt": run_stats(ms_i, npoly),
                  "roundtrip": {"mean_poly_s": float(rt.mean()),
                                "median_poly_s": float(np.median(rt))},
                  "roofline_frac_rt_median": float(np.median(rt)) * 2 * BYTES_PER_POLY_PASS /
                  1e9 / HBM_PEAK_GBS,
                  "bit_exact": bool(np.array_equal(fwd, want_f) and
                                    np.array_equal(buf.cpu().numpy(), want_i))}
    impl = Reference() if reference_available() else Oracle()
    cpu = {"kind": "reference" if reference_available() else "port"}
    for label, nthreads in (("all_threads", impl.max_threads()), ("one_thread", 1)):
        a = x.copy()
        reps = 3 if nthreads > 1 else 1
        ms_f, ms_i = [], []
        for _ in range(reps):
            a[:] = x
            t0 = time.perf_counter()
            impl.ntt_inplace(a, nthreads)
            t1 = time.perf_counter()
            impl.invntt_inplace(a, nthreads)
            t2 = time.perf_counter()
            ms_f.append((t1 - t0) * 1e3)
            ms_i.append((t2 - t1) * 1e3)
        rt = npoly / ((np.asarray(ms_f) + np.asarray(ms_i)) / 1e3)
        cpu[label] = {"threads": nthreads, "runs": reps, "ntt": run_stats(ms_f, npoly),
                      "invntt": run_stats(ms_i, npoly),
                      "roundtrip_median_poly_s": float(np.median(rt))}
    out["cpu_reference"] = cpu
    return out


def bench_ragged(tk, torch, steps):
    """Configuration 4: 2^20 polys in ragged ML-KEM requests (seed 4), forward
    and inverse plans (autotuned persistent kernel over a per-tile table).

    * ``rt_ms``: tk_ragged_plan_execute of the forward then the inverse plan
      (one dispatch each), `steps` round trips between two events;
    * ``fwd_ms`` / ``inv_ms``: each plan alone, per launch;
    * ``rt_ms_graph_replay``: each plan's graph (tk_ragged_plan_graph)
      instantiated and launched, two graph launches per round trip;
    * ``rt_ms_one_graph``: both plans as child nodes of one graph, one launch
      per round trip (the cost of the graph launch paid once per batch)."""
    import ctypes
    N = 256
    reqs, pool = tk.wl_ragged_layout(1 << 20, 4)
    x = tk.wl_uniform(pool * N, 4).reshape(pool, N)
    d = torch.from_numpy(x).to("cuda:0")
    L = tk.lib()
    plans, create_s = [], []
    for direction in (0, 1):
        table = (tk.TkRequest * len(reqs))()
        for r, (off, n, k, kind) in enumerate(reqs):
            table[r].base = d.data_ptr() + int(off) * N * 2
            table[r].npoly, table[r].k, table[r].kind = int(n), int(k), int(kind)
        p = ctypes.c_void_p()
        t0 = time.perf_counter()
        rc = L.tk_ragged_plan_create(ctypes.byref(p), ctypes.addressof(table), len(reqs),
                                     direction, tk.METHOD_AUTO, None)
        create_s.append(time.perf_counter() - t0)
        if rc != 0:
            raise SystemExit(f"tk_ragged_plan_create failed: {rc}")
        plans.append(p)
    shapes = []
    for p in plans:
        info = [ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()]
        L.tk_ragged_plan_info(p, *[ctypes.byref(v) for v in info])
        shapes.append({3: "mfma16", 4: "mfma32"}.get(info[0].value, info[0].value) +
                      f" at {info[2].value}x")
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)

    def timed(fn):
        warm(torch, fn)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    hip = hip_graph_runtime()
    graphs = []
    for p in plans:
        g = ctypes.c_void_p()
        L.tk_ragged_plan_graph(p, ctypes.byref(g))
        graphs.append(g.value)

    def instantiate(g):
        ex = ctypes.c_void_p()
        assert hip.hipGraphInstantiate(ctypes.byref(ex), g, None, None, 0) == 0
        return ex.value
    execs = [instantiate(g) for g in graphs]
    parent = ctypes.c_void_p()
    hip.hipGraphCreate(ctypes.byref(parent), 0)
    n0, n1 = ctypes.c_void_p(), ctypes.c_void_p()
    hip.hipGraphAddChildGraphNode(ctypes.byref(n0), parent, None, 0, graphs[0])
    dep = (ctypes.c_void_p * 1)(n0.value)
    hip.hipGraphAddChildGraphNode(ctypes.byref(n1), parent, dep, 1, graphs[1])
    both = instantiate(parent.value)

    ms = timed(lambda: (L.tk_ragged_plan_execute(plans[0], sp),
                        L.tk_ragged_plan_execute(plans[1], sp)))
    fwd = timed(lambda: L.tk_ragged_plan_execute(plans[0], sp))
    inv = timed(lambda: L.tk_ragged_plan_execute(plans[1], sp))
    ms_graphs = timed(lambda: (hip.hipGraphLaunch(execs[0], sp), hip.hipGraphLaunch(execs[1], sp)))
    ms_one = timed(lambda: hip.hipGraphLaunch(both, sp))
    for ex in execs + [both]:
        hip.hipGraphExecDestroy(ex)
    hip.hipGraphDestroy(parent.value)
    for p in plans:
        L.tk_ragged_plan_destroy(p)
    npoly = int(reqs[:, 1].sum())
    return {"workload": f"{len(reqs)} requests of 1..4096 polys (rank-k vectors, 25% k x k "
                        f"matrices), {npoly} polys in a {pool}-poly pool with gaps, seed 4",
            "rt_poly_s": npoly / (ms / 1e3), "rt_ms": ms, "fwd_ms": fwd, "inv_ms": inv,
            "rt_fraction_of_roofline": npoly * 2 * BYTES_PER_POLY_PASS / (ms / 1e3) / 1e9 /
            HBM_PEAK_GBS,
            "rt_ms_graph_replay": ms_graphs, "rt_ms_one_graph": ms_one,
            "plan_create_s": create_s, "shape_fwd": shapes[0], "shape_inv": shapes[1],
            "launch": "tk_ragged_plan_execute: one dispatch of the persistent kernel per plan"}


def bench_basemul(tk, torch, steps, npoly=1 << 20):
    """basemul_batch on 2^20 NTT-domain poly pairs (uniform [0, q), seeds 6
    and 7): the reference's basemul on all 128 pairs of each poly, raw ref
    values out.  HBM-bound: 2 x 512 B read + 512 B written per poly."""
    N = 256
    a = torch.from_numpy(tk.wl_uniform(npoly * N, 6)).to("cuda:0")
    b = torch.from_numpy(tk.wl_uniform(npoly * N, 7)).to("cuda:0")
    r = torch.empty_like(a)
    stream = torch.cuda.current_stream()
    fn = lambda: tk.basemul(a, b, out=r)  # noqa: E731
    warm(torch, fn)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    nbytes = npoly * 3 * 512
    return {"workload": f"{npoly} poly pairs, uniform [0,q), seeds 6 and 7", "us": ms * 1e3,
            "poly_s": npoly / (ms / 1e3), "bytes_per_poly": 3 * 512,
            "GBs": nbytes / (ms / 1e3) / 1e9,
            "roofline_frac": nbytes / (ms / 1e3) / 1e9 / HBM_PEAK_GBS}


def bench_matvec(tk, torch, steps, lat_launches=1000):
    """Configuration 5: fused NTT -> basemul -> iNTT, 65,536 instances of rank
    3 (seed 5), plus the launch latency at batch 1, 16, 256, 4096."""
    N, k, ninst = 256, 3, 65536
    A, s = tk.matvec_workload(ninst, k)
    dA = torch.from_numpy(A.reshape(-1)).to("cuda:0")
    ds = torch.from_numpy(s.reshape(-1)).to("cuda:0")
    dt = torch.empty_like(ds)
    stream = torch.cuda.current_stream()
    warm(torch, lambda: tk.matvec(dA, ds, k, out=dt))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        tk.matvec(dA, ds, k, out=dt)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    bytes_per = (k * k + 2 * k) * N * 2
    lat = {}
    for b in (1, 16, 256, 4096):
        a_b, s_b, t_b = dA[: b * k * k * N], ds[: b * k * N], dt[: b * k * N]
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(lat_launches)]
        for x0, x1 in evs:
            x0.record(stream)
            tk.matvec(a_b, s_b, k, out=t_b)
            x1.record(stream)
        torch.cuda.synchronize()
        t = np.array([x0.elapsed_time(x1) * 1e3 for x0, x1 in evs])
        lat[str(b)] = {"p50_us": float(np.percentile(t, 50)), "p99_us": float(np.percentile(t, 99))}
    inst_s = ninst / (ms / 1e3)
    return {"workload": "65,536 instances, k = 3, A_hat uniform [0,q), s CBD eta=2, seed 5",
            "inst_s": inst_s, "ms": ms, "hbm_bytes_per_inst": bytes_per,
            "roofline_frac": inst_s * bytes_per / 1e9 / HBM_PEAK_GBS,
            "latency_us": lat, "latency_launches": lat_launches}


def bench_packed(tk, torch, steps, npoly=1 << 20):
    """Packed small-coefficient input (the TC-PACK analogue): forward NTT of
    2^20 packed polys at 2 bits (ternary) and 3 bits (CBD eta = 2), against
    the int16 forward on the same values; and configuration 5 with its CBD
    secrets packed in 3 bits."""
    stream = torch.cuda.current_stream()
    out = {}

    def per_launch(fn):
        warm(torch, fn)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    for bits in (2, 3):
        x = torch.from_numpy(tk.packed_workload(bits, npoly)).to("cuda:0")
        packed = tk.pack_small(x, bits)
        dst = torch.empty_like(x)
        ms = per_launch(lambda: tk.ntt_packed(packed, bits, out=dst))
        ms16 = per_launch(lambda: tk.ntt(x, out=dst))
        ms_pack = per_launch(lambda: tk.pack_small(x, bits, out=packed))
        nbytes = tk.packed_poly_bytes(bits) + 512
        out[f"bits{bits}"] = {
            "npoly": npoly, "fwd_us": ms * 1e3, "poly_s": npoly / (ms / 1e3),
            "bytes_per_poly": nbytes,
            "roofline_frac": npoly * nbytes / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
            "int16_fwd_us_same_values": ms16 * 1e3,
            "speedup_vs_int16_fwd": ms16 / ms,
            "pack_kernel_us": ms_pack * 1e3,
        }
    A, s = tk.matvec_workload()
    k, ninst = 3, s.shape[0]
    dA = torch.from_numpy(A).to("cuda:0")
    sp = tk.pack_small(torch.from_numpy(s.reshape(-1, 256)).to("cuda:0"), 3)
    t = torch.empty((ninst * k, 256), dtype=torch.int16, device="cuda:0")
    ms = per_launch(lambda: tk.matvec_packed(dA, sp, k, 3, out=t))
    nbytes = (k * k + k) * 512 + k * tk.packed_poly_bytes(3)
    out["config5_s_packed3"] = {
        "ninst": ninst, "ms": ms, "inst_s": ninst / (ms / 1e3), "bytes_per_inst": nbytes,
        "roofline_frac": ninst * nbytes / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
    }
    return out


def bench_kpke(tk, torch, steps, ninst=65536, k=3):
    """K-PKE arithmetic cores (keygen, encryption, decryption), one fused
    kernel each, on 65,536 rank-3 instances (kpke_workload, seed 9)."""
    w = tk.kpke_workload(ninst, k)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")  # noqa: E731
    A = dev(w["A_hat"])
    At = dev(w["A_hat"].transpose(0, 2, 1, 3))
    s, e, r, e1, e2, m = (dev(w[n]) for n in ("s", "e", "r", "e1", "e2", "m"))
    t_hat, s_hat = torch.empty_like(s), torch.empty_like(s)
    u, v = torch.empty_like(r), torch.empty_like(e2)
    wout = torch.empty_like(e2)
    stream = torch.cuda.current_stream()

    def per_launch(fn):
        warm(torch, fn)
        e0, e1_ = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1_.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1_) / steps

    polys = {"keygen": k * k + 4 * k, "encrypt": k * k + 4 * k + 3, "decrypt": 2 * k + 2}
    fns = {
        "keygen": lambda: tk.kpke_keygen(A, s, e, k, t_hat=t_hat, s_hat=s_hat),
        "encrypt": lambda: tk.kpke_encrypt(At, t_hat, r, e1, e2, m, k, u=u, v=v),
        "decrypt": lambda: tk.kpke_decrypt(s_hat, u, v, k, w=wout),
    }
    out = {"workload": f"{ninst} instances, k = {k}, A_hat uniform [0,q), CBD eta=2 noise, "
                       "random message bits, seed 9"}
    for name, fn in fns.items():
        ms = per_launch(fn)
        nbytes = polys[name] * 512
        out[name] = {"us": ms * 1e3, "inst_s": ninst / (ms / 1e3), "bytes_per_inst": nbytes,
                     "roofline_frac": ninst * nbytes / (ms / 1e3) / 1e9 / HBM_PEAK_GBS}
    return out


def bench_host(tk, torch, npoly=1 << 20, reps=3):
    """Host-resident batches (PCIe-inclusive): 2^20 polys of the edge-value
    mix through tk_host_pipeline (forward NTT, out of place): pinned host
    buffers (transformed over PCIe by the kernel) and pageable ones over a
    few chunk/depth shapes.  Wall-clock per run, best of `reps`."""
    x = tk.workload(3, npoly)
    hin = torch.from_numpy(x).pin_memory()
    hout = torch.empty_like(hin).pin_memory()
    res = {"npoly": npoly, "bytes_each_way": npoly * 512}

    def run(p, a, b):
        best = 1e30
        p.run(a, b, direction=0)  # warm
        for _ in range(reps):
            t0 = time.perf_counter()
            p.run(a, b, direction=0)
            best = min(best, time.perf_counter() - t0)
        return best

    # pinned buffers: the kernel reads and writes them over PCIe (no chunks)
    p = tk.HostPipeline()
    t = run(p, hin, hout)
    p.close()
    res["pinned"] = {"ms": t * 1e3, "poly_s": npoly / t, "GBs_each_way": npoly * 512 / t / 1e9}
    # pageable buffers: staged by the pipeline through pinned chunks
    xo = np.empty_like(x)
    shapes = {}
    for chunk, depth in ((16384, 3), (65536, 2), (131072, 2)):
        p = tk.HostPipeline(chunk, depth)
        t = run(p, x, xo)
        p.close()
        shapes[f"chunk{chunk}_depth{depth}"] = {"ms": t * 1e3, "poly_s": npoly / t,
                                                "GBs_each_way": npoly * 512 / t / 1e9}
    res["pageable"] = shapes

    # PCIe reference: the same bytes between the same pinned buffers and
    # device memory with no kernel, one way at a time and both ways at once
    # (two streams): the bound the pipeline's each-way rate is read against.
    dev_a = torch.empty(hin.shape, dtype=torch.int16, device="cuda:0")
    dev_b = torch.empty(hin.shape, dtype=torch.int16, device="cuda:0")
    s_in, s_out = torch.cuda.Stream(), torch.cuda.Stream()

    def copies(h2d, d2h):
        best = 1e30
        for _ in range(reps + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if h2d:
                with torch.cuda.stream(s_in):
                    dev_a.copy_(hin, non_blocking=True)
            if d2h:
                with torch.cuda.stream(s_out):
                    hout.copy_(dev_b, non_blocking=True)
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return npoly * 512 / best / 1e9

    res["pcie_reference_GBs"] = {"h2d_only": copies(True, False), "d2h_only": copies(False, True),
                                 "both_each_way": copies(True, True)}
    del dev_a, dev_b
    return res


def bench_sched(tk, torch, steps, nreq=3000, seed=11):
    """Serving mix for the rank-aware fused plan: nreq encryption requests of
    1..8 instances at ranks 2, 3, 4 (ML-KEM-512/768/1024) drawn with seed 11,
    each with its own buffers (slices of per-rank pools).  One plan (graph
    replay) against one kpke_encrypt_core call per request."""
    import ctypes
    ranks = tk.wl_uniform(nreq, seed, 2, 5)
    sizes = tk.wl_uniform(nreq, seed + 1, 1, 9)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")  # noqa: E731
    pools = {}
    for k in (2, 3, 4):
        n = int(sizes[ranks == k].sum())
        w = tk.kpke_workload(n, k, seed=seed + k)
        t_hat, _ = tk.kpke_keygen(dev(w["A_hat"]), dev(w["s"]), dev(w["e"]), k)
        pools[k] = {"mat": dev(w["A_hat"].transpose(0, 2, 1, 3)), "mat2": t_hat,
                    "x": dev(w["r"]), "add0": dev(w["e1"]), "add1": dev(w["e2"]),
                    "add2": dev(w["m"]), "out0": torch.empty_like(dev(w["r"])),
                    "out1": torch.empty_like(dev(w["e2"])), "next": 0}
    polys = {"mat": lambda k: k * k, "mat2": lambda k: k, "x": lambda k: k, "add0": lambda k: k,
             "add1": lambda k: 1, "add2": lambda k: 1, "out0": lambda k: k, "out1": lambda k: 1}
    table = (tk.TkFusedRequest * nreq)()
    calls = []
    for r in range(nreq):
        k, n = int(ranks[r]), int(sizes[r])
        pool = pools[k]
        off = pool["next"]
        pool["next"] += n
        t = table[r]
        t.mode, t.k, t.ninst = tk.FUSED_ENCRYPT, k, n
        ptrs = {}
        for name, per in polys.items():
            ptrs[name] = pool[name].data_ptr() + off * per(k) * 512
            setattr(t, name, ptrs[name])
        calls.append((k, n, ptrs))
    plan = ctypes.c_void_p()
    rc = tk.lib().tk_fused_plan_create(ctypes.byref(plan), ctypes.addressof(table), nreq)
    if rc != 0:
        raise SystemExit(f"tk_fused_plan_create failed: {rc}")
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    L = tk.lib()

    def singles():
        for k, n, p in calls:
            L.kpke_encrypt_core(p["out0"], p["out1"], p["mat"], p["mat2"], p["x"], p["add0"],
                                p["add1"], p["add2"], k, n, sp)

    def timed(fn, reps):
        warm(torch, fn)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps, (time.perf_counter() - t0) * 1e3 / reps

    ninst = int(sizes.sum())
    g_ms, g_wall = timed(lambda: L.tk_fused_plan_execute(plan, sp), steps)
    d_ms, d_wall = timed(lambda: L.tk_fused_plan_launch(plan, sp), steps)
    s_ms, s_wall = timed(singles, max(steps // 10, 2))
    L.tk_fused_plan_destroy(plan)
    out = {"workload": f"{nreq} encryption requests of 1..8 instances, ranks 2/3/4, seed {seed}",
           "requests": nreq, "instances": ninst,
           "plan_graph": {"ms": g_ms, "wall_ms": g_wall, "inst_s": ninst / (g_ms / 1e3)},
           "plan_direct": {"ms": d_ms, "wall_ms": d_wall},
           "one_call_per_request": {"ms": s_ms, "wall_ms": s_wall,
                                    "inst_s": ninst / (max(s_ms, s_wall) / 1e3)}}
    out["speedup_vs_one_call_per_request"] = max(s_ms, s_wall) / max(g_ms, g_wall)
    return out


def check_parity(tk, torch, method, x, buf_fwd, buf_rt):
    """Canonical equality of the GPU forward and inverse against the oracle."""
    from oracle.oracle import Oracle
    o = Oracle()
    fwd = buf_fwd.cpu().numpy()
    ok_f = np.array_equal(fwd, np.mod(o.ntt(x).astype(np.int32), 3329).astype(np.int16))
    inv = buf_rt.cpu().numpy()
    want = np.mod(o.invntt(fwd).astype(np.int32), 3329)
    want = np.where(want > 1664, want - 3329, want).astype(np.int16)
    return bool(ok_f and np.array_equal(inv, want))


TRAFFIC_JSON = os.path.join(ROOT, "profiles", "traffic.json")
CSRC = os.path.join(ROOT, "tensorkyber_amd", "csrc")
# A plain device copy of the batch (torch's copy_ of 512 MiB) calibrates the
# counters: its bytes read and written are known exactly.
COPY_TAG = "copy"


# The sources the traffic-measured kernels (the linear MFMA16 / MFMA32 and
# butterfly transforms, their launch shapes and tables) are built from.
TRAFFIC_SOURCES = ("tk_api.hip", "tk_butterfly.hip", "tk_device.h", "tk_internal.h",
                   "tk_launch.h", "tk_mfma.hip", "tk_mfma_core.h", "tk_tables.cpp")


def csrc_digest():
    """SHA-256 (first 16 hex digits) of the traffic-measured kernels' sources,
    file name and contents in name order: ties a traffic measurement to the
    code measured."""
    import hashlib
    h = hashlib.sha256()
    for name in TRAFFIC_SOURCES:
        h.update(name.encode() + b"\0")
        with open(os.path.join(CSRC, name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def load_traffic(kernel, npoly):
    """HBM bytes per launch of `kernel` at 2^20 polys from profiles/traffic.json
    (two rocprofv3 --pmc runs, FETCH_SIZE and WRITE_SIZE, calibrated by the
    copy), and a note; None when the file was measured on other sources."""
    if npoly != 1 << 20:
        return None, "traffic is measured at 2^20 polys only"
    try:
        with open(TRAFFIC_JSON) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None, "profiles/traffic.json missing"
    if t.get("csrc_digest") != csrc_digest():
        return None, (f"profiles/traffic.json was measured on kernel sources "
                      f"{t.get('csrc_digest')}, these are {csrc_digest()}")
    v = t.get("kernels", {}).get(kernel)
    return (v["bytes_per_launch"] if v else None), f"rocprofv3 FETCH_SIZE/WRITE_SIZE, sources " \
                                                    f"{t['csrc_digest']}"


# Kernel name -> bench kernel label (forward canonical, inverse centred, as
# ntt_batch / invntt_batch launch them).
# The 2^20-poly batch gives each MFMA16 wave two tiles: the tile-swap
# instance (fifth template argument true) runs, with nontemporal stores (the
# last argument false: the batch is larger than the Infinity Cache).
KERNEL_LABELS = (("k_mfma16<1, false, 4, tk::LinearTiles, true, false, false>", "mfma16_fwd"),
                 ("k_mfma16<2, true, 4, tk::LinearTiles, true, false, false>", "mfma16_inv"),
                 ("k_mfma32<1, false, 4, tk::LinearTiles, false>", "mfma32_fwd"),
                 ("k_mfma32<2, true, 4, tk::LinearTiles, false>", "mfma32_inv"),
                 ("k_ntt_butterfly", "butterfly_fwd"), ("k_invntt_butterfly", "butterfly_inv"))


def kernel_label(name):
    for tag, label in KERNEL_LABELS:
        if tag in name:
            return label
    return COPY_TAG if COPY_TAG in name.lower() else None


def profile_launches(tk, torch, npoly, reps):
    """A short program for rocprofv3: `reps` launches of each transform kernel
    and of a plain copy of the batch, nothing else timed or checked."""
    x = tk.workload(3, npoly)
    src = torch.from_numpy(x).to("cuda:0")
    buf = torch.empty_like(src)
    for m in ("mfma16", "mfma32", "butterfly"):
        for _ in range(reps):
            tk.ntt(src, out=buf, method=m)
        for _ in range(reps):
            tk.invntt(buf, method=m)
        torch.cuda.synchronize()
        print(f"[bench] profiled {m} x {reps}", file=sys.stderr, flush=True)
    for _ in range(reps):
        buf.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"profiled": npoly, "reps": reps, "csrc_digest": csrc_digest()}))


def counter_rows(d, counter):
    """(kernel name, value) of every dispatch's `counter` in a rocprofv3
    output directory: its SQLite databases (counters_collection view) or,
    with --output-format csv, its counter_collection CSVs."""
    import csv
    import glob
    import sqlite3
    rows = []
    for path in glob.glob(os.path.join(d, "**", "*.db"), recursive=True):
        con = sqlite3.connect(path)
        try:
            rows += con.execute("select kernel_name, value from counters_collection "
                                "where counter_name = ?", (counter,)).fetchall()
        finally:
            con.close()
    for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as f:
            rows += [(r["Kernel_Name"], r["Counter_Value"]) for r in csv.DictReader(f)
                     if r.get("Counter_Name") == counter]
    return [(name, float(v)) for name, v in rows]


def traffic_from(fetch_dir, write_dir, npoly):
    """profiles/traffic.json from two rocprofv3 --pmc runs of profile_launches
    (FETCH_SIZE, then WRITE_SIZE, in KiB per dispatch).  Each counter is
    scaled by what it reads for the plain copy of the batch, whose bytes are
    known (npoly * 512 read and written)."""
    def per_kernel(d, counter):
        vals = {}
        for name, v in counter_rows(d, counter):
            label = kernel_label(name)
            if label:
                vals.setdefault(label, []).append(v)
        return {k: float(np.median(v)) for k, v in vals.items()}

    fetch, write = per_kernel(fetch_dir, "FETCH_SIZE"), per_kernel(write_dir, "WRITE_SIZE")
    kib = npoly * 512 / 1024
    f_scale, w_scale = kib / fetch[COPY_TAG], kib / write[COPY_TAG]
    kernels = {}
    for label in sorted(set(fetch) & set(write) - {COPY_TAG}):
        b = (fetch[label] * f_scale + write[label] * w_scale) * 1024
        kernels[label] = {"fetch_kib": fetch[label], "write_kib": write[label],
                          "bytes_per_launch": int(round(b)),
                          "vs_algorithmic": b / (npoly * BYTES_PER_POLY_PASS)}
    out = {"csrc_digest": csrc_digest(), "npoly": npoly,
           "copy_calibration": {"fetch_kib": fetch[COPY_TAG], "write_kib": write[COPY_TAG],
                                "fetch_scale": f_scale, "write_scale": w_scale},
           "kernels": kernels}
    with open(TRAFFIC_JSON, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out))


def run_extras(tk, torch, args):
    """The sections beyond the headline, as selected by --extras."""
    sections = {
        "c1": ("config1", "configuration 1 rows", lambda: bench_config_rows(
            tk, torch, 1, 30, ("butterfly", "mfma16"))),
        "c2": ("config2", "configuration 2 rows", lambda: bench_config_rows(tk, torch, 2, 30)),
        "basemul": ("basemul", "basemul_batch", lambda: bench_basemul(tk, torch, args.steps)),
        "ragged": ("config4_ragged", "config 4 ragged plan", lambda: bench_ragged(tk, torch, args.steps)),
        "matvec": ("config5_matvec", "config 5 fused matvec", lambda: bench_matvec(tk, torch, args.steps)),
        "packed": ("packed_input", "packed input", lambda: bench_packed(tk, torch, args.steps)),
        "kpke": ("kpke_cores", "K-PKE cores", lambda: bench_kpke(tk, torch, args.steps)),
        "sched": ("fused_plan_mix", "rank-aware fused plan", lambda: bench_sched(tk, torch, args.steps)),
        "host": ("host_pipeline", "host-resident batches (PCIe)", lambda: bench_host(tk, torch)),
    }
    want = [e.strip() for e in args.extras.split(",") if e.strip()]
    unknown = [e for e in want if e not in sections]
    if unknown:
        raise SystemExit(f"unknown --extras {unknown}; choose from {list(sections)}")
    out = {}
    for name in want:
        key, label, fn = sections[name]
        print(f"[bench] {label}", file=sys.stderr, flush=True)
        out[key] = fn()
    return out


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the batch shards by independent polys with no "
              "exchange, so an N-GPU figure is N single-GPU runs; run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.traffic_from:
        traffic_from(*args.traffic_from, args.npoly)
        return
    import torch
    import tensorkyber_amd as tk
    tk.load_library()
    assert torch.cuda.is_available(), "no GPU visible"
    dev = torch.device("cuda:0")
    rc = tk.lib().tk_init()
    if rc != 0:
        raise SystemExit(f"tk_init failed: {rc}")

    if args.profile_launches:
        profile_launches(tk, torch, args.npoly, args.steps)
        return

    if args.extras_only:
        print(json.dumps(run_extras(tk, torch, args)))
        return

    npoly = args.npoly
    x = tk.workload(3, npoly)
    src = torch.from_numpy(x).to(dev)
    buf = torch.empty_like(src)
    torch.cuda.synchronize()

    if args.kernels_only:
        out = {"npoly": npoly}
        for m in ([args.only] if args.only else ["mfma16", "mfma32", "butterfly"]):
            print(f"[bench] kernels {m}", file=sys.stderr, flush=True)
            out[m] = time_method(tk, torch, m, src, buf, args.steps, args.warmup, npoly)
        out["copy"] = copy_bandwidth(torch, src, buf, args.steps)
        print(json.dumps(out))
        return

    methods = [args.only] if args.only else [args.method, "butterfly"]
    res = {}
    for m in methods:
        print(f"[bench] timing {m} npoly={npoly} steps={args.steps}", file=sys.stderr, flush=True)
        res[m] = time_method(tk, torch, m, src, buf, args.steps, args.warmup, npoly)

    main_m = methods[0]
    r = res[main_m]
    value = r["rt_poly_s"]
    dom = "fwd" if r["fwd_ms"] >= r["inv_ms"] else "inv"
    dom_ms = r[f"{dom}_ms"]
    achieved = BYTES_PER_POLY_PASS * npoly / (dom_ms / 1e3) / 1e9
    kname = f"{main_m}_{dom}"

    bit_exact = None
    if not args.no_check:
        print("[bench] parity check", file=sys.stderr, flush=True)
        fwd = torch.empty_like(src)
        tk.ntt(src, out=fwd, method=main_m)
        rt = fwd.clone()
        tk.invntt(rt, method=main_m)
        torch.cuda.synchronize()
        bit_exact = check_parity(tk, torch, main_m, x, fwd, rt)

    extras = {} if args.no_extras else run_extras(tk, torch, args)

    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(x)

    traffic, traffic_note = load_traffic(kname, npoly)
    line = {
        "metric": "NTT+iNTT poly/s at 2^20 polys (MFMA vs butterfly, bit-exact); % HBM roofline",
        "value": value,
        "unit": "poly/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["rt_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "i8xi8->i32 (int16 coefficients)",
        "data": "synthetic",
        "config": {
            "workload": f"edge-value mix, {npoly} polys x 256 int16 (seed 3), "
                        "ntt (out of place) + invntt (in place), inputs resident in HBM",
            "npoly": npoly,
            "method": main_m,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kname,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_note": traffic_note,
        },
        "rt_fraction_of_roofline": value * 2 * BYTES_PER_POLY_PASS / 1e9 / HBM_PEAK_GBS,
        "bit_exact": bit_exact,
        "kernels": res,
        "cpu_baseline": cpu,
        **extras,
    }
    line["copy_reference"] = copy_bandwidth(torch, src, buf, max(args.steps, 5))
    if "butterfly" in res and main_m != "butterfly":
        b = res["butterfly"]
        line["butterfly_poly_s"] = b["rt_poly_s"]
        line["speedup_vs_butterfly"] = value / line["butterfly_poly_s"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
