#!/usr/bin/env python3
"""Headline benchmark: baryon two-point sites/s and % of HBM roofline on one MI355X.

Workload (one step): the proton two-point function on a 24^3 x 64 lattice
(V = 884,736 sites).  Three FP64 propagators u, d, u' (iid N(0,1) entries
scaled by exp(-0.5 |t - t0|), t0 = 0, resident in HBM before timing) are
contracted with Cg5 = G(5) at source and sink, both Wick terms (direct
IK_JL_NM + exchange IM_JL_NK), all 16 insertions Gamma_mn = G(0..15), and
projected onto the 27 momenta p in {-1,0,1}^3 x 2pi/L per timeslice, giving
complex128[16][27][64].  One step = one call of
contract_baryon_twopoint_project (contraction kernel + phase table + the
deterministic per-timeslice reduction).

Also measured in the same run: the FP32-storage form of the same step
(complex64 propagators, FP32 contraction, FP64 accumulation) and its
per-timeslice relative error against the FP64 result, and the drop-in
per-site baryon_two_point kernel.

    python bench.py [--gpus 1] [--steps 50] [--warmup 5]

Prints one JSON line.  `value` is sites/s of the FP64 step; roofline.achieved
is the contraction kernel's algorithmic HBM bytes (3 x 2304 B per site) over
its own duration, timed with HIP events recorded around that kernel on its
stream.
"""

from __future__ import annotations

import argparse
import ctypes
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

DIMS = (24, 24, 24, 64)
ALPHA = 0.5
GAMMA_CG5 = 5
TERMS = [(0, 1.0), (5, 1.0)]  # IK_JL_NM (direct) + IM_JL_NK (exchange) on (u, d, u')
MOM27 = [(x, y, z) for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1)]
BYTES_PER_SITE_FP64 = 3 * 2304
BYTES_PER_SITE_FP32 = 3 * 1152
HBM_PEAK_GBS = 8000.0


class HipEvents:
    """Minimal hipEvent_t wrapper over libamdhip64 (no torch types)."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.h.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.h.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def create(self):
        e = ctypes.c_void_p()
        rc = self.h.hipEventCreate(ctypes.byref(e))
        if rc != 0:
            raise RuntimeError(f"hipEventCreate failed ({rc})")
        return e.value

    def elapsed_ms(self, a, b):
        self.h.hipEventSynchronize(ctypes.c_void_p(b))
        ms = ctypes.c_float()
        rc = self.h.hipEventElapsedTime(ctypes.byref(ms), ctypes.c_void_p(a), ctypes.c_void_p(b))
        if rc != 0:
            raise RuntimeError(f"hipEventElapsedTime failed ({rc})")
        return ms.value

    def destroy(self, e):
        self.h.hipEventDestroy(ctypes.c_void_p(e))

    def record(self, e, stream):
        self.h.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        rc = self.h.hipEventRecord(ctypes.c_void_p(e), ctypes.c_void_p(stream))
        if rc != 0:
            raise RuntimeError(f"hipEventRecord failed ({rc})")


def make_propagators(torch, dims, seeds, alpha, device, t0=0):
    Lx, Ly, Lz, Lt = dims
    V = Lx * Ly * Lz * Lt
    prof = torch.exp(-alpha * torch.abs(torch.arange(Lt, dtype=torch.float64, device=device) - t0))
    out = []
    for s in seeds:
        g = torch.Generator(device=device)
        g.manual_seed(s)
        p = torch.randn((V, 144, 2), dtype=torch.float64, device=device, generator=g)
        p = torch.view_as_complex(p).reshape(Lt, -1, 144)
        p *= prof[:, None, None]
        out.append(p.reshape(V, 4, 4, 3, 3).contiguous())
    return out


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on the bench workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (profiles/traffic.json,
    fetch scaled by 2.0 and write by 1.0 against a plain copy kernel)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            k = json.load(f)["kernels"][kernel]
        return k["traffic_bytes"], "profiles/traffic.json: " + k["workload"]
    except (OSError, KeyError, ValueError):
        return None, None


PREWARM_S = 0.25  # seconds of continuous load before each leg's warmup steps (--prewarm)


def progress(leg):
    """One line per benchmark leg on stderr (the JSON line goes to stdout)."""
    print(f"bench: {leg}", file=sys.stderr, flush=True)


def time_steps(torch, fn, steps, warmup, hip=None, timer=None, prewarm_s=None):
    """Run warmup + timed steps; returns (wall ms/step, kernel ms/step or None).

    Before the `warmup` steps the step runs back to back for `prewarm_s`
    seconds: the part's power management boosts the clock after idle and
    pulls it back during the first ~20 ms of heavy load (the 24^3x64 proton
    kernel measured 0.94 ms for the first two dispatches after 10 ms idle,
    up to 1.44 ms for the next ones, ~0.95-1.0 ms once settled), so a short
    timed window that starts from idle measures the transient, not the
    sustained rate.  Python's garbage collector is paused over the timed
    steps (as timeit does): a collection while the host enqueues
    sub-millisecond steps left the GPU idle and doubled a step's wall time in
    some runs."""
    prewarm_s = PREWARM_S if prewarm_s is None else prewarm_s
    t_end = time.perf_counter() + prewarm_s
    while time.perf_counter() < t_end:
        for _ in range(4):  # a few steps in flight, so the queue ends with the prewarm
            fn()
        torch.cuda.synchronize()
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev = []
    if hip is not None:
        ev = [(hip.create(), hip.create()) for _ in range(steps)]
    gc_on = gc.isenabled()
    gc.collect()
    gc.disable()
    try:
        t0 = time.perf_counter()
        for k in range(steps):
            if ev:
                timer(ev[k][0], ev[k][1])
            fn()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) * 1e3 / steps
    finally:
        if gc_on:
            gc.enable()
    kern = None
    if ev:
        timer(None, None)
        kern = sum(hip.elapsed_ms(a, b) for a, b in ev) / steps
        for a, b in ev:
            hip.destroy(a)
            hip.destroy(b)
    return wall, kern


def per_timeslice_error(got, want):
    """max over t of  max_{insertion, momentum} |got - want| / max_{insertion, momentum} |want|
    for arrays shaped [insertion][momentum][t]."""
    num = np.abs(got - want).max(axis=(0, 1))
    den = np.abs(want).max(axis=(0, 1))
    return float((num / np.maximum(den, 1e-300)).max())


CHANNEL_FLOOR = 1e-3  # channels below this fraction of their timeslice's largest entry are not rated


def per_channel_error(got, want, floor=CHANNEL_FLOOR):
    """Relative error of every channel (insertion, momentum, t) on its own,
    |got - want| / |want|, over the channels with |want| at least `floor`
    times the largest |want| of their timeslice (a channel far below that is
    dominated by cancellation, and its relative error says nothing about the
    contraction).  Returns max, median and the fraction of channels rated."""
    a = np.abs(want)
    keep = a >= floor * a.max(axis=(0, 1), keepdims=True)
    rel = np.abs(got - want)[keep] / a[keep]
    return {"max": float(rel.max()), "median": float(np.median(rel)), "rated": float(keep.mean()),
            "floor": floor}


def cpu_description():
    """Host CPU as /proc/cpuinfo names it (model name, family/model/stepping,
    logical CPUs)."""
    info = {}
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ":" not in ln:
                    if info:
                        break
                    continue
                k, v = (x.strip() for x in ln.split(":", 1))
                if k in ("vendor_id", "model name", "cpu family", "model", "stepping", "cpu MHz") and k not in info:
                    info[k] = v
    except OSError:
        pass
    return {"model_name": info.get("model name", ""), "vendor": info.get("vendor_id", ""),
            "family_model_stepping": "/".join(info.get(k, "?") for k in ("cpu family", "model", "stepping")),
            "mhz": info.get("cpu MHz", ""), "logical_cpus": os.cpu_count()}


class Probe:
    """The HBM ceiling probes of contract_amd/lib/libcontract_probe.so."""

    def __init__(self):
        path = os.path.join(ROOT, "contract_amd", "lib", "libcontract_probe.so")
        self.L = ctypes.CDLL(path)
        P, S, I = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        self.L.probe_read_shape.argtypes = [P, S, P, P, I, I]
        self.L.probe_copy_shape.argtypes = [P, P, S, P, I, I]
        self.L.probe_read_blocks.restype = ctypes.c_int


def hbm_ceilings(torch, hip, nbytes=4 << 30, reps=20):
    """Measured HBM ceilings on this device, each timed with HIP events on
    the stream its kernel runs on: a read-only stream (16-byte non-temporal
    loads) and a device copy (bytes read + bytes written).  The rate depends
    on the launch shape (workgroups x loads in flight per thread), so each
    ceiling is the best median over a few persistent shapes; every shape's
    rate is kept."""
    pr = Probe()
    dev = torch.device("cuda:0")
    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    src = torch.ones(nbytes // 8, dtype=torch.int64, device=dev)
    dst = torch.empty_like(src)
    sink = torch.empty(pr.L.probe_read_blocks(), dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    cap = pr.L.probe_read_blocks()
    shapes = {
        "read": [(cus, 8), (3 * cus, 2), (3 * cus, 4), (6 * cus, 1), (cap, 8)],
        "copy": [(2 * cus, 2), (3 * cus, 1), (4 * cus, 2), (8 * cus, 1), (cap, 1)],
    }
    out = {}
    for name, moved in (("read", nbytes), ("copy", 2 * nbytes)):
        rates = {}
        for blocks, unroll in shapes[name]:
            blocks = min(blocks, cap)
            if name == "read":
                fn = lambda: pr.L.probe_read_shape(src.data_ptr(), nbytes, sink.data_ptr(), stream, blocks, unroll)
            else:
                fn = lambda: pr.L.probe_copy_shape(dst.data_ptr(), src.data_ptr(), nbytes, stream, blocks, unroll)
            for _ in range(3):
                if fn() != 0:
                    raise RuntimeError(f"probe_{name}_shape({blocks}, {unroll}) launch failed")
            torch.cuda.synchronize()
            ev = [(hip.create(), hip.create()) for _ in range(reps)]
            for a, b in ev:
                hip.record(a, stream)
                fn()
                hip.record(b, stream)
            ms = sorted(hip.elapsed_ms(a, b) for a, b in ev)
            for a, b in ev:
                hip.destroy(a)
                hip.destroy(b)
            rates[f"{blocks}x{unroll}"] = moved / (ms[len(ms) // 2] * 1e-3) / 1e9
        best = max(rates, key=rates.get)
        out[f"{name}_GBs"] = rates[best]
        out[f"{name}_shape"] = best
        out[f"{name}_GBs_by_shape"] = {k: round(v, 1) for k, v in rates.items()}
    if not torch.equal(dst[:4096], src[:4096]) or not torch.equal(dst[-4096:], src[-4096:]):
        raise RuntimeError("probe_copy did not copy")
    del src, dst, sink
    torch.cuda.empty_cache()
    out["bytes"] = nbytes
    out["shape_note"] = "workgroups of 256 threads x 16-byte loads in flight per thread"
    return out


def cpu_baseline(torch, props, gpu_fn, target_s=10.0):
    """Time the FP64 oracle (oracle/, C + OpenMP over sites) on a bounded sample
    of the same workload -- the first nt timeslices, nt chosen for about
    `target_s` seconds -- and check the GPU result on that sample against it."""
    import oracle
    Lx, Ly, Lz, Lt = DIMS
    V3 = Lx * Ly * Lz
    gammas = list(range(16))

    def run(nt):
        dims = (Lx, Ly, Lz, nt)
        sample = [p.reshape(-1, 144)[: V3 * nt].cpu().numpy() for p in props]
        t0 = time.perf_counter()
        out = oracle.baryon_twopoint_project(*sample, TERMS, GAMMA_CG5, GAMMA_CG5, gammas, dims, MOM27)
        return time.perf_counter() - t0, out, dims, sample

    dt, _, _, _ = run(1)
    nt = int(max(1, min(Lt, round(target_s / max(dt, 1e-3)))))
    dt, want, dims, sample = run(nt)
    sites = V3 * nt
    dev = [torch.from_numpy(np.ascontiguousarray(s)).to(props[0].device) for s in sample]
    got = gpu_fn(dev, dims).cpu().numpy()
    cores = oracle.num_threads()
    # one core on one timeslice (a few seconds): the per-core rate
    oracle.set_num_threads(1)
    dt1, _, _, _ = run(1)
    oracle.set_num_threads(0)
    cpu = cpu_description()
    model = f"{cpu['model_name']} ({cpu['vendor']} family/model/stepping {cpu['family_model_stepping']})"
    return {
        "value": sites / dt,
        "unit": "sites/s",
        "cores": cores,
        "kind": "port",
        "sample": f"first {nt} timeslices ({sites} sites) of the same 24^3x64 proton workload (16 insertions, "
                  f"27 momenta), FP64 oracle, {dt:.1f} s wall",
        "single_core_value": V3 / dt1,
        "single_core_sample": f"first timeslice ({V3} sites), 1 thread, {dt1:.1f} s wall",
        "cpu_model": model,
        "nproc": os.cpu_count(),
        "gpu_vs_oracle_rel_err": per_timeslice_error(got, want),
    }


def shared_u(torch, pc, ps, dims, out, ws, steps, warmup, hip, timer, wall3, kern3, bps3):
    """The same step with P_n = P_i, one buffer for both u quarks (u, d, u):
    the library reads that propagator once per site (4608 B/site FP64,
    2306 B/site site-scaled FP32).  Checked bit-identical against the
    three-propagator kernel on (u, d, copy of u)."""
    V = int(np.prod(dims))
    u, d = ps[0], ps[1]
    out_s = torch.empty_like(out)

    def step():
        pc.baryonTwoPointProject(u, d, u, TERMS, GAMMA_CG5, GAMMA_CG5, list(range(16)), MOM27, dims, out=out_s,
                                 workspace=ws)

    wall, kern = time_steps(torch, step, steps, warmup, hip, timer)
    u_copy = u.clone() if isinstance(u, torch.Tensor) else pc.ScaledPropagator(u.data.clone(), u.volume)
    out_c = torch.empty_like(out)
    pc.baryonTwoPointProject(u, d, u_copy, TERMS, GAMMA_CG5, GAMMA_CG5, list(range(16)), MOM27, dims, out=out_c,
                             workspace=ws)
    torch.cuda.synchronize()
    same = bool(torch.equal(out_s, out_c))
    del u_copy
    bps = bps3 - (bps3 // 3)  # one propagator (and its exponent byte) fewer
    gbs = bps * V / (kern * 1e-3) / 1e9
    return {"ms_per_step": wall, "kernel_ms": kern, "sites_per_s": V / (wall * 1e-3), "bytes_per_site": bps,
            "achieved_GBs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS, "speedup_vs_three_propagators": wall3 / wall,
            "kernel_speedup_vs_three_propagators": kern3 / kern, "identical_to_three_propagator_kernel": same}


def proton_32c64(torch, pc, dev, steps, warmup, hip, timer):
    """The proton step of the headline workload (Cg5, direct + exchange, 16
    insertions, 27 momenta) on a 32^3x64 lattice, FP64 and FP32 storage."""
    dims = (32, 32, 32, 64)
    V = 32 ** 3 * 64
    props = make_propagators(torch, dims, (4001, 4002, 4003), ALPHA, dev)
    ws = torch.empty(pc.project_workspace_bytes(dims, 27), dtype=torch.uint8, device=dev)
    out = torch.empty((16, 27, 64), dtype=torch.complex128, device=dev)
    res = {"workload": "proton 2pt 32^3x64, Cg5, direct+exchange, 16 insertions, 27 momenta", "sites": V}
    scaled = [pc.toScaledFP32(p) for p in props]
    for name, ps, bps in (("fp64", props, BYTES_PER_SITE_FP64), ("fp32", None, BYTES_PER_SITE_FP32),
                          ("fp32_scaled", scaled, BYTES_PER_SITE_FP32 + 3)):
        if ps is None:
            ps = [p.to(torch.complex64) for p in props]
            del props[:]

        def step():
            pc.baryonTwoPointProject(*ps, TERMS, GAMMA_CG5, GAMMA_CG5, list(range(16)), MOM27, dims, out=out,
                                     workspace=ws)

        wall, kern = time_steps(torch, step, steps, warmup, hip, timer)
        gbs = bps * V / (kern * 1e-3) / 1e9
        res[name] = {"ms_per_step": wall, "kernel_ms": kern, "sites_per_s": V / (wall * 1e-3),
                     "achieved_GBs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS}
        if name == "fp64":
            ref = out.cpu().numpy()
        else:
            got = out.cpu().numpy()
            res[name]["per_timeslice_rel_err_vs_fp64"] = per_timeslice_error(got, ref)
            res[name]["per_channel_rel_err_vs_fp64"] = per_channel_error(got, ref)
        if name != "fp32":
            res[name]["shared_u"] = shared_u(torch, pc, ps, dims, out, ws, steps, warmup, hip, timer, wall, kern,
                                             bps)
        del ps
    res["fp32_speedup"] = res["fp64"]["ms_per_step"] / res["fp32"]["ms_per_step"]
    res["fp32_kernel_speedup"] = res["fp64"]["kernel_ms"] / res["fp32"]["kernel_ms"]
    res["fp32_scaled_speedup"] = res["fp64"]["ms_per_step"] / res["fp32_scaled"]["ms_per_step"]
    res["fp32_scaled_kernel_speedup"] = res["fp64"]["kernel_ms"] / res["fp32_scaled"]["kernel_ms"]
    return res


def baryon_batch(torch, pc, dev, steps, warmup, hip=None, timer=None, nreq=16, dims=(16, 16, 16, 32)):
    """Batching leg: the proton two-point (16 insertions, 27 momenta) for
    `nreq` source positions on a 16^3x32 lattice -- each with its own u, d, u'
    propagators -- as one batched launch versus `nreq` single calls."""
    V = dims[0] * dims[1] * dims[2] * dims[3]
    reqs = []
    for n in range(nreq):
        ps = make_propagators(torch, dims, (7000 + 3 * n, 7001 + 3 * n, 7002 + 3 * n), ALPHA, dev)
        reqs.append((*ps, dims))
    gammas = list(range(16))
    outs = [torch.empty((16, 27, dims[3]), dtype=torch.complex128, device=dev) for _ in range(nreq)]
    ws = torch.empty(pc.baryon_batch_workspace_bytes(reqs, 27), dtype=torch.uint8, device=dev)

    def batched():
        pc.baryonTwoPointProjectBatch(reqs, TERMS, GAMMA_CG5, GAMMA_CG5, gammas, MOM27, outs=outs, workspace=ws)

    def singles():
        for (pi, pj, pn, d), o in zip(reqs, outs):
            pc.baryonTwoPointProject(pi, pj, pn, TERMS, GAMMA_CG5, GAMMA_CG5, gammas, MOM27, d, out=o, workspace=ws)

    wall_b, kern_b = time_steps(torch, batched, steps, warmup, hip, timer)
    got = [o.cpu() for o in outs]
    # the same batch captured once in a HIP graph (it allocates nothing with a
    # caller workspace) and replayed
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.graph(graph, stream=side):
        pc.baryonTwoPointProjectBatch(reqs, TERMS, GAMMA_CG5, GAMMA_CG5, gammas, MOM27, outs=outs, workspace=ws,
                                      stream=side.cuda_stream)
    wall_g, _ = time_steps(torch, graph.replay, steps, warmup)
    same_graph = all(torch.equal(a, o.cpu()) for a, o in zip(got, outs))
    del graph
    wall_s, _ = time_steps(torch, singles, steps, warmup)
    same = all(torch.equal(a, o.cpu()) for a, o in zip(got, outs))
    gbs = BYTES_PER_SITE_FP64 * V * nreq / (wall_b * 1e-3) / 1e9
    return {"workload": f"proton 2pt for {nreq} source positions, {dims[0]}^3x{dims[3]} each, Cg5, 16 insertions, "
                        "27 momenta",
            "requests": nreq, "batched_ms": wall_b, "batched_kernel_ms": kern_b, "batched_graph_ms": wall_g,
            "graph_replay_bit_identical": same_graph,
            "kernel_GBs": None if kern_b is None else BYTES_PER_SITE_FP64 * V * nreq / (kern_b * 1e-3) / 1e9,
            "singles_ms": wall_s, "speedup": wall_s / wall_b,
            "sites_per_s": V * nreq / (wall_b * 1e-3), "achieved_GBs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS,
            "bit_identical_to_singles": same}


def precision_study(torch, pc, dev, steps, warmup, hip, timer):
    """Sequential-source precision study: 32^3x64, propagators decaying as
    exp(-0.8|t|) (~20 decades over the lattice), stored in FP32.  Two forward
    propagators and a sequential propagator are contracted with Cg5 and all 16
    insertions at p = 0; the per-timeslice sum of (sequential source x
    sequential propagator) equals this two-point contraction with the
    sequential propagator in the n slot (sum_x P_n . dC/dP_n = C).  FP32
    storage + FP32 contraction + FP64 accumulation vs the FP64 oracle run on
    the same FP32-stored values."""
    import oracle
    dims = (32, 32, 32, 64)
    V = 32 ** 3 * 64
    terms = [(0, 1.0)]
    mom = [(0, 0, 0)]
    f32 = [p.to(torch.complex64) for p in make_propagators(torch, dims, (2001, 2002, 2003), 0.8, dev)]
    torch.cuda.synchronize()
    out32 = torch.empty((16, 1, 64), dtype=torch.complex128, device=dev)
    ws = torch.empty(pc.project_workspace_bytes(dims, 1), dtype=torch.uint8, device=dev)

    def step32():
        pc.baryonTwoPointProject(*f32, terms, GAMMA_CG5, GAMMA_CG5, list(range(16)), mom, dims, out=out32,
                                 workspace=ws)

    wall32, kern32 = time_steps(torch, step32, steps, warmup, hip, timer)
    got = out32.cpu().numpy()
    f64 = [p.to(torch.complex128) for p in f32]
    out64 = torch.empty_like(out32)

    def step64():
        pc.baryonTwoPointProject(*f64, terms, GAMMA_CG5, GAMMA_CG5, list(range(16)), mom, dims, out=out64,
                                 workspace=ws)

    wall64, kern64 = time_steps(torch, step64, steps, warmup, hip, timer)
    # the FP32 site-scaled form of the same stored values (exact: the
    # scaling is by powers of two)
    scaled = [pc.toScaledFP32(p) for p in f64]
    out32s = torch.empty_like(out32)

    def step32s():
        pc.baryonTwoPointProject(*scaled, terms, GAMMA_CG5, GAMMA_CG5, list(range(16)), mom, dims, out=out32s,
                                 workspace=ws)

    wall32s, kern32s = time_steps(torch, step32s, steps, warmup, hip, timer)
    got_s = out32s.cpu().numpy()
    del scaled
    host = [p.cpu().numpy() for p in f64]
    del f64
    t0 = time.perf_counter()
    want = oracle.baryon_twopoint_project(*host, terms, GAMMA_CG5, GAMMA_CG5, list(range(16)), dims, mom)
    t_oracle = time.perf_counter() - t0
    per_t = np.abs(got - want).max(axis=(0, 1)) / np.abs(want).max(axis=(0, 1))
    per_t_s = np.abs(got_s - want).max(axis=(0, 1)) / np.abs(want).max(axis=(0, 1))
    got64 = out64.cpu().numpy()
    return {
        "workload": "sequential-source precision study, 32^3x64, exp(-0.8|t|), FP32 storage, Cg5, 16 insertions, p=0",
        "max_per_timeslice_rel_err": float(per_t.max()),
        "median_per_timeslice_rel_err": float(np.median(per_t)),
        "dynamic_range_decades": float(np.log10(np.abs(want[:, 0, 0]).max() / np.abs(want[:, 0, -1]).max())),
        "fp64_gpu_vs_oracle_max_rel_err": per_timeslice_error(got64, want),
        "per_channel_rel_err": per_channel_error(got, want),
        "fp64_gpu_vs_oracle_per_channel": per_channel_error(got64, want),
        "fp32_kernel_ms": kern32,
        "fp64_kernel_ms": kern64,
        "fp32_sites_per_s": V / (wall32 * 1e-3),
        "fp64_sites_per_s": V / (wall64 * 1e-3),
        "fp32_speedup": wall64 / wall32,
        "fp32_kernel_speedup": kern64 / kern32 if kern32 and kern64 else None,
        "fp32_achieved_GBs": BYTES_PER_SITE_FP32 * V / (kern32 * 1e-3) / 1e9,
        "fp64_achieved_GBs": BYTES_PER_SITE_FP64 * V / (kern64 * 1e-3) / 1e9,
        "fp32_scaled": {"max_per_timeslice_rel_err": float(per_t_s.max()),
                        "median_per_timeslice_rel_err": float(np.median(per_t_s)),
                        "per_channel_rel_err": per_channel_error(got_s, want),
                        "kernel_ms": kern32s, "sites_per_s": V / (wall32s * 1e-3), "speedup": wall64 / wall32s,
                        "kernel_speedup": kern64 / kern32s if kern32s and kern64 else None},
        "oracle_s": t_oracle,
        "sequential_chain": sequential_chain(torch, pc, dev, f32, steps, warmup),
    }


def sequential_chain(torch, pc, dev, f32, steps, warmup):
    """The sequential-source chain of the precision study with every
    propagator in FP32 storage: the sequential source of the n quark
    (IK_JL_NM, Cg5, insertion G(8)) from the two forward propagators, its
    unit-Dirac sequential propagator, and the three-point step against the
    forward n propagator for 16 insertions at p = 0 -- site-scaled FP32
    throughout, against the same chain in FP64 on the same stored values
    (the FP64 path is pinned to the oracle at 1e-12 in the tests)."""
    dims = (32, 32, 32, 64)
    V = 32 ** 3 * 64
    mom = [(0, 0, 0)]
    ins = list(range(16))
    f64 = [p.to(torch.complex128) for p in f32]
    fwd64 = f64[2].clone()
    seq64 = f64[2]
    pc.baryonSequential(f64[0], f64[1], seq64, 0, 2, GAMMA_CG5, GAMMA_CG5, 8)
    want = pc.threePointProject([(fwd64, pc.sequentialSource(seq64), dims)], ins, mom)[0].cpu().numpy()
    sc = [pc.toScaledFP32(p) for p in (f64[0], f64[1], fwd64)]
    del f64, seq64
    fwd32 = sc[2].clone()

    def seq_step32():
        pc.baryonSequential(sc[0], sc[1], sc[2], 0, 2, GAMMA_CG5, GAMMA_CG5, 8)

    wall_seq32, _ = time_steps(torch, seq_step32, steps, warmup)
    got = pc.threePointProject([(fwd32, pc.sequentialSource(sc[2]), dims)], ins, mom)[0].cpu().numpy()
    del sc, fwd32
    p64 = [torch.empty((V, 4, 4, 3, 3), dtype=torch.complex128, device=dev) for _ in range(3)]

    def seq_step64():
        pc.baryonSequential(p64[0], p64[1], p64[2], 0, 2, GAMMA_CG5, GAMMA_CG5, 8)

    wall_seq64, _ = time_steps(torch, seq_step64, steps, warmup)
    del p64, fwd64
    torch.cuda.empty_cache()
    return {"workload": "sequential source (n, IK_JL_NM, Cg5, G(8)) -> unit-Dirac sequential propagator -> "
                        "three-point step, 16 insertions, p=0; FP32 site-scaled vs FP64",
            "max_per_timeslice_rel_err": per_timeslice_error(got, want),
            "per_channel_rel_err": per_channel_error(got, want),
            "sequential_fp64_ms": wall_seq64, "sequential_fp32_scaled_ms": wall_seq32,
            "sequential_GBs_fp64": 3 * 2304 * V / (wall_seq64 * 1e-3) / 1e9,
            "sequential_GBs_fp32_scaled": 3 * 1153 * V / (wall_seq32 * 1e-3) / 1e9,
            "sequential_fp32_speedup": wall_seq64 / wall_seq32}


GRID_LATTICES = [(16, 16, 16, 32)] * 8 + [(24, 24, 24, 48)] * 8 + [(32, 32, 32, 64)] * 8
MOM7 = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]


def meson_grid_batch(torch, pc, dev, steps, warmup, hip, timer, check_oracle, fixed_tuning=None, precisions=("fp64", "fp32")):
    """Batched meson all-to-all: 24 uneven requests (8 each on 16^3x32,
    24^3x48, 32^3x64), each a pair of FP64 propagators (exp(-0.25|t - t0|),
    t0 drawn per request from a seeded RNG), full 16x16 Gamma x Gamma' grid at
    7 momenta (origin and the six unit vectors), all in one persistent launch.
    The autotuner picks the work decomposition per precision; the tuned batch
    is then timed eagerly and as a captured HIP graph, in FP64 and in the
    FP32-contraction / FP64-accumulation form.  The FP64 batch is also timed
    at the 27 momenta of {-1,0,1}^3 under the library's default plan."""
    from contract_amd import autotune
    rng = np.random.default_rng(2024)
    reqs = []
    for k, dims in enumerate(GRID_LATTICES):
        t0 = int(rng.integers(0, dims[3]))
        a, b = make_propagators(torch, dims, (3000 + 2 * k, 3001 + 2 * k), 0.25, dev, t0=t0)
        reqs.append((a, b, dims, 0))
    sites = sum(int(np.prod(d)) for d in GRID_LATTICES)
    res = {"workload": "24 requests: 8 x 16^3x32, 8 x 24^3x48, 8 x 32^3x64; 16x16 Gamma grid, 7 momenta",
           "sites": sites}
    outs = {}
    for name, batch, bps in (("fp64", reqs, 2 * 2304), ("fp32_scaled", None, 2 * 1153), ("fp32", None, 2 * 1152)):
        if name not in precisions and not (name == "fp32_scaled" and "fp32" in precisions):
            continue
        if name == "fp32_scaled":
            batch = [(pc.toScaledFP32(a), pc.toScaledFP32(b), d, o) for a, b, d, o in reqs]
        elif batch is None:
            batch = [(a.to(torch.complex64), b.to(torch.complex64), d, o) for a, b, d, o in reqs]
            del reqs[:]
        torch.cuda.synchronize()
        t_tune = time.perf_counter()
        if fixed_tuning is None:
            best, table = autotune.tune_meson_grid(batch, MOM7, reps=2)
        else:
            best, table = tuple(fixed_tuning), {}
        t_tune = time.perf_counter() - t_tune
        ws = torch.empty(pc.meson_grid_workspace_bytes(batch, 7, best), dtype=torch.uint8, device=dev)
        out = [torch.empty((16, 16, 7, d[3]), dtype=torch.complex128, device=dev) for _, _, d, _ in batch]

        def step():
            pc.mesonGridProject(batch, MOM7, tuning=best, outs=out, workspace=ws)

        wall, kern = time_steps(torch, step, steps, warmup, hip, timer)
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        torch.cuda.synchronize()
        with torch.cuda.graph(graph, stream=side):
            pc.mesonGridProject(batch, MOM7, tuning=best, outs=out, workspace=ws, stream=side.cuda_stream)
        wall_g, _ = time_steps(torch, graph.replay, steps, warmup)
        gbs = bps * sites / (kern * 1e-3) / 1e9
        res[name] = {
            "tuning": {"sites_per_item": best[0], "tile_sites": best[1], "blocks_per_cu": best[2],
                       "trace_unit": {0: "library default", 1: "vector ALUs", 2: "matrix cores",
                                      3: "matrix cores, wave-local K"}[
                           best[3] if len(best) > 3 else 0]},
            "autotune_s": t_tune,
            "autotune_ms": {"/".join(str(x) for x in k): round(v, 3) for k, v in sorted(table.items())},
            "ms_per_batch": wall,
            "ms_per_batch_graph": wall_g,
            "kernel_ms": kern,
            "sites_per_s": sites / (wall * 1e-3),
            "achieved_GBs": gbs,
            "hbm_frac": gbs / HBM_PEAK_GBS,
            "mfma_TFLOPs": 16384.0 * sites / (kern * 1e-3) / 1e12,
        }
        if name == "fp64":
            res[name]["traffic"], res[name]["traffic_source"] = measured_traffic("meson_grid_kernel<double,4,q4>")
            # the same batch at the 27 momenta of {-1,0,1}^3 (14 +-k slots: one
            # pass with two projection row tiles), library default plan
            ws27 = torch.empty(pc.meson_grid_workspace_bytes(batch, 27), dtype=torch.uint8, device=dev)
            out27 = [torch.empty((16, 16, 27, d[3]), dtype=torch.complex128, device=dev) for _, _, d, _ in batch]

            def step27():
                pc.mesonGridProject(batch, MOM27, outs=out27, workspace=ws27)

            wall27, kern27 = time_steps(torch, step27, max(steps // 2, 5), warmup, hip, timer)
            res["fp64_mom27"] = {"momenta": 27, "tuning": "library default", "ms_per_batch": wall27,
                                 "kernel_ms": kern27, "hbm_frac": bps * sites / (kern27 * 1e-3) / 1e9 / HBM_PEAK_GBS}
            del ws27, out27
        outs[name] = [o.cpu().numpy() for o in out]
        if name == "fp64" and check_oracle:
            import oracle
            a, b, d, o = batch[0]
            want = oracle.meson_grid_project(a.cpu().numpy(), b.cpu().numpy(), d, MOM7, o)
            num = np.abs(outs[name][0] - want).max(axis=2)
            res[name]["gpu_vs_oracle_rel_err_first_request"] = float((num / np.abs(want).max(axis=2)).max())
        del batch, out, ws, graph
    if "fp64" not in outs:
        return res
    for name in ("fp32", "fp32_scaled"):
        if name not in outs:
            continue
        err = 0.0
        for got, want in zip(outs[name], outs["fp64"]):
            num = np.abs(got - want).max(axis=(0, 1, 2))
            den = np.abs(want).max(axis=(0, 1, 2))
            err = max(err, float((num / den).max()))
        res[name]["per_timeslice_rel_err_vs_fp64"] = err
        res[name]["speedup_vs_fp64"] = res["fp64"]["ms_per_batch"] / res[name]["ms_per_batch"]
    return res


def threepoint_leg(torch, pc, dev, steps, warmup, hip, timer):
    """Three-point step after the sequential source on 32^3x64: a forward
    propagator F and a sequential propagator Sigma (FP64, synthetic, decaying
    from t = 0 and from the sink t = 63), all 16 insertions, 7 momenta
    (origin and unit axes) -> complex128[16][7][64]; FP64 and the FP32
    site-scaled form.  Algorithmic bytes: 2 x 2304 B per site."""
    dims = (32, 32, 32, 64)
    V = 32 ** 3 * 64
    f, = make_propagators(torch, dims, (5001,), ALPHA, dev)
    s, = make_propagators(torch, dims, (5002,), ALPHA, dev, t0=63)
    req = [(f, s, dims)]
    ws = torch.empty(pc.threepoint_workspace_bytes(req, 7), dtype=torch.uint8, device=dev)
    out = [torch.empty((16, 7, 64), dtype=torch.complex128, device=dev)]
    res = {"workload": "three-point step 32^3x64, F and Sigma, 16 insertions, 7 momenta", "sites": V}
    for name, r, bps in (("fp64", req, 2 * 2304), ("fp32_scaled", None, 2 * 1153)):
        if r is None:
            r = [(pc.toScaledFP32(f), pc.toScaledFP32(s), dims)]

        def step():
            pc.threePointProject(r, list(range(16)), MOM7, outs=out, workspace=ws)

        wall, kern = time_steps(torch, step, steps, warmup, hip, timer)
        gbs = bps * V / (kern * 1e-3) / 1e9
        res[name] = {"ms_per_step": wall, "kernel_ms": kern, "achieved_GBs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS}
        if name == "fp64":
            ref = out[0].cpu().numpy()
        else:
            got = out[0].cpu().numpy()
            res[name]["per_timeslice_rel_err_vs_fp64"] = per_timeslice_error(got, ref)
            res[name]["per_channel_rel_err_vs_fp64"] = per_channel_error(got, ref)
        del r
    res["summary"] = {"fp64_kernel_ms": _r(res["fp64"]["kernel_ms"]), "fp64_hbm_frac": _r(res["fp64"]["hbm_frac"]),
                      "fp32_scaled_kernel_speedup": _r(res["fp64"]["kernel_ms"] / res["fp32_scaled"]["kernel_ms"]),
                      "fp32_scaled_err_per_t": _r(res["fp32_scaled"]["per_timeslice_rel_err_vs_fp64"], 2)}
    return res


def abi_entry_points(torch, pc, u, d, u2, V, steps, warmup):
    """Every drop-in entry point once per step on the 24^3x64 FP64 propagators:
    ms per call and HBM GB/s of its algorithmic bytes (inputs read once,
    outputs written once)."""
    out = {}
    correl = torch.empty(V, dtype=torch.complex128, device=dev_of(u))
    c16 = [torch.empty(V, dtype=torch.complex128, device=dev_of(u)) for _ in range(16)]
    scratch = torch.empty_like(u)
    proj = pc.gamma_matrix(8)
    cases = {
        "meson_two_point": (lambda: pc.meson_two_point(correl, u, d, V, 5, 5), 2 * 2304 + 16),
        "meson_all_sink_two_point": (lambda: pc.meson_all_sink_two_point(c16, u, d, V, 5), 2 * 2304 + 256),
        "baryon_diquark": (lambda: pc.baryon_diquark(scratch, u, d, V, 5, 5), 3 * 2304),
        "baryon_two_point IK_JL_NM": (lambda: pc.baryon_two_point(correl, u, d, u2, 0, V, 5, 5, 8), 3 * 2304 + 16),
        "baryon_two_point IM_JL_NK": (lambda: pc.baryon_two_point(correl, u, d, u2, 5, V, 5, 5, 8), 3 * 2304 + 16),
        "baryon_general_two_point IK_JL_NM": (
            lambda: pc.baryon_general_two_point(correl, u, d, u2, 0, V, 5, 5, proj), 3 * 2304 + 16),
        "baryon_sequential_two_point IK_JL_NM / N": (
            lambda: pc.baryon_sequential_two_point(u, d, scratch, 0, 2, V, 5, 5, 8), 3 * 2304),
    }
    for name, (fn, bps) in cases.items():
        wall, _ = time_steps(torch, fn, steps, warmup)
        gbs = bps * V / (wall * 1e-3) / 1e9
        out[name] = {"ms": round(wall, 4), "GBs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 3),
                     "bytes_per_site": bps}
    return out


def dev_of(t):
    return t.device


def serving_leg(torch, args):
    """Serving loop: Poisson arrivals of mixed requests (50% meson grid, 30%
    proton two-point, 20% diquark) on lattices drawn from 16^3x32, 24^3x64,
    32^3x64 and 48^3x96, 1-16 gammas and 1-27 momenta each, against a
    catalogue of 4 FP64 propagators per lattice held in an LRU pool capped at
    200 GB; one stream of `serving_duration` seconds per offered rate."""
    from contract_amd import serving

    def memory(what):
        free, total = torch.cuda.mem_get_info()
        print(f"bench: {what}: allocated {torch.cuda.memory_allocated() / 1e9:.1f} GB, reserved "
              f"{torch.cuda.memory_reserved() / 1e9:.1f} GB, device free {free / 1e9:.1f} of {total / 1e9:.1f} GB",
              file=sys.stderr, flush=True)

    gc.collect()
    torch.cuda.empty_cache()
    memory("before the serving sweep")
    rates = [float(r) for r in args.serving_rates.split(",")]
    res = serving.sweep(rates, args.serving_duration, seed=7, cap_gb=200.0, streams=4, max_inflight=8)
    res["workload"] = ("Poisson arrivals, 50% meson / 30% baryon / 20% diquark, lattices 16^3x32, 24^3x64, "
                       "32^3x64, 48^3x96, 1-16 gammas, 1-27 momenta, LRU pool 200 GB, 4 streams")
    gc.collect()  # the sweep's engine (its ~190 GB pool) sits in reference cycles until collected
    torch.cuda.empty_cache()
    if args.serving_scaled_rates:
        # the same request sequence against a pool of FP32 site-scaled
        # propagators (packed from FP64 as they enter the pool)
        memory("before the site-scaled serving sweep")
        progress("serving sweep, site-scaled FP32 pool")
        rs = [float(r) for r in args.serving_scaled_rates.split(",")]
        res["fp32_scaled"] = serving.sweep(rs, args.serving_duration, seed=7, cap_gb=200.0, streams=4,
                                           max_inflight=8, dtype="fp32_scaled")
        torch.cuda.empty_cache()
    return res


def main():
    global PREWARM_S
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-precision", action="store_true", help="skip the 32^3x64 FP32 precision study")
    ap.add_argument("--no-grid", action="store_true", help="skip the batched meson grid leg")
    ap.add_argument("--no-big", action="store_true", help="skip the 32^3x64 proton leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-baryon leg")
    ap.add_argument("--only-big", action="store_true", help="run only the 32^3x64 proton leg")
    ap.add_argument("--only-batch", action="store_true",
                    help="run only the batched-baryon and 32^3x64 proton legs")
    ap.add_argument("--only-grid", action="store_true", help="run only the batched meson grid leg")
    ap.add_argument("--only-threepoint", action="store_true", help="run only the three-point leg")
    ap.add_argument("--grid-tuning", default=None,
                    help="sites_per_item,tile_sites,blocks_per_cu for the grid leg instead of autotuning")
    ap.add_argument("--grid-precision", default="fp64,fp32", help="precisions of the grid leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop sweep")
    ap.add_argument("--only-serving", action="store_true", help="run only the serving-loop sweep")
    ap.add_argument("--serving-rates", default="10,25,50,100,200,300", help="offered request rates (req/s)")
    ap.add_argument("--serving-scaled-rates", default="100,200,300,400",
                    help="offered rates of the sweep against a site-scaled FP32 pool ('' to skip)")
    ap.add_argument("--no-ceilings", action="store_true", help="skip the measured HBM ceilings")
    ap.add_argument("--prewarm", type=float, default=PREWARM_S,
                    help="seconds each timed leg runs back to back before its warmup steps")
    ap.add_argument("--no-threepoint", action="store_true", help="skip the three-point leg")
    ap.add_argument("--detail", default=os.path.join("gpurun_out", "bench_detail.json"),
                    help="file for the full per-leg results ('' to skip)")
    ap.add_argument("--serving-duration", type=float, default=3.0, help="seconds of arrivals per rate")
    ap.add_argument("--copy-ref", action="store_true",
                    help="also run a plain 2 GiB device copy, the reference kernel for scaling HBM counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: the contraction has no multi-GPU path in this build "
              f"(requested --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)

    PREWARM_S = max(0.0, args.prewarm)
    import torch
    from contract_amd import pycontract as pc

    dev = torch.device("cuda:0")
    pc.init(0)
    hip = HipEvents()
    timer = pc.lib().contract_set_kernel_timer

    if args.only_serving:
        print(json.dumps({"serving": serving_leg(torch, args)}))
        return

    if args.only_big:
        print(json.dumps({"proton_32c64": proton_32c64(torch, pc, dev, args.steps, args.warmup, hip, timer)}))
        return

    if args.only_batch:
        print(json.dumps({"baryon_batch": baryon_batch(torch, pc, dev, args.steps, args.warmup, hip, timer),
                          "proton_32c64": proton_32c64(torch, pc, dev, args.steps, args.warmup, hip, timer)}))
        return

    if args.only_threepoint:
        print(json.dumps({"threepoint": threepoint_leg(torch, pc, dev, args.steps, args.warmup, hip, timer)}))
        return

    if args.only_grid:
        tuning = [int(v) for v in args.grid_tuning.split(",")] if args.grid_tuning else None
        print(json.dumps({"meson_grid_batch": meson_grid_batch(torch, pc, dev, args.steps, args.warmup, hip, timer,
                                                                not args.no_cpu, tuning,
                                                                args.grid_precision.split(","))}))
        return

    Lx, Ly, Lz, Lt = DIMS
    V = Lx * Ly * Lz * Lt
    u, d, u2 = make_propagators(torch, DIMS, (1001, 1002, 1003), ALPHA, dev)
    gammas = list(range(16))
    out64 = torch.empty((16, 27, Lt), dtype=torch.complex128, device=dev)
    ws = torch.empty(pc.project_workspace_bytes(DIMS, 27), dtype=torch.uint8, device=dev)

    def step64():
        pc.baryonTwoPointProject(u, d, u2, TERMS, GAMMA_CG5, GAMMA_CG5, gammas, MOM27, DIMS, out=out64,
                                 workspace=ws)

    if args.copy_ref:
        src = torch.empty(2 << 30, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    progress("hbm ceilings")
    ceil = None if args.no_ceilings else hbm_ceilings(torch, hip)
    progress("24^3x64 proton FP64")
    wall64, kern64 = time_steps(torch, step64, args.steps, args.warmup, hip, timer)
    ref64 = out64.cpu().numpy()

    # FP32 storage form of the same step
    u_f, d_f, u2_f = (p.to(torch.complex64) for p in (u, d, u2))
    out32 = torch.empty_like(out64)

    def step32():
        pc.baryonTwoPointProject(u_f, d_f, u2_f, TERMS, GAMMA_CG5, GAMMA_CG5, gammas, MOM27, DIMS, out=out32,
                                 workspace=ws)

    wall32, kern32 = time_steps(torch, step32, args.steps, args.warmup, hip, timer)
    err32 = per_timeslice_error(out32.cpu().numpy(), ref64)
    err32c = per_channel_error(out32.cpu().numpy(), ref64)
    del u_f, d_f, u2_f

    # FP32 site-scaled storage form of the same step (conversion timed apart)
    conv = [None, None, None]

    def convert():
        for k, p in enumerate((u, d, u2)):
            conv[k] = pc.toScaledFP32(p, out=None if conv[k] is None else conv[k].data)

    wall_conv, _ = time_steps(torch, convert, 5, 1)
    out32s = torch.empty_like(out64)

    def step32s():
        pc.baryonTwoPointProject(*conv, TERMS, GAMMA_CG5, GAMMA_CG5, gammas, MOM27, DIMS, out=out32s, workspace=ws)

    wall32s, kern32s = time_steps(torch, step32s, args.steps, args.warmup, hip, timer)
    err32s = per_timeslice_error(out32s.cpu().numpy(), ref64)
    err32sc = per_channel_error(out32s.cpu().numpy(), ref64)
    del conv

    # drop-in per-site kernel (one insertion, one term), reference-style output
    correl = torch.empty(V, dtype=torch.complex128, device=dev)

    def step_abi():
        pc.baryon_two_point(correl, u, d, u2, 5, V, GAMMA_CG5, GAMMA_CG5, 8)

    wall_abi, _ = time_steps(torch, step_abi, args.steps, args.warmup)
    progress("drop-in entry points")
    abi_all = abi_entry_points(torch, pc, u, d, u2, V, max(5, args.steps // 5), args.warmup)

    precision = None
    if not args.no_precision:
        del ws
        progress("precision study")
        precision = precision_study(torch, pc, dev, args.steps, args.warmup, hip, timer)
    progress("32^3x64 proton")
    big = None if args.no_big else proton_32c64(torch, pc, dev, args.steps, args.warmup, hip, timer)
    progress("baryon batch")
    batch = None if args.no_batch else baryon_batch(torch, pc, dev, min(args.steps, 20), args.warmup, hip, timer)

    grid = None
    if not args.no_grid:
        progress("meson grid batch")
        grid = meson_grid_batch(torch, pc, dev, min(args.steps, 20), args.warmup, hip, timer, not args.no_cpu)

    threept = None
    if not args.no_threepoint:
        progress("three-point")
        threept = threepoint_leg(torch, pc, dev, min(args.steps, 20), args.warmup, hip, timer)

    serving = None
    if not args.no_serving:
        torch.cuda.empty_cache()
        progress("serving sweep")
        serving = serving_leg(torch, args)

    cpu = None
    if not args.no_cpu:
        def gpu_sample(devp, dims):
            return pc.baryonTwoPointProject(*devp, TERMS, GAMMA_CG5, GAMMA_CG5, gammas, MOM27, dims)
        progress("cpu baseline")
        cpu = cpu_baseline(torch, (u, d, u2), gpu_sample)

    achieved = BYTES_PER_SITE_FP64 * V / (kern64 * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic("baryon_project_kernel<1,double,5,nucleon>")
    detail = {
        "fp32": {
            "sites_per_s": V / (wall32 * 1e-3),
            "ms_per_step": wall32,
            "kernel_ms": kern32,
            "achieved_GBs": BYTES_PER_SITE_FP32 * V / (kern32 * 1e-3) / 1e9,
            "speedup_vs_fp64": wall64 / wall32,
            "kernel_speedup_vs_fp64": kern64 / kern32,
            "per_timeslice_rel_err_vs_fp64": err32,
            "per_channel_rel_err_vs_fp64": err32c,
        },
        "fp32_scaled": {
            "note": "FP32 site-scaled storage (contract_propagator_to_scaled_fp32): no per-site rescaling in "
                    "the kernel; conversion of the three propagators timed apart",
            "sites_per_s": V / (wall32s * 1e-3),
            "ms_per_step": wall32s,
            "kernel_ms": kern32s,
            "achieved_GBs": (BYTES_PER_SITE_FP32 + 3) * V / (kern32s * 1e-3) / 1e9,
            "speedup_vs_fp64": wall64 / wall32s,
            "kernel_speedup_vs_fp64": kern64 / kern32s,
            "per_timeslice_rel_err_vs_fp64": err32s,
            "per_channel_rel_err_vs_fp64": err32sc,
            "convert_ms_three_propagators": wall_conv,
        },
        "abi_baryon_two_point": {
            "ms_per_call": wall_abi,
            "sites_per_s": V / (wall_abi * 1e-3),
            "achieved_GBs": (BYTES_PER_SITE_FP64 + 16) * V / (wall_abi * 1e-3) / 1e9,
            "note": "one type (IM_JL_NK), one insertion, per-site output, like one reference call",
            "entry_points": abi_all,
        },
        "precision_study": precision,
        "proton_32c64": big,
        "baryon_batch": batch,
        "meson_grid_batch": grid,
        "threepoint": threept,
        "serving": serving,
        "hbm_ceilings": ceil,
    }
    result = {
        "metric": "baryon 2pt sites/s & % HBM roofline, FP32 vs FP64; p50/p99 request latency",
        "value": V / (wall64 * 1e-3),
        "unit": "sites/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: iid N(0,1) propagator entries x exp(-0.5|t|), resident in HBM",
        "config": {
            "workload": "proton 2pt, 24^3x64, u/d/u' FP64, Cg5, direct+exchange, 16 Gamma, 27 momenta "
                        "-> complex128[16][27][64]",
            "prewarm_s": PREWARM_S,
            "sites": V,
            "insertions": 16,
            "momenta": 27,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "baryon_project_kernel<1, double, 5, false, 1, false>",
            "kernel_ms": kern64,
            "bytes_per_site": BYTES_PER_SITE_FP64,
            "measured_read_ceiling_GBs": ceil and round(ceil["read_GBs"], 1),
            "measured_copy_ceiling_GBs": ceil and round(ceil["copy_GBs"], 1),
            "ceiling_shapes": ceil and {"read": ceil["read_shape"], "copy": ceil["copy_shape"]},
            "frac_of_read_ceiling": ceil and achieved / ceil["read_GBs"],
        },
        "cpu_baseline": cpu,
        "detail_file": None,
        "north_star": north_star(kern64, achieved, kern32, kern32s, err32, err32s, err32c, err32sc, big, precision,
                                 grid, batch, threept, serving),
    }
    if args.detail:
        try:
            os.makedirs(os.path.dirname(os.path.abspath(args.detail)), exist_ok=True)
            with open(args.detail, "w") as f:
                json.dump(dict(result, detail=detail), f, indent=1)
            result["detail_file"] = args.detail
        except OSError as e:
            print(f"bench: could not write {args.detail}: {e}", file=sys.stderr)
    print(json.dumps(result, separators=(",", ":")))


def _r(x, n=4):
    """x to n significant digits (None stays None)."""
    return None if x is None else float(f"{x:.{n}g}")


def north_star(kern64, achieved, kern32, kern32s, err32, err32s, err32c, err32sc, big, precision, grid, batch,
               threept, serving):
    """The figures the targets are stated in, compact enough to survive the
    tail of a log: FP64 kernel time and HBM fraction at 24^3x64 and 32^3x64,
    the FP32 forms' kernel speedups and errors, the precision study's
    errors, the meson grid's HBM fraction, and serving latency per rate."""
    ns = {"fp64_24c64": {"kernel_ms": _r(kern64), "hbm_frac": _r(achieved / HBM_PEAK_GBS)},
          "fp32_24c64": {"scaled_kernel_speedup": _r(kern64 / kern32s), "plain_kernel_speedup": _r(kern64 / kern32),
                         "max_err_per_t": _r(max(err32, err32s), 2),
                         "max_err_per_channel": _r(max(err32c["max"], err32sc["max"]), 2)}}
    if big:
        ns["fp64_32c64"] = {"kernel_ms": _r(big["fp64"]["kernel_ms"]), "hbm_frac": _r(big["fp64"]["hbm_frac"]),
                            "target_hbm_frac": 0.70}
        ns["fp32_32c64"] = {"scaled_kernel_speedup": _r(big["fp32_scaled_kernel_speedup"]),
                            "plain_kernel_speedup": _r(big["fp32_kernel_speedup"]), "target_speedup": 1.8,
                            "max_err_per_t": _r(max(big["fp32"]["per_timeslice_rel_err_vs_fp64"],
                                                    big["fp32_scaled"]["per_timeslice_rel_err_vs_fp64"]), 2),
                            "max_err_per_channel": _r(max(big["fp32"]["per_channel_rel_err_vs_fp64"]["max"],
                                                          big["fp32_scaled"]["per_channel_rel_err_vs_fp64"]["max"]),
                                                      2),
                            "shared_u_fp64_kernel_ms": _r(big["fp64"]["shared_u"]["kernel_ms"])}
    if precision:
        ns["precision_study"] = {"max_err_per_t": _r(precision["max_per_timeslice_rel_err"], 2),
                                 "scaled_max_err_per_t": _r(precision["fp32_scaled"]["max_per_timeslice_rel_err"], 2),
                                 "max_err_per_channel": _r(precision["per_channel_rel_err"]["max"], 2),
                                 "fp64_vs_oracle": _r(precision["fp64_gpu_vs_oracle_max_rel_err"], 2),
                                 "decades": _r(precision["dynamic_range_decades"], 3),
                                 "scaled_kernel_speedup": _r(precision["fp32_scaled"]["kernel_speedup"]),
                                 "fp32_chain_err_per_t": _r(precision["sequential_chain"]["max_per_timeslice_rel_err"],
                                                            2),
                                 "fp32_sequential_speedup": _r(precision["sequential_chain"]["sequential_fp32_speedup"])}
    if grid and "fp64" in grid:
        ns["meson_grid"] = {"fp64_hbm_frac": _r(grid["fp64"]["hbm_frac"]),
                            "fp64_kernel_ms": _r(grid["fp64"]["kernel_ms"])}
        if "fp64_mom27" in grid:
            ns["meson_grid"]["fp64_mom27_kernel_ms"] = _r(grid["fp64_mom27"]["kernel_ms"])
        for k in ("fp32_scaled", "fp32"):
            if k in grid:
                ns["meson_grid"][k + "_speedup"] = _r(grid[k].get("speedup_vs_fp64"))
    if batch:
        ns["baryon_batch16"] = {"kernel_GBs": _r(batch["batched_kernel_GBs"] if "batched_kernel_GBs" in batch
                                                 else batch["kernel_GBs"]),
                                "speedup_vs_singles": _r(batch["speedup"])}
    if threept:
        ns["threepoint"] = threept.get("summary")
    if serving:
        def rows(rates):
            return {r: [_r(v.get("arrival_rps") or float(r)), _r(v["throughput_rps"]), _r(v["p50_ms"], 3),
                        _r(v["p99_ms"], 3), _r(v["meson_batch_mean"], 3), _r(v["baryon_batch_mean"], 3)]
                    for r, v in rates.items()}

        ns["serving"] = rows(serving["rates"])
        ns["serving_columns"] = ("nominal rate: [realised arrivals req/s, served req/s, p50 ms, p99 ms, "
                                 "meson batch mean, baryon batch mean]")
        ns["serving_saturation"] = serving.get("saturation")
        sc = serving.get("fp32_scaled")
        if sc:
            ns["serving_fp32_scaled"] = rows(sc["rates"])
            ns["serving_fp32_scaled_saturation"] = sc.get("saturation")
    return ns


if __name__ == "__main__":
    main()
