"""Benchmark of the micro-image PatchMatch depth path on one MI355X.

A step = one full frame of the hot path on inputs already resident in HBM: random init,
6 red-black propagation+refinement iterations, final disparity record, fused CheckModule pass
and a 32-slice refocus, on the 4080x3068 synthetic frame (pitch 23, 27,258 MIs, 18 neighbours,
15x15 window).  `--mode plain` mirrors the reference's launch pattern (one launch per phase,
stream sync after each).  Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "plenoptic-for-depth_amd"))

METRIC = "depth frames/s + p99 ms at 4080x3068 raw, batched vs plain GPU, <=0.05px err"
PEAK_FP32 = 157.3e12      # MI355X published FP32 vector peak (FLOP/s)
PEAK_HBM = 8.0e12         # MI355X HBM3E (B/s)
FLOP_PER_SAMPLE = 22      # per evaluated (pixel, window offset, view, hypothesis) sample
FLOP_PER_TRIPLE = 16      # per evaluated (pixel, view, hypothesis) finalize


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", default="plain", choices=["plain"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-count", action="store_true", help="skip the work-counting pass")
    ap.add_argument("--raw", default="4080x3068")
    ap.add_argument("--calib-copy", action="store_true",
                    help="after the steps, run one 1 GiB device-to-device copy (the known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE/WRITE_SIZE counters)")
    return ap.parse_args()


class Timer:
    """HIP events on the stream the kernels are launched on (torch's current stream)."""

    def __init__(self, torch):
        self.torch = torch
        self.ev = {}

    def span(self, name):
        t = self.torch
        s, e = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        self.ev.setdefault(name, []).append((s, e))
        return s, e

    def ms(self, name):
        return [s.elapsed_time(e) for s, e in self.ev.get(name, [])]


def run_step(lfpm, df, bufs, fo, rf, timer, torch, plain=True):
    L = lfpm.lib()
    wl = df.wl
    pm = wl.pm
    V = wl.V
    n = V * wl.crop ** 2
    st = lfpm._stream()
    ws = bufs.ws
    state_bytes = n * 24
    pb = ws[:n * 16].view(torch.float32)
    cb = ws[n * 16:n * 20].view(torch.float32)
    sb = ws[n * 20:n * 24].view(torch.int32)
    scratch_off = (state_bytes + 255) // 256 * 256
    scr = ws[scratch_off:]
    scr_bytes = C.c_size_t(bufs.ws_bytes - scratch_off)
    dp = lfpm._dp

    def sync():
        if plain:
            torch.cuda.current_stream().synchronize()

    s, e = timer.span("init")
    s.record()
    lfpm._check(L.lfpm_patchmatch_init(dp(df.tiles), dp(df.centers), dp(df.nbr), dp(df.cnt), V,
                                       C.byref(pm), dp(bufs.plane), dp(bufs.cost), dp(bufs.sel),
                                       dp(scr), scr_bytes, st))
    e.record()
    sync()
    for it in range(pm.iterations):
        for ph in (0, 1):
            src = (bufs.plane, bufs.cost, bufs.sel) if ph == 0 else (pb, cb, sb)
            dst = (pb, cb, sb) if ph == 0 else (bufs.plane, bufs.cost, bufs.sel)
            s, e = timer.span("phase")
            s.record()
            lfpm._check(L.lfpm_patchmatch_phase(
                dp(df.tiles), dp(df.centers), dp(df.nbr), dp(df.cnt), V, C.byref(pm), it, ph,
                dp(src[0]), dp(src[1]), dp(src[2]), dp(dst[0]), dp(dst[1]), dp(dst[2]), dp(scr),
                scr_bytes, st))
            e.record()
            sync()
    s, e = timer.span("finalize")
    s.record()
    lfpm._check(L.lfpm_patchmatch_finalize(dp(df.tiles), dp(df.centers), dp(df.nbr), dp(df.cnt),
                                           V, C.byref(pm), dp(bufs.plane), dp(bufs.disp), st))
    e.record()
    sync()
    s, e = timer.span("filter")
    s.record()
    lfpm._check(L.lfpm_filter_frame(
        dp(df.gray), dp(df.centers), dp(df.nbr), dp(df.cnt), V, wl.crop, pm.max_neighbors,
        C.c_float(pm.base), C.byref(wl.fp), dp(bufs.plane), dp(bufs.cost), dp(bufs.disp),
        dp(bufs.sel), dp(fo["plane"]), dp(fo["cost"]), dp(fo["disp"]), dp(fo["sel"]),
        dp(fo["reason"]), st))
    e.record()
    sync()
    s, e = timer.span("refocus")
    s.record()
    lfpm.refocus(df, out=rf)
    e.record()
    torch.cuda.current_stream().synchronize()


def count_work(lfpm, df, bufs, fo, rf, torch):
    """Untimed pass with counting kernels: evaluated samples / triples per kernel kind."""
    L = lfpm.lib()
    cnt = torch.zeros(2, dtype=torch.int64, device="cuda")
    out = {}
    timer = Timer(torch)
    # split the per-kind counts by running the pieces with the counters attached
    L.lfpm_set_work_counters(C.c_void_p(cnt.data_ptr()))
    try:
        run_step(lfpm, df, bufs, fo, rf, timer, torch)
    finally:
        L.lfpm_set_work_counters(None)
    torch.cuda.synchronize()
    tot = cnt.cpu().numpy().astype(np.int64)
    out["all"] = dict(samples=int(tot[0]), triples=int(tot[1]))
    # phase-only counts: rerun phases alone from the same init state
    cnt.zero_()
    pm = df.wl.pm
    V = df.wl.V
    n = V * df.wl.crop ** 2
    ws = bufs.ws
    st = lfpm._stream()
    dp = lfpm._dp
    pb = ws[:n * 16].view(torch.float32)
    cb = ws[n * 16:n * 20].view(torch.float32)
    sb = ws[n * 20:n * 24].view(torch.int32)
    off = (n * 24 + 255) // 256 * 256
    scr, scr_bytes = ws[off:], C.c_size_t(bufs.ws_bytes - off)
    lfpm._check(L.lfpm_patchmatch_init(dp(df.tiles), dp(df.centers), dp(df.nbr), dp(df.cnt), V,
                                       C.byref(pm), dp(bufs.plane), dp(bufs.cost), dp(bufs.sel),
                                       dp(scr), scr_bytes, st))
    L.lfpm_set_work_counters(C.c_void_p(cnt.data_ptr()))
    try:
        for it in range(pm.iterations):
            for ph in (0, 1):
                src = (bufs.plane, bufs.cost, bufs.sel) if ph == 0 else (pb, cb, sb)
                dst = (pb, cb, sb) if ph == 0 else (bufs.plane, bufs.cost, bufs.sel)
                lfpm._check(L.lfpm_patchmatch_phase(
                    dp(df.tiles), dp(df.centers), dp(df.nbr), dp(df.cnt), V, C.byref(pm), it, ph,
                    dp(src[0]), dp(src[1]), dp(src[2]), dp(dst[0]), dp(dst[1]), dp(dst[2]),
                    dp(scr), scr_bytes, st))
    finally:
        L.lfpm_set_work_counters(None)
    torch.cuda.synchronize()
    ph = cnt.cpu().numpy().astype(np.int64)
    out["phase"] = dict(samples=int(ph[0]), triples=int(ph[1]), launches=2 * pm.iterations)
    return out


def measured_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (scaled by the copy-kernel calibration), or None if absent."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_plain_4080x3068.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0].startswith("lfpm::" + kernel):
            try:
                return float(cells[4])
            except ValueError:
                return None
    return None


def cpu_baseline(lfpm):
    """FP32 CPU restatement (oracle, OpenMP) of the same solve on a bounded sample: the same
    parameters on a 400x320 sub-frame, scaled to a 4080x3068 frame by MI count."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    wl = lfpm.workload_c2(raw_w=400, raw_h=320)
    tiles = oracle.slice_u8(wl.raw(), wl.L["origins"], wl.crop)
    pm = oracle.PmParams(*[getattr(wl.pm, f[0]) for f in oracle.PmParams._fields_])
    threads = max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    oracle.solve(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], pm, dtype=np.float32,
                 threads=threads)
    dt = time.perf_counter() - t0
    full = lfpm.workload_c2()
    s_per_frame = dt * full.V / wl.V
    return {"value": 1.0 / s_per_frame, "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": "FP32 oracle solve (init + 6 iterations, 18 neighbours, 15x15) of a 400x320 "
                      "sub-frame (%d MIs) in %.1f s, scaled by MI count to the %d-MI frame"
                      % (wl.V, dt, full.V)}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU; --gpus %d is not supported "
              "(multi-GPU sharding is a design note only)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    import lfpm
    rw, rh = (int(x) for x in args.raw.split("x"))
    log("bench: synthesising %dx%d frame" % (rw, rh))
    wl = lfpm.workload_c2(raw_w=rw, raw_h=rh)
    df = lfpm.DeviceFrame(wl)
    bufs = lfpm.SolveBuffers(wl.V, wl.crop, wl.pm)
    n = wl.V * wl.crop ** 2
    fo = dict(plane=torch.empty((n, 4), device="cuda"), cost=torch.empty(n, device="cuda"),
              disp=torch.empty((n, 4), device="cuda"),
              sel=torch.empty(n, dtype=torch.int32, device="cuda"),
              reason=torch.empty(n, dtype=torch.uint8, device="cuda"))
    rf = torch.empty((wl.slices, rh, rw), device="cuda")
    torch.cuda.synchronize()
    log("bench: V=%d crop=%d K=%d iters=%d" % (wl.V, wl.crop, wl.pm.max_neighbors,
                                                wl.pm.iterations))
    for i in range(args.warmup):
        run_step(lfpm, df, bufs, fo, rf, Timer(torch), torch)
        log("bench: warmup %d done" % i)
    timer = Timer(torch)
    lat = []
    torch.cuda.synchronize()
    t_all0 = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        run_step(lfpm, df, bufs, fo, rf, timer, torch)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t0) * 1e3)
        log("bench: step %d %.1f ms" % (i, lat[-1]))
    torch.cuda.synchronize()
    t_all = time.perf_counter() - t_all0
    if args.calib_copy:
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    fps = args.steps / t_all
    kt = {k: float(np.mean(timer.ms(k))) for k in ("init", "phase", "finalize", "filter",
                                                    "refocus")}
    roof = None
    work = None
    if not args.no_count:
        log("bench: counting pass")
        work = count_work(lfpm, df, bufs, fo, rf, torch)
        ph = work["phase"]
        flop_per_launch = (ph["samples"] * FLOP_PER_SAMPLE + ph["triples"] * FLOP_PER_TRIPLE) / \
            ph["launches"]
        achieved = flop_per_launch / (kt["phase"] * 1e-3) / 1e12
        roof = {"bound": "valu", "kernel": "k_phase", "achieved": achieved,
                "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s", "frac": achieved / (PEAK_FP32 / 1e12),
                "traffic": measured_traffic("k_phase"), "avg_launch_ms": kt["phase"],
                "flop_per_launch": flop_per_launch}
    cpu = None
    if not args.no_cpu_baseline:
        log("bench: cpu baseline")
        cpu = cpu_baseline(lfpm)
    line = {
        "metric": METRIC, "value": fps, "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t_all * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "plain GPU single-frame solve (init + %d red-black iterations + "
                               "disp record) + CheckModule filter + %d-slice refocus"
                               % (wl.pm.iterations, wl.slices),
                   "raw": "%dx%d" % (rw, rh), "pitch": 23.0, "mis": wl.V,
                   "neighbors": wl.pm.max_neighbors, "window": 2 * wl.pm.radius + 1,
                   "iterations": wl.pm.iterations, "slices": wl.slices, "mode": args.mode},
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "kernel_ms": kt, "work": work, "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
