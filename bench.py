"""Benchmark of the micro-image PatchMatch depth path on one MI355X.

The hot path per frame, on inputs already resident in HBM: random init, 6 red-black
propagation+refinement iterations, final disparity record, fused CheckModule pass and a 32-slice
refocus, on the 4080x3068 synthetic frame (pitch 23, 27,258 MIs, 18 neighbours, 15x15 window).
Modes:
  plain    one frame per step, one launch per phase with a stream sync after each (the
           reference's launch pattern); value = frames/s
  batched  (default) the plain measurement, then --batch frames of the same size per step in one
           captured hipGraph; value = batched frames/s, with the plain figures beside it and a
           bit-for-bit check that the batched frame 0 equals its plain solve
  mixed    --batch uneven frames (sizes, pitches and neighbour counts mixed) in one graph
  serve    Poisson request stream through the continuous-batching server; value = completed
           frames/s, p50/p99 = request latency
Prints one JSON line.
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
    ap.add_argument("--mode", default="batched",
                    choices=["plain", "batched", "mixed", "serve", "precision", "parity",
                             "ingest", "products", "composites", "filter"])
    ap.add_argument("--hypotheses", type=int, default=1000000, help="precision: hypotheses")
    ap.add_argument("--batch", type=int, default=4, help="frames per batch (batched / mixed)")
    ap.add_argument("--requests", type=int, default=24, help="serve: number of requests")
    ap.add_argument("--rate", type=float, default=40.0, help="serve: Poisson arrival rate (1/s)")
    ap.add_argument("--hbm-budget-gb", type=float, default=200.0, help="serve: admission budget")
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
    lfpm._check(L.lfpm_patchmatch_finalize_ws(dp(df.tiles), dp(df.centers), dp(df.nbr),
                                              dp(df.cnt), V, C.byref(pm), dp(bufs.plane),
                                              dp(bufs.disp), dp(scr), scr_bytes, st))
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
    cnt = torch.zeros(4, dtype=torch.int64, device="cuda")
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
    out["all"] = dict(samples=int(tot[0]), triples=int(tot[1]), samples_skipped=int(tot[2]),
                      triples_skipped=int(tot[3]))
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
    out["phase"] = dict(samples=int(ph[0]), triples=int(ph[1]), samples_skipped=int(ph[2]),
                        triples_skipped=int(ph[3]), launches=2 * pm.iterations)
    return out


def measured_traffic(row):
    """HBM bytes per dispatch of the k_stage summary row of stage `row` ("init", "phase", "final")
    from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (scaled by the copy-kernel calibration), or None if
    absent."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_plain_4080x3068.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0].startswith("lfpm::") and "k_stage" in cells[0] and \
                cells[0].endswith(" " + row):
            try:
                return float(cells[4])
            except ValueError:
                return None
    return None


def filter_traffic():
    """HBM bytes per lfpm_filter_frame dispatch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (calibrated), or None if absent."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_plain_4080x3068.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0] == "lfpm::k_filter<false>":
            try:
                return float(cells[4])
            except ValueError:
                return None
    return None


def cpu_model():
    """The host CPU's model name (first processor in /proc/cpuinfo), or None."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(lfpm):
    """FP32 CPU restatement (oracle, OpenMP) of the same solve on a bounded sample: the same
    parameters on an 800x640 sub-frame (1088 MIs, about 10 s on 16 cores), scaled to a 4080x3068
    frame by MI count."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    wl = lfpm.workload_c2(raw_w=800, raw_h=640)
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
            "cpu": cpu_model(),
            "sample": "FP32 oracle solve (init + 6 iterations, 18 neighbours, 15x15) of an 800x640 "
                      "sub-frame (%d MIs) in %.1f s, scaled by MI count to the %d-MI frame"
                      % (wl.V, dt, full.V)}


def plain_section(args, lfpm, torch, wl):
    """One frame at a time, one launch per phase with a stream sync after each."""
    rw, rh = wl.sp.raw_w, wl.sp.raw_h
    df = lfpm.DeviceFrame(wl)
    bufs = lfpm.SolveBuffers(wl.V, wl.crop, wl.pm)
    n = wl.V * wl.crop ** 2
    fo = dict(plane=torch.empty((n, 4), device="cuda"), cost=torch.empty(n, device="cuda"),
              disp=torch.empty((n, 4), device="cuda"),
              sel=torch.empty(n, dtype=torch.int32, device="cuda"),
              reason=torch.empty(n, dtype=torch.uint8, device="cuda"))
    rf = torch.empty((wl.slices, rh, rw), device="cuda")
    torch.cuda.synchronize()
    for i in range(args.warmup):
        run_step(lfpm, df, bufs, fo, rf, Timer(torch), torch)
        log("bench: plain warmup %d done" % i)
    timer = Timer(torch)
    lat = []
    torch.cuda.synchronize()
    t_all0 = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        run_step(lfpm, df, bufs, fo, rf, timer, torch)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t0) * 1e3)
        log("bench: plain step %d %.1f ms" % (i, lat[-1]))
    t_all = time.perf_counter() - t_all0
    kt = {k: float(np.mean(timer.ms(k))) for k in ("init", "phase", "finalize", "filter",
                                                    "refocus")}
    out = dict(fps=args.steps / t_all, ms_per_step=t_all * 1e3 / args.steps, lat=lat, kt=kt,
               result=dict(plane=bufs.plane.clone(), cost=bufs.cost.clone(),
                           reason=fo["reason"].clone(), refocus=rf[-1].clone()))
    if not args.no_count:
        log("bench: counting pass")
        work = count_work(lfpm, df, bufs, fo, rf, torch)
        ph = work["phase"]
        flop_per_launch = (ph["samples"] * FLOP_PER_SAMPLE + ph["triples"] * FLOP_PER_TRIPLE) / \
            ph["launches"]
        achieved = flop_per_launch / (kt["phase"] * 1e-3) / 1e12
        # the same launches credited with the samples a single-round scoring evaluates too
        ref_flop = flop_per_launch + (ph["samples_skipped"] * FLOP_PER_SAMPLE +
                                      ph["triples_skipped"] * FLOP_PER_TRIPLE) / ph["launches"]
        ref_achieved = ref_flop / (kt["phase"] * 1e-3) / 1e12
        out["work"] = work
        out["roofline"] = {
            "bound": "valu", "pipe": "FP32 VALU (compute-bound; the NCC is not a contraction, "
                                     "so no MFMA instructions are issued; peak = MI355X FP32 "
                                     "vector)",
            "kernel": "k_stage_w (red/black phase)", "achieved": achieved,
            "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s", "frac": achieved / (PEAK_FP32 / 1e12),
            "traffic": measured_traffic("phase"),
            "traffic_source": "profiles/hbm_traffic_plain_4080x3068.md (rocprofv3 FETCH_SIZE / "
                              "WRITE_SIZE, committed summary)",
            "avg_launch_ms": kt["phase"],
            "flop_per_launch": flop_per_launch,
            "per_unit": "22 FLOP per evaluated window sample + 16 per evaluated "
                        "(pixel, view, hypothesis) triple",
            "single_round_work": {
                "flop_per_launch": ref_flop, "achieved": ref_achieved,
                "frac": ref_achieved / (PEAK_FP32 / 1e12),
                "note": "FLOP of every sample the reference's one-round scoring evaluates; the "
                        "two-round scoring leaves out the views of candidates whose cost over "
                        "the first half of the selection already reaches the best (same "
                        "results), so these launches do that work in the time above"}}
    del df, bufs, fo, rf
    return out


def batched_section(args, lfpm, torch, wls):
    """All frames in one captured hipGraph (class branches, persistent k_stage launches)."""
    slots = [lfpm.FrameSlot(wl) for wl in wls]
    batch = lfpm.Batch(slots)
    batch.capture()
    torch.cuda.synchronize()
    for i in range(args.warmup):
        batch.replay()
        torch.cuda.synchronize()
        log("bench: batched warmup %d done" % i)
    lat = []
    t_all0 = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        batch.replay()
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t0) * 1e3)
        log("bench: batched step %d (%d frames) %.1f ms" % (i, len(slots), lat[-1]))
    t_all = time.perf_counter() - t_all0
    s0 = slots[0]
    res = dict(plane=s0.plane, cost=s0.cost, reason=s0.reason, refocus=s0.refocus[-1])
    batch.close()
    return dict(fps=len(slots) * args.steps / t_all, ms_per_step=t_all * 1e3 / args.steps,
                lat=lat, frames=len(slots), result=res, mis=sum(w.V for w in wls))


def plain_frames_section(args, lfpm, torch, wls):
    """The same frames solved one at a time by the plain path (launch per phase, sync after
    each, then filter and refocus), all frames' buffers resident: frames/s over args.steps
    passes."""
    frames = []
    for wl in wls:
        rw, rh = wl.sp.raw_w, wl.sp.raw_h
        df = lfpm.DeviceFrame(wl)
        bufs = lfpm.SolveBuffers(wl.V, wl.crop, wl.pm)
        n = wl.V * wl.crop ** 2
        fo = dict(plane=torch.empty((n, 4), device="cuda"), cost=torch.empty(n, device="cuda"),
                  disp=torch.empty((n, 4), device="cuda"),
                  sel=torch.empty(n, dtype=torch.int32, device="cuda"),
                  reason=torch.empty(n, dtype=torch.uint8, device="cuda"))
        rf = torch.empty((wl.slices, rh, rw), device="cuda")
        frames.append((df, bufs, fo, rf))
    for i in range(args.warmup):
        for f in frames:
            run_step(lfpm, *f, Timer(torch), torch)
        log("bench: plain-frames warmup %d done" % i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        for f in frames:
            run_step(lfpm, *f, Timer(torch), torch)
        torch.cuda.synchronize()
        log("bench: plain-frames pass %d done" % i)
    t = time.perf_counter() - t0
    del frames
    return dict(fps=len(wls) * args.steps / t, ms_per_pass=1e3 * t / args.steps)


def same_results(a, b, torch):
    return all(torch.equal(a[k], b[k]) for k in a)


def mixed_workloads(lfpm, n, seed=11):
    """The uneven batch of lfpm.mixed_workloads."""
    return lfpm.mixed_workloads(n, seed)


def serve_section(args, lfpm, torch):
    import serving
    sizes = [(1024, 800), (2048, 1536), (4080, 3068)]
    radius = {15.0: 5, 23.0: 7, 31.0: 9}

    def make(w, h, p, planes, s):
        wl = lfpm.workload_c2(raw_w=w, raw_h=h, pitch=p, radius=radius[p], planes=planes,
                              scene_seed=s, noise_seed=s + 1, seed=s, frame_uid=s)
        wl.raw()
        return wl
    def one_run(rate, seed):
        log("bench: synthesising %d requests at %.2f req/s" % (args.requests, rate))
        reqs = serving.poisson_requests(args.requests, rate, seed, sizes, [15.0, 23.0, 31.0],
                                        make)
        srv = serving.Server(hbm_budget_bytes=int(args.hbm_budget_gb * 2 ** 30), max_slots=16,
                             keep_results=False)
        try:
            t0 = time.perf_counter()
            done = srv.run(reqs, log=log)
            span = time.perf_counter() - t0
            ss = srv.stats()
        finally:
            srv.close()
        st = serving.latency_stats(done)
        return dict(rate_offered=rate, fps=len(done) / span, span_s=span, requests=len(done),
                    steps=ss["steps"], launches=ss["launches"], max_active=ss["max_active"],
                    classes=ss["classes"],
                    mean_queue_ms=float(np.mean([r.queue_ms for r in done])), **st)

    r = one_run(args.rate, 7)
    # the same stream shape at an offered rate the server sustains (0.7 of the rate it served
    # the first stream at), where latency is service time rather than queueing
    r["sustainable"] = one_run(0.7 * r["fps"], 8)
    return r


def parity_section(args, lfpm, torch, rw, rh):
    """The whole frame at the full-frame parameters solved on the GPU and by the FP64 oracle on
    the host: the fraction of valid pixels (oracle cost <= 1.8) whose disparity is within
    0.05 px, the bar the solve is held to; the oracle prints a line per stage."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    wl = lfpm.workload_c2(raw_w=rw, raw_h=rh)
    log("bench: parity %dx%d, V=%d, K=%d, window %d, %d iterations"
        % (rw, rh, wl.V, wl.pm.max_neighbors, 2 * wl.pm.radius + 1, wl.pm.iterations))
    df = lfpm.DeviceFrame(wl)
    t0 = time.perf_counter()
    bufs = lfpm.patchmatch_frame(df)
    torch.cuda.synchronize()
    gpu_s = time.perf_counter() - t0
    d_g = bufs.plane.cpu().numpy()[:, 3].astype(np.float64)
    sel_g = bufs.sel.cpu().numpy().view(np.uint32)
    log("bench: GPU solve %.2f s" % gpu_s)
    tiles = oracle.slice_u8(wl.raw(), wl.L["origins"], wl.crop)
    pm = oracle.PmParams(*[getattr(wl.pm, f[0]) for f in oracle.PmParams._fields_])
    threads = max(1, min(16, os.cpu_count() or 1))
    oracle.set_progress(True)
    t0 = time.perf_counter()
    o = oracle.solve(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], pm, threads=threads)
    orc_s = time.perf_counter() - t0
    oracle.set_progress(False)
    log("bench: FP64 oracle solve %.1f s on %d threads" % (orc_s, threads))
    valid = o["cost"] <= 1.8
    err = np.abs(d_g - o["plane"][:, 3])
    return {"pixels": int(err.size), "valid_frac": float(valid.mean()),
            "frac_le_005": float((err[valid] <= 0.05).mean()),
            "frac_all_le_005": float((err <= 0.05).mean()),
            "max_err_px": float(err[valid].max()),
            "sel_equal": float((sel_g == o["sel"]).mean()),
            "gpu_s": gpu_s, "oracle_s": orc_s, "oracle_threads": threads, "cpu": cpu_model()}


def precision_section(args, lfpm, torch):
    """FP32 vs FP16 LDS storage against the FP64 oracle: 10^6 seeded hypotheses (normals within
    60 degrees of the optical axis, virtual depth in [2, 8]) on the 4080x3068 frame, and the
    full-solve disparity agreement on a 1024x800 frame with the same solve parameters."""
    import copy
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    wl = lfpm.workload_c2()
    hyp = lfpm.cone_hypotheses(wl, args.hypotheses, seed=2024)
    tiles = oracle.slice_u8(wl.raw(), wl.L["origins"], wl.crop)
    opm = oracle.PmParams(*[getattr(wl.pm, f[0]) for f in oracle.PmParams._fields_])
    log("bench: oracle FP64 costs for %d hypotheses" % len(hyp))
    t0 = time.perf_counter()
    ref = oracle.eval_costs(tiles, wl.L["centers"], wl.L["nbr"], wl.L["cnt"], opm, hyp,
                            threads=threads)
    t_orc = time.perf_counter() - t0
    df = lfpm.DeviceFrame(wl)
    hyp_d = torch.from_numpy(hyp).cuda()
    out = {"hypotheses": len(hyp), "oracle_s": t_orc, "oracle_threads": threads}
    for name, storage in (("fp32", 0), ("fp16", 1)):
        pm = copy.copy(wl.pm)
        pm.storage = storage
        lfpm.eval_costs(df, hyp_d, pm=pm)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g = lfpm.eval_costs(df, hyp_d, pm=pm)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        g = g.cpu().numpy().astype(np.float64)
        both = (g < 2) & (ref < 2)
        e = np.abs(g[both] - ref[both]) / np.maximum(np.abs(ref[both]), 1e-3)
        out[name] = {"views_scored": int(both.sum()),
                     "flip_frac": float(((g < 2) != (ref < 2)).mean()),
                     "rel_err_median": float(np.median(e)), "rel_err_p99": float(np.quantile(e, 0.99)),
                     "rel_err_p999": float(np.quantile(e, 0.999)), "rel_err_max": float(e.max()),
                     "frac_le_1e5": float((e <= 1e-5).mean()), "frac_le_2e3": float((e <= 2e-3).mean()),
                     "gpu_ms": dt * 1e3, "hyp_views_per_s": both.size / dt}
        log("bench: %s %s" % (name, out[name]))
    # full-solve disparity agreement on a 1024x800 frame with the 4080x3068 solve parameters
    sub = lfpm.workload_c2(raw_w=1024, raw_h=800)
    stiles = oracle.slice_u8(sub.raw(), sub.L["origins"], sub.crop)
    log("bench: oracle FP64 solve of the 1024x800 frame (%d MIs)" % sub.V)
    t0 = time.perf_counter()
    o = oracle.solve(stiles, sub.L["centers"], sub.L["nbr"], sub.L["cnt"],
                     oracle.PmParams(*[getattr(sub.pm, f[0]) for f in oracle.PmParams._fields_]),
                     threads=threads)
    out["oracle_solve_s"] = time.perf_counter() - t0
    sdf = lfpm.DeviceFrame(sub)
    valid = o["cost"] <= 1.8
    for name, storage in (("fp32", 0), ("fp16", 1)):
        pm = copy.copy(sub.pm)
        pm.storage = storage
        bufs = lfpm.patchmatch_frame(sdf, pm=pm)
        torch.cuda.synchronize()
        d = bufs.plane[:, 3].cpu().numpy().astype(np.float64)
        err = np.abs(d - o["plane"][:, 3])
        out[name]["solve_frac_le_005px"] = float((err[valid] <= 0.05).mean())
        out[name]["solve_median_err_px"] = float(np.median(err[valid]))
    return out


def point_cloud_traffic():
    """HBM bytes per lfpm_vd_point_cloud call from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (flag, scan and emit kernels, calibrated), or None if absent."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_products_4080x3068.md")
    if not os.path.exists(path):
        return None
    tot = 0.0
    for row in open(path):
        cells = [c.strip() for c in row.strip().strip("|").split("|")]
        if len(cells) == 5 and ("k_pt_" in cells[0] or "k_scan_" in cells[0]):
            try:
                tot += float(cells[4])
            except ValueError:
                return None
    return tot or None


def products_section(args, lfpm, torch, wl):
    """Per-frame products of one solved frame: disparity mosaic, VD map with its gray colour
    image, and the VD point cloud (PLY records).  The frame is solved once; each step runs the
    three products on the launch stream, timed with HIP events per product (no host sync
    inside a step)."""
    import ctypes as C
    df = lfpm.DeviceFrame(wl)
    bufs = lfpm.patchmatch_frame(df)
    torch.cuda.synchronize()
    h, w = df.raw.shape
    N = h * w
    ws = lfpm.products_workspace(df)
    vd = torch.empty((h, w), dtype=torch.float32, device="cuda")
    col = torch.empty((h, w), dtype=torch.uint8, device="cuda")
    mos = torch.empty((h, w), dtype=torch.float32, device="cuda")
    rng = torch.empty(2, dtype=torch.float32, device="cuda")
    recs = torch.empty(N * 27, dtype=torch.uint8, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    L, st, dp = lfpm.lib(), lfpm._stream(), lfpm._dp
    V, W = wl.V, wl.crop
    prm = (C.c_float(W * 0.5), C.c_float(W * 0.5), C.c_float(0.0055), C.c_float(35.0),
           C.c_float(0.8), C.c_float(2.5))

    def ev():
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        return e

    def step():
        e0 = ev()
        lfpm._check(L.lfpm_disparity_mosaic(dp(bufs.plane), dp(bufs.cost), dp(df.centers), V, W,
                                            C.c_float(0.2), w, h, dp(mos), dp(rng), dp(ws),
                                            C.c_size_t(ws.numel()), st))
        e1 = ev()
        lfpm._check(L.lfpm_virtual_depth_map(dp(bufs.plane), dp(df.centers), dp(df.gray), V, W,
                                             C.c_float(wl.L["pitch"]), w, h, dp(vd), dp(col),
                                             dp(rng), dp(ws), C.c_size_t(ws.numel()), st))
        e2 = ev()
        lfpm._check(L.lfpm_vd_point_cloud(dp(vd), dp(col), w, h, *prm, dp(recs),
                                          C.c_size_t(recs.numel()), dp(cnt), dp(ws),
                                          C.c_size_t(ws.numel()), st))
        e3 = ev()
        return e0, e1, e2, e3

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    marks = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        marks.append(step())
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    ms = {k: float(np.mean([m[i].elapsed_time(m[i + 1]) for m in marks]))
          for i, k in enumerate(("mosaic", "vd_map", "point_cloud"))}
    n = int(cnt.item())
    pc_bytes = 5.0 * N + 27.0 * n  # vd + colour read, records written
    gbs = pc_bytes / (ms["point_cloud"] * 1e-3) / 1e9
    # CPU baseline: the oracle's three products of the same solved frame
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    p = bufs.plane.cpu().numpy()
    c = bufs.cost.cpu().numpy()
    g = df.gray.cpu().numpy().reshape(-1)
    t1 = time.perf_counter()
    oracle.disparity_mosaic(p, c, wl.L["centers"], W, 0.2, w, h)
    ovd, ocol, _ = oracle.vd_map(p, wl.L["centers"], g, W, wl.L["pitch"], w, h)
    orec = oracle.vd_point_cloud(ovd, ocol, (W * 0.5, W * 0.5), 0.0055, 35.0, 0.8, 2.5)
    cpu_s = time.perf_counter() - t1
    same = bool(np.array_equal(recs[:n * 27].cpu().numpy().reshape(-1, 27), orec))
    return {"fps": 1e3 / sum(ms.values()), "ms": ms, "wall_ms": 1e3 * wall, "points": n,
            "records_equal_oracle": same,
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s",
                         "frac": gbs / 8000.0, "traffic": point_cloud_traffic(),
                         "traffic_source": "profiles/hbm_traffic_products_4080x3068.md "
                                           "(committed rocprofv3 FETCH_SIZE / WRITE_SIZE "
                                           "summary, not measured in this run)",
                         "kernel": "lfpm_vd_point_cloud (flag, scan, emit)",
                         "avg_launch_ms": ms["point_cloud"],
                         "per_unit": "5 B per raw pixel read (vd, colour) + 27 B per point "
                                     "written"},
            "cpu_baseline": {"value": 1.0 / cpu_s, "unit": "frames/s", "cores": 1,
                             "kind": "port", "cpu": cpu_model(),
                             "sample": "oracle mosaic + VD map + point cloud of the same solved "
                                       "%dx%d frame, one frame in %.1f s" % (w, h, cpu_s)}}


def filter_section(args, lfpm, torch, wl):
    """The CheckModule pass over one solved frame in both orders: lfpm_filter_frame (snapshot,
    one launch) and lfpm_filter_frame_ordered (the reference's MI order, one launch per level of
    the schedule), timed with HIP events on the launch stream, with the share of pixels whose
    reason differs between the two.  CPU baseline: the oracle's FP32 twins of both orders on
    the same solved frame."""
    df = lfpm.DeviceFrame(wl)
    bufs = lfpm.patchmatch_frame(df)
    torch.cuda.synchronize()
    V, W = wl.V, wl.crop
    px = V * W * W
    levels = len(lfpm._df_levels(df)[1]) - 1

    def ev():
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        return e

    out = {}
    ms = {}
    for order in ("snapshot", "reference"):
        for _ in range(args.warmup):
            lfpm.filter_frame(df, bufs, order=order)
        torch.cuda.synchronize()
        marks = []
        for _ in range(args.steps):
            e0 = ev()
            out[order] = lfpm.filter_frame(df, bufs, order=order)
            marks.append((e0, ev()))
        torch.cuda.synchronize()
        ms[order] = float(np.mean([a.elapsed_time(b) for a, b in marks]))
    rs = out["snapshot"]["reason"].cpu().numpy()
    rr = out["reference"]["reason"].cpu().numpy()
    # compulsory bytes per pixel: plane 16 + cost 4 + disparity record 16 + selection 4 + gray 4
    # read, the same four outputs (40 B) and the reason byte written
    per_px = 44.0 + 41.0
    gbs = per_px * px / (ms["snapshot"] * 1e-3) / 1e9
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    host = dict(gray=df.gray.cpu().numpy(), plane=bufs.plane.cpu().numpy(),
                cost=bufs.cost.cpu().numpy(), disp=bufs.disp.cpu().numpy(),
                sel=bufs.sel.cpu().numpy().view(np.uint32))
    args_o = (host["gray"], wl.L["centers"], wl.L["nbr"], wl.L["cnt"], W, wl.pm.max_neighbors,
              wl.pm.base, oracle.FilterParams(*[getattr(wl.fp, f[0]) for f in
                                                oracle.FilterParams._fields_]),
              host["plane"], host["cost"],
              host["disp"], host["sel"])
    threads = max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    os_ = oracle.filter_f32(*args_o, threads=threads)
    snap_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    or_ = oracle.filter_f32_inplace(*args_o)
    ord_s = time.perf_counter() - t0
    same = bool(np.array_equal(os_["reason"], rs) and np.array_equal(or_["reason"], rr))
    return {"ms": ms, "levels": levels, "us_per_level": 1e3 * ms["reference"] / levels,
            "pixels": px, "kept_snapshot": float((rs == 0).mean()),
            "kept_reference": float((rr == 0).mean()), "reason_differs": float((rr != rs).mean()),
            "reasons_equal_oracle": same,
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s",
                         "frac": gbs / 8000.0, "traffic": filter_traffic(),
                         "traffic_source": "profiles/hbm_traffic_plain_4080x3068.md "
                                           "(committed rocprofv3 FETCH_SIZE / WRITE_SIZE "
                                           "summary, not measured in this run)",
                         "kernel": "k_filter<false> (lfpm_filter_frame)",
                         "avg_launch_ms": ms["snapshot"],
                         "per_unit": "85 B per MI pixel compulsory (44 B read, 41 B written)"},
            "cpu_baseline": {"value": 1.0 / ord_s, "unit": "frames/s", "cores": 1,
                             "kind": "port", "cpu": cpu_model(),
                             "sample": "oracle FP32 twin in the reference's order (one thread, "
                                       "%.2f s) on the same solved frame; the snapshot twin on "
                                       "%d threads took %.2f s" % (ord_s, threads, snap_s)}}


def composites_section(args, lfpm, torch, wl):
    """Refocus-side composites of one solved frame with a colour raw image: the all-in-focus
    composite with its depth maps, hole filling and denoising (lfpm_aif_composite) and the
    back-projection consistency fusion with its point records (lfpm_fuse_backproject).  The
    frame is solved once and the fusion's level schedule built once (host work per lens
    layout); each step runs both entry points on the launch stream, timed with HIP events."""
    import ctypes as C
    df = lfpm.DeviceFrame(wl)
    bufs = lfpm.patchmatch_frame(df)
    h, w = df.raw.shape
    V, W = wl.V, wl.crop
    raw = wl.raw().astype(np.int64)
    rng = np.random.default_rng(35)
    bgr = np.stack([raw, (raw * 7 + 11) % 256, rng.integers(0, 256, raw.shape)],
                   -1).astype(np.uint8)
    tiles = lfpm.slice_bgr_tiles(df, torch.from_numpy(bgr).cuda())
    rich = np.where(rng.random((V, W, W)) < 0.4, 255, 0).astype(np.uint8)
    blur = rng.integers(0, 256, (V, W, W)).astype(np.uint8)
    blur[rng.random((V, W, W)) < 0.25] = 255
    rich_d, blur_d = torch.from_numpy(rich).cuda(), torch.from_numpy(blur).cuda()
    K = wl.L["nbr"].size // V
    B, bl0, f = 0.9, 42.0, 30.0
    baseline_bp = wl.L["pitch"] * (bl0 + B) / bl0
    order, start, nl = lfpm.fuse_schedule(wl.L["nbr"].reshape(V, K), wl.L["cnt"])
    order_d = torch.from_numpy(order).cuda()
    L, st, dp = lfpm.lib(), lfpm._stream(), lfpm._dp
    na = L.lfpm_aif_workspace_size(V, W, w, h)
    nf = L.lfpm_fuse_workspace_size(V, W, w, h)
    ws_a = torch.empty(na, dtype=torch.uint8, device="cuda")
    ws_f = torch.empty(nf, dtype=torch.uint8, device="cuda")
    f32 = lambda *s: torch.empty(s, dtype=torch.float32, device="cuda")  # noqa: E731
    u8 = lambda *s: torch.empty(s, dtype=torch.uint8, device="cuda")  # noqa: E731
    a_out = [f32(h, w), f32(h, w), f32(h, w), u8(h, w, 3), u8(h, w, 3), u8(h, w, 3)]
    a_stats = torch.empty(8, dtype=torch.int32, device="cuda")
    f_vd, f_aif, f_conf = f32(h, w), u8(h, w, 3), u8(h, w)
    pts = f32(V * W * W, 9)
    counts = torch.zeros(2, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()

    def ev():
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        return e

    def step():
        e0 = ev()
        lfpm._check(L.lfpm_aif_composite(
            dp(bufs.plane), dp(df.centers), dp(tiles), dp(rich_d), dp(blur_d), None, V, W,
            C.c_float(wl.L["pitch"]), C.c_float(B), C.c_float(bl0), C.c_float(f), w, h,
            *[dp(t) for t in a_out], dp(a_stats), dp(ws_a), C.c_size_t(na), st))
        e1 = ev()
        lfpm._check(L.lfpm_fuse_backproject(
            dp(bufs.plane), dp(df.centers), dp(tiles), dp(df.nbr), dp(df.cnt), K, None,
            dp(order_d), lfpm._np(start), nl, V, W, C.c_float(baseline_bp), C.c_float(B),
            C.c_float(bl0), w, h, dp(f_vd), dp(f_aif), dp(f_conf), dp(pts),
            C.c_size_t(pts.numel() * 4), dp(counts), dp(ws_f), C.c_size_t(nf), st))
        e2 = ev()
        return e0, e1, e2

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    marks = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        marks.append(step())
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    ms = {k: float(np.mean([m[i].elapsed_time(m[i + 1]) for m in marks]))
          for i, k in enumerate(("aif_composite", "fuse_backproject"))}
    ast = a_stats.cpu().numpy()
    c = counts.cpu().numpy()
    # algorithmic HBM bytes of the two calls: inputs read once (plane 16 B + colour 3 B per MI
    # pixel; richness + blur 2 B per MI pixel for the composite), outputs written once (21 B per
    # raw pixel for the composite's maps and images, 8 B for the fusion's, 36 B per point)
    M, N = V * W * W, w * h
    alg = 2 * 19.0 * M + 2.0 * M + 29.0 * N + 36.0 * int(c[0])
    gbs = alg / (sum(ms.values()) * 1e-3) / 1e9
    res = {"fps": 1e3 / sum(ms.values()), "ms": ms, "wall_ms": 1e3 * wall,
           "levels": nl, "fused": int(c[1]), "points": int(c[0]),
           "roofline": {"bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s",
                        "frac": gbs / 8000.0, "traffic": None,
                        "kernel": "lfpm_aif_composite + lfpm_fuse_backproject (all launches)",
                        "avg_launch_ms": sum(ms.values()),
                        "per_unit": "19 B per MI pixel per call + 2 B (score maps) + 29 B per "
                                    "raw pixel + 36 B per fused point",
                        "note": "dependency-bound, not bandwidth-bound: the reference's fusion "
                                "walk is sequential within each MI (solved here as a fixed "
                                "point, one workgroup per MI) and across %d schedule levels; "
                                "hole filling iterates rows to a fixed point" % nl},
           "aif_stats": {"collected": int(ast[0]), "kept": int(ast[1]), "passes": int(ast[2]),
                         "fills": int(ast[3])}, "cpu_baseline": None}
    if args.no_cpu_baseline:
        return res
    # CPU baseline: the oracle's two composites of the same frame
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    p = bufs.plane.cpu().numpy()
    tl = tiles.cpu().numpy()
    t1 = time.perf_counter()
    oa = oracle.aif_composite(p, wl.L["centers"], tl, W, wl.L["pitch"], B, bl0, f, w, h,
                              rich=rich, blur=blur)
    of = oracle.fuse_backproject(p, wl.L["centers"], tl, wl.L["nbr"].reshape(V, K), wl.L["cnt"],
                                 W, baseline_bp, B, bl0, w, h)
    cpu_s = time.perf_counter() - t1
    same = bool(np.array_equal(a_out[4].cpu().numpy(), oa["filled"]) and
                np.array_equal(f_vd.cpu().numpy().view(np.uint32), of["vd"].view(np.uint32)))
    res.update(equal_oracle=same,
               cpu_baseline={"value": 1.0 / cpu_s, "unit": "frames/s", "cores": 1,
                             "kind": "port", "cpu": cpu_model(),
                             "sample": "oracle AIF composite + back-projection fusion of the "
                                       "same solved %dx%d frame, one frame in %.1f s"
                                       % (w, h, cpu_s)})
    return res


def calib_copy(torch):
    """One 1 GiB device-to-device copy: the known-size kernel that scales the rocprofv3
    FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    src = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def ingest_section(args, lfpm, torch, rw, rh):
    """Colour-frame ingest: lfpm_bgr_to_gray over rw x rh BGR frames.  Eight input/output pairs
    (8 x 4 B/px, more than the 256 MB last-level cache for 4080x3068) are converted in turn, so
    every launch reads its frame from HBM; 64 launches are captured in one graph so that the
    host's launch rate does not limit the 50 MB launches.  Timed with HIP events around the
    graph replays on the capture stream."""
    n_buf, per_graph = 8, 64
    gen = torch.Generator(device="cuda").manual_seed(3)
    bgr = [torch.randint(0, 256, (rh, rw, 3), dtype=torch.uint8, device="cuda", generator=gen)
           for _ in range(n_buf)]
    gray = [torch.empty((rh, rw), dtype=torch.uint8, device="cuda") for _ in range(n_buf)]
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        for i in range(n_buf):  # warm-up outside the capture (module load, first-touch)
            lfpm.bgr_to_gray(bgr[i], out=gray[i])
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        for i in range(per_graph):
            lfpm.bgr_to_gray(bgr[i % n_buf], out=gray[i % n_buf])
    with torch.cuda.stream(s):
        for _ in range(max(1, args.warmup)):
            g.replay()
    s.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(s):
        e0.record(s)
        for _ in range(args.steps):
            g.replay()
        e1.record(s)
    s.synchronize()
    launches = args.steps * per_graph
    ms = e0.elapsed_time(e1) / launches
    # every replayed frame equals a fresh eager conversion
    check = torch.empty_like(gray[0])
    lfpm.bgr_to_gray(bgr[0], out=check)
    torch.cuda.synchronize()
    same = bool(torch.equal(check, gray[0]))
    alg_bytes = 4.0 * rw * rh  # 3 B read + 1 B written per pixel
    gbs = alg_bytes / (ms * 1e-3) / 1e9
    # CPU baseline: the oracle's numpy rule on one frame (whole frame, one core)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    host = bgr[0].cpu().numpy()
    t0 = time.perf_counter()
    reps = 0
    while time.perf_counter() - t0 < 10.0:
        oracle.bgr_to_gray(host)
        reps += 1
    cpu_fps = reps / (time.perf_counter() - t0)
    traffic = None
    path = os.path.join(ROOT, "profiles", "hbm_traffic_ingest_4080x3068.md")
    if os.path.exists(path):
        for row in open(path):
            cells = [c.strip() for c in row.strip().strip("|").split("|")]
            if len(cells) == 5 and "k_bgr_gray" in cells[0]:
                try:
                    traffic = float(cells[4])
                except ValueError:
                    pass
    return {"fps": 1e3 / ms, "ms": ms, "launches": launches, "replay_equals_eager": same,
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s",
                         "frac": gbs / 8000.0, "traffic": traffic, "kernel": "k_bgr_gray",
                         "avg_launch_ms": ms,
                         "per_unit": "4 B per pixel (3 B of B, G, R read, 1 B gray written)"},
            "cpu_baseline": {"value": cpu_fps, "unit": "frames/s", "cores": 1, "kind": "port",
                             "cpu": cpu_model(),
                             "sample": "oracle rule (numpy, int64) on one whole %dx%d frame, "
                                       "%d frames in %.0f s" % (rw, rh, reps, 10.0)}}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU; --gpus %d is not supported "
              "(multi-GPU sharding is a design note only)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    import lfpm
    rw, rh = (int(x) for x in args.raw.split("x"))
    line = {"metric": METRIC, "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32", "data": "synthetic"}
    if args.mode == "precision":
        r = precision_section(args, lfpm, torch)
        line.update(value=r["fp16"]["hyp_views_per_s"], unit="hypothesis-views/s", steps=1,
                    ms_per_step=r["fp16"]["gpu_ms"], precision=r,
                    config={"workload": "10^6 seeded hypotheses on the 4080x3068 frame (18 views, "
                                        "15x15 window), FP32 and FP16 LDS tiles vs the FP64 "
                                        "oracle; full-solve agreement on a 1024x800 frame"})
        line["dtype"] = "fp16 storage / fp32 arithmetic"
        print(json.dumps(line), flush=True)
        return
    if args.mode == "parity":
        r = parity_section(args, lfpm, torch, rw, rh)
        line.update(metric="valid pixels within 0.05 px of the FP64 oracle (%dx%d, full solve)"
                    % (rw, rh), value=r["frac_le_005"], unit="fraction", steps=1,
                    ms_per_step=1e3 * r["gpu_s"], parity=r,
                    config={"workload": "one %dx%d frame at the full-frame parameters (18 views, "
                                        "15x15 window, 6 iterations) solved on the GPU and by the "
                                        "FP64 oracle on the host" % (rw, rh)})
        print(json.dumps(line), flush=True)
        return
    if args.mode == "serve":
        r = serve_section(args, lfpm, torch)
        line.update(value=r["fps"], ms_per_step=1e3 * r["span_s"] / max(1, r["requests"]),
                    p50_ms=r["p50_ms"], p99_ms=r["p99_ms"], serve=r,
                    config={"workload": "continuous-batching server: Poisson %.0f req/s (seed 7), "
                                        "raw from {1024x800, 2048x1536, 4080x3068}, pitch from "
                                        "{15, 23, 31}, 6 iterations + cycle check + 32-slice "
                                        "refocus per request" % args.rate,
                            "requests": args.requests, "hbm_budget_gb": args.hbm_budget_gb})
        print(json.dumps(line), flush=True)
        return
    if args.mode == "ingest":
        r = ingest_section(args, lfpm, torch, rw, rh)
        if args.calib_copy:
            calib_copy(torch)
        line.update(metric="colour raw frames converted to gray per s (%dx%d BGR)" % (rw, rh),
                    value=r["fps"], ms_per_step=r["ms"], dtype="u8", roofline=r["roofline"],
                    cpu_baseline=r["cpu_baseline"], ingest=r,
                    config={"workload": "cv::cvtColor(COLOR_BGR2GRAY) of %dx%d colour raw "
                                        "frames on the device, 8 frames in rotation, 64 launches "
                                        "per captured graph" % (rw, rh)})
        print(json.dumps(line), flush=True)
        return
    if args.mode == "products":
        wl = lfpm.workload_c2(raw_w=rw, raw_h=rh)
        r = products_section(args, lfpm, torch, wl)
        if args.calib_copy:
            calib_copy(torch)
        line.update(metric="per-frame products per s (%dx%d: mosaic, VD map, point cloud)"
                    % (rw, rh), value=r["fps"], ms_per_step=sum(r["ms"].values()),
                    dtype="fp32", roofline=r["roofline"], cpu_baseline=r["cpu_baseline"],
                    products=r,
                    config={"workload": "disparity mosaic + VD median splat with gray colour "
                                        "image + VD point cloud of one solved %dx%d frame "
                                        "(%d MIs)" % (rw, rh, wl.V)})
        print(json.dumps(line), flush=True)
        return
    if args.mode == "filter":
        wl = lfpm.workload_c2(raw_w=rw, raw_h=rh)
        r = filter_section(args, lfpm, torch, wl)
        roof, cpu = r.pop("roofline"), r.pop("cpu_baseline")
        line.update(metric="CheckModule passes per s (%dx%d, the reference's MI order)"
                    % (rw, rh), value=1e3 / r["ms"]["reference"],
                    ms_per_step=r["ms"]["reference"], dtype="fp32", roofline=roof,
                    cpu_baseline=cpu, filter=r,
                    config={"workload": "cycle-consistency / photometric / spike filter of one "
                                        "solved %dx%d frame (%d MIs, 18 neighbours, at most 4 "
                                        "views per pixel), snapshot and reference order"
                                        % (rw, rh, wl.V)})
        print(json.dumps(line), flush=True)
        return
    if args.mode == "composites":
        wl = lfpm.workload_c2(raw_w=rw, raw_h=rh)
        r = composites_section(args, lfpm, torch, wl)
        if args.calib_copy:
            calib_copy(torch)
        line.update(metric="refocus composites per s (%dx%d: AIF composite + back-projection "
                           "fusion)" % (rw, rh), value=r["fps"], ms_per_step=sum(r["ms"].values()),
                    dtype="fp32", roofline=r["roofline"], cpu_baseline=r["cpu_baseline"],
                    composites=r,
                    config={"workload": "all-in-focus composite (depth maps, hole filling, "
                                        "denoising) + back-projection consistency fusion (18 "
                                        "neighbours) of one solved %dx%d frame (%d MIs)"
                                        % (rw, rh, wl.V)})
        print(json.dumps(line), flush=True)
        return
    log("bench: synthesising %dx%d frame" % (rw, rh))
    wl = lfpm.workload_c2(raw_w=rw, raw_h=rh)
    log("bench: V=%d crop=%d K=%d iters=%d" % (wl.V, wl.crop, wl.pm.max_neighbors,
                                                wl.pm.iterations))
    plain = plain_section(args, lfpm, torch, wl)
    if args.calib_copy:
        calib_copy(torch)
    cfg = {"raw": "%dx%d" % (rw, rh), "pitch": 23.0, "mis": wl.V,
           "neighbors": wl.pm.max_neighbors, "window": 2 * wl.pm.radius + 1,
           "iterations": wl.pm.iterations, "slices": wl.slices, "mode": args.mode}
    line.update(plain_fps=plain["fps"], plain_p50_ms=float(np.percentile(plain["lat"], 50)),
                plain_p99_ms=float(np.percentile(plain["lat"], 99)), kernel_ms=plain["kt"],
                work=plain.get("work"), roofline=plain.get("roofline"))
    if args.mode == "plain":
        line.update(value=plain["fps"], ms_per_step=plain["ms_per_step"],
                    p50_ms=line["plain_p50_ms"], p99_ms=line["plain_p99_ms"])
        cfg["workload"] = ("plain GPU single-frame solve (init + %d red-black iterations + disp "
                           "record) + CheckModule filter (snapshot order) + %d-slice refocus, one "
                           "launch per phase with a sync after each"
                           % (wl.pm.iterations, wl.slices))
    else:
        if args.mode == "mixed":
            log("bench: synthesising %d mixed frames" % args.batch)
            wls = mixed_workloads(lfpm, args.batch)
            cfg.update(raw="mixed {1024x800, 2048x1536, 4080x3068}", pitch="mixed {15, 23, 31}",
                       neighbors="6/12/18 by lens type", mis=sum(w.V for w in wls))
            cfg["workload"] = ("%d uneven frames in one captured hipGraph: solve + filter + "
                               "refocus per frame" % args.batch)
        else:
            log("bench: synthesising %d frames" % args.batch)
            wls = [wl] + [lfpm.workload_c2(raw_w=rw, raw_h=rh, scene_seed=4321 + 17 * i,
                                           noise_seed=4322 + 17 * i, frame_uid=i)
                          for i in range(1, args.batch)]
            cfg["workload"] = ("%d frames of %dx%d in one captured hipGraph (persistent k_stage_w "
                               "launches over all frames' MIs) + per-frame CheckModule filter "
                               "(snapshot order) + %d-slice refocus"
                               % (args.batch, rw, rh, wl.slices))
        bat = batched_section(args, lfpm, torch, wls)
        if args.mode == "mixed":
            pf = plain_frames_section(args, lfpm, torch, wls)
            line.update(plain_same_frames_fps=pf["fps"],
                        speedup_vs_plain_same_frames=bat["fps"] / pf["fps"])
        line.update(value=bat["fps"], ms_per_step=bat["ms_per_step"],
                    p50_ms=float(np.percentile(bat["lat"], 50)),
                    p99_ms=float(np.percentile(bat["lat"], 99)),
                    batched_frames=bat["frames"], speedup_vs_plain=bat["fps"] / plain["fps"])
        if args.mode == "batched":
            line["batched_equals_plain"] = same_results(plain["result"], bat["result"], torch)
        cfg["mode"] = args.mode
    line["config"] = {"workload": cfg.pop("workload"), **cfg}
    if not args.no_cpu_baseline:
        log("bench: cpu baseline")
        line["cpu_baseline"] = cpu_baseline(lfpm)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
