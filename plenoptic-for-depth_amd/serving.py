"""Served form: Poisson request streams through the continuous-batching frame server of the C ABI
(lfpm_server_create / lfpm_frame_submit / lfpm_frame_wait).

The server (csrc/lfpm_serve.hip) admits submitted frames by HBM budget, keeps their state
resident, advances every admitted frame one stage per step with one launch per parameter class
(classes on their own streams, overlapping), and returns each frame's results on wait.  This
module submits requests at their arrival times and collects the results; latency is measured by
the server from submit to completion (queueing included).
"""
import ctypes as C
import threading
import time

import numpy as np

import lfpm


class ServeOpts(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_size_t), ("max_slots", C.c_int), ("streams", C.c_int),
                ("filter_order", C.c_int)]


class FrameRequest(C.Structure):
    _fields_ = [("layout", C.POINTER(lfpm.LayoutParams)), ("raw_host", C.c_void_p),
                ("raw_pitch", C.c_size_t), ("params", C.POINTER(lfpm.PmParams)),
                ("filter", C.POINTER(lfpm.FilterParams)), ("refocus_slices", C.c_int),
                ("refocus_vmin", C.c_float), ("refocus_vmax", C.c_float)]


class FrameResult(C.Structure):
    _fields_ = [("plane", C.c_void_p), ("cost", C.c_void_p), ("disp", C.c_void_p),
                ("sel", C.c_void_p), ("reason", C.c_void_p), ("refocus", C.c_void_p),
                ("num_mis", C.c_int), ("crop", C.c_int), ("latency_ms", C.c_double),
                ("queue_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_longlong), ("completed", C.c_longlong),
                ("steps", C.c_longlong), ("launches", C.c_longlong), ("max_active", C.c_int),
                ("classes", C.c_int), ("in_use_bytes", C.c_size_t),
                ("budget_bytes", C.c_size_t)]


class Request:
    """One frame to serve: a workload (layout, raw frame, solve / filter / refocus settings)
    and its arrival time in seconds from the start of the stream."""

    def __init__(self, rid, arrival_s, wl, filt=True, refocus=True):
        self.rid = rid
        self.arrival = arrival_s
        self.wl = wl
        self.filt = filt
        self.refocus = refocus
        self.ticket = None
        self.result = None
        self.latency_ms = None
        self.queue_ms = None


class Server:
    """A frame server on one device (lfpm_server): submit() returns at once, wait() blocks for
    that request's results."""

    def __init__(self, hbm_budget_bytes=0, max_slots=16, streams=4, device=0, keep_results=True,
                 filter_order="snapshot"):
        codes = {"snapshot": 0, "reference": 1}
        if filter_order not in codes:
            raise ValueError("Server: filter_order must be 'snapshot' or 'reference'")
        L = lfpm.lib()
        self._L = L
        self.keep_results = keep_results
        h = C.c_void_p()
        opts = ServeOpts(int(hbm_budget_bytes), int(max_slots), int(streams),
                         codes[filter_order])
        lfpm._check(L.lfpm_server_create(int(device), C.byref(opts), C.byref(h)))
        self._h = h
        self.done = []

    def close(self):
        if self._h:
            self._L.lfpm_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, r):
        wl = r.wl
        raw = np.ascontiguousarray(wl.raw())
        r._raw = raw
        r._req = FrameRequest(C.pointer(wl.lp), raw.ctypes.data_as(C.c_void_p), raw.shape[1],
                              C.pointer(wl.pm), C.pointer(wl.fp) if r.filt else None,
                              wl.slices if r.refocus else 0, wl.vmin, wl.vmax)
        t = C.c_uint64()
        lfpm._check(self._L.lfpm_frame_submit(self._h, C.byref(r._req), C.byref(t)))
        r.ticket = t.value
        return r.ticket

    def wait(self, r):
        wl = r.wl
        out = FrameResult()
        res = None
        if self.keep_results:
            n = wl.V * wl.crop ** 2
            h, w = wl.raw().shape
            res = dict(plane=np.zeros((n, 4), np.float32), cost=np.zeros(n, np.float32),
                       disp=np.zeros((n, 4), np.float32), sel=np.zeros(n, np.uint32))
            if r.filt:
                res["reason"] = np.zeros(n, np.uint8)
            if r.refocus and wl.slices > 0:
                res["refocus"] = np.zeros((wl.slices, h, w), np.float32)
            for k, a in res.items():
                setattr(out, k, a.ctypes.data_as(C.c_void_p))
        lfpm._check(self._L.lfpm_frame_wait(self._h, C.c_uint64(r.ticket), C.byref(out)))
        r.result = res
        r.latency_ms = out.latency_ms
        r.queue_ms = out.queue_ms
        return r

    def stats(self):
        st = ServerStats()
        lfpm._check(self._L.lfpm_server_get_stats(self._h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def run(self, requests, clock=time.perf_counter, log=None):
        """Submit `requests` at their arrival times (seconds from now) while a waiter thread
        collects them in submit order; returns the requests, each with its result and the
        server's latency."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        submitted = []
        cv = threading.Condition()
        errors = []

        def waiter():
            i = 0
            while i < len(reqs):
                with cv:
                    cv.wait_for(lambda: len(submitted) > i)
                    r = submitted[i]
                try:
                    self.wait(r)
                except Exception as e:  # noqa: BLE001 - reported after the run
                    errors.append(e)
                if log:
                    log("serve: request %d done (%dx%d pitch %g) latency %.1f ms"
                        % (r.rid, r.wl.sp.raw_w, r.wl.sp.raw_h, r.wl.sp.pitch,
                           r.latency_ms or -1.0))
                i += 1

        th = threading.Thread(target=waiter)
        th.start()
        t0 = clock()
        for r in reqs:
            wait = r.arrival - (clock() - t0)
            if wait > 0:
                time.sleep(wait)
            self.submit(r)
            with cv:
                submitted.append(r)
                cv.notify()
        th.join()
        if errors:
            raise errors[0]
        self.done = reqs
        return reqs


def poisson_requests(n, rate_hz, seed, sizes, pitches, make_workload):
    """n requests with exponential inter-arrival times (rate_hz) and per-request draws."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(n):
        t += rng.exponential(1.0 / rate_hz)
        w, h = sizes[int(rng.integers(len(sizes)))]
        p = pitches[int(rng.integers(len(pitches)))]
        planes = int(rng.integers(6, 13))
        reqs.append(Request(i, t, make_workload(w, h, p, planes, 5000 + i)))
    return reqs


def latency_stats(done):
    lat = np.array([r.latency_ms for r in done])
    return dict(p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                mean_ms=float(lat.mean()), max_ms=float(lat.max()))
