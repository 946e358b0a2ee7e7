"""Continuous-batching depth server over the lfpm C ABI.

Requests (one raw frame each) arrive on a wall clock.  A request is admitted into a device slot
when its bytes fit the HBM budget and a slot is free; its raw frame is uploaded and its plane /
cost / selection state stays device-resident until it completes.  Every server step advances all
in-flight frames by one stage (init, one red/black phase, or final record + filter + refocus)
through one `lfpm_batch_step` call, which groups frames of the same parameter class (same lens
pitch / window / neighbours) into one persistent kernel launch — requests that arrive while
others are on the GPU join at the next step.  A frame's results do not depend on what it is
batched with.  Latency = completion time - arrival time (queueing included).
"""
import time

import numpy as np

import lfpm


class Request:
    def __init__(self, rid, arrival_s, wl):
        self.rid = rid
        self.arrival = arrival_s
        self.wl = wl
        self.slot = None
        self.schedule = None
        self.pos = 0
        self.done_s = None
        self.admit_s = None


def schedule_for(pm):
    q = [(lfpm.STAGE_INIT, 0, 0)]
    for it in range(pm.iterations):
        q += [(lfpm.STAGE_PHASE, it, 0), (lfpm.STAGE_PHASE, it, 1)]
    q.append((lfpm.STAGE_FINAL, 0, 0))
    return q


def slot_bytes(wl, slices):
    """Device bytes one frame occupies while in flight (inputs, state, products)."""
    n = wl.V * wl.crop ** 2
    h, w = wl.sp.raw_h, wl.sp.raw_w
    return (h * w + n * (4 + 4 + 16 + 4 + 4 + 16 + 16 + 4 + 4 + 16 + 4 + 16 + 4 + 1)
            + slices * h * w * 4 + wl.L["V"] * (8 + 8 + 4 * wl.pm.max_neighbors + 4))


class Server:
    def __init__(self, hbm_budget_bytes, max_slots=16, device="cuda"):
        self.budget = int(hbm_budget_bytes)
        self.max_slots = max_slots
        self.device = device
        self.in_use = 0
        self.active = []
        self.queue = []
        self.done = []
        self.steps = 0
        self.batch_sizes = []

    def _admit(self, now):
        while self.queue and len(self.active) < self.max_slots:
            r = self.queue[0]
            need = slot_bytes(r.wl, r.wl.slices)
            if self.in_use + need > self.budget:
                if not self.active:
                    raise lfpm.LfpmError(-3, "request %d needs %d B > budget %d B"
                                         % (r.rid, need, self.budget))
                break
            self.queue.pop(0)
            r.slot = lfpm.FrameSlot(r.wl, device=self.device)
            r.schedule = schedule_for(r.wl.pm)
            r.pos = 0
            r.admit_s = now
            r.bytes = need
            self.in_use += need
            self.active.append(r)

    def run(self, requests, clock=time.perf_counter, log=None):
        """Serve `requests` (sorted by arrival, seconds from start); returns the done list."""
        import torch
        pending = sorted(requests, key=lambda r: r.arrival)
        t0 = clock()
        batch = None
        batch_key = None
        while pending or self.queue or self.active:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                self.queue.append(pending.pop(0))
            self._admit(now)
            if not self.active:
                wait = pending[0].arrival - (clock() - t0)
                if wait > 0:
                    time.sleep(wait)
                continue
            key = tuple(id(r) for r in self.active)
            if key != batch_key:
                batch = lfpm.Batch([r.slot for r in self.active], device=self.device)
                batch_key = key
            st, it, ph = [], [], []
            for r in self.active:
                s = r.schedule[r.pos]
                st.append(s[0]); it.append(s[1]); ph.append(s[2])
            batch.step(st, it, ph)
            torch.cuda.current_stream().synchronize()
            self.steps += 1
            self.batch_sizes.append(len(self.active))
            now = clock() - t0
            still = []
            for r in self.active:
                r.pos += 1
                if r.pos == len(r.schedule):
                    r.done_s = now
                    self.in_use -= r.bytes
                    self.done.append(r)
                    if log:
                        log("serve: request %d done (%dx%d pitch %g) latency %.1f ms"
                            % (r.rid, r.wl.sp.raw_w, r.wl.sp.raw_h, r.wl.sp.pitch,
                               1e3 * (r.done_s - r.arrival)))
                else:
                    still.append(r)
            if len(still) != len(self.active):
                for r in self.active:
                    if r.pos == len(r.schedule):
                        r.slot = None  # release device buffers
                self.active = still
        return self.done


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
    lat = np.array([1e3 * (r.done_s - r.arrival) for r in done])
    return dict(p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                mean_ms=float(lat.mean()), max_ms=float(lat.max()))
