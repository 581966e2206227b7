"""Serving loop for the tx-type search stage.

Requests (images of the serving mix: log-uniform area, 8 or 10 bits, a qp
each) arrive at Poisson times.  A request is admitted when one of a fixed
pool of contexts is free -- each context is one stream and one HBM arena of
a fixed budget, so the pool bounds the HBM the stage may hold (admission
control).  An admitted request is copied into its context's pinned host
staging buffer, uploaded into the context's arena (an asynchronous DMA),
and searched over every tx size and legal type at its qindex by one call
(zav1_txsearch_frame: grids and outputs from the arena, all on the
context's stream); its results stay resident in HBM for the next stage.
Its latency runs from its arrival to the completion of its last kernel,
observed by polling the completion event.
"""
import ctypes
import time

import numpy as np


def request_bytes(h, w, tsm):
    """Arena bytes one request takes: the u16 plane and, per tx size, the
    block grid and the per-candidate outputs (each allocation rounded up to
    the arena's 256-byte alignment)."""
    def r(n):
        return (n + 255) // 256 * 256
    tot = r(2 * h * w)
    for ts in range(19):
        bw, bh = tsm.TX_DIMS[ts]
        nb = (w // bw) * (h // bh)
        n = nb * len(tsm.legal_types(ts))
        tot += r(8 * nb) + r(8 * n) + 2 * r(4 * n)
    return tot


class ServingLoop:
    def __init__(self, zav1, tsm, synth, n_ctx=4, ctx_budget=None, max_hw=(4096, 3072)):
        self.zav1, self.tsm, self.synth = zav1, tsm, synth
        self.L = zav1.lib()
        budget = ctx_budget or request_bytes(max_hw[0], max_hw[1], tsm)
        self.ctx = [zav1.Context(0, budget) for _ in range(n_ctx)]
        self.budget = budget
        # pinned staging per context, large enough for any request the
        # budget admits (u16 samples)
        self.stage_bytes = 2 * max_hw[0] * max_hw[1] if ctx_budget is None else budget
        self.stage = []
        for _ in self.ctx:
            p = ctypes.c_void_p()
            zav1.check(self.L.zav1_host_alloc(ctypes.byref(p), self.stage_bytes), 'host_alloc')
            self.stage.append(p.value)
        self.ev = []
        for _ in self.ctx:
            e = ctypes.c_void_p()
            zav1.check(self.L.zav1_event_create(ctypes.byref(e)), 'event_create')
            self.ev.append(e.value)
        self.types = [np.array(tsm.legal_types(t), np.uint8) for t in range(19)]

    def submit(self, c, img, bd, qindex):
        """Enqueue one request on context c (staging copy, upload, search of
        every size); returns the FrameSearch record of its outputs."""
        ctx, L = self.ctx[c], self.L
        h, w = img.shape
        if 2 * h * w > self.stage_bytes:
            raise self.zav1.Zav1Error('request larger than the staging buffer')
        dst = np.ctypeslib.as_array((ctypes.c_uint16 * (h * w)).from_address(self.stage[c])).reshape(h, w)
        np.copyto(dst, img)
        ctx.reset()
        p = ctypes.c_void_p()
        self.zav1.check(L.zav1_image_upload(ctx.h, self.stage[c], w, w, h, 10, ctypes.byref(p)), 'image_upload')
        fs = ctx.txsearch_frame(p.value, w, h, bd, [qindex])
        self.zav1.check(L.zav1_event_record(self.ev[c], ctx.stream()), 'event_record')
        return fs

    def run(self, requests, images, rate_hz, seed=5001):
        """Serve `requests` [(h, w, bd, qp)] with host images `images` at
        Poisson arrivals of `rate_hz`; returns the per-request latencies (s),
        the megapixels served and the wall time from the first arrival to
        the last completion."""
        rng = np.random.default_rng(seed)
        gaps = rng.exponential(1.0 / rate_hz, len(requests))
        busy = [None] * len(self.ctx)   # request index on each context
        done_t = [0.0] * len(requests)
        arrive_t = [0.0] * len(requests)

        def poll():
            now = time.perf_counter()
            for c, k in enumerate(busy):
                if k is not None and self.L.zav1_event_query(self.ev[c]) == 1:
                    done_t[k] = now
                    busy[c] = None

        for c in range(len(self.ctx)):
            self.zav1.sync(self.ctx[c].stream())
        t0 = time.perf_counter()
        t_next = t0
        for k, (h, w, bd, qp) in enumerate(requests):
            t_next += gaps[k]
            while time.perf_counter() < t_next:
                poll()
            arrive_t[k] = t_next
            while None not in busy:  # admission: wait for a free context
                poll()
            c = busy.index(None)
            busy[c] = k
            self.submit(c, images[k], bd, self.synth.qp_to_qindex(qp))
        while any(b is not None for b in busy):
            poll()
        lat = np.array([done_t[k] - arrive_t[k] for k in range(len(requests))])
        mp = sum(h * w for h, w, _, _ in requests) / 1e6
        return lat, mp, max(done_t) - t0

    def close(self):
        for c in self.ctx:
            self.zav1.sync(c.stream())
        for p in self.stage:
            self.L.zav1_host_free(p)
        for e in self.ev:
            self.L.zav1_event_destroy(e)
        for c in self.ctx:
            c.close()
