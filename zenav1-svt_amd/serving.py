"""Serving loop over the superblock-wavefront mode decision (zav1_wf_*).

Requests -- still luma planes of mixed size, bit depth, qp and preset --
arrive at Poisson times.  A request is admitted when its device footprint
(zav1_wf_footprint: source, reconstruction, decisions and its step chain's
buffers, all in one allocation) fits what is left of the HBM budget and an
active slot is free; its plane is uploaded and its decision planned
(zav1_wf_create) on a stream of its own.

Continuous batching: every active request keeps up to `depth` chunks of
`chunk` wavefront steps in flight (zav1_wf_run_steps, an event after each
chunk).  A request admitted while others run starts at once, and its chunks
interleave with theirs on the device's hardware queues -- no request waits
for another to finish.  A request completes when its last chunk's event has
fired; its reconstruction and decisions are read back, its device memory is
released, and the next arrivals are admitted against the freed budget.
Latency runs from a request's arrival to its completion.  Requests share no
device state, so each result is the one the request gives alone.

Encode mode (`encode=True`): every plan keeps its winners' levels and uses
the fewest uniform tile columns within AV1's tile-area limit; when a
request's decision completes, its AV1 still stream is written on a host
thread pool (zav1_wf_write_stills reads the decisions back on the
request's stream, then codes the tiles) while the loop keeps feeding the
device; the request completes -- and its device memory is released -- when
its stream is written.  Latency then runs from arrival to the stream.
"""
import collections
import importlib.util
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np


def _bitstream_module():
    name = 'zav1_bitstream'
    if name not in sys.modules:
        spec = importlib.util.spec_from_file_location(name, os.path.join(os.path.dirname(__file__), 'bitstream.py'))
        m = importlib.util.module_from_spec(spec)
        sys.modules[name] = m
        spec.loader.exec_module(m)
    return sys.modules[name]


class Request:
    __slots__ = ('idx', 'arrival', 'img', 'bd', 'qp', 'preset', 'cfg', 'footprint', 'md', 'stream', 'next_step',
                 'events', 't_admit', 't_done', 'recon', 'decisions', 'deferred', 'write', 'stream_bytes')

    def __init__(self, idx, arrival, img, bd, qp, preset):
        self.idx, self.arrival, self.img, self.bd, self.qp, self.preset = idx, arrival, img, bd, qp, preset
        self.md = self.stream = self.t_admit = self.t_done = self.recon = self.decisions = None
        self.write = self.stream_bytes = None
        self.next_step, self.events, self.deferred = 0, collections.deque(), False


class WavefrontServer:
    def __init__(self, zav1, wf, synth, budget_bytes, max_active=8, chunk=64, depth=2, leaf=16, encode=False,
                 writers=4, write_threads=4):
        self.z, self.wf, self.synth = zav1, wf, synth
        self.L = zav1.lib()
        self.budget, self.max_active, self.chunk, self.depth, self.leaf = budget_bytes, max_active, chunk, depth, leaf
        self.pool_ev = []
        self.encode, self.write_threads = encode, write_threads
        self.writers = ThreadPoolExecutor(writers) if encode else None
        self.bs = _bitstream_module() if encode else None

    def request_cfg(self, h, w, bd, qp, preset):
        qidx = self.synth.qp_to_qindex(qp)
        lam = int(self.L.zav1_kf_full_lambda(qidx, qp, bd))
        if not self.encode:
            return self.wf.funnel_cfg(w, h, bd, qidx, lam, preset, qp, leaf=self.leaf)
        cfg = self.wf.funnel_cfg(w, h, bd, qidx, lam, preset, qp, leaf=self.leaf,
                                 tiles=(self.bs.legal_tile_cols(w, h, 64), 1))
        cfg[0]['keep_levels'] = 1
        return cfg

    def _write(self, r):
        return r.md.write_stills(0, 1, threads=self.write_threads, stream=r.stream)[0]

    def _event(self):
        if self.pool_ev:
            return self.pool_ev.pop()
        return self.z.event_create()

    def run(self, specs, arrivals, keep_results=True):
        """specs: (img, bd, qp, preset) per request; arrivals: seconds after
        the start.  Returns (requests, stats)."""
        reqs = [Request(i, float(a), img, bd, qp, p) for i, ((img, bd, qp, p), a) in enumerate(zip(specs, arrivals))]
        for r in reqs:
            h, w = r.img.shape
            r.cfg = self.request_cfg(h, w, r.bd, r.qp, r.preset)
            r.footprint = self.wf.WavefrontMD.footprint(self.z, r.cfg)
            if r.footprint > self.budget:
                raise ValueError('request %d needs %d bytes, over the budget %d' % (r.idx, r.footprint, self.budget))
        pending = collections.deque(sorted(reqs, key=lambda r: r.arrival))
        active, used, peak_used, peak_active, deferred = [], 0, 0, 0, 0
        t0 = time.perf_counter()
        while pending or active:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now and len(active) < self.max_active:
                r = pending[0]
                if used + r.footprint > self.budget:
                    if not r.deferred:
                        r.deferred = True
                        deferred += 1
                    break
                pending.popleft()
                r.stream = self.z.stream_create()
                r.md = self.wf.WavefrontMD(self.z, r.cfg)
                r.md.upload(r.img[None], r.stream)
                r.t_admit = time.perf_counter() - t0
                used += r.footprint
                active.append(r)
            peak_used, peak_active = max(peak_used, used), max(peak_active, len(active))
            progressed = False
            for r in active:
                while r.events and self.z.event_query(r.events[0]):
                    self.pool_ev.append(r.events.popleft())
                    progressed = True
                n = r.md.info.n_steps
                while len(r.events) < self.depth and r.next_step < n:
                    k = min(self.chunk, n - r.next_step)
                    r.md.run_steps(r.next_step, k, r.stream)
                    r.next_step += k
                    e = self._event()
                    self.z.event_record(e, r.stream)
                    r.events.append(e)
                    progressed = True
            decided = [r for r in active if r.next_step == r.md.info.n_steps and not r.events]
            done = []
            for r in decided:
                if not self.encode:
                    done.append(r)
                elif r.write is None:
                    r.write = self.writers.submit(self._write, r)  # the stream, off the feeding thread
                    progressed = True
                elif r.write.done():
                    r.stream_bytes = r.write.result()
                    done.append(r)
            for r in done:
                r.t_done = time.perf_counter() - t0
                if keep_results:
                    r.recon, r.decisions = r.md.recon()[0], r.md.decisions()[0]
                r.md.close()
                self.z.stream_destroy(r.stream)
                r.md = r.stream = None
                used -= r.footprint
                active.remove(r)
                progressed = True
            if not progressed:
                time.sleep(50e-6)
        wall = max(r.t_done for r in reqs) - min(r.arrival for r in reqs)
        lat = np.array([r.t_done - r.arrival for r in reqs])
        mp = sum(r.img.size for r in reqs) / 1e6
        stats = {'wall_s': wall, 'megapixels': mp, 'latency_s': lat, 'peak_active': peak_active,
                 'peak_bytes': peak_used, 'deferred_by_memory': deferred}
        if self.encode:
            stats['stream_bytes'] = sum(len(r.stream_bytes) for r in reqs)
        return reqs, stats

    def close(self):
        for e in self.pool_ev:
            self.z.event_destroy(e)
        self.pool_ev = []
        if self.writers:
            self.writers.shutdown()
            self.writers = None
