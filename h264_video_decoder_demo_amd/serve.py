"""Serving: a continuous-batching request loop over many streams on one GPU.

Frames of many independent streams arrive over time (Poisson arrivals per stream).
The loop admits a stream when its first frame arrives and its DPB fits the HBM
budget (h264g_stream_open refuses otherwise; a refused stream waits until another
stream closes), keeps each stream's reference pictures resident in its device DPB,
and every `window_ms` enqueues one batched launch per stage holding the next
arrived frames of every admitted stream (those of a stream that do not depend on
each other, e.g. two B pictures and the next P picture).  Batches run back to back on the
context's HIP stream while the host stages the next one; completion of each batch
is observed through a fence.  A frame's latency is its batch's completion time
minus its arrival time.

No counterpart in the reference, which decodes one file synchronously.
"""
from __future__ import annotations

import collections
import concurrent.futures
import dataclasses
import os
import time

import numpy as np

from ._lib import H264G_E_BUDGET, H264GError
from .engine import MAX_STREAM_PICS, Engine, Generator, slot_use

# resolution mix of the serving workload: (share, width MBs, height MBs)
SERVE_MIX = ((0.25, 22, 18), (0.35, 80, 45), (0.30, 120, 68), (0.10, 240, 135))


@dataclasses.dataclass
class StreamSpec:
    width_mbs: int
    height_mbs: int
    seed: int
    n_frames: int
    start_s: float          # first arrival
    rate_fps: float         # Poisson rate of this stream's frames
    arrivals: tuple = ()    # the frames' arrival times, when fixed in advance (else drawn)


@dataclasses.dataclass
class ServeResult:
    frames: int
    mpix: float
    wall_s: float
    latencies_ms: list
    admitted_late: int      # streams that waited for DPB budget
    max_batch: int
    arrival_span_s: float = 0.0  # first to last frame arrival

    @property
    def offered_fps(self) -> float:
        """Frames that arrived over the span of their arrival times."""
        return self.frames / self.arrival_span_s if self.arrival_span_s > 0 else 0.0

    @property
    def served_fps(self) -> float:
        """Frames completed over the loop's wall time (first arrival to last completion)."""
        return self.frames / self.wall_s if self.wall_s > 0 else 0.0

    @property
    def mpix_per_s(self) -> float:
        return self.mpix / self.wall_s

    def pct(self, q: float) -> float:
        v = sorted(self.latencies_ms)
        return v[min(len(v) - 1, max(0, int(round(q / 100.0 * (len(v) - 1)))))]


def make_specs(n_streams: int, total_fps: float, duration_s: float, seed: int = 0x5EED0005, gop: int = 32):
    """Streams of the serving mix: resolution by the mix shares; each stream's frames are
    a Poisson process of rate total_fps / n_streams on [0, duration_s), so all frames
    together arrive at total_fps over the whole window (no ramp at the start, no tail
    after it).  A stream whose process has no arrival in the window is left out."""
    rng = np.random.default_rng(seed)
    shares = np.array([m[0] for m in SERVE_MIX])
    specs = []
    rate = total_fps / n_streams
    for i in range(n_streams):
        k = int(rng.choice(len(SERVE_MIX), p=shares / shares.sum()))
        arr, t = [], 0.0
        while True:
            t += float(rng.exponential(1.0 / rate))
            if t >= duration_s:
                break
            arr.append(t)
        if arr:
            specs.append(StreamSpec(SERVE_MIX[k][1], SERVE_MIX[k][2], seed + 1 + i, len(arr), arr[0], rate,
                                    tuple(arr)))
    del gop
    return specs


def arrival_times(spec: StreamSpec, rng: np.random.Generator):
    """The spec's fixed arrival times, or n_frames Poisson arrivals from start_s."""
    if spec.arrivals:
        return np.asarray(spec.arrivals, dtype=np.float64)
    gaps = rng.exponential(1.0 / spec.rate_fps, size=spec.n_frames)
    gaps[0] = 0.0
    return spec.start_s + np.cumsum(gaps)


class _Stream:
    def __init__(self, idx, spec, gen, arrivals):
        self.idx, self.spec, self.gen, self.arrivals = idx, spec, gen, arrivals
        self.sid = None
        self.next_arrival = 0   # frames that have arrived
        self.next_submit = 0    # frames submitted
        self.done = 0           # frames completed
        self.waited = False


class ServeLoop:
    """Runs streams through the engine as they arrive.

    on_frame(stream_index, frame_index, engine, stream_id, desc), if given, is called
    after the frame's batch has completed and before the stream's next frame is
    submitted (the loop then keeps one frame per stream in flight, so the frame's DPB
    slot can be read back).  on_close(stream_index, engine, stream_id), if given, is
    called after a stream's last frame has completed, before the stream is closed (its
    DPB can be read back then)."""

    def __init__(self, engine: Engine, specs, gop: int = 32, window_ms: float = 2.0, seed: int = 1,
                 on_frame=None, gens=None, on_close=None):
        self.eng = engine
        self.window = window_ms / 1e3
        self.on_frame = on_frame
        self.on_close = on_close
        rng = np.random.default_rng(seed)
        self.streams = []
        for i, sp in enumerate(specs):
            g = gens[i] if gens is not None else Generator(2, width_mbs=sp.width_mbs, height_mbs=sp.height_mbs,
                                                          n_frames=sp.n_frames, seed=sp.seed, gop=gop)
            self.streams.append(_Stream(i, sp, g, arrival_times(sp, rng)))

    def run(self) -> ServeResult:
        eng = self.eng
        inflight = collections.deque()   # (fence, [(stream, frame, arrival_s, pic)])
        lat, frames, mpix = [], 0, 0.0
        waiting = []                     # streams refused by the DPB budget, in arrival order
        admitted_late = 0
        max_batch = 0
        total = sum(s.gen.num_pictures for s in self.streams)
        t0 = time.perf_counter()
        next_window = 0.0
        stagers = concurrent.futures.ThreadPoolExecutor(min(16, os.cpu_count() or 1))
        while frames < total:
            now = time.perf_counter() - t0
            # arrivals and admission
            for s in self.streams:
                while s.next_arrival < s.gen.num_pictures and s.arrivals[s.next_arrival] <= now:
                    s.next_arrival += 1
                if s.sid is None and s.next_arrival > 0 and s not in waiting and s.done < s.gen.num_pictures:
                    waiting.append(s)
            for s in list(waiting):
                try:
                    s.sid = eng.open_stream(s.spec.width_mbs, s.spec.height_mbs, s.gen.dpb_slots)
                    waiting.remove(s)
                    if s.waited:
                        admitted_late += 1
                except H264GError as e:
                    if e.code != H264G_E_BUDGET:
                        raise
                    if not inflight and all(t.sid is None for t in self.streams):
                        raise RuntimeError(f"stream {s.idx} does not fit the DPB budget even alone") from e
                    s.waited = True
                    break  # admit in arrival order
            # completions
            while inflight and eng.done(inflight[0][0]):
                f, items = inflight.popleft()
                t_done = time.perf_counter() - t0
                eng.drop_fence(f)
                for s, k, arr, pic in items:
                    lat.append(1e3 * (t_done - arr))
                    frames += 1
                    mpix += s.spec.width_mbs * s.spec.height_mbs * 256 / 1e6
                    desc = pic.desc
                    eng.release(pic)
                    s.done += 1
                    if self.on_frame is not None:
                        self.on_frame(s.idx, k, eng, s.sid, desc)
                    if s.done == s.gen.num_pictures:
                        if self.on_close is not None:
                            self.on_close(s.idx, eng, s.sid)
                        eng.close_stream(s.sid)
                        s.sid = None
            # a batch window
            if now >= next_window:
                todo = []
                for s in self.streams:
                    # the stream's arrived frames that do not depend on each other (up to
                    # MAX_STREAM_PICS, e.g. B1 B2 P6) share the window's batch; read-back
                    # mode keeps one frame per stream in flight
                    taken = []
                    while (s.sid is not None and s.next_submit < s.next_arrival and len(taken) < MAX_STREAM_PICS
                           and (self.on_frame is None or (not taken and s.next_submit == s.done))):
                        dst, refs = slot_use(s.gen.picture_raw(s.next_submit)[0])
                        if any(dst == d or d in refs or dst in r for d, r in taken):
                            break
                        taken.append((dst, refs))
                        todo.append((s, s.next_submit))
                        s.next_submit += 1
                # host staging of the window's pictures on several threads (the library
                # serialises only the device block allocation)
                pics = list(stagers.map(lambda t: eng.begin_raw(t[0].sid, *t[0].gen.picture_raw(t[1])), todo))
                batch = [(s, k, float(s.arrivals[k]), pic) for (s, k), pic in zip(todo, pics)]
                if batch:
                    eng.batch_submit([p for _, _, _, p in batch])
                    inflight.append((eng.fence(), batch))
                    max_batch = max(max_batch, len(batch))
                next_window = max(next_window + self.window, now)
            else:
                time.sleep(min(2e-4, max(0.0, next_window - now)))
        eng.sync()
        stagers.shutdown()
        arr = [float(s.arrivals[k]) for s in self.streams for k in range(s.gen.num_pictures)]
        return ServeResult(frames, mpix, time.perf_counter() - t0, lat, admitted_late, max_batch,
                           max(arr) - min(arr) if arr else 0.0)
