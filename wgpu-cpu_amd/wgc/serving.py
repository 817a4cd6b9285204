"""Serving loop: render requests arrive over time and are batched as they
arrive (continuous batching), with meshes and textures resident on the device.

Request stream: Poisson arrivals at a fixed rate; each request is a render
pass of the batched workload's distribution (target W and H log-uniform in
[64, 1024], one of a pool of resident displaced spheres of 100-50k triangles,
half of them textured with the shared 1024x1024 checker), with its own camera
uniform for its aspect ratio.

Admission: every request's targets and uniform are sub-allocated from one
device heap whose size is the HBM budget; a request waits in the queue while
the heap cannot hold it, and is admitted as soon as completed requests give
their memory back.

Continuous batching: at most ``max_inflight`` submissions are on the device
at once; whenever a slot is free, every queued request (up to ``max_batch``)
is recorded into one command buffer and submitted, so requests that arrive
while others are executing join the next launch instead of waiting for the
queue to drain.  The plain form submits every request on its own the moment
it arrives.

The reference has no counterpart: it executes one pass at a time on one
engine thread (wgpu-cpu/src/engine.rs:26-46).
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import (CULL_BACK, DEPTH32_FLOAT, FILTER_LINEAR, FILTER_NEAREST, FLOAT32X2, FLOAT32X3, FLOAT32X4,
               FRONT_CCW, INDEX_UINT32, INTERP_LINEAR, LESS, PROGRAM_MVP_COLOR, PROGRAM_MVP_TEXTURE,
               REPEAT, RGBA8_UNORM, RGBA8_UNORM_SRGB, STEP_VERTEX, Device)
from . import scenes


@dataclass
class Mesh:
    textured: bool
    triangles: int
    vertex_buffer: object
    index_buffer: object
    index_count: int
    pipeline: object
    bind_group1: Optional[object]


@dataclass
class Request:
    rid: int
    arrival: float           # seconds after the start of the run
    width: int
    height: int
    mesh: int
    uniform: bytes


@dataclass
class ServeStats:
    mode: str
    requests: int
    completed: int
    seconds: float
    fragments: int
    submissions: int
    latencies_ms: np.ndarray
    max_queue: int
    max_heap_used: int
    admission_waits: int

    @property
    def mfrag_per_s(self) -> float:
        return self.fragments / self.seconds / 1e6

    def summary(self) -> dict:
        lat = self.latencies_ms
        return {
            "mode": self.mode,
            "requests": self.requests,
            "completed": self.completed,
            "seconds": round(self.seconds, 3),
            "requests_per_s": round(self.completed / self.seconds, 1),
            "Mfragments_per_s": round(self.mfrag_per_s, 2),
            "p50_ms": round(float(np.percentile(lat, 50)), 3) if len(lat) else None,
            "p99_ms": round(float(np.percentile(lat, 99)), 3) if len(lat) else None,
            "mean_ms": round(float(lat.mean()), 3) if len(lat) else None,
            "submissions": self.submissions,
            "mean_batch": round(self.completed / max(1, self.submissions), 2),
            "max_queue": self.max_queue,
            "max_heap_used_MB": round(self.max_heap_used / 2**20, 1),
            "admission_waits": self.admission_waits,
        }


class MeshPool:
    """Resident meshes: `n` displaced spheres with triangle counts log-uniform
    in [100, 50k], every other one textured."""

    def __init__(self, device: Device, n: int, seed: int):
        self.device = device
        self._keep: list = []
        u = scenes.splitmix_f32(seed, n).astype(np.float64)
        tris = np.exp(np.log(100) + u * (np.log(50_000) - np.log(100)))
        self.meshes: List[Mesh] = []
        tex = None
        samp = None
        for i in range(n):
            textured = (i % 2) == 1
            segs, rings = scenes.sphere_for_triangles(float(tris[i]))
            pos, uv, idx = scenes.displaced_sphere(segs, rings, scenes.sub_seed(seed, i))
            if textured:
                vb = np.zeros((len(pos), 6), np.float32)
                vb[:, 0:3] = pos
                vb[:, 3] = 1.0
                vb[:, 4:6] = uv
                attrs = [(FLOAT32X4, 0, 0), (FLOAT32X2, 16, 1)]
                program, stride = PROGRAM_MVP_TEXTURE, 24
            else:
                col = scenes.splitmix_f32(scenes.sub_seed(seed, 1000 + i), len(pos) * 3).reshape(-1, 3)
                vb = np.zeros((len(pos), 7), np.float32)
                vb[:, 0:3] = pos
                vb[:, 3] = 1.0
                vb[:, 4:7] = col
                attrs = [(FLOAT32X4, 0, 0), (FLOAT32X3, 16, 1)]
                program, stride = PROGRAM_MVP_COLOR, 28
            module = device.create_shader_module_builtin(program, INTERP_LINEAR)
            pipeline = device.create_render_pipeline(module, [(stride, STEP_VERTEX, attrs)], front_face=FRONT_CCW,
                                                     cull_mode=CULL_BACK, depth=(LESS, True),
                                                     target_format=RGBA8_UNORM)
            bg1 = None
            if textured:
                if tex is None:
                    tex = device.create_texture(1024, 1024, RGBA8_UNORM_SRGB)
                    device.write_texture(tex, scenes.shared_texture())
                    samp = device.create_sampler(REPEAT, REPEAT, FILTER_NEAREST, FILTER_NEAREST)
                    self._keep += [tex, samp]
                bg1 = device.create_bind_group([(0, tex), (1, samp)])
            vbuf = device.create_buffer_init(vb.tobytes())
            ibuf = device.create_buffer_init(idx.astype(np.uint32).tobytes())
            self._keep += [module]
            self.meshes.append(Mesh(textured, int(len(idx) // 3), vbuf, ibuf, int(len(idx)), pipeline, bg1))
        device.poll()


def make_trace(seed: int, rate: float, seconds: float, n_meshes: int) -> List[Request]:
    """Poisson arrivals at `rate` per second over `seconds`; sizes of the
    batched workload; each request picks a resident mesh."""
    n_max = int(rate * seconds * 1.5 + 64)
    u = scenes.splitmix_f32(seed, n_max * 4).astype(np.float64).reshape(n_max, 4)
    gaps = -np.log(np.maximum(1.0 - u[:, 0], 2.0 ** -24)) / rate
    t = np.cumsum(gaps)
    n = int(np.searchsorted(t, seconds))
    w = np.round(np.exp(np.log(64) + u[:n, 1] * (np.log(1024) - np.log(64)))).astype(int)
    h = np.round(np.exp(np.log(64) + u[:n, 2] * (np.log(1024) - np.log(64)))).astype(int)
    mesh = np.minimum((u[:n, 3] * n_meshes).astype(int), n_meshes - 1)
    out = []
    cams: Dict[tuple, bytes] = {}
    for i in range(n):
        key = (int(w[i]), int(h[i]))
        if key not in cams:
            cams[key] = scenes.perspective_view(60.0, key[0] / key[1], 0.1, 10.0, 3.0)
        out.append(Request(i, float(t[i]), key[0], key[1], int(mesh[i]), cams[key]))
    return out


class _Live:
    """A request's device resources while it is queued or executing."""
    __slots__ = ("req", "color", "depth", "ubuf", "bg0")

    def __init__(self, req, color, depth, ubuf, bg0):
        self.req, self.color, self.depth, self.ubuf, self.bg0 = req, color, depth, ubuf, bg0

    def release(self):
        for h in (self.bg0, self.ubuf, self.color, self.depth):
            if h is not None:
                h.release()


class Server:
    """Executes a request trace against one device and one heap."""

    def __init__(self, device: Device, pool: MeshPool, heap_bytes: int, max_batch: int = 256,
                 max_inflight: int = 2):
        self.device = device
        self.pool = pool
        self.heap = device.create_heap(heap_bytes)
        self.max_batch = max_batch
        self.max_inflight = max_inflight

    def request_bytes(self, r: Request) -> int:
        gran = 256
        return 2 * ((r.width * r.height * 4 + gran - 1) // gran * gran) + gran

    def _admit(self, r: Request) -> Optional[_Live]:
        color = self.heap.create_texture(r.width, r.height, RGBA8_UNORM)
        if color is None:
            return None
        depth = self.heap.create_texture(r.width, r.height, DEPTH32_FLOAT)
        if depth is None:
            color.release()
            return None
        ubuf = self.heap.create_buffer(64)
        if ubuf is None:
            color.release()
            depth.release()
            return None
        self.device.write_buffer(ubuf, 0, r.uniform)
        bg0 = self.device.create_bind_group([(0, ubuf)])
        return _Live(r, color, depth, ubuf, bg0)

    def _record(self, enc, live: _Live):
        m = self.pool.meshes[live.req.mesh]
        rp = enc.begin_render_pass(color=live.color, clear_color=(0.1, 0.2, 0.3, 1.0), depth=live.depth,
                                   clear_depth=1.0)
        rp.set_pipeline(m.pipeline)
        rp.set_bind_group(0, live.bg0)
        if m.bind_group1 is not None:
            rp.set_bind_group(1, m.bind_group1)
        rp.set_vertex_buffer(0, m.vertex_buffer)
        rp.set_index_buffer(m.index_buffer, INDEX_UINT32)
        rp.draw_indexed(range(0, m.index_count))
        rp.end()

    def run(self, trace: Sequence[Request], batched: bool = True, keep: Sequence[int] = (),
            time_scale: float = 1.0) -> ServeStats:
        """Replays `trace` in real time.  `keep`: request ids whose colour and
        depth are read back on completion (stored in ``self.kept``)."""
        dev = self.device
        keep = set(keep)
        self.kept: Dict[int, tuple] = {}
        f0, s0 = dev.totals()
        queue: deque = deque()
        inflight: deque = deque()  # (submission index, [lives], command buffer)
        lat = np.zeros(len(trace))
        done_n = 0
        nxt = 0
        max_queue = 0
        max_heap = 0
        waits = 0
        max_inflight = self.max_inflight if batched else 1 << 30
        max_batch = self.max_batch if batched else 1
        t0 = time.perf_counter()
        while done_n < len(trace):
            now = (time.perf_counter() - t0) / time_scale
            while nxt < len(trace) and trace[nxt].arrival <= now:
                queue.append(trace[nxt])
                nxt += 1
            max_queue = max(max_queue, len(queue))
            # completions (in submission order)
            while inflight and dev.is_done(inflight[0][0]):
                _, lives, cb = inflight.popleft()
                t_done = (time.perf_counter() - t0) / time_scale
                cb.release()  # the recorded passes hold the targets until then
                for lv in lives:
                    lat[lv.req.rid] = (t_done - lv.req.arrival) * 1e3
                    if lv.req.rid in keep:
                        self.kept[lv.req.rid] = (lv.color.read(), lv.depth.read())
                    lv.release()
                done_n += len(lives)
            # launch when a slot is free
            while queue and len(inflight) < max_inflight:
                lives = []
                while queue and len(lives) < max_batch:
                    lv = self._admit(queue[0])
                    if lv is None:  # heap full: wait for completions
                        waits += 1
                        break
                    queue.popleft()
                    lives.append(lv)
                if not lives:
                    if not inflight:
                        raise RuntimeError(f"request {queue[0].rid} ({queue[0].width}x{queue[0].height}) "
                                           f"does not fit in an idle heap of {self.heap.size} bytes")
                    break
                used = self.heap.stats()[0]
                max_heap = max(max_heap, used)
                enc = dev.create_command_encoder()
                for lv in lives:
                    self._record(enc, lv)
                cb = enc.finish()
                idx = dev.submit([cb])
                inflight.append((idx, lives, cb))
            if not queue and not inflight and nxt < len(trace):
                # idle until the next arrival
                wait = trace[nxt].arrival * time_scale - (time.perf_counter() - t0)
                if wait > 0.0002:
                    time.sleep(wait - 0.0001)
        seconds = (time.perf_counter() - t0) / time_scale
        dev.poll()
        f1, s1 = dev.totals()
        return ServeStats("batched" if batched else "plain", len(trace), done_n, seconds, f1 - f0, s1 - s0,
                          lat, max_queue, max_heap, waits)
