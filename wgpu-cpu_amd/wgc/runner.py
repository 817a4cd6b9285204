"""Records and submits :class:`wgc.scenes.PassSpec` render passes through libwgc.

``PreparedPasses`` uploads every resource once, records all passes into one
command buffer and can submit it repeatedly; with several independent passes
the runtime merges them into shared launches (the batched form), launched
directly or replayed as a hipGraph (``OPTION_GRAPHS``).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import DEPTH32_FLOAT, RGBA8_UNORM, RGBA8_UNORM_SRGB, STEP_INSTANCE, STEP_VERTEX, Device
from .scenes import PassSpec


@dataclass
class PassResult:
    color: Optional[np.ndarray]      # (H, W, 4) uint8
    depth: Optional[np.ndarray]      # (H, W) float32
    fragments: int
    frag_count: Optional[np.ndarray] = None


class _PassResources:
    def __init__(self, device: Device, spec: PassSpec, keep: list):
        w, h = spec.width, spec.height
        self.spec = spec
        self.color = device.create_texture(w, h, spec.color_format) if spec.has_color else None
        self.depth = device.create_texture(w, h, DEPTH32_FLOAT) if spec.has_depth else None
        self.device = device
        self.init_attachments()

    def init_attachments(self):
        if self.color is not None and self.spec.color_init is not None:
            self.device.write_texture(self.color, self.spec.color_init)
        if self.depth is not None and self.spec.depth_init is not None:
            self.device.write_texture(self.depth, self.spec.depth_init)

    def read(self):
        w, h = self.spec.width, self.spec.height
        color = np.frombuffer(self.color.read(), np.uint8).reshape(h, w, 4) if self.color else None
        depth = np.frombuffer(self.depth.read(), np.float32).reshape(h, w) if self.depth else None
        return color, depth


def record_pass(device: Device, enc, spec: PassSpec, res: _PassResources, keep: list, cache: dict):
    """Records one pass; pipelines, textures and buffers are shared through
    `cache` when passes use identical content (resident resources)."""
    rp = enc.begin_render_pass(color=res.color, clear_color=spec.clear_color,
                               depth=res.depth, clear_depth=spec.clear_depth)

    def cached(key, make):
        if key not in cache:
            cache[key] = make()
            keep.append(cache[key])
        return cache[key]

    for d in spec.draws:
        if d.wgsl is not None:
            module = cached(("wgsl", d.wgsl), lambda: device.create_shader_module_wgsl(d.wgsl))
        else:
            module = cached(("module", d.program, d.interp),
                            lambda: device.create_shader_module_builtin(d.program, d.interp))
        layouts = [(d.vertex_stride, STEP_VERTEX, d.attributes)] if d.vertex_data else []
        if d.instance_buffer is not None:
            layouts.append((d.instance_buffer[1], STEP_INSTANCE, d.instance_buffer[2]))
        pkey = ("pipeline", d.program, d.interp, d.wgsl, d.vertex_stride, tuple(d.attributes), d.front_face,
                d.cull_mode, d.depth, spec.color_format, d.blend, d.topology, d.strip_index_format,
                None if d.instance_buffer is None else (d.instance_buffer[1], tuple(d.instance_buffer[2])))
        pipeline = cached(pkey, lambda: device.create_render_pipeline(
            module, layouts, front_face=d.front_face, cull_mode=d.cull_mode, depth=d.depth,
            target_format=spec.color_format, blend=d.blend, topology=d.topology,
            strip_index_format=d.strip_index_format))
        rp.set_pipeline(pipeline)
        if d.uniform is not None:
            ub = cached(("buf", d.uniform), lambda: device.create_buffer_init(d.uniform))
            bg0 = cached(("bg0", d.uniform), lambda: device.create_bind_group([(0, ub)]))
            rp.set_bind_group(0, bg0)
        if d.texture is not None:
            tw, th, tbytes, srgb = d.texture

            def make_tex():
                t = device.create_texture(tw, th, RGBA8_UNORM_SRGB if srgb else RGBA8_UNORM)
                device.write_texture(t, tbytes)
                return t
            tex = cached(("tex", tw, th, tbytes, srgb), make_tex)
            samp = cached(("samp", d.address_u, d.address_v, d.filter),
                          lambda: device.create_sampler(d.address_u, d.address_v, d.filter, d.filter))
            bg1 = cached(("bg1", id(tex), id(samp)), lambda: device.create_bind_group([(0, tex), (1, samp)]))
            rp.set_bind_group(1, bg1)
        if d.vertex_data:
            vb = cached(("buf", d.vertex_data), lambda: device.create_buffer_init(d.vertex_data))
            rp.set_vertex_buffer(0, vb)
        if d.instance_buffer is not None:
            ib_data = d.instance_buffer[0]
            instb = cached(("buf", ib_data), lambda: device.create_buffer_init(ib_data))
            rp.set_vertex_buffer(len(layouts) - 1, instb)
        if d.viewport is not None:
            rp.set_viewport(*d.viewport)
        if d.scissor is not None:
            rp.set_scissor_rect(*d.scissor)
        if d.index_data is not None:
            ib = cached(("buf", d.index_data), lambda: device.create_buffer_init(d.index_data))
            rp.set_index_buffer(ib, d.index_format)
            rp.draw_indexed(range(d.first, d.first + d.count), d.base_vertex,
                            range(d.first_instance, d.first_instance + d.instance_count))
        else:
            rp.draw(range(d.first, d.first + d.count),
                    range(d.first_instance, d.first_instance + d.instance_count))
    rp.end()


class PreparedPasses:
    """Device-resident resources and one finished command buffer holding all
    `specs` in order; `submit` may be called repeatedly."""

    def __init__(self, device: Device, specs: Sequence[PassSpec]):
        self.device = device
        self.specs = list(specs)
        self._keep: List[object] = []
        cache: dict = {}
        self.res = [_PassResources(device, s, self._keep) for s in self.specs]
        enc = device.create_command_encoder()
        for s, r in zip(self.specs, self.res):
            record_pass(device, enc, s, r, self._keep, cache)
        self.command_buffer = enc.finish()

    def reset_attachments(self):
        for r in self.res:
            r.init_attachments()

    def submit(self) -> int:
        return self.device.submit([self.command_buffer])

    def results(self) -> List[PassResult]:
        self.device.poll()
        out = []
        for r in self.res:
            c, d = r.read()
            out.append(PassResult(c, d, -1))
        return out

    def total_fragments(self) -> int:
        return self.device.last_pass_stats().fragments


class PreparedPass(PreparedPasses):
    def __init__(self, device: Device, spec: PassSpec):
        super().__init__(device, [spec])

    def result(self) -> PassResult:
        r = self.results()[0]
        r.fragments = self.total_fragments()
        return r


def render(spec: PassSpec, device: Optional[Device] = None, fragment_counts: bool = False) -> PassResult:
    """Executes one pass on the GPU and reads the attachments back."""
    dev = device if device is not None else Device(0, spec.semantics)
    if fragment_counts:
        dev.enable_fragment_counts(spec.width, spec.height)
    pp = PreparedPass(dev, spec)
    pp.submit()
    res = pp.result()
    if fragment_counts:
        res.frag_count = dev.fragment_counts(spec.width, spec.height).reshape(spec.height, spec.width)
    return res


def render_batch(specs: Sequence[PassSpec], device: Device) -> List[PassResult]:
    """Executes all passes in one submission (the batched form)."""
    pp = PreparedPasses(device, specs)
    pp.submit()
    res = pp.results()
    total = pp.total_fragments()
    for r in res:
        r.fragments = -1
    res_total = total
    return res, res_total
