"""Records and submits a :class:`wgc.scenes.PassSpec` through libwgc."""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import (DEPTH32_FLOAT, STEP_VERTEX, Device, PROGRAM_MVP_COLOR, PROGRAM_MVP_TEXTURE,
               RGBA8_UNORM, RGBA8_UNORM_SRGB)
from .scenes import PassSpec


@dataclass
class PassResult:
    color: Optional[np.ndarray]      # (H, W, 4) uint8
    depth: Optional[np.ndarray]      # (H, W) float32
    fragments: int
    frag_count: Optional[np.ndarray] = None


class PreparedPass:
    """Device-resident resources and a finished command buffer for a pass;
    `submit` may be called repeatedly (the attachments are re-cleared by the
    pass's load ops, or re-initialised from the spec for LoadOp::Load)."""

    def __init__(self, device: Device, spec: PassSpec):
        self.device = device
        self.spec = spec
        self._keep: List[object] = []
        w, h = spec.width, spec.height
        self.color = device.create_texture(w, h, spec.color_format) if spec.has_color else None
        self.depth = device.create_texture(w, h, DEPTH32_FLOAT) if spec.has_depth else None
        self._init_attachments()
        enc = device.create_command_encoder()
        rp = enc.begin_render_pass(color=self.color, clear_color=spec.clear_color,
                                   depth=self.depth, clear_depth=spec.clear_depth)
        for d in spec.draws:
            module = device.create_shader_module_builtin(d.program, d.interp)
            layouts = []
            if d.vertex_data:
                layouts = [(d.vertex_stride, STEP_VERTEX, d.attributes)]
            pipeline = device.create_render_pipeline(
                module, layouts, front_face=d.front_face, cull_mode=d.cull_mode, depth=d.depth,
                target_format=spec.color_format, blend=d.blend)
            rp.set_pipeline(pipeline)
            self._keep += [module, pipeline]
            if d.uniform is not None:
                ub = device.create_buffer_init(d.uniform)
                bg0 = device.create_bind_group([(0, ub)])
                rp.set_bind_group(0, bg0)
                self._keep += [ub, bg0]
            if d.texture is not None:
                tw, th, tbytes, srgb = d.texture
                tex = device.create_texture(tw, th, RGBA8_UNORM_SRGB if srgb else RGBA8_UNORM)
                device.write_texture(tex, tbytes)
                samp = device.create_sampler(d.address_u, d.address_v, d.filter, d.filter)
                bg1 = device.create_bind_group([(0, tex), (1, samp)])
                rp.set_bind_group(1, bg1)
                self._keep += [tex, samp, bg1]
            if d.vertex_data:
                vb = device.create_buffer_init(d.vertex_data)
                rp.set_vertex_buffer(0, vb)
                self._keep.append(vb)
            if d.viewport is not None:
                rp.set_viewport(*d.viewport)
            if d.scissor is not None:
                rp.set_scissor_rect(*d.scissor)
            if d.index_data is not None:
                ib = device.create_buffer_init(d.index_data)
                rp.set_index_buffer(ib, d.index_format)
                self._keep.append(ib)
                rp.draw_indexed(range(d.first, d.first + d.count), d.base_vertex,
                                range(d.first_instance, d.first_instance + d.instance_count))
            else:
                rp.draw(range(d.first, d.first + d.count),
                        range(d.first_instance, d.first_instance + d.instance_count))
        rp.end()
        self.command_buffer = enc.finish()

    def _init_attachments(self):
        spec = self.spec
        if self.color is not None and spec.color_init is not None:
            self.device.write_texture(self.color, spec.color_init)
        if self.depth is not None and spec.depth_init is not None:
            self.device.write_texture(self.depth, spec.depth_init)

    def submit(self) -> int:
        return self.device.submit([self.command_buffer])

    def result(self) -> PassResult:
        self.device.poll()
        w, h = self.spec.width, self.spec.height
        color = np.frombuffer(self.color.read(), np.uint8).reshape(h, w, 4) if self.color else None
        depth = np.frombuffer(self.depth.read(), np.float32).reshape(h, w) if self.depth else None
        frags = self.device.last_pass_stats().fragments
        return PassResult(color, depth, frags)


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
