"""wgc -- Python mirror of wgpu's device / queue / render-pass interface over
libwgc, the MI355X render-pass executor.

The object model follows the wgpu custom-backend interface that
jgraef/wgpu-cpu implements (wgpu-cpu/src/device.rs:82-478,
render_pass/mod.rs:73-329): a Device creates buffers, textures, samplers,
bind groups, shader modules and render pipelines; a CommandEncoder records
render passes; ``Device.submit`` enqueues them and ``Device.poll`` waits.
Errors that the reference reports by panicking raise :class:`WgcError`.

libwgc.so must have been built (``wgc._build.build_library()``); importing
this module without it raises, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import Optional, Sequence

# WGC_LIBRARY names another build of the library (for example an
# instrumented one); the default is the in-tree build next to this file.
_LIB_PATH = Path(os.environ.get("WGC_LIBRARY") or Path(__file__).resolve().parent / "libwgc.so")

# ---- enums (include/wgc.h) ----
SEMANTICS_REFERENCE = 0
SEMANTICS_EXTENDED = 1

OK = 0
ERR_INVALID = -1
ERR_UNSUPPORTED = -2
ERR_OUT_OF_MEMORY = -3
ERR_DEVICE = -4
ERR_TIMEOUT = -5
ERR_EXECUTION = -6

# WGSL entry-point runs (wgc_shader_module_run)
RUN_WORDS = 69
ABORT_KILL = 3
ABORT_OUT_OF_BOUNDS = 4
ABORT_DIVISION_BY_ZERO = 5
ABORT_OVERFLOW = 6
PROGRAM_WGSL = 5

RGBA8_UNORM = 0
RGBA8_UNORM_SRGB = 1
BGRA8_UNORM = 2
BGRA8_UNORM_SRGB = 3
DEPTH32_FLOAT = 4

FLOAT32 = 1
FLOAT32X2 = 2
FLOAT32X3 = 3
FLOAT32X4 = 4

STEP_VERTEX = 0
STEP_INSTANCE = 1
INDEX_UINT16 = 2
INDEX_UINT32 = 4
POINT_LIST = 0
LINE_LIST = 1
LINE_STRIP = 2
TRIANGLE_LIST = 3
TRIANGLE_STRIP = 4
FRONT_CCW = 0
FRONT_CW = 1
CULL_NONE = 0
CULL_FRONT = 1
CULL_BACK = 2

NEVER, LESS, EQUAL, LESS_EQUAL, GREATER, NOT_EQUAL, GREATER_EQUAL, ALWAYS = range(1, 9)

CLAMP_TO_EDGE = 0
REPEAT = 1
MIRROR_REPEAT = 2
FILTER_NEAREST = 0
FILTER_LINEAR = 1
LOAD = 0
CLEAR = 1
BLEND_REPLACE = 0
BLEND_ALPHA = 1
INTERP_LINEAR = 0
INTERP_PERSPECTIVE = 1
INTERP_FLAT = 2

PROGRAM_MVP_COLOR = 1
PROGRAM_MVP_TEXTURE = 2
PROGRAM_COLORED_TRIANGLE = 3
PROGRAM_CLIP_COLOR = 4

OPTION_RASTER_TILE = 1
OPTION_BATCHING = 2
OPTION_GRAPHS = 3
# raster tile shapes (tile side, lanes per workgroup)
RASTER_TILES = {0: (16, 256), 1: (8, 64), 2: (16, 64), 3: (32, 256), 4: (16, 128), 5: (32, 128), 6: (32, 512),
                7: (32, 64)}

BINDING_BUFFER = 1
BINDING_TEXTURE = 2
BINDING_SAMPLER = 3


class WgcError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"wgc error {code}: {message}")
        self.code = code


# ---- C structures ----
class DeviceDescriptor(C.Structure):
    _fields_ = [("hip_device", C.c_int32), ("semantics", C.c_uint32)]


class SamplerDescriptor(C.Structure):
    _fields_ = [("address_mode_u", C.c_uint32), ("address_mode_v", C.c_uint32),
                ("mag_filter", C.c_uint32), ("min_filter", C.c_uint32)]


class BindGroupEntry(C.Structure):
    _fields_ = [("binding", C.c_uint32), ("kind", C.c_uint32), ("buffer", C.c_void_p),
                ("offset", C.c_uint64), ("size", C.c_uint64), ("texture", C.c_void_p),
                ("sampler", C.c_void_p)]


class VertexAttribute(C.Structure):
    _fields_ = [("format", C.c_uint32), ("offset", C.c_uint32), ("shader_location", C.c_uint32)]


class VertexBufferLayout(C.Structure):
    _fields_ = [("array_stride", C.c_uint32), ("step_mode", C.c_uint32),
                ("attribute_count", C.c_uint32), ("attributes", C.POINTER(VertexAttribute))]


class RenderPipelineDescriptor(C.Structure):
    _fields_ = [("vertex_module", C.c_void_p), ("vertex_buffer_count", C.c_uint32),
                ("vertex_buffers", C.POINTER(VertexBufferLayout)), ("topology", C.c_uint32),
                ("front_face", C.c_uint32), ("cull_mode", C.c_uint32),
                ("has_depth_stencil", C.c_uint32), ("depth_format", C.c_uint32),
                ("depth_write_enabled", C.c_uint32), ("depth_compare", C.c_uint32),
                ("fragment_module", C.c_void_p), ("target_format", C.c_uint32),
                ("blend", C.c_uint32), ("vertex_entry", C.c_char_p), ("fragment_entry", C.c_char_p),
                ("strip_index_format", C.c_uint32)]


class ColorAttachment(C.Structure):
    _fields_ = [("view", C.c_void_p), ("load_op", C.c_uint32), ("clear_value", C.c_double * 4)]


class DepthStencilAttachment(C.Structure):
    _fields_ = [("view", C.c_void_p), ("has_depth_ops", C.c_uint32), ("depth_load_op", C.c_uint32),
                ("depth_clear_value", C.c_float)]


class RenderPassDescriptor(C.Structure):
    _fields_ = [("color_attachment_count", C.c_uint32),
                ("color_attachments", C.POINTER(ColorAttachment)),
                ("depth_stencil_attachment", C.POINTER(DepthStencilAttachment))]


class TexelCopyBuffer(C.Structure):
    _fields_ = [("buffer", C.c_void_p), ("offset", C.c_uint64), ("bytes_per_row", C.c_uint32)]


class TexelCopyTexture(C.Structure):
    _fields_ = [("texture", C.c_void_p), ("origin_x", C.c_uint32), ("origin_y", C.c_uint32)]


MAP_CALLBACK = C.CFUNCTYPE(None, C.c_int32, C.c_void_p)
_MAP_THUNKS: dict = {}   # callbacks of pending mapping requests
MAP_READ = 1
MAP_WRITE = 2


class PassStats(C.Structure):
    _fields_ = [("fragments", C.c_uint64), ("primitives", C.c_uint64), ("kernel_launches", C.c_uint64)]


# name -> (restype, argtypes); every symbol declared in include/wgc.h
_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
SIGNATURES = {
    "wgc_abi_version": (C.c_uint32, []),
    "wgc_last_error": (C.c_char_p, []),
    "wgc_device_create": (C.c_int32, [C.POINTER(DeviceDescriptor), _PP]),
    "wgc_device_release": (None, [_P]),
    "wgc_buffer_create": (C.c_int32, [_P, C.c_uint64, _PP]),
    "wgc_queue_write_buffer": (C.c_int32, [_P, _P, C.c_uint64, C.c_void_p, C.c_uint64]),
    "wgc_buffer_read": (C.c_int32, [_P, C.c_uint64, C.c_void_p, C.c_uint64]),
    "wgc_buffer_release": (None, [_P]),
    "wgc_buffer_map_async": (C.c_int32, [_P, C.c_uint32, C.c_uint64, C.c_uint64, MAP_CALLBACK, C.c_void_p]),
    "wgc_buffer_get_mapped_range": (C.c_int32, [_P, C.c_uint64, C.c_uint64, C.POINTER(C.c_void_p)]),
    "wgc_buffer_unmap": (C.c_int32, [_P]),
    "wgc_command_encoder_copy_buffer_to_buffer": (C.c_int32, [_P, _P, C.c_uint64, _P, C.c_uint64, C.c_uint64]),
    "wgc_command_encoder_copy_texture_to_buffer": (C.c_int32, [_P, C.POINTER(TexelCopyTexture),
                                                               C.POINTER(TexelCopyBuffer), C.c_uint32, C.c_uint32]),
    "wgc_command_encoder_copy_buffer_to_texture": (C.c_int32, [_P, C.POINTER(TexelCopyBuffer),
                                                               C.POINTER(TexelCopyTexture), C.c_uint32, C.c_uint32]),
    "wgc_command_encoder_copy_texture_to_texture": (C.c_int32, [_P, C.POINTER(TexelCopyTexture),
                                                                C.POINTER(TexelCopyTexture), C.c_uint32, C.c_uint32]),
    "wgc_command_encoder_clear_buffer": (C.c_int32, [_P, _P, C.c_uint64, C.c_uint64]),
    "wgc_command_encoder_clear_texture": (C.c_int32, [_P, _P]),
    "wgc_texture_create": (C.c_int32, [_P, C.c_uint32, C.c_uint32, C.c_uint32, _PP]),
    "wgc_queue_write_texture": (C.c_int32, [_P, _P, C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint32]),
    "wgc_texture_read": (C.c_int32, [_P, C.c_void_p, C.c_uint64]),
    "wgc_texture_release": (None, [_P]),
    "wgc_heap_create": (C.c_int32, [_P, C.c_uint64, _PP]),
    "wgc_heap_release": (None, [_P]),
    "wgc_heap_stats": (C.c_int32, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "wgc_buffer_create_in_heap": (C.c_int32, [_P, C.c_uint64, _PP]),
    "wgc_texture_create_in_heap": (C.c_int32, [_P, C.c_uint32, C.c_uint32, C.c_uint32, _PP]),
    "wgc_sampler_create": (C.c_int32, [_P, C.POINTER(SamplerDescriptor), _PP]),
    "wgc_sampler_release": (None, [_P]),
    "wgc_bind_group_create": (C.c_int32, [_P, C.c_uint32, C.POINTER(BindGroupEntry), _PP]),
    "wgc_bind_group_release": (None, [_P]),
    "wgc_shader_module_create_builtin": (C.c_int32, [_P, C.c_uint32, C.c_uint32, _PP]),
    "wgc_shader_module_release": (None, [_P]),
    "wgc_shader_module_create_wgsl": (C.c_int32, [_P, C.c_char_p, C.c_size_t, _PP]),
    "wgc_shader_module_run": (C.c_int32, [_P, _P, C.c_char_p, C.c_uint32, C.c_uint32, _PP, C.c_uint32,
                                          C.POINTER(C.c_uint32), C.c_size_t]),
    "wgc_wgsl_translate": (C.c_int32, [C.c_char_p, C.c_size_t, C.c_uint32, C.c_char_p, C.c_size_t,
                                       C.POINTER(C.c_size_t)]),
    "wgc_wgsl_check_pipeline": (C.c_int32, [C.c_char_p, C.c_size_t, C.c_char_p, C.c_char_p, C.c_uint32]),
    "wgc_render_pipeline_create": (C.c_int32, [_P, C.POINTER(RenderPipelineDescriptor), _PP]),
    "wgc_render_pipeline_release": (None, [_P]),
    "wgc_command_encoder_create": (C.c_int32, [_P, _PP]),
    "wgc_command_encoder_begin_render_pass": (C.c_int32, [_P, C.POINTER(RenderPassDescriptor), _PP]),
    "wgc_render_pass_set_pipeline": (C.c_int32, [_P, _P]),
    "wgc_render_pass_set_bind_group": (C.c_int32, [_P, C.c_uint32, _P]),
    "wgc_render_pass_set_vertex_buffer": (C.c_int32, [_P, C.c_uint32, _P, C.c_uint64, C.c_uint64]),
    "wgc_render_pass_set_index_buffer": (C.c_int32, [_P, _P, C.c_uint32, C.c_uint64, C.c_uint64]),
    "wgc_render_pass_set_viewport": (C.c_int32, [_P] + [C.c_float] * 6),
    "wgc_render_pass_set_scissor_rect": (C.c_int32, [_P] + [C.c_uint32] * 4),
    "wgc_render_pass_set_blend_constant": (C.c_int32, [_P, C.POINTER(C.c_double)]),
    "wgc_render_pass_set_stencil_reference": (C.c_int32, [_P, C.c_uint32]),
    "wgc_render_pass_draw": (C.c_int32, [_P] + [C.c_uint32] * 4),
    "wgc_render_pass_draw_indexed": (C.c_int32, [_P, C.c_uint32, C.c_uint32, C.c_int32, C.c_uint32, C.c_uint32]),
    "wgc_render_pass_end": (C.c_int32, [_P]),
    "wgc_command_encoder_finish": (C.c_int32, [_P, _PP]),
    "wgc_command_encoder_release": (None, [_P]),
    "wgc_command_buffer_release": (None, [_P]),
    "wgc_queue_submit": (C.c_int32, [_P, C.c_uint32, _PP, C.POINTER(C.c_uint64)]),
    "wgc_queue_submit_noalloc": (C.c_int32, [_P, C.c_uint32, _PP, C.POINTER(C.c_uint64)]),
    "wgc_device_poll": (C.c_int32, [_P, C.c_uint64, C.c_uint64]),
    "wgc_device_last_pass_stats": (C.c_int32, [_P, C.POINTER(PassStats)]),
    "wgc_device_totals": (C.c_int32, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "wgc_device_set_option": (C.c_int32, [_P, C.c_uint32, C.c_uint32]),
    "wgc_device_get_option": (C.c_int32, [_P, C.c_uint32, C.POINTER(C.c_uint32)]),
    "wgc_device_debug_fragment_counts": (C.c_int32, [_P, C.c_uint32, C.c_uint32, C.c_void_p]),
    "wgc_device_timestamp": (C.c_int32, [_P, C.c_uint32]),
    "wgc_device_elapsed_ms": (C.c_int32, [_P, C.c_uint32, C.c_uint32, C.POINTER(C.c_float)]),
    "wgc_device_set_kernel_timing": (C.c_int32, [_P, C.c_uint32]),
    "wgc_device_kernel_time": (C.c_int32, [_P, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
}

_lib: Optional[C.CDLL] = None


def wgsl_translate(source: str, what: int = 0) -> str:
    """Device-less translation: what=0 the emitted C++, what=1 the interface
    as JSON (raises WgcError for invalid or unsupported WGSL)."""
    lib = load_library()
    data = source.encode()
    need = C.c_size_t()
    _check(lib.wgc_wgsl_translate(data, len(data), what, None, 0, C.byref(need)))
    buf = C.create_string_buffer(need.value)
    _check(lib.wgc_wgsl_translate(data, len(data), what, buf, need.value, C.byref(need)))
    return buf.value.decode()


def wgsl_check_pipeline(source: str, vertex_entry: Optional[str] = None, fragment_entry: Optional[str] = None,
                        semantics: int = SEMANTICS_REFERENCE) -> None:
    """Compiles a pipeline's kernels from the module for gfx950 without a
    device (raises WgcError on failure)."""
    lib = load_library()
    data = source.encode()
    _check(lib.wgc_wgsl_check_pipeline(data, len(data), vertex_entry.encode() if vertex_entry else None,
                                       fragment_entry.encode() if fragment_entry else None, semantics))


def library_path() -> Path:
    return _LIB_PATH


def load_library() -> C.CDLL:
    """Loads libwgc.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise ImportError(f"{_LIB_PATH} is missing: build it with wgc._build.build_library()")
        lib = C.CDLL(str(_LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _check(code: int) -> None:
    if code != 0:
        msg = load_library().wgc_last_error()
        raise WgcError(code, msg.decode() if msg else "")


def _out() -> C.c_void_p:
    return C.c_void_p()


class _Handle:
    _release = ""

    def __init__(self, ptr: C.c_void_p):
        self._ptr = ptr

    @property
    def ptr(self) -> C.c_void_p:
        return self._ptr

    def release(self) -> None:
        if self._ptr and self._ptr.value:
            getattr(load_library(), self._release)(self._ptr)
            self._ptr = C.c_void_p()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Buffer(_Handle):
    _release = "wgc_buffer_release"

    def __init__(self, ptr, size):
        super().__init__(ptr)
        self.size = size

    def read(self, offset: int = 0, size: Optional[int] = None) -> bytes:
        n = self.size - offset if size is None else size
        buf = (C.c_uint8 * n)()
        _check(load_library().wgc_buffer_read(self._ptr, offset, buf, n))
        return bytes(buf)

    # mapping (BufferInterface::map_async / get_mapped_range / unmap)
    def map_async(self, mode: int, offset: int = 0, size: Optional[int] = None, callback=None):
        """callback(status) runs from Device.poll once the range is mapped."""
        n = self.size - offset if size is None else size

        key = object()

        def trampoline(status, _ud):
            _MAP_THUNKS.pop(key, None)
            if callback is not None:
                callback(status)
        thunk = MAP_CALLBACK(trampoline)
        _check(load_library().wgc_buffer_map_async(self._ptr, mode, offset, n, thunk, None))
        # the library calls the thunk exactly once (mapped or aborted), maybe
        # after this Buffer object is gone: keep it alive until then
        _MAP_THUNKS[key] = thunk

    def get_mapped_range(self, offset: int = 0, size: Optional[int] = None):
        """A writable view (ctypes array) of the mapped bytes."""
        n = self.size - offset if size is None else size
        p = C.c_void_p()
        _check(load_library().wgc_buffer_get_mapped_range(self._ptr, offset, n, C.byref(p)))
        return (C.c_uint8 * n).from_address(p.value)

    def unmap(self):
        _check(load_library().wgc_buffer_unmap(self._ptr))


class Texture(_Handle):
    _release = "wgc_texture_release"

    def __init__(self, ptr, width, height, format):
        super().__init__(ptr)
        self.width, self.height, self.format = width, height, format

    @property
    def size_bytes(self) -> int:
        return self.width * self.height * 4

    def read(self) -> bytes:
        """wgpu_cpu::image::rgba_texture_image (lib.rs:158-173): waits, copies."""
        buf = (C.c_uint8 * self.size_bytes)()
        _check(load_library().wgc_texture_read(self._ptr, buf, self.size_bytes))
        return bytes(buf)


class Sampler(_Handle):
    _release = "wgc_sampler_release"


class Heap(_Handle):
    """Device memory pool for resources created at a high rate (no reference
    counterpart).  Creation returns None when the heap is full, so that a
    caller can hold work back until memory frees up."""
    _release = "wgc_heap_release"

    def __init__(self, ptr, size):
        super().__init__(ptr)
        self.size = size

    def stats(self):
        used, cap, largest = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(load_library().wgc_heap_stats(self._ptr, C.byref(used), C.byref(cap), C.byref(largest)))
        return used.value, cap.value, largest.value

    def create_texture(self, width: int, height: int, format: int) -> Optional["Texture"]:
        out = _out()
        code = load_library().wgc_texture_create_in_heap(self._ptr, width, height, format, C.byref(out))
        if code == ERR_OUT_OF_MEMORY:
            return None
        _check(code)
        return Texture(out, width, height, format)

    def create_buffer(self, size: int) -> Optional["Buffer"]:
        out = _out()
        code = load_library().wgc_buffer_create_in_heap(self._ptr, size, C.byref(out))
        if code == ERR_OUT_OF_MEMORY:
            return None
        _check(code)
        return Buffer(out, size)


class BindGroup(_Handle):
    _release = "wgc_bind_group_release"


class ShaderModule(_Handle):
    _release = "wgc_shader_module_release"


class RenderPipeline(_Handle):
    _release = "wgc_render_pipeline_release"


class CommandBuffer(_Handle):
    _release = "wgc_command_buffer_release"


class RenderPass:
    """RenderPassInterface (render_pass/mod.rs:73-322)."""

    def __init__(self, ptr):
        self._ptr = ptr
        self._lib = load_library()

    def set_pipeline(self, pipeline: RenderPipeline):
        _check(self._lib.wgc_render_pass_set_pipeline(self._ptr, pipeline.ptr))

    def set_bind_group(self, index: int, group: Optional[BindGroup]):
        _check(self._lib.wgc_render_pass_set_bind_group(self._ptr, index, group.ptr if group else None))

    def set_vertex_buffer(self, slot: int, buffer: Buffer, offset: int = 0, size: int = 0):
        _check(self._lib.wgc_render_pass_set_vertex_buffer(self._ptr, slot, buffer.ptr, offset, size))

    def set_index_buffer(self, buffer: Buffer, format: int, offset: int = 0, size: int = 0):
        _check(self._lib.wgc_render_pass_set_index_buffer(self._ptr, buffer.ptr, format, offset, size))

    def set_viewport(self, x, y, w, h, min_depth=0.0, max_depth=1.0):
        _check(self._lib.wgc_render_pass_set_viewport(self._ptr, x, y, w, h, min_depth, max_depth))

    def set_scissor_rect(self, x, y, w, h):
        _check(self._lib.wgc_render_pass_set_scissor_rect(self._ptr, x, y, w, h))

    def draw(self, vertices: range, instances: range = range(0, 1)):
        _check(self._lib.wgc_render_pass_draw(self._ptr, vertices.start, len(vertices),
                                              instances.start, len(instances)))

    def draw_indexed(self, indices: range, base_vertex: int = 0, instances: range = range(0, 1)):
        _check(self._lib.wgc_render_pass_draw_indexed(self._ptr, indices.start, len(indices), base_vertex,
                                                      instances.start, len(instances)))

    def end(self):
        if self._ptr:
            p, self._ptr = self._ptr, None
            _check(self._lib.wgc_render_pass_end(p))


class CommandEncoder:
    def __init__(self, ptr):
        self._ptr = ptr
        self._lib = load_library()
        self._keep = []

    def begin_render_pass(self, color: Optional[Texture] = None, clear_color=None,
                          depth: Optional[Texture] = None, clear_depth: Optional[float] = None,
                          depth_ops: bool = True) -> RenderPass:
        """color/depth attachments; clear_* None means LoadOp::Load."""
        ca = ColorAttachment()
        desc = RenderPassDescriptor()
        if color is not None:
            ca.view = color.ptr
            ca.load_op = LOAD if clear_color is None else CLEAR
            for k, v in enumerate(clear_color or (0, 0, 0, 0)):
                ca.clear_value[k] = v
            desc.color_attachment_count = 1
            desc.color_attachments = C.pointer(ca)
        da = DepthStencilAttachment()
        if depth is not None:
            da.view = depth.ptr
            da.has_depth_ops = 1 if depth_ops else 0
            da.depth_load_op = LOAD if clear_depth is None else CLEAR
            da.depth_clear_value = 0.0 if clear_depth is None else clear_depth
            desc.depth_stencil_attachment = C.pointer(da)
        out = _out()
        _check(self._lib.wgc_command_encoder_begin_render_pass(self._ptr, C.byref(desc), C.byref(out)))
        return RenderPass(out)

    def finish(self) -> CommandBuffer:
        out = _out()
        _check(self._lib.wgc_command_encoder_finish(self._ptr, C.byref(out)))
        return CommandBuffer(out)

    # copies and clears, executed in order with the passes
    def copy_buffer_to_buffer(self, src: Buffer, src_offset: int, dst: Buffer, dst_offset: int, size: int):
        _check(self._lib.wgc_command_encoder_copy_buffer_to_buffer(self._ptr, src.ptr, src_offset, dst.ptr,
                                                                   dst_offset, size))

    def copy_texture_to_buffer(self, src: Texture, origin, dst: Buffer, offset: int, bytes_per_row: int,
                               width: int, height: int):
        t = TexelCopyTexture(src.ptr, origin[0], origin[1])
        b = TexelCopyBuffer(dst.ptr, offset, bytes_per_row)
        _check(self._lib.wgc_command_encoder_copy_texture_to_buffer(self._ptr, C.byref(t), C.byref(b), width,
                                                                    height))

    def copy_buffer_to_texture(self, src: Buffer, offset: int, bytes_per_row: int, dst: Texture, origin,
                               width: int, height: int):
        b = TexelCopyBuffer(src.ptr, offset, bytes_per_row)
        t = TexelCopyTexture(dst.ptr, origin[0], origin[1])
        _check(self._lib.wgc_command_encoder_copy_buffer_to_texture(self._ptr, C.byref(b), C.byref(t), width,
                                                                    height))

    def copy_texture_to_texture(self, src: Texture, src_origin, dst: Texture, dst_origin, width: int, height: int):
        a = TexelCopyTexture(src.ptr, src_origin[0], src_origin[1])
        b = TexelCopyTexture(dst.ptr, dst_origin[0], dst_origin[1])
        _check(self._lib.wgc_command_encoder_copy_texture_to_texture(self._ptr, C.byref(a), C.byref(b), width,
                                                                     height))

    def clear_buffer(self, buffer: Buffer, offset: int = 0, size: int = 0):
        _check(self._lib.wgc_command_encoder_clear_buffer(self._ptr, buffer.ptr, offset, size))

    def clear_texture(self, texture: Texture):
        _check(self._lib.wgc_command_encoder_clear_texture(self._ptr, texture.ptr))

    def __del__(self):
        try:
            if self._ptr:
                self._lib.wgc_command_encoder_release(self._ptr)
                self._ptr = None
        except Exception:
            pass


class Device:
    """Instance + adapter + device + queue (lib.rs:25-27, device.rs:40-74)."""

    def __init__(self, hip_device: int = 0, semantics: int = SEMANTICS_REFERENCE):
        self._lib = load_library()
        desc = DeviceDescriptor(hip_device, semantics)
        out = _out()
        _check(self._lib.wgc_device_create(C.byref(desc), C.byref(out)))
        self._ptr = out
        self.semantics = semantics

    @property
    def ptr(self):
        return self._ptr

    def close(self):
        if self._ptr:
            self._lib.wgc_device_release(self._ptr)
            self._ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # resources
    def create_buffer(self, size: int) -> Buffer:
        out = _out()
        _check(self._lib.wgc_buffer_create(self._ptr, size, C.byref(out)))
        return Buffer(out, size)

    def create_heap(self, size: int) -> Heap:
        out = _out()
        _check(self._lib.wgc_heap_create(self._ptr, size, C.byref(out)))
        return Heap(out, size)

    def create_buffer_init(self, data: bytes) -> Buffer:
        b = self.create_buffer(len(data))
        self.write_buffer(b, 0, data)
        return b

    def write_buffer(self, buffer: Buffer, offset: int, data: bytes):
        _check(self._lib.wgc_queue_write_buffer(self._ptr, buffer.ptr, offset, data, len(data)))

    def create_texture(self, width: int, height: int, format: int) -> Texture:
        out = _out()
        _check(self._lib.wgc_texture_create(self._ptr, width, height, format, C.byref(out)))
        return Texture(out, width, height, format)

    def write_texture(self, texture: Texture, data: bytes, bytes_per_row: Optional[int] = None, offset: int = 0):
        bpr = texture.width * 4 if bytes_per_row is None else bytes_per_row
        _check(self._lib.wgc_queue_write_texture(self._ptr, texture.ptr, data, len(data), offset, bpr))

    def create_sampler(self, address_mode_u=CLAMP_TO_EDGE, address_mode_v=CLAMP_TO_EDGE,
                       mag_filter=FILTER_NEAREST, min_filter=FILTER_NEAREST) -> Sampler:
        desc = SamplerDescriptor(address_mode_u, address_mode_v, mag_filter, min_filter)
        out = _out()
        _check(self._lib.wgc_sampler_create(self._ptr, C.byref(desc), C.byref(out)))
        return Sampler(out)

    def create_bind_group(self, entries: Sequence[tuple]) -> BindGroup:
        """entries: (binding, Buffer | Texture | Sampler[, offset, size])."""
        arr = (BindGroupEntry * max(1, len(entries)))()
        for i, e in enumerate(entries):
            binding, res = e[0], e[1]
            arr[i].binding = binding
            if isinstance(res, Buffer):
                arr[i].kind = BINDING_BUFFER
                arr[i].buffer = res.ptr
                arr[i].offset = e[2] if len(e) > 2 else 0
                arr[i].size = e[3] if len(e) > 3 else 0
            elif isinstance(res, Texture):
                arr[i].kind = BINDING_TEXTURE
                arr[i].texture = res.ptr
            elif isinstance(res, Sampler):
                arr[i].kind = BINDING_SAMPLER
                arr[i].sampler = res.ptr
            else:
                raise TypeError(res)
        out = _out()
        _check(self._lib.wgc_bind_group_create(self._ptr, len(entries), arr, C.byref(out)))
        return BindGroup(out)

    def create_shader_module_builtin(self, program: int, interpolation: int = INTERP_LINEAR) -> ShaderModule:
        out = _out()
        _check(self._lib.wgc_shader_module_create_builtin(self._ptr, program, interpolation, C.byref(out)))
        return ShaderModule(out)

    def create_shader_module_wgsl(self, source: str) -> ShaderModule:
        """A module from WGSL source (raises WgcError on invalid or unsupported WGSL)."""
        data = source.encode()
        out = _out()
        _check(self._lib.wgc_shader_module_create_wgsl(self._ptr, data, len(data), C.byref(out)))
        return ShaderModule(out)

    def run_entry_point(self, module: ShaderModule, entry: Optional[str] = None, first_index: int = 0,
                        count: int = 1, groups: Sequence[Optional["BindGroup"]] = ()):
        """Runs a WGSL entry point for `count` invocations on the device; returns
        (count, RUN_WORDS) uint32: abort code, position / frag depth, 16 x 4
        location words."""
        import numpy as np
        res = np.zeros((count, RUN_WORDS), np.uint32)
        arr = (C.c_void_p * 4)(*[(g.ptr.value if g is not None else None) for g in groups] + [None] * (4 - len(groups)))
        _check(self._lib.wgc_shader_module_run(self._ptr, module.ptr, entry.encode() if entry else None, first_index,
                                               count, arr, len(groups),
                                               res.ctypes.data_as(C.POINTER(C.c_uint32)), res.size))
        return res

    def create_render_pipeline(self, module: ShaderModule, vertex_buffers: Sequence[tuple] = (),
                               front_face=FRONT_CCW, cull_mode=CULL_NONE, depth: Optional[tuple] = None,
                               target_format=RGBA8_UNORM, blend=BLEND_REPLACE, fragment=True,
                               vertex_entry: Optional[str] = None, fragment_entry: Optional[str] = None,
                               fragment_module: Optional[ShaderModule] = None, topology: int = TRIANGLE_LIST,
                               strip_index_format: int = 0) -> RenderPipeline:
        """vertex_buffers: (array_stride, step_mode, [(format, offset, location), ...]);
        depth: (compare, write_enabled) or None.  WGSL modules: entry point
        names (None: the module's only one of that stage); `fragment_module`
        defaults to `module`."""
        layouts = (VertexBufferLayout * max(1, len(vertex_buffers)))()
        keep = []
        for i, (stride, step, attrs) in enumerate(vertex_buffers):
            a = (VertexAttribute * max(1, len(attrs)))()
            for j, (fmt, off, loc) in enumerate(attrs):
                a[j] = VertexAttribute(fmt, off, loc)
            keep.append(a)
            layouts[i].array_stride = stride
            layouts[i].step_mode = step
            layouts[i].attribute_count = len(attrs)
            layouts[i].attributes = C.cast(a, C.POINTER(VertexAttribute))
        desc = RenderPipelineDescriptor()
        desc.vertex_module = module.ptr
        desc.vertex_buffer_count = len(vertex_buffers)
        desc.vertex_buffers = C.cast(layouts, C.POINTER(VertexBufferLayout))
        desc.topology = topology
        desc.strip_index_format = strip_index_format
        desc.front_face = front_face
        desc.cull_mode = cull_mode
        if depth is not None:
            desc.has_depth_stencil = 1
            desc.depth_format = DEPTH32_FLOAT
            desc.depth_compare = depth[0]
            desc.depth_write_enabled = 1 if depth[1] else 0
        desc.fragment_module = (fragment_module or module).ptr if fragment else None
        desc.target_format = target_format
        desc.blend = blend
        desc.vertex_entry = vertex_entry.encode() if vertex_entry else None
        desc.fragment_entry = fragment_entry.encode() if fragment_entry else None
        out = _out()
        _check(self._lib.wgc_render_pipeline_create(self._ptr, C.byref(desc), C.byref(out)))
        return RenderPipeline(out)

    def create_command_encoder(self) -> CommandEncoder:
        out = _out()
        _check(self._lib.wgc_command_encoder_create(self._ptr, C.byref(out)))
        return CommandEncoder(out)

    # queue
    def submit(self, buffers: Sequence[CommandBuffer]) -> int:
        arr = (C.c_void_p * max(1, len(buffers)))(*[b.ptr.value for b in buffers])
        idx = C.c_uint64()
        _check(self._lib.wgc_queue_submit(self._ptr, len(buffers), arr, C.byref(idx)))
        return idx.value

    def submit_noalloc(self, buffers: Sequence[CommandBuffer]) -> int:
        """The allocation-free submission (wgc_queue_submit_noalloc): same
        results, scratch from the library's static per-GPU array."""
        arr = (C.c_void_p * max(1, len(buffers)))(*[b.ptr.value for b in buffers])
        idx = C.c_uint64()
        _check(self._lib.wgc_queue_submit_noalloc(self._ptr, len(buffers), arr, C.byref(idx)))
        return idx.value

    def poll(self, submission_index: Optional[int] = None, timeout_ns: int = 0):
        idx = (1 << 64) - 1 if submission_index is None else submission_index
        _check(self._lib.wgc_device_poll(self._ptr, idx, timeout_ns))

    def is_done(self, submission_index: int) -> bool:
        """Non-blocking: whether the submission has executed (raises if one of
        the submissions it retires failed)."""
        code = self._lib.wgc_device_poll(self._ptr, submission_index, 1)
        if code == ERR_TIMEOUT:
            return False
        _check(code)
        return True

    def set_option(self, option: int, value: int):
        _check(self._lib.wgc_device_set_option(self._ptr, option, value))

    def get_option(self, option: int) -> int:
        v = C.c_uint32()
        _check(self._lib.wgc_device_get_option(self._ptr, option, C.byref(v)))
        return v.value

    def last_pass_stats(self) -> PassStats:
        s = PassStats()
        _check(self._lib.wgc_device_last_pass_stats(self._ptr, C.byref(s)))
        return s

    def totals(self):
        """(fragments, submissions) over every executed submission; no wait."""
        f, n = C.c_uint64(), C.c_uint64()
        _check(self._lib.wgc_device_totals(self._ptr, C.byref(f), C.byref(n)))
        return f.value, n.value

    def enable_fragment_counts(self, width: int, height: int):
        _check(self._lib.wgc_device_debug_fragment_counts(self._ptr, width, height, None))

    def fragment_counts(self, width: int, height: int):
        import numpy as np
        out = np.zeros(width * height, dtype=np.uint32)
        _check(self._lib.wgc_device_debug_fragment_counts(self._ptr, width, height,
                                                          out.ctypes.data_as(C.c_void_p)))
        return out

    # measurement
    def timestamp(self, slot: int):
        _check(self._lib.wgc_device_timestamp(self._ptr, slot))

    def elapsed_ms(self, a: int, b: int) -> float:
        ms = C.c_float()
        _check(self._lib.wgc_device_elapsed_ms(self._ptr, a, b, C.byref(ms)))
        return ms.value

    def set_kernel_timing(self, enabled: bool):
        _check(self._lib.wgc_device_set_kernel_timing(self._ptr, 1 if enabled else 0))

    def kernel_time(self, name: str):
        ms = C.c_double()
        n = C.c_uint64()
        _check(self._lib.wgc_device_kernel_time(self._ptr, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value
