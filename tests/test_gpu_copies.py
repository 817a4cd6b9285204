"""Copies, clears and mapped readback on the GPU, through the C ABI
(CommandEncoderInterface::copy_* / clear_*, command.rs:37-73 and 102-116;
BufferInterface::map_async / get_mapped_range / unmap, buffer.rs:112-167).

The copies are todo!() in the reference, so their results are pinned by
what the operations mean: bytes moved to the places wgpu's copy layout
names (checked against numpy), and readback of rendered passes equal to the
oracle's render of the same pass.
"""
import sys

import numpy as np
import pytest

import wgc
from oracle import bridge
from wgc import runner, scenes

pytestmark = pytest.mark.gpu


def announce(name):
    print(f"\n[gpu case] {name}", flush=True)
    sys.stdout.flush()


def map_read(dev, buf, offset=0, size=None):
    """map_async(READ) + poll + get_mapped_range + unmap; checks the callback ran with status 0."""
    seen = []
    buf.map_async(wgc.MAP_READ, offset, size, callback=seen.append)
    dev.poll()
    assert seen == [0]
    data = bytes(buf.get_mapped_range(offset, size))
    buf.unmap()
    return data


def record(dev, enc, spec, keep, cache):
    res = runner._PassResources(dev, spec, keep)
    runner.record_pass(dev, enc, spec, res, keep, cache)
    return res


def pitch_for(width):
    return (width * 4 + 255) // 256 * 256


def unpitch(data, offset, pitch, width, height):
    rows = [data[offset + r * pitch: offset + r * pitch + width * 4] for r in range(height)]
    return np.frombuffer(b"".join(rows), np.uint8).reshape(height, width, 4)


def test_render_then_copy_and_map_readback(gpu_device):
    """A rendered pass read back the wgpu way (texture -> buffer with 256-byte
    row pitch, then a mapped read) equals the oracle's render."""
    announce("copy + map readback")
    dev = gpu_device
    spec = scenes.config1_plain_small()      # 256x256, colour + depth
    w, h = spec.width, spec.height
    ref = bridge.render(spec)
    keep, cache = [], {}
    enc = dev.create_command_encoder()
    res = record(dev, enc, spec, keep, cache)
    pitch = pitch_for(w)
    depth_off = pitch * h + 256
    buf = dev.create_buffer(depth_off + pitch * h)
    enc.copy_texture_to_buffer(res.color, (0, 0), buf, 0, pitch, w, h)
    enc.copy_texture_to_buffer(res.depth, (0, 0), buf, depth_off, pitch, w, h)
    dev.submit([enc.finish()])
    data = map_read(dev, buf)
    color = unpitch(data, 0, pitch, w, h)
    depth = unpitch(data, depth_off, pitch, w, h).view(np.float32).reshape(h, w)
    assert np.array_equal(color, ref.color)
    assert np.array_equal(depth.view(np.uint32), ref.depth.view(np.uint32))


def test_copies_run_in_order_with_passes(gpu_device):
    """pass A, copy, pass B (clear only, same target), copy: the first copy
    holds A's image, the second B's clear colour."""
    announce("copy ordering")
    dev = gpu_device
    spec = scenes.config1_plain_small()
    w, h = spec.width, spec.height
    ref = bridge.render(spec)
    keep, cache = [], {}
    enc = dev.create_command_encoder()
    res = record(dev, enc, spec, keep, cache)
    pitch = pitch_for(w)
    b1 = dev.create_buffer(pitch * h)
    b2 = dev.create_buffer(pitch * h)
    enc.copy_texture_to_buffer(res.color, (0, 0), b1, 0, pitch, w, h)
    rp = enc.begin_render_pass(color=res.color, clear_color=(1.0, 0.0, 0.0, 1.0))
    rp.end()
    enc.copy_texture_to_buffer(res.color, (0, 0), b2, 0, pitch, w, h)
    dev.submit([enc.finish()])
    assert np.array_equal(unpitch(map_read(dev, b1), 0, pitch, w, h), ref.color)
    red = unpitch(map_read(dev, b2), 0, pitch, w, h)
    assert (red == np.array([255, 0, 0, 255], np.uint8)).all()


def test_texture_copies_and_clears_move_the_named_texels(gpu_device):
    announce("texture copies")
    dev = gpu_device
    rng = np.random.default_rng(7)
    w, h = 300, 200
    t1 = dev.create_texture(w, h, wgc.RGBA8_UNORM)
    src = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    dev.write_texture(t1, src.tobytes())
    t2 = dev.create_texture(w, h, wgc.RGBA8_UNORM)
    t3 = dev.create_texture(w, h, wgc.RGBA8_UNORM_SRGB)
    t4 = dev.create_texture(w, h, wgc.RGBA8_UNORM)
    dev.write_texture(t4, rng.integers(0, 256, (h, w, 4), dtype=np.uint8).tobytes())
    init2 = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    init3 = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    dev.write_texture(t2, init2.tobytes())
    dev.write_texture(t3, init3.tobytes())
    # sub-rectangle 123x77 at (17, 29) -> buffer (pitch 512, offset 1024)
    sx, sy, cw, ch = 17, 29, 123, 77
    buf = dev.create_buffer(1024 + 512 * ch)
    enc = dev.create_command_encoder()
    enc.copy_texture_to_buffer(t1, (sx, sy), buf, 1024, 512, cw, ch)
    enc.copy_buffer_to_texture(buf, 1024, 512, t2, (150, 100), cw, ch)
    enc.copy_texture_to_texture(t1, (0, 0), t3, (40, 50), 200, 120)   # Rgba8Unorm -> Rgba8UnormSrgb: bytes
    enc.clear_texture(t4)
    dev.submit([enc.finish()])
    dev.poll()
    got_buf = unpitch(map_read(dev, buf), 1024, 512, cw, ch)
    assert np.array_equal(got_buf, src[sy:sy + ch, sx:sx + cw])
    want2 = init2.copy()
    want2[100:100 + ch, 150:150 + cw] = src[sy:sy + ch, sx:sx + cw]
    assert np.array_equal(np.frombuffer(t2.read(), np.uint8).reshape(h, w, 4), want2)
    want3 = init3.copy()
    want3[50:170, 40:240] = src[0:120, 0:200]
    assert np.array_equal(np.frombuffer(t3.read(), np.uint8).reshape(h, w, 4), want3)
    assert not np.frombuffer(t4.read(), np.uint8).any()


def test_buffer_copy_and_clear(gpu_device):
    announce("buffer copy / clear")
    dev = gpu_device
    rng = np.random.default_rng(11)
    a = rng.integers(0, 256, 100_000, dtype=np.uint8).tobytes()
    src = dev.create_buffer_init(a)
    dst = dev.create_buffer_init(bytes(100_000))
    enc = dev.create_command_encoder()
    enc.copy_buffer_to_buffer(src, 4, dst, 1000, 50_000)        # unaligned to 16: 4-byte path
    enc.copy_buffer_to_buffer(src, 16, dst, 60_000, 32_000)     # 16-byte path
    enc.clear_buffer(src, 8, 400)
    dev.submit([enc.finish()])
    want = bytearray(100_000)
    want[1000:51_000] = a[4:50_004]
    want[60_000:92_000] = a[16:32_016]
    assert map_read(dev, dst) == bytes(want)
    want_src = bytearray(a)
    want_src[8:408] = bytes(400)
    assert map_read(dev, src, 0, 1024) == bytes(want_src[:1024])


def test_map_write_fills_a_vertex_buffer(gpu_device):
    """Vertex data written through a write mapping renders like uploaded data."""
    announce("map write")
    dev = gpu_device
    spec = scenes.config1_plain_small()
    ref = bridge.render(spec)
    d = spec.draws[0]
    vb = dev.create_buffer(len(d.vertex_data))
    vb.map_async(wgc.MAP_WRITE)
    dev.poll()
    view = vb.get_mapped_range()
    view[:] = d.vertex_data
    vb.unmap()
    # the runner creates buffers from bytes; render by hand with the mapped buffer
    keep, cache = [vb], {("buf", d.vertex_data): vb}
    enc = dev.create_command_encoder()
    res = record(dev, enc, spec, keep, cache)
    dev.submit([enc.finish()])
    color, depth = res.read()
    assert np.array_equal(color, ref.color)
    assert np.array_equal(depth.view(np.uint32), ref.depth.view(np.uint32))


def test_mapping_and_copy_validation(gpu_device):
    announce("map / copy validation")
    dev = gpu_device
    buf = dev.create_buffer(4096)
    tex = dev.create_texture(64, 64, wgc.RGBA8_UNORM)
    enc = dev.create_command_encoder()
    with pytest.raises(wgc.WgcError):
        enc.copy_texture_to_buffer(tex, (0, 0), buf, 0, 100, 16, 16)      # pitch not a multiple of 256
    with pytest.raises(wgc.WgcError):
        enc.copy_texture_to_buffer(tex, (60, 0), buf, 0, 256, 16, 4)      # outside the texture
    with pytest.raises(wgc.WgcError):
        enc.copy_texture_to_buffer(tex, (0, 0), buf, 0, 256, 64, 64)      # outside the buffer
    with pytest.raises(wgc.WgcError):
        enc.copy_buffer_to_buffer(buf, 2, buf, 8, 4)                     # same buffer, unaligned
    with pytest.raises(wgc.WgcError):
        enc.copy_texture_to_texture(tex, (0, 0), tex, (8, 8), 4, 4)       # same texture
    # a submission that uses a mapped buffer is rejected
    enc.clear_buffer(buf)
    cb = enc.finish()
    statuses = []
    buf.map_async(wgc.MAP_READ, callback=statuses.append)
    with pytest.raises(wgc.WgcError):
        buf.map_async(wgc.MAP_READ)                                      # already mapped
    with pytest.raises(wgc.WgcError):
        dev.submit([cb])
    dev.poll()
    assert statuses == [0]
    buf.unmap()
    with pytest.raises(wgc.WgcError):
        buf.unmap()                                                      # not mapped
    with pytest.raises(wgc.WgcError):
        buf.get_mapped_range()
    # unmapped before the request completed: its callback reports an error
    big = dev.create_buffer(64 << 20)
    statuses = []
    big.map_async(wgc.MAP_READ, callback=statuses.append)
    big.unmap()
    assert statuses and statuses[0] != 0
    dev.submit([cb])
    dev.poll()
    assert map_read(dev, buf) == bytes(4096)
