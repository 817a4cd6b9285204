"""Serving loop on the GPU: heap admission, continuous batching, and every
request's result equal alone and in the mix (bit-exact colour and depth)."""
import sys

import numpy as np
import pytest

import wgc
from wgc import serving

pytestmark = pytest.mark.gpu


def announce(name):
    print(f"\n[gpu case] {name}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pool(gpu_device):
    return serving.MeshPool(gpu_device, 6, seed=0x5EED_0005)


def test_heap_admission_and_reuse(gpu_device):
    announce("heap")
    heap = gpu_device.create_heap(4 << 20)
    a = heap.create_texture(512, 512, wgc.RGBA8_UNORM)      # 1 MiB
    b = heap.create_texture(512, 512, wgc.DEPTH32_FLOAT)    # 1 MiB
    c = heap.create_texture(1024, 512, wgc.RGBA8_UNORM)     # 2 MiB
    assert a is not None and b is not None and c is not None
    assert heap.create_buffer(256) is None                   # full
    used, cap, largest = heap.stats()
    assert used == cap == 4 << 20 and largest == 0
    b.release()
    d = heap.create_buffer(1 << 20)                          # reuses b's range
    assert d is not None
    # zero-filled on creation, like every resource
    assert not any(d.read())
    for h in (a, c, d):
        h.release()
    assert heap.stats()[0] == 0
    assert heap.stats()[2] == 4 << 20                        # coalesced back to one range


def _alone(device, pool, req):
    srv = serving.Server(device, pool, 64 << 20)
    req0 = serving.Request(0, 0.0, req.width, req.height, req.mesh, req.uniform)
    srv.run([req0], batched=False, keep=[0])
    return srv.kept[0]


def test_alone_equals_in_the_mix(gpu_device, pool):
    announce("alone_vs_mix")
    trace = serving.make_trace(0x5EED_0006, rate=20000.0, seconds=0.004, n_meshes=len(pool.meshes))
    assert len(trace) >= 40
    keep = [1, 7, 15, 22, 31, len(trace) - 1]
    srv = serving.Server(gpu_device, pool, 1 << 30, max_batch=16, max_inflight=2)
    st = srv.run(trace, batched=True, keep=keep)
    assert st.completed == len(trace)
    assert st.submissions < len(trace)           # requests were batched
    for rid in keep:
        color, depth = _alone(gpu_device, pool, trace[rid])
        assert srv.kept[rid][0] == color, f"colour of request {rid} differs in the mix"
        assert srv.kept[rid][1] == depth, f"depth of request {rid} differs in the mix"


def test_admission_waits_for_memory(gpu_device, pool):
    announce("admission")
    trace = serving.make_trace(0x5EED_0007, rate=20000.0, seconds=0.003, n_meshes=len(pool.meshes))
    big = max(trace, key=lambda r: r.width * r.height)
    # room for about two of the largest requests at a time
    heap = 2 * (2 * big.width * big.height * 4 + 512) + (1 << 20)
    keep = [big.rid, 0]
    srv = serving.Server(gpu_device, pool, heap, max_batch=64, max_inflight=2)
    st = srv.run(trace, batched=True, keep=keep)
    assert st.completed == len(trace)
    assert st.admission_waits > 0
    assert st.max_heap_used <= heap
    for rid in keep:
        color, depth = _alone(gpu_device, pool, trace[rid])
        assert srv.kept[rid] == (color, depth)


def test_plain_and_batched_render_the_same_fragments(gpu_device, pool):
    announce("plain_vs_batched_fragments")
    trace = serving.make_trace(0x5EED_0008, rate=20000.0, seconds=0.003, n_meshes=len(pool.meshes))
    a = serving.Server(gpu_device, pool, 1 << 30).run(trace, batched=True)
    b = serving.Server(gpu_device, pool, 1 << 30).run(trace, batched=False)
    assert a.completed == b.completed == len(trace)
    assert a.fragments == b.fragments > 0
    assert b.submissions == len(trace)
