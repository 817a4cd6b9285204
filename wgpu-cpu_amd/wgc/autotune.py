"""Autotuner for the raster tile shape (tile side x lanes per workgroup).

The binning tile and the raster workgroup shape trade setup work against
per-tile parallelism: 8x8 tiles suit small targets with tiny triangles,
32x32 tiles large targets.  `tune` times each shape on a prepared workload
with device-stream events and leaves the device on the fastest one.
"""
from __future__ import annotations

from typing import Dict

from . import OPTION_RASTER_TILE, RASTER_TILES


def time_submissions(prepared, reps: int, warmup: int = 2) -> float:
    """Milliseconds per submission, event-timed on the device stream."""
    dev = prepared.device
    for _ in range(warmup):
        prepared.submit()
    dev.poll()
    dev.timestamp(0)
    for _ in range(reps):
        prepared.submit()
    dev.timestamp(1)
    dev.poll()
    return dev.elapsed_ms(0, 1) / reps


def tune(prepared, reps: int = 10, shapes=None) -> Dict[int, float]:
    """Times every raster tile shape; sets the fastest; returns ms per shape."""
    dev = prepared.device
    shapes = list(RASTER_TILES) if shapes is None else list(shapes)
    timings = {}
    for cfg in shapes:
        dev.set_option(OPTION_RASTER_TILE, cfg)
        timings[cfg] = time_submissions(prepared, reps)
    best = min(timings, key=timings.get)
    dev.set_option(OPTION_RASTER_TILE, best)
    return timings
