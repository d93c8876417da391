"""Owner-kernel tuning with a JSON cache.

vgpu_autotune times the dual apply for every owner tile shape (32x8, 16x16,
64x4 element columns) and z chunk (0 = balanced over the resident
workgroups; 4, 8, 16 and naz / k, k = 1..8, node planes per (chunk, tile)
group with one workgroup each) and keeps the fastest on the context.  The cache maps (library version, mesh shape, precision) to the
chosen parameters; it is written with sorted keys so the same choices give the
same file.  Every owner node is summed by one thread in a fixed order, so the
parameters change timing only: operator results are bitwise the same.
"""
from __future__ import annotations

import json
import os

from . import _lib


def cache_key(ctx, precision) -> str:
    ver = _lib.load().vgpu_version().decode()
    return f"{ver}|{ctx.nax}x{ctx.nay}x{ctx.naz}|prec{int(precision)}"


def tune(ctx, precision=_lib.FP64, cache_path: str | None = None, reps: int = 20) -> dict:
    """Apply the cached choice for this mesh, or time the candidates, apply the
    fastest and record it.  Returns the entry plus "cached": bool."""
    cache = {}
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
    key = cache_key(ctx, precision)
    if key in cache:
        e = cache[key]
        ctx.set_options("owner", int(e["zchunk"]), int(e["tile"]))
        return dict(e, cached=True)
    r = ctx.autotune(precision, reps)
    entry = {"tile": int(r["tile"]), "zchunk": int(r["zchunk"]), "ms": round(float(r["ms"]), 6)}
    if cache_path:
        cache[key] = entry
        tmp = cache_path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, sort_keys=True, indent=1)
            f.write("\n")
        os.replace(tmp, cache_path)
    return dict(entry, cached=False)
