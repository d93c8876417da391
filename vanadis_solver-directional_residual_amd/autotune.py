"""Launch-shape tuning with a JSON cache.

Owner scatter: vgpu_autotune times the solver's iteration launch for every
owner tile shape (32x8, 16x16, 64x4 element columns) and z chunk (0 =
balanced over the resident workgroups; 4, 8, 16 and naz / k, k = 1..8, node
planes per (chunk, tile) group with one workgroup each) and keeps the fastest
on the context.  Every owner node is summed by one thread in a fixed order,
so the parameters change timing only: operator results are bitwise the same.

Colored / atomic scatters (the reference's two solver flavours):
vgpu_autotune_bricks times the dual apply for every brick shape (the same
three) and brick depth (the library's rule, 1, 2, 3, 4, 6, 8, 12, 16 element
layers); results agree with the oracle to <= 1e-13 for every choice and are
bitwise reproducible for a fixed one.

The cache maps (library version, mesh shape, precision, scatter) to the chosen
parameters; it is written with sorted keys so the same choices give the same
file.
"""
from __future__ import annotations

import json
import os

from . import _lib


def cache_key(ctx, precision, scatter="owner") -> str:
    ver = _lib.load().vgpu_version().decode()
    key = f"{ver}|{ctx.nax}x{ctx.nay}x{ctx.naz}|prec{int(precision)}"
    return key if scatter == "owner" else f"{key}|{scatter}"


def tune(ctx, precision=_lib.FP64, cache_path: str | None = None, reps: int = 20,
         scatter: str = "owner") -> dict:
    """Apply the cached choice for this mesh and scatter, or time the
    candidates, apply the fastest and record it.  Returns the entry plus
    "cached": bool."""
    cache = {}
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
    key = cache_key(ctx, precision, scatter)
    if key in cache:
        e = cache[key]
        ctx.set_options(scatter, int(e["zchunk"]), int(e["tile"]))
        return dict(e, cached=True)
    if scatter == "owner":
        r = ctx.autotune(precision, reps)
    else:
        r = ctx.autotune_bricks(scatter, precision, reps)
    entry = {"tile": int(r["tile"]), "zchunk": int(r["zchunk"]), "ms": round(float(r["ms"]), 6)}
    if cache_path:
        cache[key] = entry
        tmp = cache_path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, sort_keys=True, indent=1)
            f.write("\n")
        os.replace(tmp, cache_path)
    return dict(entry, cached=False)
