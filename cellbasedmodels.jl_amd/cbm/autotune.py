"""Autotuner of the neighbor structure and launch shape.

The reference fixes `cellEdge`, `skin` and the launch geometry by hand
(src/neighbors.jl:587-619, src/platforms.jl:67-88).  Here, for a given model
and set of simulations, `tune` times a few real steps of trial contexts built
from a copy of the state for each candidate structure (cell-linked edge, or
Verlet list radius with its rebuild displacement), keeps the fastest, then
lets the library pick the block size (cbm_autotune, timed on a launch that
writes only scratch).  Choices are cached in a JSON file keyed by the model
family and population scale.
"""
import ctypes as C
import json
import math
import os
import time

from . import _abi as A
from .api import Context


def _key(m, sims, cutoff):
    n = sum(s.n for s in sims)
    return (f"d{m.dims} i{m.integrator} law{m.law} per{m.periodic} rule{m.rule} sims{len(sims)} "
            f"n2^{int(math.log2(max(n, 1)))} cut{cutoff:.3g}")


def candidates(cutoff, periodic_boxes=None):
    """Cell-linked edges and Verlet list radii to try for an interaction cutoff."""
    out = []
    for f in (1.0, 1.25):
        out.append(("cell", cutoff * f))
    for margin in (0.2, 0.35, 0.6):
        out.append(("verlet", cutoff + margin))
    return out


def clone(m):
    n = type(m)()
    C.memmove(C.byref(n), C.byref(m), C.sizeof(m))
    return n


def _apply(m, choice, max_neighbors):
    m = clone(m)
    kind, v = choice
    if kind == "cell":
        m.neighbors = A.CELL_LINKED
        m.cell_edge = v
    else:
        m.neighbors = A.VERLET_DISPLACEMENT
        m.skin = v
        m.max_neighbors = max_neighbors
    return m


def _time(m, sims, upload, steps, warm):
    ctx = Context(m, sims)
    upload(ctx)
    ctx.neighbors()
    ctx.step(warm)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    dt = time.perf_counter() - t0
    n = sum(s.n for s in sims)
    ctx.close()
    return n * steps / dt


def tune(m, sims, upload, cutoff, max_neighbors=48, steps=20, warm=5, cache_path=None, choices=None):
    """Return (model_desc, info) with the fastest structure for this workload.

    upload(ctx) must load the initial agent state into a fresh context.
    For Verlet candidates the rebuild displacement is half the skin margin
    (the standard safe trigger), so every choice computes the same forces.
    """
    key = _key(m, sims, cutoff)
    cache = {}
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
    if key in cache:
        ch = tuple(cache[key]["choice"])
        mm = _apply(m, ch, max_neighbors)
        if ch[0] == "verlet":
            mm.rebuild_disp = (ch[1] - cutoff) / 2
        return mm, dict(cache[key], cached=True)
    results = []
    for ch in (choices or candidates(cutoff)):
        mm = _apply(m, ch, max_neighbors)
        if ch[0] == "verlet":
            mm.rebuild_disp = (ch[1] - cutoff) / 2
        try:
            rate = _time(mm, sims, upload, steps, warm)
        except A.CBMError:
            continue            # e.g. a periodic grid too coarse for this edge, or list overflow
        results.append((rate, ch))
    if not results:
        raise A.CBMError(A.ERR_CONFIG, "no structure candidate ran")
    rate, ch = max(results)
    mm = _apply(m, ch, max_neighbors)
    if ch[0] == "verlet":
        mm.rebuild_disp = (ch[1] - cutoff) / 2
    info = {"choice": list(ch), "agent_steps_per_s": rate, "tried": [[list(c), r] for r, c in results]}
    if cache_path:
        cache[key] = info
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1)
    return mm, info


def tune_block(ctx, cache_path=None):
    """Library block-size tuning (no state change); returns the chosen block size."""
    b = C.c_int32()
    path = cache_path.encode() if cache_path else None
    A.check(ctx.lib.cbm_autotune(ctx.h, path, C.byref(b)), ctx.h)
    return b.value
