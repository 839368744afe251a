"""Autotuner for batched stepping.

Chooses, for a given set of meshes on the current device:
  - the tile shape of the element stage: workgroup threads and elements per
    tile (128/128, 256/256, 512/512, and 512-thread tiles of 384 or 256
    elements: 64, 48 or 32 elements per wavefront, the rest of the threads
    carrying only tile nodes), from the batched K*x time per tet on a
    stratified sample of the meshes (every k-th mesh by size);
  - the CG chunk length (iterations per recorded graph, between two device
    compactions of the lists of the simulations still iterating), from the
    batched step time on the same sample.
Results are cached in memory per (device, workload signature), and in a JSON
file when `cache_path` is given.
"""
import json
import os
import time

import numpy as np

from .fem import Batch, device_count

TILE_CANDIDATES = [(128, 128), (256, 256), (512, 512), (512, 384), (512, 256)]
CHUNK_CANDIDATES = [4, 8, 16, 32]
_MEMO = {}


def signature(meshes):
    """Workload signature: mesh count and a coarse histogram of tets per mesh."""
    T = np.array([m["tets"].size // 4 for m in meshes])
    hist = np.histogram(np.log2(np.maximum(T, 1)), bins=12, range=(4, 24))[0]
    return f"{len(meshes)}:{int(T.sum())}:" + ",".join(str(int(h)) for h in hist)


def _sample(meshes, n):
    order = np.argsort([m["tets"].size for m in meshes], kind="stable")
    step = max(1, len(meshes) // n)
    return [meshes[i] for i in order[::step]]


def _batch(sample, threads, elems, chunk, use_graphs):
    b = Batch(tile_threads=threads, tile_elems=elems, cg_chunk=chunk, use_graphs=use_graphs)
    mids = b.add_meshes([(m["pos"], m["tets"]) for m in sample], threads=16)
    for mid, m in zip(mids, sample):
        b.add_sim(mid, m["params"], knc=m["knc"])
    return b


def tune(meshes, sample_size=128, kx_reps=10, cache_path=None, use_graphs=True):
    """Returns {"tile_threads", "tile_elems", "cg_chunk", "kx_ns_per_tet": {...},
    "step_ms": {...}, "seconds"}."""
    if device_count() < 1:
        raise RuntimeError("autotuning needs a GPU")
    key = signature(meshes) + (":graphs" if use_graphs else ":direct")
    if key in _MEMO:
        return _MEMO[key]
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            disk = json.load(f)
        if key in disk:
            _MEMO[key] = disk[key]
            return disk[key]
    t0 = time.perf_counter()
    sample = _sample(meshes, sample_size)
    tets = sum(m["tets"].size // 4 for m in sample)
    kx = {}
    for threads, elems in TILE_CANDIDATES:
        b = _batch(sample, threads, elems, 8, use_graphs)
        b.fixed_step(stats=False)  # rotations of a real state
        ms = b.time_kx(kx_reps)
        kx[f"{threads}/{elems}"] = round(ms * 1e6 / tets, 4)
        b.close()
    best = min(kx, key=kx.get)
    threads, elems = (int(v) for v in best.split("/"))
    steps = {}
    b = _batch(sample, threads, elems, CHUNK_CANDIDATES[0], use_graphs)
    b.fixed_step(stats=False)
    for chunk in CHUNK_CANDIDATES:
        b.set_schedule(cg_chunk=chunk, use_graphs=use_graphs)
        b.fixed_step(stats=False)  # records the graphs
        b.synchronize()
        t = time.perf_counter()
        b.fixed_step(stats=False)
        b.synchronize()
        steps[str(chunk)] = round((time.perf_counter() - t) * 1e3, 3)
    b.close()
    chunk = int(min(steps, key=steps.get))
    res = {"tile_threads": threads, "tile_elems": elems, "cg_chunk": chunk, "kx_ns_per_tet": kx,
           "step_ms": steps, "sample": f"{len(sample)} of {len(meshes)} meshes", "seconds":
           round(time.perf_counter() - t0, 2)}
    _MEMO[key] = res
    if cache_path:
        disk = {}
        if os.path.exists(cache_path):
            with open(cache_path) as f:
                disk = json.load(f)
        disk[key] = res
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        with open(cache_path, "w") as f:
            json.dump(disk, f, indent=1)
    return res
