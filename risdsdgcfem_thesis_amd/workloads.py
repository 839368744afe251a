"""Seeded synthetic scenes for the batched and served forms of the step.

Every scene is built from the reference's box mesher (cube grids, 5 tets per
cube, Geo/shape/TetSolidShape3.cpp:921-999) with edge 0.01 m, rho 1000,
gravity -9.81 y, dt 1/60 and QIE-v + CG at relative tolerance 1e-4. Random
draws use numpy's Philox counter-based generator keyed by the scene seed.
"""
import numpy as np

from .fem import mesh_box, default_solid_params
from ._capi import RM_DAPD

EDGE = 0.01
FACES = [(0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1)]  # (axis, side)


def rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def box_mesh(cubes):
    """Cube grid with `cubes` cubes per axis (vertices cubes+1), edge 0.01 m,
    centred at the origin."""
    counts = tuple(int(c) + 1 for c in cubes)
    half = tuple(0.5 * EDGE * c for c in cubes)
    pos, tets = mesh_box(half, counts)
    return pos, tets, half


def face_nodes(pos, half, face):
    axis, side = face
    return np.nonzero(np.isclose(pos[axis::3], side * half[axis]))[0].astype(np.uint32)


def solid_params(E, nu, cg_max_iter, rm=RM_DAPD):
    return default_solid_params(fixed_dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), young_modulus=float(E),
                                poisson_ratio=float(nu), damping_ratio=0.0, rm=rm, cg_rel_epsilon=1e-4,
                                cg_abs_epsilon=1e-6, cg_max_iter=int(cg_max_iter), density=1000.0)


def uniform_so3(g):
    """A rotation uniform on SO(3) from three uniforms (Shoemake's quaternion method)."""
    u1, u2, u3 = g.random(3)
    x, y, z, w = (np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3))
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def large_mesh_state(pos, seed=2, edge=EDGE):
    """Initial pose of the single large mesh: x0 = Q diag(s) X with one seeded
    uniform rotation Q and stretches s uniform in [0.7, 1.3], plus a per-node
    jitter uniform within 1% of the edge length."""
    g = rng(seed)
    Q = uniform_so3(g)
    s = g.uniform(0.7, 1.3, 3)
    X = pos.reshape(-1, 3).astype(np.float64)
    x = X @ (Q @ np.diag(s)).T
    x += g.uniform(-0.01 * edge, 0.01 * edge, x.shape)
    return x.astype(np.float32).ravel()


def large_mesh(cubes=(64, 64, 256)):
    """The single large mesh (64x64x256 cubes: 5,242,880 tets, 1,085,825 nodes
    at the default size), its initial pose and its z-min face (held at that pose)."""
    pos, tets, half = box_mesh(cubes)
    x0 = large_mesh_state(pos)
    knc = face_nodes(pos, half, (2, -1))
    return pos, tets, x0, knc


def uneven_batch(count=1024, lo=4, hi=40, seed=4, cg_max_iter=50, only=None):
    """3D uneven batch: `count` cube grids with each axis' cube count uniform
    in [lo, hi], E log-uniform in [1e4, 1e7] Pa, nu uniform in [0.2, 0.45], one
    random face held at rest. Returns a list of dicts (pos, tets, knc, params,
    cubes, face, index); only: the indices of the meshes to build (the draws
    are those of the whole batch)."""
    g = rng(seed)
    cubes = g.integers(lo, hi + 1, size=(count, 3))
    E = np.exp(g.uniform(np.log(1e4), np.log(1e7), size=count))
    nu = g.uniform(0.2, 0.45, size=count)
    faces = g.integers(0, 6, size=count)
    out = []
    for i in (range(count) if only is None else only):
        pos, tets, half = box_mesh(cubes[i])
        face = FACES[int(faces[i])]
        out.append({"pos": pos, "tets": tets, "knc": face_nodes(pos, half, face),
                    "params": solid_params(E[i], nu[i], cg_max_iter), "cubes": tuple(int(c) for c in cubes[i]),
                    "face": face, "E": float(E[i]), "nu": float(nu[i]), "index": int(i)})
    return out


def serving_requests(rate, duration_s, seed=5, tets_lo=500, tets_hi=200_000, steps_lo=1, steps_hi=60,
                     cg_max_iter=50):
    """Poisson arrivals at `rate` requests/s over `duration_s` seconds. Each
    request: a cube grid of n^3 cubes with 5 n^3 tets log-uniform in
    [tets_lo, tets_hi] (n rounded to the nearest integer), 1-60 steps, E
    log-uniform in [1e4, 1e7], nu in [0.2, 0.45], one random face held.
    Returns a list of dicts sorted by arrival time (s)."""
    g = rng(seed)
    t = 0.0
    reqs = []
    while True:
        t += g.exponential(1.0 / rate)
        if t >= duration_s:
            break
        tets = np.exp(g.uniform(np.log(tets_lo), np.log(tets_hi)))
        n = max(1, int(round((tets / 5.0) ** (1.0 / 3.0))))
        reqs.append({"arrival": t, "n": n, "steps": int(g.integers(steps_lo, steps_hi + 1)),
                     "E": float(np.exp(g.uniform(np.log(1e4), np.log(1e7)))),
                     "nu": float(g.uniform(0.2, 0.45)), "face": FACES[int(g.integers(0, 6))],
                     "cg_max_iter": cg_max_iter})
    return reqs


def serving_trace(rate, duration_s, seed=5, **kw):
    """serving_requests() as (meshes, requests) for fem.serve: one mesh per
    distinct grid size, requests referencing it, the face nodes held at rest."""
    reqs = serving_requests(rate, duration_s, seed=seed, **kw)
    index, meshes, halves = {}, [], []
    for q in reqs:
        if q["n"] not in index:
            pos, tets, half = box_mesh((q["n"],) * 3)
            index[q["n"]] = len(meshes)
            meshes.append((pos, tets))
            halves.append(half)
        m = index[q["n"]]
        q["mesh"] = m
        q["knc"] = face_nodes(meshes[m][0], halves[m], q["face"])
    return meshes, reqs


def grid2(n):
    """n x n square grid (2 n^2 triangles, (n+1)^2 nodes), edge 0.01 m,
    centred at the origin; the bottom edge's nodes."""
    from .fem import mesh_box2
    half = (0.5 * EDGE * n, 0.5 * EDGE * n)
    pos, tris = mesh_box2(half, (n + 1, n + 1))
    knc = np.nonzero(np.isclose(pos[1::2], -half[1]))[0].astype(np.uint32)
    return pos, tris, knc


def batch2d(count=65536, lo=4, hi=32, seed=3, cg_max_iter=50):
    """2D batch: `count` n x n grids with n uniform in [lo, hi], E log-uniform
    in [1e4, 1e7] Pa, nu uniform in [0.2, 0.45], bottom edge held, thickness 1.
    Meshes of one size share their arrays. Returns a list of dicts."""
    g = rng(seed)
    ns = g.integers(lo, hi + 1, size=count)
    E = np.exp(g.uniform(np.log(1e4), np.log(1e7), size=count))
    nu = g.uniform(0.2, 0.45, size=count)
    cache = {}
    out = []
    for i in range(count):
        n = int(ns[i])
        if n not in cache:
            cache[n] = grid2(n)
        pos, tris, knc = cache[n]
        out.append({"n": n, "pos": pos, "tris": tris, "knc": knc,
                    "params": solid_params(E[i], nu[i], cg_max_iter), "E": float(E[i]), "nu": float(nu[i])})
    return out
