"""CPU: tet-shape baking (EditableTetSolidShape3::EndEdition,
Geo/shape/TetSolidShape3.cpp:262-268,278-307,555-845) and the TetGen importer
(examples/tetgen2geo/main.cpp:80-180), through the C-ABI (host code only).

Pinned by brute force on the definitions, not by restating the code:
- a face is a boundary face iff no other tetrahedron has the same vertex set;
  face neighbours are the tetrahedra sharing it; boundary faces point out;
- vertex layers are the fixed point of the reference's relaxation
  layer[v] = min(layer[v], layer[w] + 1) over the vertices of each tetrahedron,
  started from 0 on boundary vertices and T elsewhere (iterated here in numpy
  until nothing changes); a tetrahedron's layer is its vertices' least;
- the baked order is the input order stably sorted by layer, with negatively
  oriented tetrahedra and unused vertices dropped.
No reference shape files or TetGen files exist in the reference; the TetGen
inputs here are written by the tests.
"""
import numpy as np
import pytest


def brute_force(pos, tets):
    P = pos.reshape(-1, 3).astype(np.float64)
    T4 = tets.reshape(-1, 4)
    keep = [e for e in range(len(T4)) if np.linalg.det((P[T4[e, 1:]] - P[T4[e, 0]]).T) > 0]
    T4 = T4[keep]
    used = np.unique(T4)
    remap = -np.ones(len(P), np.int64)
    remap[used] = np.arange(len(used))
    T4 = remap[T4]
    P = P[used]
    T = len(T4)
    faces = {}
    for e in range(T):
        for n in range(4):
            key = tuple(sorted(T4[e, [(n + 1) % 4, (n + 2) % 4, (n + 3) % 4]]))
            faces.setdefault(key, []).append((e, n))
    ntid = -np.ones((T, 4), np.int64)
    boundary = set()
    for key, lst in faces.items():
        if len(lst) == 1:
            boundary.add(lst[0])
        for (e, n) in lst:
            for (f, _) in lst:
                if f != e:
                    ntid[e, n] = f
    layer = np.full(len(P), T, np.int64)
    for (e, n) in boundary:
        layer[T4[e, [(n + 1) % 4, (n + 2) % 4, (n + 3) % 4]]] = 0
    while True:
        tmin = layer[T4].min(axis=1)
        new = layer.copy()
        np.minimum.at(new, T4.ravel(), np.repeat(tmin + 1, 4))
        if np.array_equal(new, layer):
            break
        layer = new
    tl = layer[T4].min(axis=1)
    order = np.argsort(tl, kind="stable")
    return P, T4, ntid, boundary, tl, order, np.array(keep)


def check_shape(shape, pos, tets):
    P, T4, ntid, boundary, tl, order, keep = brute_force(pos, tets)
    assert np.array_equal(shape.pos.reshape(-1, 3), P.astype(np.float32))
    got = shape.tets.reshape(-1, 4)
    assert np.array_equal(got, T4[order])
    assert np.array_equal(shape.src, keep[order])
    inv = np.empty_like(order)
    inv[order] = np.arange(len(order))
    exp_ntid = np.where(ntid[order] >= 0, inv[np.maximum(ntid[order], 0)], -1)
    assert np.array_equal(shape.ntid.reshape(-1, 4), exp_ntid)
    # layers: contiguous runs of equal tet layer in baked order
    lay = tl[order]
    firsts = np.flatnonzero(np.r_[True, lay[1:] != lay[:-1]]) if len(lay) else np.array([], int)
    assert lay.tolist() == sorted(lay.tolist())
    nonempty = shape.layers[shape.layers[:, 1] > 0]
    assert np.array_equal(nonempty[:, 0], firsts)
    # boundary faces: the unshared faces, each pointing away from its tetrahedron
    bf = shape.boundary_faces.reshape(-1, 3)
    assert len(bf) == len(boundary)
    assert {tuple(sorted(f)) for f in bf} == {tuple(sorted(T4[e, [(n + 1) % 4, (n + 2) % 4, (n + 3) % 4]]))
                                               for (e, n) in boundary}
    for (e, n) in boundary:
        key = tuple(sorted(T4[e, [(n + 1) % 4, (n + 2) % 4, (n + 3) % 4]]))
        f = next(f for f in bf if tuple(sorted(f)) == key)
        normal = np.cross(P[f[1]] - P[f[0]], P[f[2]] - P[f[0]])
        assert np.dot(P[T4[e, n]] - P[f[0]], normal) < 0


@pytest.mark.parametrize("counts", [(2, 2, 2), (3, 4, 5), (6, 3, 3), (5, 5, 9)])
def test_box_bake_brute_force(cf, counts):
    pos, tets = cf.mesh_box((0.5, 0.7, 0.9), counts)
    check_shape(cf.TetShape.bake(pos, tets), pos, tets)


def test_config1_beam_layers(cf):
    s = cf.mesh_box_baked((0.05, 0.05, 0.2), (11, 11, 41))
    pos, tets = cf.mesh_box((0.05, 0.05, 0.2), (11, 11, 41))
    assert np.array_equal(s.tets.reshape(-1, 4), tets.reshape(-1, 4)[s.src])
    assert s.boundary_faces.size // 3 == 2 * (10 * 10 * 2 + 10 * 40 * 4)
    assert s.layers[:, 1].sum() == 20000 and s.layers[0, 0] == 0
    assert not np.array_equal(s.src, np.arange(20000))  # baking does reorder the box


def test_bake_drops_inverted_tets_and_unused_vertices(cf):
    pos, tets = cf.mesh_box((1.0, 1.0, 1.0), (3, 3, 3))
    P = np.concatenate([pos, np.array([5.0, 5.0, 5.0], np.float32)])  # an unused vertex at the end
    T4 = tets.reshape(-1, 4).copy()
    T4[7, [1, 2]] = T4[7, [2, 1]]  # inverted
    T4[12, [0, 3]] = T4[12, [3, 0]]  # inverted
    pre = np.array([9, 10, 11, 12], np.float32)  # an unused vertex before the others: indices shift
    P = np.concatenate([pre[:3], P])
    T4 = T4 + 1
    s = cf.TetShape.bake(P, T4.ravel())
    assert s.tets.size // 4 == len(T4) - 2 and 7 not in s.src and 12 not in s.src
    assert s.pos.size // 3 == pos.size // 3
    check_shape(s, P, T4.ravel())


def test_random_subsets_bake_brute_force(cf):
    """Irregular shapes: random subsets of a box's tetrahedra (interior holes,
    several components)."""
    rng = np.random.default_rng(5)
    pos, tets = cf.mesh_box((1.0, 1.0, 1.0), (6, 5, 5))
    T4 = tets.reshape(-1, 4)
    for _ in range(4):
        sub = T4[np.sort(rng.choice(len(T4), size=len(T4) * 2 // 3, replace=False))]
        check_shape(cf.TetShape.bake(pos, sub.ravel()), pos, sub.ravel())


def write_tetgen(path, pos, tets, base=1, attrs=1, markers=1):
    P = pos.reshape(-1, 3)
    T4 = tets.reshape(-1, 4)
    with open(path + ".node", "w") as f:
        f.write(f"{len(P)}  3  {attrs}  {markers}\n")
        for i, p in enumerate(P):
            extra = " ".join(["0.5"] * attrs + ["1"] * markers)
            f.write(f"{i + base}  {float(p[0])!r}  {float(p[1])!r}  {float(p[2])!r}  {extra}\n")
        f.write("# Generated by a test\n")
    with open(path + ".ele", "w") as f:
        f.write(f"{len(T4)}  4  {attrs}\n")
        for i, t in enumerate(T4):
            extra = " ".join(["2"] * attrs)
            f.write(f"{i + base}  {t[0] + base} {t[1] + base} {t[2] + base} {t[3] + base}  {extra}\n")
        f.write("# Generated by a test\n")


@pytest.mark.parametrize("base,attrs,markers", [(1, 1, 1), (0, 0, 0), (1, 2, 0)])
def test_tetgen_import(cf, tmp_path, base, attrs, markers):
    pos, tets = cf.mesh_box((0.3, 0.2, 0.1), (4, 3, 3))
    pos = (pos + 1e-4 * (np.arange(pos.size) % 7)).astype(np.float32)  # not all on a grid
    T4 = tets.reshape(-1, 4).copy()
    T4[3, [1, 2]] = T4[3, [2, 1]]  # one inverted tetrahedron, dropped like AddTetrahedron does
    base_name = str(tmp_path / "mesh")
    write_tetgen(base_name, pos, T4, base, attrs, markers)
    s = cf.TetShape.from_tetgen(base_name)
    ref = cf.TetShape.bake(pos, T4.ravel())
    assert np.array_equal(s.pos, ref.pos) and np.array_equal(s.tets, ref.tets) and np.array_equal(s.src, ref.src)
    assert 3 not in s.src


def test_tetgen_natural_scale(cf, tmp_path):
    pos, tets = cf.mesh_box((0.3, 0.2, 0.1), (4, 3, 3))
    pos = (pos.reshape(-1, 3) + np.array([2.0, -1.0, 0.5], np.float32)).astype(np.float32).ravel()
    base_name = str(tmp_path / "mesh")
    write_tetgen(base_name, pos, tets)
    s = cf.TetShape.from_tetgen(base_name, natural_scale=2.0)
    P = s.pos.reshape(-1, 3)
    lo, hi = P.min(axis=0), P.max(axis=0)
    assert np.allclose(0.5 * (lo + hi), 0.0, atol=1e-6)
    assert abs((hi - lo).max() - 2.0) < 1e-6
    assert np.allclose((hi - lo) / (hi - lo).max(), np.array([0.6, 0.4, 0.2]) / 0.6, atol=1e-6)


def test_tetgen_errors(cf, tmp_path):
    with pytest.raises(cf.CorofemError):
        cf.TetShape.from_tetgen(str(tmp_path / "missing"))
    base_name = str(tmp_path / "bad")
    with open(base_name + ".node", "w") as f:
        f.write("4 2 0 0\n")
    with open(base_name + ".ele", "w") as f:
        f.write("0 4 0\n")
    with pytest.raises(cf.CorofemError):
        cf.TetShape.from_tetgen(base_name)
    pos, tets = cf.mesh_box((1.0, 1.0, 1.0), (2, 2, 2))
    T4 = tets.reshape(-1, 4).copy()
    T4[0, 0] = 99
    write_tetgen(base_name, pos, T4)
    with pytest.raises(cf.CorofemError):
        cf.TetShape.from_tetgen(base_name)
