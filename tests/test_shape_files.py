"""CPU: shape library files (ShapeLibrary::Load/Save of a TetSolidShape3,
Geo/shape/ShapeLibrary.cpp:332-359,446-468, in the ItemStream container of
util/ItemStream.cpp and util/ItemStreamSerialization.cpp).

- the golden files in tests/golden/ were written by the reference's own
  ItemStream code (tests/golden/gen_shape_files.py); the C-ABI reader loads
  both forms to the baked box;
- where oracle/_ref/libitemstream_ref.so is built (from the reference sources,
  in this container), the writer's text output is byte-identical to the
  reference's, its binary output equal except for the two header padding bytes
  per item that the reference leaves uninitialised, and the reference reads
  the writer's files back to the same arrays.
"""
import os
import struct
import sys

import numpy as np
import pytest

from conftest import GOLDEN

sys.path.insert(0, GOLDEN)
import gen_shape_files as gsf  # noqa: E402


def same_shape(a, b):
    return (np.array_equal(a.pos, b.pos) and np.array_equal(a.tets, b.tets) and np.array_equal(a.ntid, b.ntid)
            and np.array_equal(a.boundary_faces, b.boundary_faces) and np.array_equal(a.layers, b.layers))


@pytest.mark.parametrize("ext", ["bin", "txt"])
def test_load_reference_golden(cf, ext):
    s = cf.TetShape.load(os.path.join(GOLDEN, f"shape_box_4x4x5.{ext}"))
    box = cf.mesh_box_baked(gsf.HALF, gsf.COUNTS)
    assert same_shape(s, box)
    assert s.layers.shape == (2, 2) and s.layers[1, 1] > 0  # an interior layer
    assert np.array_equal(s.src, np.arange(s.tets.size // 4))


def item_header_pads(data, begin, end):
    """Offsets of the two padding bytes of every item header in [begin, end)."""
    out = []
    while begin < end:
        _, ty, size = struct.unpack_from("<iHxxI", data, begin)
        out += [begin + 6, begin + 7]
        if ty & 0x8000:
            out += item_header_pads(data, begin + 12, begin + size)
        begin += size
    return out


@pytest.fixture(scope="module")
def ref():
    lib = gsf.ref_lib()
    if lib is None:
        pytest.skip("oracle/_ref/libitemstream_ref.so not built (needs the reference sources)")
    return lib


@pytest.mark.parametrize("counts", [(4, 4, 5), (2, 3, 2), (6, 5, 4)])
def test_save_matches_reference_writer(cf, ref, tmp_path, counts):
    shape = cf.mesh_box_baked((0.3, 0.4, 0.5), counts)
    for binary in (0, 1):
        mine, theirs = tmp_path / f"mine{binary}", tmp_path / f"ref{binary}"
        shape.save(mine, binary=bool(binary))
        gsf.ref_write(ref, theirs, binary, shape)
        a, b = mine.read_bytes(), theirs.read_bytes()
        if not binary:
            assert a == b
            continue
        assert len(a) == len(b) and a[:28] == b[:28]
        h = struct.unpack_from("<7I", a, 0)
        da, db = bytearray(a[28:28 + h[5]]), bytearray(b[28:28 + h[5]])
        for i in item_header_pads(da, h[2], h[3]):
            da[i] = db[i] = 0
        assert da == db
        sa, sb = a[28 + h[5]:], b[28 + h[5]:]
        assert sa[1:h[4]] == sb[1:h[4]]


@pytest.mark.parametrize("binary", [0, 1])
def test_reference_reads_saved_files(cf, ref, tmp_path, binary):
    shape = cf.mesh_box_baked((0.3, 0.4, 0.5), (5, 4, 6))
    path = tmp_path / "lib"
    shape.save(path, binary=bool(binary))
    pos, t16, bf, l16 = gsf.ref_read(ref, path, binary)
    p0, t0, b0, l0 = gsf.arrays16(shape)
    assert np.array_equal(pos, p0) and np.array_equal(t16, t0) and np.array_equal(bf, b0.ravel())
    assert np.array_equal(l16, l0)
    assert same_shape(cf.TetShape.load(path), shape)


def test_tetgen_to_library_round_trip(cf, tmp_path):
    """The tetgen2geo flow (examples/tetgen2geo/main.cpp): TetGen files in,
    baked shape, saved as a text library, loaded back."""
    from test_shape import write_tetgen
    pos, tets = cf.mesh_box((0.3, 0.2, 0.1), (4, 3, 3))
    base = str(tmp_path / "m")
    write_tetgen(base, pos, tets)
    s = cf.TetShape.from_tetgen(base)
    s.save(tmp_path / "m.txt", binary=False)
    assert same_shape(cf.TetShape.load(tmp_path / "m.txt"), s)


def test_save_refuses_more_than_16_bit_indices(cf, tmp_path):
    shape = cf.mesh_box_baked((0.5, 0.5, 50.0), (2, 2, 16385))  # 65,540 vertices
    with pytest.raises(cf.CorofemError):
        shape.save(tmp_path / "big.bin")


def test_load_errors(cf, tmp_path):
    with pytest.raises(cf.CorofemError):
        cf.TetShape.load(tmp_path / "missing.bin")
    good = open(os.path.join(GOLDEN, "shape_box_4x4x5.bin"), "rb").read()
    for bad in (good[:20], good[:200], b"#1\n{\n    shape_type 21u\n", b"#1\n{\n    shape_type 21u\n    version 3u\n}\n",
                good[:28] + b"\xff" * (len(good) - 28)):
        p = tmp_path / "bad"
        p.write_bytes(bad)
        with pytest.raises(cf.CorofemError):
            cf.TetShape.load(p)
