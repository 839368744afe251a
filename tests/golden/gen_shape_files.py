"""Writes the golden shape library files in tests/golden/ with the reference's
own ItemStream container (oracle/_ref/libitemstream_ref.so, built by
oracle/Makefile from the reference's util/ItemStream*.cpp), holding the box
mesh of Make_TetSolidShape3_Box as EndEdition bakes it. Run where the
reference is mounted:

    make -C oracle && python tests/golden/gen_shape_files.py

Fixtures:
  shape_box_4x4x5.bin, shape_box_4x4x5.txt   half sizes (0.25, 0.5, 0.75),
                                            4 x 4 x 5 vertices (binary: the
                                            reference leaves item-header padding
                                            uninitialised, so those bytes are
                                            whatever its allocator held)
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)
import risdsdgcfem_thesis_amd as cf  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
HALF, COUNTS = (0.25, 0.5, 0.75), (4, 4, 5)


def ref_lib():
    path = os.path.join(ROOT, "oracle", "_ref", "libitemstream_ref.so")
    if not os.path.exists(path):
        return None
    lib = C.CDLL(path)
    vp = C.c_void_p
    lib.ref_is_write_tetshape.argtypes = [C.c_char_p, C.c_int, C.c_uint32, vp, C.c_uint32, vp, C.c_uint32, vp,
                                          C.c_uint32, vp]
    lib.ref_is_read_tetshape.argtypes = [C.c_char_p, C.c_int] + [C.c_uint32] * 4 + [
        C.POINTER(C.c_uint32), vp, C.POINTER(C.c_uint32), vp, C.POINTER(C.c_uint32), vp, C.POINTER(C.c_uint32), vp]
    return lib


def arrays16(shape):
    T = shape.tets.size // 4
    t16 = np.zeros((T, 8), np.uint16)
    t16[:, :4] = shape.tets.reshape(-1, 4)
    nt = shape.ntid.reshape(-1, 4)
    t16[:, 4:] = np.where(nt < 0, 0xFFFF, nt)
    return (np.ascontiguousarray(shape.pos, np.float32), t16, shape.boundary_faces.astype(np.uint16),
            shape.layers.astype(np.uint16).ravel())


def ref_write(lib, path, binary, shape):
    pos, t16, bf, l16 = arrays16(shape)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    ok = lib.ref_is_write_tetshape(str(path).encode(), int(binary), pos.size // 3, p(pos), len(t16), p(t16),
                                   bf.size // 3, p(bf), l16.size // 2, p(l16))
    assert ok


def ref_read(lib, path, binary, cap=1 << 16):
    pos = np.zeros(3 * cap, np.float32)
    t16 = np.zeros(8 * cap, np.uint16)
    bf = np.zeros(3 * 4 * cap, np.uint16)
    l16 = np.zeros(2 * cap, np.uint16)
    n = [C.c_uint32() for _ in range(4)]
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    ok = lib.ref_is_read_tetshape(str(path).encode(), int(binary), cap, cap, 4 * cap, cap, C.byref(n[0]), p(pos),
                                  C.byref(n[1]), p(t16), C.byref(n[2]), p(bf), C.byref(n[3]), p(l16))
    assert ok
    nv, nt, nbf, nl = (x.value for x in n)
    return pos[:3 * nv], t16[:8 * nt].reshape(-1, 8), bf[:3 * nbf], l16[:2 * nl]


def main():
    lib = ref_lib()
    if lib is None:
        sys.exit("oracle/_ref/libitemstream_ref.so is missing: run make -C oracle where the reference is mounted")
    shape = cf.mesh_box_baked(HALF, COUNTS)
    for ext, binary in (("bin", 1), ("txt", 0)):
        ref_write(lib, os.path.join(OUT, f"shape_box_4x4x5.{ext}"), binary, shape)
    print("wrote shape_box_4x4x5.bin/.txt")


if __name__ == "__main__":
    main()
