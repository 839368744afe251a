"""Writes the golden vectors in tests/golden/ from the reference's own compiled
math (oracle/_ref/libcorofem_ref.so, built by oracle/Makefile from the
reference sources). Run in a container where the reference is mounted:

    make -C oracle && python tests/golden/gen_golden.py

Fixtures (inputs and the reference's outputs, float32, seeded):
  ke_tets.npz      256 tets: rest corners r (12), E, nu -> Ke (144), volume
  rotations.npz    F = Q diag(s) (config-2 style, Q uniform on SO(3), s in [0.7,1.3])
                   and general F -> det F, Newton polar R, PhysBAM-SVD R, U, s, Vt
  dots.npz         vectors -> SSE real_array::dot
  tris.npz         2D triangles -> Ke (36), area; 2x2 F -> polar R, Compute_R_SVD R
"""
import os
import sys

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)
from oracle import orc  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def uniform_so3(rng, n):
    """Shoemake's uniform random unit quaternions -> rotation matrices."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], 1)
    x, y, z, w = q.T
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                  2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                  2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1)
    return R.reshape(n, 3, 3)


def main():
    R = orc.ref()
    if R is None:
        raise SystemExit("oracle/_ref not built (needs the reference sources)")
    fp = orc.fp
    rng = np.random.default_rng(20240601)

    # ---- Ke
    n = 256
    base = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1], np.float32)
    r = (base[None, :] * rng.uniform(0.005, 0.02, (n, 1)) + rng.uniform(-0.003, 0.003, (n, 12))).astype(np.float32)
    r += rng.uniform(-1, 1, (n, 1)).astype(np.float32) * np.tile(rng.uniform(-1, 1, (n, 3)), 4).astype(np.float32)
    E = (10.0 ** rng.uniform(4, 7, n)).astype(np.float32)
    nu = rng.uniform(0.2, 0.45, n).astype(np.float32)
    K = np.zeros((n, 144), np.float32)
    V = np.zeros(n, np.float32)
    Dinv = np.zeros((n, 9), np.float32)
    for i in range(n):
        ri = np.ascontiguousarray(r[i])
        ki = np.zeros(144, np.float32)
        V[i] = R.ref_compute_K(fp(ri), float(E[i]), float(nu[i]), fp(ki))
        K[i] = ki
        di = np.zeros(9, np.float32)
        R.ref_inv_dm(fp(ri), fp(di))
        Dinv[i] = di
    np.savez_compressed(os.path.join(OUT, "ke_tets.npz"), r=r, E=E, nu=nu, K=K, V=V, Dinv=Dinv)

    # ---- rotations
    m = 4096
    Q = uniform_so3(rng, m // 2)
    s = rng.uniform(0.7, 1.3, (m // 2, 3))
    F1 = np.einsum("nij,nj->nij", Q, s)
    F2 = np.eye(3)[None] + rng.uniform(-0.6, 0.6, (m // 2, 3, 3))
    F = np.concatenate([F1, F2]).astype(np.float32).reshape(m, 9)
    det = np.zeros(m, np.float32)
    Rpd = np.zeros((m, 9), np.float32)
    Rsvd = np.zeros((m, 9), np.float32)
    U = np.zeros((m, 9), np.float32)
    S = np.zeros((m, 3), np.float32)
    Vt = np.zeros((m, 9), np.float32)
    for i in range(m):
        f = np.ascontiguousarray(F[i])
        det[i] = R.ref_det3(fp(f))
        o = np.zeros(9, np.float32)
        R.ref_pd3(fp(f), float(det[i]), fp(o))
        Rpd[i] = o
        o = np.zeros(9, np.float32)
        R.ref_R_svd3(fp(f), float(det[i]), 0.1, fp(o))
        Rsvd[i] = o
        u, sv, vt = np.zeros(9, np.float32), np.zeros(3, np.float32), np.zeros(9, np.float32)
        R.ref_svd3(fp(f), fp(u), fp(sv), fp(vt))
        U[i], S[i], Vt[i] = u, sv, vt
    np.savez_compressed(os.path.join(OUT, "rotations.npz"), F=F, det=det, R_pd=Rpd, R_svd=Rsvd, U=U, S=S, Vt=Vt,
                        Q=Q.reshape(-1, 9).astype(np.float32), s=s.astype(np.float32))

    # ---- dots
    lens = np.array([1, 2, 3, 4, 5, 7, 8, 13, 100, 1001, 14883], np.int64)
    a = rng.standard_normal(lens.sum()).astype(np.float32)
    b = rng.standard_normal(lens.sum()).astype(np.float32)
    res = np.zeros(len(lens), np.float32)
    o = 0
    for i, L in enumerate(lens):
        aa, bb = np.ascontiguousarray(a[o:o + L]), np.ascontiguousarray(b[o:o + L])
        res[i] = R.ref_dot(fp(aa), fp(bb), int(L))
        o += L
    np.savez_compressed(os.path.join(OUT, "dots.npz"), lens=lens, a=a, b=b, dot=res)

    # ---- 2D
    n2 = 256
    tri = (np.array([0, 0, 1, 0, 0, 1], np.float32)[None] * rng.uniform(0.005, 0.05, (n2, 1))
           + rng.uniform(-0.002, 0.002, (n2, 6))).astype(np.float32)
    E2 = (10.0 ** rng.uniform(4, 7, n2)).astype(np.float32)
    nu2 = rng.uniform(0.2, 0.45, n2).astype(np.float32)
    K2 = np.zeros((n2, 36), np.float32)
    A2 = np.zeros(n2, np.float32)
    for i in range(n2):
        ti = np.ascontiguousarray(tri[i])
        k = np.zeros(36, np.float32)
        A2[i] = R.ref_compute_K2(fp(ti), float(E2[i]), float(nu2[i]), fp(k))
        K2[i] = k
    F2d = (np.eye(2)[None] + rng.uniform(-0.8, 0.8, (1024, 2, 2))).astype(np.float32).reshape(-1, 4)
    det2 = (F2d[:, 0] * F2d[:, 3] - F2d[:, 2] * F2d[:, 1]).astype(np.float32)
    R2pd = np.zeros((1024, 4), np.float32)
    R2svd = np.zeros((1024, 4), np.float32)
    for i in range(1024):
        f = np.ascontiguousarray(F2d[i])
        o = np.zeros(4, np.float32)
        R.ref_pd2(fp(f), float(det2[i]), fp(o))
        R2pd[i] = o
        o = np.zeros(4, np.float32)
        R.ref_R_svd2(fp(f), float(det2[i]), 0.1, fp(o))
        R2svd[i] = o
    np.savez_compressed(os.path.join(OUT, "tris.npz"), r=tri, E=E2, nu=nu2, K=K2, A=A2, F=F2d, det=det2,
                        R_pd=R2pd, R_svd=R2svd)
    print("golden fixtures written to", OUT)


if __name__ == "__main__":
    main()
