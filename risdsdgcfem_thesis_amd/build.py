"""Build recipe for libcorofem.so (gfx950), in-tree.

    python risdsdgcfem_thesis_amd/build.py

compiles every source under csrc/ with hipcc for gfx950 into
risdsdgcfem_thesis_amd/libcorofem.so. Packed-FP32 instruction selection is
disabled for all device code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libcorofem.so")
ARCH = os.environ.get("COROFEM_ARCH", "gfx950")

SOURCES = ["cf_kernels.hip", "cf_batch_kernels.hip", "cf_kernels2d.hip", "cf_exact.hip", "cf_noalloc.hip", "cf_tiles.cpp", "cf_capi.cpp",
           "cf_batch.cpp", "cf_batch2.cpp", "cf_server.cpp", "cf_shape.cpp"]
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    for p in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if p and os.path.exists(p):
            return p
    return "hipcc"


def build(verbose=True, jobs=None):
    """Compiles every source (in parallel, `jobs` at once; default: the CPU
    count, at most 8) and links the library."""
    from concurrent.futures import ThreadPoolExecutor
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(bdir, os.path.splitext(src)[0] + ".o")
        cmds.append([hipcc(), "-x", "hip", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
                     "-Wall", "-Wno-unused-result", "-I", os.path.join(HERE, "..", "include"),
                     "-c", path, "-o", obj] + DEVICE_FLAGS)
        objs.append(obj)
    jobs = jobs or min(8, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", OUT] + objs
    _run(cmd, verbose)
    return OUT


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    # the host pass warns that it does not know the device-only feature flag
    noise = "'-packed-fp32-ops' is not a recognized feature"
    msgs = [ln for ln in (r.stdout + r.stderr).splitlines() if noise not in ln]
    if r.returncode != 0 or (verbose and msgs):
        print("\n".join(msgs), file=sys.stderr)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
