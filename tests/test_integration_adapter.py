"""The reference-side IForce3 adapter shown in INTEGRATION.md section 1.

CPU: the block compiles against the reference's own
ms/fem/TetrahedronElementSet3.h and Params.h (no stand-ins; the reference
tree is read in place) and links against libcorofem.so, as a concrete
IForce3 (every pure virtual overridden) with EditableTetrahedronElementSet3's
edition calls, BeginEdition(num_nodes, vec_r, num_elements, params)
(TetrahedronElementSet3.h:250).
GPU: the built checker drives the adapter with host arrays the way the leaf
does and finds every result (f_e, df_x twice, f_d, V) bit-identical to the
same operators called through the bulk C-ABI on device arrays.
"""
import os
import subprocess

import pytest

from conftest import ROOT

REFSRC = "/root/reference/src"
CHECK = os.path.join(ROOT, "oracle", "_ref", "adapter_check")


def test_adapter_block_is_the_edition_interface():
    import sys
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from extract_adapter import adapter_block
    with open(os.path.join(ROOT, "INTEGRATION.md")) as f:
        code = adapter_block(f.read())
    assert "void BeginEdition(unsigned int num_nodes, const Vec3* vec_r, unsigned int num_elements, const Params& params)" in code
    assert "cf_tetset_begin_edition(num_nodes, reinterpret_cast<const float*>(vec_r), num_elements, &q, &m_pSet)" in code
    for name in ("BeginEvaluation", "df_x", "df_v", "EndEvaluation", "CreateCache", "DestroyCache", "AddElement",
                 "EndEdition", "SetParams"):
        assert name in code


@pytest.mark.skipif(not os.path.isdir(REFSRC), reason="the reference tree is not mounted here")
def test_adapter_compiles_against_reference_headers(cf):
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "_ref/adapter_check"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert os.path.exists(CHECK)


@pytest.mark.gpu
def test_adapter_runs_bitexact(gpu):
    if not os.path.exists(CHECK):
        pytest.skip("adapter checker not built (needs the reference headers at build time)")
    r = subprocess.run([CHECK], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "adapter ok" in r.stdout
