"""include/partc_bridge.hpp: the reference's C++ names (buildHashTable,
tc_fksCuckoo, Hash_Table_SPA, matricize_spGEMM, matricize_sort_spGEMM) over
the C ABI.  tests/bridge/demo.cpp is the reference driver's sequence
(main.cpp:59-118) written against those names.

CPU: the demo compiles and links against the reference's own struct Tensor
(tensor.hpp, read in place; skipped where that tree is absent) and against a
tensor type with the same fields.  GPU: the demo runs a contraction from a
.tns file and a .meta file, and its three results (tc_fksCuckoo,
tc_fksCuckoo<Hash_Table_SPA>, matricize_spGEMM) match the oracle.
"""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

import partc_amd as P

ROOT = Path(__file__).resolve().parents[1]
REF_TENSOR = Path("/root/reference/tensor.hpp")
DEMO = ROOT / "tests" / "bridge" / "demo.cpp"
FIELDS = ROOT / "tests" / "bridge" / "tensor_fields.hpp"


def build_demo(tensor_header: Path, out: Path) -> None:
    P.lib()  # libpartc.so is built
    cmd = ["g++", "-O1", "-std=c++17", f"-DTENSOR_HEADER=\"{tensor_header}\"", f"-I{ROOT / 'include'}",
           str(DEMO), f"-L{P.LIB_PATH.parent}", "-lpartc", f"-Wl,-rpath,{P.LIB_PATH.parent}", "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


@pytest.mark.skipif(not REF_TENSOR.exists(), reason="the reference tree is not present")
def test_bridge_builds_against_reference_tensor(tmp_path):
    build_demo(REF_TENSOR, tmp_path / "demo_ref")


def test_bridge_builds_against_field_compatible_tensor(tmp_path):
    build_demo(FIELDS, tmp_path / "demo")


@pytest.mark.gpu
def test_bridge_demo_matches_oracle(tmp_path):
    from oracle import orc
    from tests.helpers import max_rel, sort_coo

    h, v = P.gen_coo(7, [60, 70, 80], 8000, P.DIST_ZIPF)
    T = P.Tensor(h, v)
    tns = tmp_path / "t.tns"
    T.write_tns(tns)
    meta = tmp_path / "t.meta"
    meta.write_text("a 3-mode tensor\ncontract modes 1 and 2\n2\n1 2 0\n")
    exe = tmp_path / "demo"
    build_demo(FIELDS, exe)
    print("[case] bridge demo: 3-mode 60x70x80, 8k Zipf nnz, contract {1,2}", flush=True)
    r = subprocess.run([str(exe), str(tns), str(meta), str(tmp_path / "out")], capture_output=True, text=True,
                       timeout=100, env=dict(os.environ))
    assert r.returncode == 0, (r.stdout, r.stderr[-2000:])
    Tr = P.read_tns(tns)  # the values as the demo read them
    A, B = P.self_contraction_inputs(Tr.hedges_array, Tr.val_array, [1, 2])
    oh, ov, _ = orc.tc_fksCuckoo(A, B)
    oh, ov = sort_coo(oh, ov)
    for tag in ("tc", "spa", "mat"):
        O = P.read_tns(tmp_path / f"out.{tag}.tns")
        gh, gv = sort_coo(O.hedges_array, O.val_array)
        assert np.array_equal(gh, oh), tag
        assert max_rel(gv, ov) <= 1e-12, tag
