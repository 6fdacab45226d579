"""The command-line driver (partc-tc, the reference main.cpp's arguments):
reads a .tns file and its .meta, contracts on the GPU, writes the output --
which must equal the oracle's contraction of the same file."""
import subprocess

import numpy as np
import pytest

import partc_amd as P
from oracle import orc
from tests.helpers import max_rel, sort_coo

pytestmark = pytest.mark.gpu
CLI = P.LIB_PATH.parent / "partc-tc"


@pytest.mark.parametrize("method", ["two_pass", "upper_bound", "learned"])
def test_cli_contracts_a_tns_file(tmp_path, method):
    if not CLI.exists():
        pytest.fail(f"{CLI} is not built")
    h, v = P.gen_coo(21, [60, 70, 80, 5], 20000, P.DIST_ZIPF)
    T = P.Tensor(h, v)
    tns, meta, out = tmp_path / "a.tns", tmp_path / "a.meta", tmp_path / "o.tns"
    T.write_tns(tns)
    meta.write_text("contract modes 1 2 3\nof a 4-mode tensor\n3\n1 2 3 0\n")
    print(f"[case] cli {method}", flush=True)
    r = subprocess.run([str(CLI), str(tns), str(meta), "0", "2", method, str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    A, B = P.self_contraction_inputs(P.read_tns(tns).hedges_array, P.read_tns(tns).val_array, [1, 2, 3])
    oh, ov, _ = orc.tc_fksCuckoo(A, B)
    oh, ov = sort_coo(oh, ov)
    assert f"nonzeros in the output tensor {oh.shape[0]}" in r.stdout
    O = P.read_tns(out)
    h2, v2 = sort_coo(O.hedges_array, O.val_array)
    assert np.array_equal(h2, oh)
    assert max_rel(v2, ov) <= 1e-12
