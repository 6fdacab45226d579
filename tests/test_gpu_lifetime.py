"""Object lifetimes around ptc_destroy, and contexts on a caller's stream.

ptc_destroy releases all device memory of a context at once; a device
tensor, key table, batch plan or row-size model freed afterwards releases
only its host record and makes no HIP call (include/partc.h, lifetime rule).
The teardown case runs in a fresh child process that calls the C ABI
directly (no Python context object to order the frees), so a use-after-free
or a HIP call after the context is gone shows as a non-zero exit.
"""
import ctypes as C
import subprocess
import sys
import textwrap
from pathlib import Path

import numpy as np
import pytest

import partc_amd as P
from oracle import orc
from tests.helpers import max_rel, sort_coo

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]

CHILD = textwrap.dedent(
    """
    import ctypes as C, sys
    sys.path.insert(0, {root!r})
    import numpy as np
    import partc_amd as P
    L = P.lib()
    P._mlp_sigs(L)
    ctx = C.c_void_p()
    assert L.ptc_create(C.byref(ctx), 0) == 0
    h, v = P.gen_coo(5, [60, 50, 40], 4000, P.DIST_UNIFORM)
    A, B = P.self_contraction_inputs(h, v, [1, 2])
    tA, tB = C.c_void_p(), C.c_void_p()
    va, vb = A.view(), B.view()
    assert L.ptc_dtensor_upload(ctx, C.byref(va), C.byref(tA)) == 0
    assert L.ptc_dtensor_upload(ctx, C.byref(vb), C.byref(tB)) == 0
    HA, HB = C.c_void_p(), C.c_void_p()
    assert L.ptc_hashtable_build(ctx, tB, C.byref(HB)) == 0
    assert L.ptc_hashtable_build(ctx, tA, C.byref(HA)) == 0
    O = C.c_void_p()
    st = P.Stats()
    assert L.ptc_tc_fksCuckoo(ctx, HA, HB, tA, tB, None, C.byref(O), C.byref(st)) == 0
    m = C.c_void_p()
    assert L.ptc_mlp_create(ctx, 3, C.byref(m)) == 0
    # contexts destroyed while everything above is still live ...
    L.ptc_destroy(ctx)
    print("destroyed", flush=True)
    # ... and every object freed afterwards: tables before tensors, the
    # output, the model
    L.ptc_hashtable_free(HA)
    L.ptc_hashtable_free(HB)
    L.ptc_dtensor_free(O)
    L.ptc_dtensor_free(tA)
    L.ptc_dtensor_free(tB)
    L.ptc_mlp_free(m)
    print("freed after destroy", flush=True)
    """
)


def test_frees_after_destroy_exit_cleanly(tmp_path):
    script = tmp_path / "teardown_child.py"
    script.write_text(CHILD.format(root=str(ROOT)))
    print("[case] child: destroy context, then free tables, tensors, output, model", flush=True)
    r = subprocess.run([sys.executable, str(script)], capture_output=True, text=True, timeout=100)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
    assert "freed after destroy" in r.stdout


def test_context_on_caller_stream():
    """A context on a caller's stream computes what a context with its own
    stream computes, and leaves the stream to the caller after ptc_destroy."""
    from tests import hiprt

    s = hiprt.stream_create()
    h, v = P.gen_coo(1, [200, 300, 400], 20000, P.DIST_UNIFORM)
    A, B = P.self_contraction_inputs(h, v, [1, 2])
    print("[case] context on a caller's HIP stream", flush=True)
    c = P.Context(0, stream=s)
    assert c.stream() == s
    O = P.tc_fksCuckoo(P.buildHashTable(A, c), P.buildHashTable(B, c)).sorted()
    c.close()
    oh, ov, _ = orc.tc_fksCuckoo(A, B)
    oh, ov = sort_coo(oh, ov)
    assert np.array_equal(O.hedges_array, oh)
    assert max_rel(O.val_array, ov) <= 1e-12
    # the stream is still the caller's: usable, then destroyed by the caller
    assert hiprt.hip().hipStreamSynchronize(C.c_void_p(s)) == 0
    hiprt.stream_destroy(s)
    with pytest.raises(P.PtcError):
        P.Context(0, stream=0)
