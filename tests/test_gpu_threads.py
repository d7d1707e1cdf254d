"""Host-array entry points called from several threads at once.

The library serialises calls that use its device workspace process-wide (one
lock around each build, the workspace release included), and every host-array
call owns its input/output buffers and transfer pipeline, so concurrent
callers must each get the answer a lone call gives.  ctypes drops the GIL
around foreign calls, so the threads below are inside the library together.
Bar: FP64 1e-12 (split int8 1e-10) against the oracle and the committed
golden fixture."""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from cutc_amd import cutc, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "chain_tiny.npz")


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def test_concurrent_host_entry_points():
    import torch
    assert torch.cuda.is_available()
    with np.load(FIX) as d:
        g = {k: d[k] for k in d.files}
    sa = synth.triatomic(n1=300, n2=1200)
    sb = synth.make_system(4, 40, 257, 999, seed=77, name="four")
    want_a = oracle.cutc_int(sa)
    want_b = oracle.cutc_int(sb)
    ne_a, ne_b = (int(v) for v in g["ne"])

    def job(kind):
        if kind == "int_a":
            X, I = cutc.cutc_int(sa)
            return max(rel(X, want_a[0]), rel(I, want_a[1])), 1e-12
        if kind == "int_b_int8":
            X, I, _ = cutc.cutc_int_ex(sb, cutc.options(precision="int8"))
            return max(rel(X, want_b[0]), rel(I, want_b[1])), 1e-10
        if kind == "mo":
            G = cutc.cutc_ao_to_mo(g["mo_l_coef"], g["mo_r_coef"], g["int2_grad1_u12_ao"])
            return rel(G, g["int2_grad1_u12"]), 1e-12
        o2, o1, o0 = cutc.cutc_no(ne_a, ne_b, g["wr1"], g["mos_l"], g["mos_r"], g["int2_grad1_u12"])
        return max(rel(o2, g["no_2e"]), rel(o1, g["no_1e"]), abs(o0 - g["no_0e"][0]) / abs(g["no_0e"][0])), 1e-12

    kinds = ["int_a", "int_b_int8", "mo", "no"] * 3
    print(f"concurrent host calls: {len(kinds)} on 4 threads", flush=True)
    with ThreadPoolExecutor(max_workers=4) as ex:
        results = list(ex.map(job, kinds))
    for kind, (err, tol) in zip(kinds, results):
        assert err <= tol, (kind, err)
