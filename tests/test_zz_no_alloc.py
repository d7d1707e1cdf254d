"""cutc_int_device_noalloc (every scratch buffer in the library's fixed
device array) against cutc_int_device and the FP64 oracle through the C
ABI: two calls in a row, the second on a smaller system, in FP64 (bar
1e-12) and int8 (bar 1e-10); a system whose working set does not fit the
array is refused with CUTC_ERR_NOMEM before any launch."""
import numpy as np
import pytest

from cutc_amd import cutc, synth
from oracle import oracle

pytestmark = pytest.mark.gpu


def relerr(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need the MI355X"
    return t


def both(torch, s, precision):
    ds = cutc.DeviceSystem.upload(s)
    X, I = cutc.alloc_outputs(ds)
    Xn, In = cutc.alloc_outputs(ds)
    Xn.fill_(float("nan"))
    In.fill_(float("nan"))
    cutc.cutc_int_device(ds, X, I, cutc.options(precision))
    cutc.release_workspace()
    cutc.cutc_int_device_noalloc(ds, Xn, In, cutc.options(precision))
    torch.cuda.synchronize()
    return X.cpu().numpy(), I.cpu().numpy(), Xn.cpu().numpy(), In.cpu().numpy()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("int8", 1e-10)])
def test_noalloc_matches_device_entry_twice(torch, precision, tol):
    # the 3-nucleus workload's basis at a quarter of its grids, then a smaller system
    for s in (synth.triatomic(n1=1000, n2=4000), synth.make_system(5, 17, 300, 1100, 71, name="smaller")):
        print(f"noalloc {precision} {s.shape_str()}", flush=True)
        X, I, Xn, In = both(torch, s, precision)
        assert np.isfinite(Xn).all() and np.isfinite(In).all()
        ex, ei = relerr(Xn, X), relerr(In, I)
        Xo = oracle.int2_grad1_u12_ao(s)
        Io = oracle.int_2e_ao(s, Xo)
        eo = max(relerr(Xn, Xo), relerr(In, Io))
        print(f"  vs cutc_int_device X {ex:.2e} int_2e_ao {ei:.2e}; vs oracle {eo:.2e}", flush=True)
        assert ex <= tol and ei <= tol and eo <= tol


def test_noalloc_holds_no_device_memory_and_refuses_large(torch):
    s = synth.make_system(5, 17, 300, 1100, 72, name="mem")
    ds = cutc.DeviceSystem.upload(s)
    X, I = cutc.alloc_outputs(ds)
    cutc.cutc_int_device_noalloc(ds, X, I)
    torch.cuda.synchronize()
    cutc.release_workspace()
    free0 = torch.cuda.mem_get_info()[0]
    cutc.cutc_int_device_noalloc(ds, X, I)
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
    # n_ao = 200 with 200000 r2 points: the pair weights alone are 32 GB
    big = synth.cube20(n1=4, n2=200000)
    dsb = cutc.DeviceSystem.upload(big)
    Xb, Ib = cutc.alloc_outputs(dsb)
    with pytest.raises(cutc.CutcError) as e:
        cutc.cutc_int_device_noalloc(dsb, Xb, Ib)
    assert e.value.code == cutc.CUTC_ERR_NOMEM
