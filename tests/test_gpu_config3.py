"""64 independent systems (n_nuc U{3..12}, n_ao U{30..150}) at a twentieth
of the benchmark's grid sizes in one batched graph launch
(cutc_int_batch_device), against one streamed call per system and, for
the smallest system, the FP64 oracle."""
import numpy as np
import pytest

from cutc_amd import cutc, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12
TOL_SPLIT = 1e-10


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need the MI355X"
    return t


def rel_dev(torch, a, b):
    """relative Frobenius error of device tensor a against device tensor b"""
    nb = float(torch.linalg.vector_norm(b))
    return float(torch.linalg.vector_norm(a - b)) / (nb if nb > 0 else 1.0)


# ------------------------------------------------------------ 64 systems

@pytest.fixture(scope="module")
def batch64(torch):
    sys_ = synth.batch_systems(64, scale=0.05)
    print(f"batch64 at scale 0.05: n_ao {min(s.n_ao for s in sys_)}..{max(s.n_ao for s in sys_)}, "
          f"n1 {min(s.n1 for s in sys_)}..{max(s.n1 for s in sys_)}, "
          f"n2 {min(s.n2 for s in sys_)}..{max(s.n2 for s in sys_)}", flush=True)
    dss = [cutc.DeviceSystem.upload(s) for s in sys_]
    yield sys_, dss
    del dss
    cutc.release_workspace()
    torch.cuda.empty_cache()


@pytest.mark.timeout(900)
@pytest.mark.parametrize("precision", ["fp64", "int8"])
def test_batch64_matches_single_calls_and_oracle(torch, batch64, precision):
    sys_, dss = batch64
    opt = cutc.options(precision)
    Ib = [torch.full((s.n_ao,) * 4, float("nan"), dtype=torch.float64, device="cuda") for s in sys_]
    st = cutc.cutc_int_batch_device(dss, Ib, opt)
    torch.cuda.synchronize()
    print(f"  batch64 {precision}: {st['n_nodes']} graph nodes, launch {st['ms_launch']:.0f} ms", flush=True)
    assert st["n_systems"] == 64
    worst = 0.0
    for s, ds, I in zip(sys_, dss, Ib):
        Is = torch.full_like(I, float("nan"))
        cutc.cutc_int_streamed_device(ds, Is, opt)
        torch.cuda.synchronize()
        assert bool(torch.isfinite(I).all()), s.name
        worst = max(worst, rel_dev(torch, I, Is))
        del Is
    print(f"  batch vs single calls: max {worst:.2e}", flush=True)
    assert worst <= 1e-13
    # the smallest basis against the oracle
    i = min(range(64), key=lambda k: (sys_[k].n_ao, sys_[k].n1))
    s = sys_[i]
    Io = oracle.int_2e_ao(s, oracle.int2_grad1_u12_ao(s))
    e = rel_dev(torch, Ib[i], torch.from_numpy(Io).cuda())
    print(f"  {s.name} ({s.shape_str()}) vs oracle {e:.2e}", flush=True)
    assert e <= (TOL_FP64 if precision == "fp64" else TOL_SPLIT)
    del Ib
    cutc.release_workspace()
