"""Batched build (cutc_int_batch_device): systems of different sizes in one
recorded hipGraph launch, against one streamed call per system and the
FP64 oracle."""
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
    assert t.cuda.is_available()
    return t


def systems():
    return [synth.make_system(3, 24, 700, 3000, 1101, name="b0"),
            synth.make_system(5, 41, 1300, 2500, 1102, name="b1"),
            synth.make_system(2, 7, 0, 500, 1103, name="b2-empty-r1"),
            synth.make_system(4, 30, 300, 0, 1104, name="b3-empty-r2"),
            synth.make_system(7, 55, 5000, 1800, 1105, name="b4")]


@pytest.mark.parametrize("precision,streams", [("fp64", 1), ("fp64", 0), ("fp64", 3), ("int8", 0), ("bf16", 2),
                                               ("bf16crt", 2)])
def test_batch_matches_single_calls(torch, precision, streams):
    """streams 0: the default, four graph branches for the modular forms, one for the others"""
    sys_ = systems()
    print(f"batch {precision} streams={streams}: {[s.shape_str() for s in sys_]}", flush=True)
    dss = [cutc.DeviceSystem.upload(s) for s in sys_]
    Ib = [torch.full((s.n_ao,) * 4, float("nan"), dtype=torch.float64, device="cuda") for s in sys_]
    opt = cutc.options(precision, n1_chunk=2048, streams=streams)
    st = cutc.cutc_int_batch_device(dss, Ib, opt)
    torch.cuda.synchronize()
    print(f"  graph nodes {st['n_nodes']}, branches {st['n_branches']}, launch {st['ms_launch']:.1f} ms, "
          f"setup {st['ms_setup']:.1f} ms", flush=True)
    assert st["n_systems"] == len(sys_) and st["n_nodes"] > len(sys_)
    want = streams or (4 if precision in ("int8", "bf16crt", "fp16crt") else 1)
    assert st["n_branches"] == min(want, len(sys_))
    for s, ds, I in zip(sys_, dss, Ib):
        Is = torch.full_like(I, float("nan"))
        cutc.cutc_int_streamed_device(ds, Is, cutc.options(precision, n1_chunk=2048))
        torch.cuda.synchronize()
        a, b = I.cpu().numpy(), Is.cpu().numpy()
        assert np.isfinite(a).all()
        assert relerr(a, b) <= 1e-13, s.name  # same chunks, same kernels
    # one system against the oracle
    s = sys_[1]
    Io = oracle.int_2e_ao(s, oracle.int2_grad1_u12_ao(s))
    e = relerr(Ib[1].cpu().numpy(), Io)
    print(f"  {s.name} vs oracle {e:.2e}", flush=True)
    assert e <= {"fp64": 1e-12, "int8": 1e-10, "bf16": 1e-6, "bf16crt": 1e-10}[precision]


def test_batch_replays_and_empty(torch):
    assert cutc.cutc_int_batch_device([], [])["n_systems"] == 0
    sys_ = systems()[:2]
    dss = [cutc.DeviceSystem.upload(s) for s in sys_]
    I1 = [torch.zeros((s.n_ao,) * 4, dtype=torch.float64, device="cuda") for s in sys_]
    I2 = [torch.zeros_like(I) for I in I1]
    cutc.cutc_int_batch_device(dss, I1)
    x = torch.ones(1 << 20, device="cuda").sum()  # eager work between the launches
    cutc.cutc_int_batch_device(dss, I2)
    torch.cuda.synchronize()
    assert float(x) == float(1 << 20)
    for a, b in zip(I1, I2):
        assert torch.equal(a, b)
