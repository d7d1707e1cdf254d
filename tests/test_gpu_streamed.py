"""Streamed whole path (cutc_int_streamed_device): int_2e_ao without storing
X, against the stored path and the FP64 oracle, over several r1 chunks."""
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


def stored(torch, s, **opt):
    ds = cutc.DeviceSystem.upload(s)
    X, I = cutc.alloc_outputs(ds)
    cutc.cutc_int_device(ds, X, I, cutc.options(**opt))
    torch.cuda.synchronize()
    return I.cpu().numpy()


def streamed(torch, s, **opt):
    ds = cutc.DeviceSystem.upload(s)
    I = torch.full((s.n_ao,) * 4, float("nan"), dtype=torch.float64, device="cuda")
    st = cutc.cutc_int_streamed_device(ds, I, cutc.options(**opt))
    torch.cuda.synchronize()
    return I.cpu().numpy(), st


@pytest.mark.parametrize("precision", ["fp64", "int8", "bf16crt", "bf16"])
def test_streamed_matches_stored_and_oracle(torch, precision):
    s = synth.triatomic(n1=1500, n2=6000)
    print(f"streamed {precision} triatomic n1=1500 n2=6000 chunk 448", flush=True)
    Is, st = streamed(torch, s, precision=precision, n1_chunk=448)
    assert st["n1_chunks"] == 4
    Id = stored(torch, s, precision=precision)
    Xo = oracle.int2_grad1_u12_ao(s)
    Io = oracle.int_2e_ao(s, Xo)
    e_st, e_or = relerr(Is, Id), relerr(Is, Io)
    print(f"  vs stored {e_st:.2e} vs oracle {e_or:.2e}", flush=True)
    # vs stored: the chunks differ, so two split builds differ by up to twice
    # their own error
    assert e_st <= {"fp64": 1e-13, "int8": 1e-10, "bf16crt": 1e-10, "bf16": 1e-8}[precision]
    assert e_or <= {"fp64": 1e-12, "int8": 1e-10, "bf16crt": 1e-10, "bf16": 1e-8}[precision]
    # symmetrised: (kl|ab) = (ab|kl) in the n_ao^2 x n_ao^2 view
    n2 = s.n_ao ** 2
    M = Is.reshape(n2, n2)
    assert np.allclose(M, M.T, rtol=0, atol=1e-12 * np.abs(M).max())


def test_streamed_edge_cases(torch):
    s = synth.make_system(2, 7, 0, 50, 901, name="empty-r1")
    I, _ = streamed(torch, s)
    assert np.array_equal(I, np.zeros_like(I))
    s = synth.make_system(2, 7, 70, 0, 902, name="empty-r2")
    I, _ = streamed(torch, s, n1_chunk=16)
    assert np.array_equal(I, np.zeros_like(I))
    s = synth.make_system(3, 11, 133, 257, 903, name="ragged")
    I1, st = streamed(torch, s, n1_chunk=40)
    assert st["n1_chunks"] == 4
    I2 = stored(torch, s)
    assert relerr(I1, I2) <= 1e-13
