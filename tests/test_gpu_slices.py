"""Slice (Ozaki scheme I) split precision: bf16 / fp16 digit planes with
FP32 MFMA accumulation summed into FP64.  The slice GEMM is checked against
the same digit products computed exactly in float64, and the whole path
against the FP64 oracle; errors must fall with every extra slice, and bf16
and fp16 digits (the same integers) must give identical results."""
import ctypes

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


def digits(X, S):
    """row exponents e (max |x| < 2^e) and S integer digit planes, the same
    steps as oz_store_digits in ozaki1.hip"""
    mx = np.abs(X).max(axis=1)
    e = np.where(mx > 0, np.frexp(mx)[1], 0)
    y = np.ldexp(X, (8 - e)[:, None])
    D = []
    for _ in range(S):
        r = np.rint(y)
        D.append(r)
        y = (y - r) * 256.0
    return e, D


def pack16(D, rows_pad):
    """slice-packed layout of a (rows, kpad) plane of 16-bit values: the 2 KB
    tile (row block, 32-wide K step) holds element k = 16 c + 8 h + e of row
    r at c 512 + (32 h + r) 8 + e"""
    rows, kpad = D.shape
    P = np.zeros((rows_pad, kpad), D.dtype)
    P[:rows] = D
    t = P.reshape(rows_pad // 32, 32, kpad // 32, 2, 2, 8)  # rb, r, kt, c, h, e
    return np.ascontiguousarray(t.transpose(0, 2, 3, 4, 1, 5)).reshape(-1)


def to16(D, f16):
    if f16:
        return D.astype(np.float16).view(np.uint16)
    return (D.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)


@pytest.mark.parametrize("S", [2, 3, 4])
@pytest.mark.parametrize("f16", [0, 1])
def test_oz_gemm_matches_digit_products(torch, S, f16):
    M, N, K = 70, 200, 1000
    print(f"oz_gemm S={S} f16={f16} M={M} N={N} K={K}", flush=True)
    kpad = 1536
    rng = np.random.default_rng(S + 10 * f16)
    A = rng.standard_normal((M, K)) * np.exp(2 * rng.standard_normal((M, 1)))
    B = rng.standard_normal((N, K)) * np.exp(rng.standard_normal((N, K)))
    A[5] = 0.0
    ea, DA = digits(A, S)
    eb, DB = digits(B, S)
    assert max(np.abs(d).max() for d in DA + DB) <= 256

    def planes(D, rows):
        rp = (rows + 127) // 128 * 128
        out = []
        for d in D:
            full = np.zeros((rows, kpad))
            full[:, :K] = d
            out.append(pack16(to16(full, f16), rp))
        return torch.from_numpy(np.concatenate(out).view(np.int16)).cuda()

    pA, pB = planes(DA, M), planes(DB, N)
    dEA = torch.from_numpy(ea.astype(np.int32)).cuda()
    dEB = torch.from_numpy(eb.astype(np.int32)).cuda()
    C = torch.zeros(M * N, dtype=torch.float64, device="cuda")
    L = cutc.lib()
    L.cutc_oz_gemm.restype = ctypes.c_int
    L.cutc_oz_gemm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                               ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_void_p]
    torch.cuda.synchronize()
    rc = L.cutc_oz_gemm(S, f16, pA.data_ptr(), M, pB.data_ptr(), N, kpad, dEA.data_ptr(), dEB.data_ptr(),
                        C.data_ptr(), None)
    torch.cuda.synchronize()
    assert rc == 0
    got = C.cpu().numpy().reshape(M, N)
    want = np.zeros((M, N))
    for s in range(S):
        for t in range(S - s):
            want += np.ldexp(DA[s] @ DB[t].T, -8 * (s + t))  # exact integer products in float64
    want = np.ldexp(want, (ea[:, None] + eb[None, :]) - 16)
    scale = np.abs(want).max()
    assert np.abs(got - want).max() <= 1e-14 * scale
    assert np.array_equal(got[5], np.zeros(N))
    # the dropped slice products bound the error against the true product
    rel = relerr(got, A @ B.T)
    print(f"  vs A B^T: {rel:.2e}", flush=True)
    assert rel < 2.0 ** (-8 * S + 6)


def run(torch, s, **opt):
    ds = cutc.DeviceSystem.upload(s)
    X, I = cutc.alloc_outputs(ds)
    st = cutc.cutc_int_device(ds, X, I, cutc.options(**opt))
    torch.cuda.synchronize()
    return X.cpu().numpy(), I.cpu().numpy(), st


def test_slices_whole_path_against_oracle(torch):
    s = synth.triatomic(n1=600, n2=5000)
    Xo = oracle.int2_grad1_u12_ao(s)
    Io = oracle.int_2e_ao(s, Xo)
    ex, ei = {}, {}
    for prec in ("bf16", "fp16"):
        for S in (2, 3, 4):
            print(f"whole path {prec} S={S}", flush=True)
            X, I, _ = run(torch, s, precision=prec, slices=S)
            ex[prec, S], ei[prec, S] = relerr(X, Xo), relerr(I, Io)
            print(f"  X {ex[prec, S]:.2e} int_2e_ao {ei[prec, S]:.2e}", flush=True)
            if prec == "fp16":
                Xb, Ib, _ = run(torch, s, precision="bf16", slices=S)
                assert np.array_equal(X, Xb) and np.array_equal(I, Ib)
    for prec in ("bf16", "fp16"):
        assert ex[prec, 2] > ex[prec, 3] > ex[prec, 4]
        assert ex[prec, 4] < 1e-7 and ei[prec, 4] < 1e-7
        assert ex[prec, 3] < 1e-5 and ex[prec, 2] < 1e-3


def test_slices_chunking_and_bad_slices(torch):
    s = synth.triatomic(n1=700, n2=3000)
    X1, I1, st1 = run(torch, s, precision="bf16", slices=3)
    X2, I2, st2 = run(torch, s, precision="bf16", slices=3, n1_chunk=256)
    assert st2["n1_chunks"] > 1
    assert relerr(X2, X1) <= 1e-14 and relerr(I2, I1) <= 1e-13
    with pytest.raises(cutc.CutcError):
        run(torch, s, precision="bf16", slices=5)
