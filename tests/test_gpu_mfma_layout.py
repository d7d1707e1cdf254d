"""Lane layouts of the MFMA instructions the contraction kernels use, checked
on the GPU with one instruction on known operands (in-bounds probe kernels in
tests/hip/mfma_probe.hip, a test-only library built beside the tests, not part
of libcutcint.so).  The kernels' epilogues and loaders assume exactly these
mappings."""
import ctypes
import os

import numpy as np
import pytest

from cutc_amd import cutc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


PROBE_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hip", "libmfmaprobe.so")
_P = None


def probe_lib():
    """tests/hip/libmfmaprobe.so, loaded after libcutcint.so (and torch), so
    that the process keeps one HIP runtime"""
    global _P
    if _P is None:
        cutc.lib()
        assert os.path.exists(PROBE_LIB), f"{PROBE_LIB} missing: run __graft_entry__.build()"
        _P = ctypes.CDLL(PROBE_LIB)
    return _P


def probe(torch, kind, a, b, out_dtype, nout):
    L = probe_lib()
    L.cutc_mfma_probe.restype = ctypes.c_int
    L.cutc_mfma_probe.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    da = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    db = torch.from_numpy(np.ascontiguousarray(b)).cuda()
    d = torch.zeros(64 * nout, dtype=out_dtype, device="cuda")
    assert L.cutc_mfma_probe(kind, ctypes.c_void_p(da.data_ptr()), ctypes.c_void_p(db.data_ptr()),
                             ctypes.c_void_p(d.data_ptr())) == 0
    return d.cpu().numpy().reshape(64, nout)


def check(D, C, row, col, label):
    bad = []
    for l in range(64):
        for r in range(D.shape[1]):
            i, j = row(l, r), col(l, r)
            if D[l, r] != C[i, j]:
                cand = np.argwhere(C == D[l, r])[:3].tolist()
                bad.append((l, r, (i, j), cand))
    assert not bad, f"{label}: {len(bad)} mismatches, first {bad[:6]}"


def test_f64_16x16x4(torch):
    rng = np.random.default_rng(0)
    A = rng.integers(-8, 9, (16, 4)).astype(np.float64)
    B = rng.integers(-8, 9, (4, 16)).astype(np.float64)
    a = np.array([A[l % 16, l // 16] for l in range(64)])
    b = np.array([B[l // 16, l % 16] for l in range(64)])
    D = probe(torch, 0, a, b, torch.float64, 4)
    # result element r of lane l: row 4 r + l / 16, column l % 16
    check(D, A @ B, lambda l, r: 4 * r + l // 16, lambda l, r: l % 16, "f64 16x16x4")


def _dlayout32(l, r):
    # 32x32 results: lane l, element r -> row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32
    return 8 * (r // 4) + 4 * (l // 32) + r % 4


def test_i8_32x32x32(torch):
    rng = np.random.default_rng(1)
    A = rng.integers(-128, 128, (32, 32)).astype(np.int64)
    B = rng.integers(-128, 128, (32, 32)).astype(np.int64)
    # lane l holds 16 consecutive k of row/column l % 32, starting at k = 16 (l / 32)
    a = np.zeros((64, 16), np.int8)
    b = np.zeros((64, 16), np.int8)
    for l in range(64):
        a[l] = A[l % 32, 16 * (l // 32):16 * (l // 32) + 16]
        b[l] = B[16 * (l // 32):16 * (l // 32) + 16, l % 32]
    D = probe(torch, 1, a, b, torch.int32, 16)
    check(D, A @ B, _dlayout32, lambda l, r: l % 32, "i8 32x32x32")


@pytest.mark.parametrize("kind,dt", [(2, "bf16"), (3, "f16")])
def test_16bit_32x32x16(torch, kind, dt):
    rng = np.random.default_rng(2 + kind)
    A = rng.integers(-64, 65, (32, 16)).astype(np.float32)
    B = rng.integers(-64, 65, (16, 32)).astype(np.float32)
    a = np.zeros((64, 8), np.float32)
    b = np.zeros((64, 8), np.float32)
    for l in range(64):
        a[l] = A[l % 32, 8 * (l // 32):8 * (l // 32) + 8]
        b[l] = B[8 * (l // 32):8 * (l // 32) + 8, l % 32]
    tdt = torch.bfloat16 if dt == "bf16" else torch.float16
    ta = torch.from_numpy(a).to(tdt).view(torch.int16).numpy()
    tb = torch.from_numpy(b).to(tdt).view(torch.int16).numpy()
    D = probe(torch, kind, ta, tb, torch.float32, 16)
    check(D, A @ B, _dlayout32, lambda l, r: l % 32, f"{dt} 32x32x16")
