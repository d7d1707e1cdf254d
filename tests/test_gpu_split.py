"""Split-precision modular (CRT) contractions against the FP64 oracle and
the FP64 GPU path: int8 residues (default 11 moduli) and bf16 / fp16
residues (default 10 moduli).  Bar: relative Frobenius error <= 1e-10 on
int2_grad1_u12_ao and int_2e_ao with the default moduli."""
import ctypes

import numpy as np
import pytest

from cutc_amd import cutc, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-10


def relerr(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


def run(torch, s, **opt):
    ds = cutc.DeviceSystem.upload(s)
    X, I = cutc.alloc_outputs(ds)
    st = cutc.cutc_int_device(ds, X, I, cutc.options(**opt))
    torch.cuda.synchronize()
    return X.cpu().numpy(), I.cpu().numpy(), st


@pytest.mark.parametrize("precision", ["int8", "bf16crt", "fp16crt"])
@pytest.mark.parametrize("shape", [(1, 1, 1, 1), (2, 5, 3, 7), (3, 24, 37, 129), (7, 61, 77, 2049),
                                   (33, 40, 65, 333), (4, 30, 300, 5000)])
def test_modular_matches_oracle_small(torch, shape, precision):
    n_nuc, n_ao, n1, n2 = shape
    print(f"{precision} shape n_nuc={n_nuc} n_ao={n_ao} n1={n1} n2={n2}", flush=True)
    s = synth.make_system(n_nuc, n_ao, n1, n2, 300 + n_ao, name="int8")
    X, I, _ = run(torch, s, precision=precision)
    Xo = oracle.int2_grad1_u12_ao(s)
    Io = oracle.int_2e_ao(s, Xo)
    ex, ei = relerr(X, Xo), relerr(I, Io)
    print(f"  relerr X {ex:.3e} int_2e_ao {ei:.3e}", flush=True)
    assert ex <= TOL and ei <= TOL


@pytest.mark.parametrize("precision", ["int8", "bf16crt"])
def test_modular_triatomic_and_chunking(torch, precision):
    s = synth.triatomic()
    X, I, st = run(torch, s, precision=precision)
    Xo = oracle.int2_grad1_u12_ao(s)
    Io = oracle.int_2e_ao(s, Xo)
    ex, ei = relerr(X, Xo), relerr(I, Io)
    print(f"triatomic {precision}: relerr X {ex:.3e} int_2e_ao {ei:.3e}", flush=True)
    assert ex <= TOL and ei <= TOL
    Xc, Ic, stc = run(torch, s, precision=precision, n1_chunk=640)
    assert stc["n1_chunks"] > 1
    assert relerr(Xc, X) <= 1e-13 and relerr(Ic, I) <= 1e-12


def k_pass_limit(s, precision):
    """a device memory limit under which the planner cuts stage 1's K (r2)
    into several passes, found on the host"""
    big = cutc.make_plan(s.n1, s.n2, s.n_ao, s.n_nuc, 1 << 40, precision=precision)
    assert big.k_chunk == big.ldk
    for frac in (0.8, 0.7, 0.6, 0.5, 0.45, 0.4, 0.35, 0.3):
        lim = int(big.bytes_stage1 * frac)
        try:
            p = cutc.make_plan(s.n1, s.n2, s.n_ao, s.n_nuc, lim, precision=precision)
        except cutc.CutcError:
            continue
        if p.k_chunk < p.ldk and p.bytes_stage2 <= lim:
            return lim, p
    raise AssertionError("no limit gives K passes")


@pytest.mark.parametrize("precision", ["int8", "bf16crt", "fp16crt"])
def test_modular_k_passes_match_one_pass(torch, precision):
    """Under a memory limit the modular stage 1 runs K in several passes
    whose residue products are added modulo each modulus: X and int_2e_ao
    are bit-identical to the one-pass build (exact integer arithmetic), and
    the passes' list-driven contractions to the same passes run dense."""
    s = synth.make_system(6, 40, 300, 5000, 77, name="kpass")
    lim, p = k_pass_limit(s, precision)
    print(f"{precision} K passes: limit {lim / 1e6:.0f} MB, k_chunk {p.k_chunk} of ldk {p.ldk}", flush=True)
    X1, I1, _ = run(torch, s, precision=precision)
    Xk, Ik, st = run(torch, s, precision=precision, mem_limit=lim)
    assert np.array_equal(Xk, X1) and np.array_equal(Ik, I1)
    assert st["workspace_bytes"] <= lim
    Xd, Id, std = run(torch, s, precision=precision, mem_limit=lim, dense=1)
    print(f"  k_density of the passes {st['k_density']:.3f} (dense {std['k_density']:.3f})", flush=True)
    assert np.array_equal(Xd, X1) and np.array_equal(Id, I1)
    assert 0.0 < st["k_density"] <= 1.0 and std["k_density"] == 1.0


def test_int8_fewer_moduli_degrade_gracefully(torch):
    s = synth.triatomic(n1=200, n2=3000)
    Xo = oracle.int2_grad1_u12_ao(s)
    errs = []
    for n in (8, 10, 12, 14):
        X, I, _ = run(torch, s, precision="int8", slices=n)
        errs.append(relerr(X, Xo))
    print("moduli 8/10/12/14 relerr X:", [f"{e:.2e}" for e in errs], flush=True)
    assert errs[0] < 1e-4 and errs[1] <= TOL and errs[2] <= TOL
    assert errs[0] > errs[1] > errs[2]


@pytest.mark.parametrize("precision", ["int8", "bf16crt"])
def test_modular_cube20_full_basis_subset(torch, precision):
    """20 nuclei, n_ao = 200, all 200000 r2 points, 24 r1 points."""
    s = synth.cube20(n1=24, n2=200000)
    X, I, st = run(torch, s, precision=precision)
    Xf, If, _ = run(torch, s)
    Xo = oracle.int2_grad1_u12_ao(s)
    ex, exf = relerr(X, Xo), relerr(X, Xf)
    print(f"cube20 r1<24 {precision}: relerr X vs oracle {ex:.3e} vs FP64 GPU {exf:.3e}", flush=True)
    assert ex <= TOL
    Io = oracle.int_2e_ao(s, Xo)
    ei, eif = relerr(I, Io), relerr(I, If)
    print(f"cube20 r1<24 {precision}: relerr int_2e_ao vs oracle {ei:.3e} vs FP64 GPU {eif:.3e}", flush=True)
    assert ei <= TOL


# ------------------------------------------------------------------ packed GEMM

class Crt(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("bits", ctypes.c_int), ("m", ctypes.c_int * 16), ("y", ctypes.c_int * 16),
                ("inv_m", ctypes.c_double * 16), ("Mi", (ctypes.c_double * 3) * 16), ("M", ctypes.c_double * 3),
                ("Minv", ctypes.c_double), ("kind", ctypes.c_int)]


def pack(X, rows_pad):
    """Fragment-packed layout of the int8 matrix X (rows, kpad): the 2048-byte
    tile (row block rb, K step kt) at (rb KT + kt) 2048 holds byte
    k = 64 kt + 32 h + 16 c + e of row 32 rb + r at c 1024 + (32 h + r) 16 + e."""
    rows, kpad = X.shape
    Xp = np.zeros((rows_pad, kpad), np.int8)
    Xp[:rows] = X
    RB, KT = rows_pad // 32, kpad // 64
    t = Xp.reshape(RB, 32, KT, 2, 2, 16)  # rb, r, kt, h, c, e
    return np.ascontiguousarray(t.transpose(0, 2, 4, 3, 1, 5)).reshape(-1)


@pytest.mark.parametrize("M,N,K", [(37, 5, 100), (1100, 300, 5000), (512, 130, 131328), (600, 128, 262000)])
def test_gemm_pk_matches_exact_products(torch, M, N, K):
    """cutc_crt_gemm_pk against float64 products (exact: |sum| < 2^53) mod m,
    with rows of -128 (the largest |product|) at K beyond one int32 segment."""
    print(f"gemm_pk M={M} N={N} K={K}", flush=True)
    L = cutc.lib()
    L.cutc_crt_setup.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Crt)]
    L.cutc_crt_gemm_pk.argtypes = [ctypes.POINTER(Crt), ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                   ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    c = Crt()
    assert L.cutc_crt_setup(K, 6, 0, ctypes.byref(c)) == 0
    n = 2  # moduli 256 and 255
    c.n = n
    kpad = (K + 767) // 768 * 768
    rng = np.random.default_rng(K)
    ra = rng.integers(-128, 128, (n, M, kpad)).astype(np.int8)
    rb = rng.integers(-128, 128, (n, N, kpad)).astype(np.int8)
    ra[:, :, K:] = 0
    rb[:, :, K:] = 0
    ra[0, 0, :K] = -128
    rb[0, 0, :K] = -128
    ma, nb = (M + 511) // 512 * 512, (N + 255) // 256 * 256
    A = torch.from_numpy(np.concatenate([pack(ra[i], ma) for i in range(n)])).cuda()
    B = torch.from_numpy(np.concatenate([pack(rb[i], nb) for i in range(n)])).cuda()
    C = torch.zeros(n * M * N, dtype=torch.int8, device="cuda")
    torch.cuda.synchronize()
    rc = L.cutc_crt_gemm_pk(ctypes.byref(c), A.data_ptr(), M, B.data_ptr(), N, kpad, C.data_ptr(), 0, None)
    torch.cuda.synchronize()
    assert rc == 0
    got = C.cpu().numpy().reshape(n, M, N).astype(np.int64)
    # a second K pass of the same products added to the residues in C
    rc = L.cutc_crt_gemm_pk(ctypes.byref(c), A.data_ptr(), M, B.data_ptr(), N, kpad, C.data_ptr(), 1, None)
    torch.cuda.synchronize()
    assert rc == 0
    got2 = C.cpu().numpy().reshape(n, M, N).astype(np.int64)
    for i in range(n):
        m = c.m[i]
        exact = (ra[i].astype(np.float64) @ rb[i].astype(np.float64).T).astype(np.int64)
        assert np.array_equal(got2[i] % m, (2 * exact) % m), f"modulus {m}, accumulated"
        want = exact % m
        assert np.array_equal(got[i] % m, want), f"modulus {m}"
        assert got[i].min() >= -128 and got[i].max() <= 127


def pack16(D, rows_pad):
    """16-bit fragment-packed layout of a (rows, kpad) plane: the 2048-byte
    tile (row block, 32-wide K step) holds element k = 16 c + 8 h + e of row
    r at element c 512 + (32 h + r) 8 + e"""
    rows, kpad = D.shape
    Pl = np.zeros((rows_pad, kpad), D.dtype)
    Pl[:rows] = D
    t = Pl.reshape(rows_pad // 32, 32, kpad // 32, 2, 2, 8)  # rb, r, kt, c, h, e
    return np.ascontiguousarray(t.transpose(0, 2, 3, 4, 1, 5)).reshape(-1)


def to16(D, f16):
    """bf16 / fp16 bit patterns of small integers (exact in both)"""
    if f16:
        return D.astype(np.float16).view(np.uint16)
    return (D.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)


@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("M,N,K,accum", [(37, 5, 100, 0), (1100, 300, 5000, 0), (512, 130, 131328, 0),
                                         (600, 128, 262000, 0), (600, 130, 5000, 1)])
def test_gemm16_matches_exact_products(torch, kind, M, N, K, accum):
    """The bf16 / fp16 modular GEMM against float64 products (exact) mod m,
    with rows of +-127 (the largest |product|, summed far beyond 2^24) and
    K beyond several FP32 reduction windows; accum: the residues already in
    C are added (one K pass of several)."""
    print(f"gemm16 kind={kind} M={M} N={N} K={K} accum={accum}", flush=True)
    L = cutc.lib()
    L.cutc_crt_setup.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Crt)]
    L.cutc_crt_gemm_pk.argtypes = [ctypes.POINTER(Crt), ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                   ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    c = Crt()
    assert L.cutc_crt_setup(K, 6, kind, ctypes.byref(c)) == 0
    n = 2  # moduli 255 and 253
    c.n = n
    kpad = (K + 767) // 768 * 768
    rng = np.random.default_rng(K + kind)
    ra = rng.integers(-127, 128, (n, M, kpad))
    rb = rng.integers(-127, 128, (n, N, kpad))
    ra[:, :, K:] = 0
    rb[:, :, K:] = 0
    ra[0, 0, :K] = -127
    rb[0, 0, :K] = -127
    ra[1, 0, :K] = 127
    rb[1, 0, :K] = -127
    ma, nb = (M + 511) // 512 * 512, (N + 255) // 256 * 256
    A = torch.from_numpy(np.concatenate([pack16(to16(ra[i], kind == 2), ma) for i in range(n)]).view(np.int16)).cuda()
    B = torch.from_numpy(np.concatenate([pack16(to16(rb[i], kind == 2), nb) for i in range(n)]).view(np.int16)).cuda()
    C0 = np.zeros((n, M, N), np.int64)
    if accum:
        for i in range(n):
            C0[i] = rng.integers(-(c.m[i] // 2), (c.m[i] - 1) // 2 + 1, (M, N))
    C = torch.from_numpy(C0.astype(np.int8).reshape(-1)).cuda()
    torch.cuda.synchronize()
    rc = L.cutc_crt_gemm_pk(ctypes.byref(c), A.data_ptr(), M, B.data_ptr(), N, kpad, C.data_ptr(), accum, None)
    torch.cuda.synchronize()
    assert rc == 0
    got = C.cpu().numpy().reshape(n, M, N).astype(np.int64)
    for i in range(n):
        m = c.m[i]
        exact = (ra[i].astype(np.float64) @ rb[i].astype(np.float64).T).astype(np.int64) + C0[i]
        assert np.array_equal(got[i] % m, exact % m), f"modulus {m}"
        assert got[i].min() >= -(m // 2) and got[i].max() <= (m - 1) // 2
