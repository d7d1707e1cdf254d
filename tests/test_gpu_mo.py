"""AO -> MO transform (cutc_ao_to_mo) on the GPU against the FP64 oracle
(oracle/oracle_mo.c), and the whole chain cutc_int -> cutc_ao_to_mo ->
cutc_no on device against the oracle's chain.  Bar: relative Frobenius error
<= 1e-12 (FP64 throughout; only the summation order differs)."""
import ctypes

import numpy as np
import pytest

from cutc_amd import cutc, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def inputs(n1, na, nm, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((3, n1, na, na))
    X = X + X.transpose(0, 1, 3, 2)  # symmetric in (k, l), as cutc_int returns it
    return rng.standard_normal((nm, na)), rng.standard_normal((nm, na)), X


def device_run(torch, cl, cr, X, mo_chunk=0, ncomp=3):
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    nm, n1 = cl.shape[0], X.shape[1]
    Xd = t(X)
    if ncomp == 4:  # cutc_int's 4-component buffer: the 4th is not read
        Xd = torch.cat([Xd, torch.full_like(Xd[:1], float("nan"))])
    out = torch.full((nm, nm, 3, n1), float("nan"), dtype=torch.float64, device="cuda")
    st = cutc.cutc_ao_to_mo_device(t(cl), t(cr), Xd, out, mo_chunk=mo_chunk)
    torch.cuda.synchronize()
    return out.cpu().numpy(), st


# K = n_ao: multiples of 16 take the 16-byte-load GEMM, others the 8-byte form
@pytest.mark.parametrize("n1,na,nm", [(300, 24, 24), (257, 32, 20), (100, 17, 13), (50, 9, 140), (40, 200, 200),
                                      (33, 64, 1), (7, 1, 5)])
def test_ao_to_mo_matches_oracle(torch, n1, na, nm):
    print(f"ao_to_mo n1={n1} n_ao={na} n_mo={nm}", flush=True)
    cl, cr, X = inputs(n1, na, nm, n1 + na + nm)
    want = oracle.ao_to_mo(cl, cr, X)
    got, st = device_run(torch, cl, cr, X, ncomp=4 if na == 24 else 3)
    print(f"  relerr {rel(got, want):.2e}", flush=True)
    assert rel(got, want) <= TOL
    assert st["n_chunks"] == 1 and st["flops"] == 6.0 * n1 * na * nm * (na + nm)


@pytest.mark.parametrize("chunk", [5, 128])
def test_ao_to_mo_chunked(torch, chunk):
    """MO indices in several passes (the last one short)"""
    print(f"ao_to_mo chunk={chunk}", flush=True)
    nm = 13 if chunk == 5 else 300
    cl, cr, X = inputs(23, 19, nm, 9)
    want = oracle.ao_to_mo(cl, cr, X)
    got, st = device_run(torch, cl, cr, X, mo_chunk=chunk)
    assert st["n_chunks"] == -(-nm // chunk)
    assert rel(got, want) <= TOL


def test_ao_to_mo_host_entry_and_edges(torch):
    print("ao_to_mo host entry, edges", flush=True)
    cl, cr, X = inputs(31, 11, 7, 4)
    assert rel(cutc.cutc_ao_to_mo(cl, cr, X), oracle.ao_to_mo(cl, cr, X)) <= TOL
    cl, cr, X = inputs(5, 0, 3, 4)  # no AOs: zero output
    got, _ = device_run(torch, cl, cr, X)
    assert not got.any()
    L = cutc.lib()
    assert L.cutc_ao_to_mo_device(-1, 2, 2, None, None, None, None, 0, 0, None, None) == -2
    assert L.cutc_ao_to_mo(2, 2, 2, None, None, None, None) == -2


@pytest.mark.parametrize("M,N,K,nsplit", [(130, 77, 200, 1), (64, 300, 201, 1), (300, 40, 1001, 3), (5, 3, 7, 1),
                                         (130, 50, 1000, 4), (257, 129, 96, 1)])
def test_gemm_tn_any_k(torch, M, N, K, nsplit):
    """cutc_gemm_tn_f64 in its three load forms: K a multiple of 16 (96), even K
    (200, 1000: 16-byte pairs up to K), odd K and leading dimensions (8-byte
    elements); one slice and split K"""
    print(f"gemm_tn any-K M={M} N={N} K={K} nsplit={nsplit}", flush=True)
    L = cutc.lib()
    L.cutc_gemm_tn_f64.restype = ctypes.c_int
    L.cutc_gemm_tn_f64.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                   ctypes.c_int64, ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
                                   ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    L.cutc_gemm_tn_split_count.restype = ctypes.c_int
    L.cutc_gemm_tn_split_count.argtypes = [ctypes.c_int64, ctypes.c_int]
    rng = np.random.default_rng(M + K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((N, K))
    ns = L.cutc_gemm_tn_split_count(K, nsplit)
    dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    C0 = rng.standard_normal((N, M))
    C = torch.from_numpy(C0).cuda()
    work = torch.zeros(ns * M * N, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    if nsplit == 1:
        rc = L.cutc_gemm_tn_f64(dA.data_ptr(), K, dB.data_ptr(), K, M, N, K, 0.5, -2.0, C.data_ptr(), M, 1, None, None)
        torch.cuda.synchronize()
        assert rc == 0
        got, want = C.cpu().numpy().T, 0.5 * A @ B.T - 2.0 * C0.T
    else:
        rc = L.cutc_gemm_tn_f64(dA.data_ptr(), K, dB.data_ptr(), K, M, N, K, 1.0, 0.0, None, M, nsplit,
                                work.data_ptr(), None)
        torch.cuda.synchronize()
        assert rc == 0
        got, want = work.cpu().numpy().reshape(ns, N, M).sum(0).T, A @ B.T
    assert rel(got, want) < 1e-14


def test_chain_int_mo_no(torch):
    """cutc_int_device -> cutc_ao_to_mo_device -> cutc_no_device, all on the
    device, against oracle.cutc_int -> oracle.ao_to_mo -> oracle.cutc_no"""
    print("chain cutc_int -> ao_to_mo -> cutc_no", flush=True)
    s = synth.triatomic(n1=150, n2=900, seed=21)
    nm = 16
    rng = np.random.default_rng(5)
    cl = rng.standard_normal((nm, s.n_ao)) / np.sqrt(s.n_ao)
    cr = cl + 0.05 * rng.standard_normal((nm, s.n_ao)) / np.sqrt(s.n_ao)
    chi = s.aos1[0]  # (n_ao, n1): AO values on grid 1
    mol, mor = cl @ chi, cr @ chi  # mos_l_in_r, mos_r_in_r as (n_mo, n1)
    # oracle chain
    Xo, _ = oracle.cutc_int(s)
    Go = oracle.ao_to_mo(cl, cr, Xo)
    want = oracle.cutc_no(4, 3, s.wr1, mol, mor, Go)
    # device chain
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    ds = cutc.DeviceSystem.upload(s)
    X, I = cutc.alloc_outputs(ds)
    cutc.cutc_int_device(ds, X, I)
    G = torch.empty((nm, nm, 3, s.n1), dtype=torch.float64, device="cuda")
    cutc.cutc_ao_to_mo_device(t(cl), t(cr), X, G)
    o2 = torch.empty((nm,) * 4, dtype=torch.float64, device="cuda")
    o1 = torch.empty((nm, nm), dtype=torch.float64, device="cuda")
    o0 = torch.empty(1, dtype=torch.float64, device="cuda")
    cutc.cutc_no_device(4, 3, t(s.wr1), t(mol), t(mor), G, o2, o1, o0)
    torch.cuda.synchronize()
    assert rel(G.cpu().numpy(), Go) <= TOL
    e = [rel(o2.cpu().numpy(), want[0]), rel(o1.cpu().numpy(), want[1]),
         abs(float(o0.cpu()[0]) - want[2]) / abs(want[2])]
    print(f"  relerr no_2e {e[0]:.2e} no_1e {e[1]:.2e} no_0e {e[2]:.2e}", flush=True)
    assert max(e) <= TOL
