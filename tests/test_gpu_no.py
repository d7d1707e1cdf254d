"""Normal ordering (cutc_no) on the GPU against the FP64 oracle
(oracle/oracle_no.c).  Bar: relative Frobenius error <= 1e-12 for no_2e and
no_1e, relative error <= 1e-12 for no_0e (FP64 throughout; only the summation
order differs)."""
import numpy as np
import pytest

from cutc_amd import cutc
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


def inputs(ng, nm, seed):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((nm, nm, 3, ng))
    w = rng.random(ng) * np.exp(rng.standard_normal(ng))
    L = rng.standard_normal((nm, ng))
    R = L + 0.1 * rng.standard_normal((nm, ng))
    return w, L, R, G


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def device_run(torch, na, nb, w, L, R, G, grid_chunk=0):
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    nm = G.shape[0]
    o2 = torch.full((nm,) * 4, float("nan"), dtype=torch.float64, device="cuda")
    o1 = torch.full((nm, nm), float("nan"), dtype=torch.float64, device="cuda")
    o0 = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    st = cutc.cutc_no_device(na, nb, t(w), t(L), t(R), t(G), o2, o1, o0, grid_chunk=grid_chunk)
    torch.cuda.synchronize()
    return o2.cpu().numpy(), o1.cpu().numpy(), float(o0.cpu()[0]), st


def check(got, want):
    e2, e1 = rel(got[0], want[0]), rel(got[1], want[1])
    e0 = abs(got[2] - want[2]) / max(abs(want[2]), 1e-300)
    print(f"  relerr no_2e {e2:.2e} no_1e {e1:.2e} no_0e {e0:.2e}", flush=True)
    assert e2 <= TOL and e1 <= TOL and e0 <= TOL


@pytest.mark.parametrize("ng,nm,na,nb", [(300, 8, 3, 3), (517, 12, 5, 3), (1000, 20, 6, 6), (64, 4, 1, 3),
                                         (257, 9, 9, 9), (130, 33, 11, 7), (3001, 40, 10, 10)])
def test_no_matches_oracle(torch, ng, nm, na, nb):
    print(f"cutc_no ng={ng} nm={nm} na={na} nb={nb}", flush=True)
    args = inputs(ng, nm, ng + nm)
    want = oracle.cutc_no(na, nb, *args)
    got = device_run(torch, na, nb, *args)
    check(got[:3], want)
    assert got[3]["n_chunks"] == 1


@pytest.mark.parametrize("na,nb", [(4, 4), (6, 2)])
def test_no_chunked_grid(torch, na, nb):
    """several grid chunks (the last one short) sum to the one-chunk result"""
    print(f"cutc_no chunks na={na} nb={nb}", flush=True)
    args = inputs(1111, 10, 77)
    want = oracle.cutc_no(na, nb, *args)
    got = device_run(torch, na, nb, *args, grid_chunk=250)
    assert got[3]["n_chunks"] == 5
    check(got[:3], want)


def test_no_host_entry_point(torch):
    """cutc_no on host arrays (the reference's entry point)"""
    print("cutc_no host entry", flush=True)
    args = inputs(400, 7, 5)
    want = oracle.cutc_no(3, 2, *args)
    got = cutc.cutc_no(3, 2, *args)
    check(got, want)


def test_no_host_entry_pipelined_upload(torch):
    """cutc_no on host arrays uploads int2_grad1_u12 in 8 grid parts through
    the transfer pipeline (here 4 staging pieces per part) while the build
    runs on the parts already in place; same result as the device entry point
    on the whole array"""
    print("cutc_no host entry, pipelined upload", flush=True)
    args = inputs(20000, 60, 31)
    got = cutc.cutc_no(8, 6, *args)
    want = device_run(torch, 8, 6, *args)
    check(got, want[:3])


def test_no_edge_cases(torch):
    print("cutc_no edge cases", flush=True)
    w, L, R, G = inputs(50, 5, 9)
    got = device_run(torch, 0, 0, w, L, R, G)  # no electrons
    assert not got[0].any() and not got[1].any() and got[2] == 0.0
    e = np.zeros(0)
    got = device_run(torch, 2, 2, e, L[:, :0], R[:, :0], G[..., :0])  # empty grid
    assert not got[0].any() and not got[1].any() and got[2] == 0.0
    with pytest.raises(cutc.CutcError):
        device_run(torch, 6, 2, w, L, R, G)  # more electrons than orbitals


@pytest.mark.parametrize("M,N,K,nsplit", [(12, 12, 2640, 2), (200, 130, 5 * 6016, 40), (300, 300, 4096, 1)])
def test_gemm_tn_split(torch, M, N, K, nsplit):
    """cutc_gemm_tn_f64 on K-contiguous rows: one slice (dense result) and K
    slices (partial sums added here) against numpy."""
    import ctypes
    print(f"gemm_tn M={M} N={N} K={K} nsplit={nsplit}", flush=True)
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
    work = torch.zeros(ns * M * N, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = L.cutc_gemm_tn_f64(dA.data_ptr(), K, dB.data_ptr(), K, M, N, K, 1.0, 0.0, work.data_ptr(), M, nsplit,
                            work.data_ptr(), None)
    torch.cuda.synchronize()
    assert rc == 0
    got = work.cpu().numpy().reshape(ns, N, M).sum(0).T  # [i, j]
    want = A @ B.T
    assert rel(got, want) < 1e-14
