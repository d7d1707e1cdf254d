"""ctypes front end of the FP64 CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker.  The product library
never calls it.
"""
from __future__ import annotations

import ctypes
import glob
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i64 = ctypes.c_int64
_L = None


def build():
    r = subprocess.run(["make", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stdout + r.stderr)


def _find_openblas():
    import numpy
    d = os.path.join(os.path.dirname(os.path.dirname(numpy.__file__)), "numpy.libs")
    c = sorted(glob.glob(os.path.join(d, "libscipy_openblas64_*.so")))
    return c[0] if c else None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.oracle_use_blas.restype = ctypes.c_int
        L.oracle_use_blas.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        L.oracle_blas_loaded.restype = ctypes.c_int
        L.oracle_threads.restype = ctypes.c_int
        L.oracle_powd_int.restype = ctypes.c_double
        L.oracle_powd_int.argtypes = [ctypes.c_double, ctypes.c_int]
        L.oracle_dgemm.restype = None
        L.oracle_dgemm.argtypes = [ctypes.c_char, ctypes.c_char, _i64, _i64, _i64, ctypes.c_double, _dp, _i64,
                                   _dp, _i64, ctypes.c_double, _dp, _i64]
        L.oracle_jast_bh.restype = None
        L.oracle_jast_bh.argtypes = [_i64, _i64, _i64, _i64, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _ip,
                                     _ip, _ip, _dp]
        L.oracle_long_range.restype = None
        L.oracle_long_range.argtypes = [_i64, _i64, _i64, _i64, ctypes.c_int, _dp, _dp, _dp]
        L.oracle_herm.restype = None
        L.oracle_herm.argtypes = [_i64, ctypes.c_int, _dp, _dp, _dp]
        L.oracle_nonherm.restype = None
        L.oracle_nonherm.argtypes = [_i64, ctypes.c_int, _dp, _dp, _dp]
        L.oracle_trans_inplace.restype = None
        L.oracle_trans_inplace.argtypes = [_dp, _i64]
        L.oracle_int2_grad1_u12_ao.restype = ctypes.c_int
        L.oracle_int2_grad1_u12_ao.argtypes = [_i64, _i64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                               _dp, _dp, _dp, _ip, _ip, _ip, _dp, _i64, _i64]
        L.oracle_int_2e_ao.restype = ctypes.c_int
        L.oracle_int_2e_ao.argtypes = [_i64, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.oracle_int_2e_ao_cols.restype = ctypes.c_int
        L.oracle_int_2e_ao_cols.argtypes = [_i64, ctypes.c_int, _dp, _dp, _dp, _i64, _i64, _dp]
        L.oracle_dd_2e.restype = ctypes.c_int
        L.oracle_dd_2e.argtypes = [_i64, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.oracle_cutc_no.restype = ctypes.c_int
        L.oracle_cutc_no.argtypes = [_i64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.oracle_aos.restype = ctypes.c_int
        L.oracle_aos.argtypes = [_i64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _ip, _ip, _ip, _dp, _dp,
                                 _dp]
        L.oracle_ao_to_mo.restype = ctypes.c_int
        L.oracle_ao_to_mo.argtypes = [_i64, _i64, _i64, _dp, _dp, _dp, _dp]
        ob = _find_openblas()
        if ob and os.environ.get("CUTC_ORACLE_BUILTIN_DGEMM", "0") != "1":
            L.oracle_use_blas(ob.encode(), b"scipy_dgemm_64_")
        _L = L
    return _L


def blas_name():
    return "openblas" if lib().oracle_blas_loaded() else "builtin"


def threads():
    return lib().oracle_threads()


def _p(a, t=_dp):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def powd_int(a, b):
    return lib().oracle_powd_int(float(a), int(b))


def dgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    """Column-major BLAS dgemm on flat float64 arrays (C updated in place)."""
    lib().oracle_dgemm(ta.encode(), tb.encode(), m, n, k, alpha, _p(A), lda, _p(B), ldb, beta, _p(C), ldc)


def jast_bh(sys, ii0=0, n1=None, jj0=0, n2=None):
    """G with shape (4, n1, n2): grad_1 u (x, y, z) and -|grad_1 u|^2 / 2."""
    n1 = sys.n1 - ii0 if n1 is None else n1
    n2 = sys.n2 - jj0 if n2 is None else n2
    G = np.zeros((4, n1, n2))
    lib().oracle_jast_bh(ii0, n1, jj0, n2, sys.n_nuc, sys.size_bh, _p(_f64(sys.r1)), _p(_f64(sys.r2)),
                         _p(_f64(sys.rn)), _p(_f64(sys.c_bh)), _p(_i32(sys.m_bh), _ip), _p(_i32(sys.n_bh), _ip),
                         _p(_i32(sys.o_bh), _ip), _p(G))
    return G


def long_range(sys):
    """A with numpy shape (n_ao, n_ao, n2): A[i, j, r2] = w2 chi_i chi_j
    (Fortran A(r2, j + n_ao i))."""
    A = np.zeros((sys.n_ao, sys.n_ao, sys.n2))
    lib().oracle_long_range(0, sys.n2, sys.n2, sys.n2, sys.n_ao, _p(_f64(sys.wr2)), _p(_f64(sys.aos2)), _p(A))
    return A


def herm(sys):
    """H with numpy shape (n_ao, n_ao, n1): H[i, j, r1] = w1 chi_i chi_j."""
    H = np.zeros((sys.n_ao, sys.n_ao, sys.n1))
    lib().oracle_herm(sys.n1, sys.n_ao, _p(_f64(sys.wr1)), _p(_f64(sys.aos1)), _p(H))
    return H


def nonherm(sys):
    """NH with numpy shape (n_ao, n_ao, 3, n1):
    NH[i, j, m, r1] = w1 (chi_j d_m chi_i - chi_i d_m chi_j)."""
    NH = np.zeros((sys.n_ao, sys.n_ao, 3, sys.n1))
    lib().oracle_nonherm(sys.n1, sys.n_ao, _p(_f64(sys.wr1)), _p(_f64(sys.aos1)), _p(NH))
    return NH


def trans_inplace(M):
    """M + M^T of a square matrix (returns a new array)."""
    M = np.array(M, dtype=np.float64, order="C", copy=True)
    lib().oracle_trans_inplace(_p(M), M.shape[0])
    return M


def default_n2_chunk(sys, max_bytes=1 << 31):
    """r2 chunk keeping the long-range block (n2c x n_ao^2 doubles) under max_bytes."""
    return int(max(1, min(sys.n2, max_bytes // (8 * max(1, sys.n_ao) ** 2))))


def int2_grad1_u12_ao(sys, n1c=0, n2c=None):
    """X with numpy shape (4, n1, n_ao, n_ao): X[m, r1, l, k] is Fortran
    X(k, l, r1, m).  n2c None: chunk r2 so that the long-range block stays
    under 2 GiB."""
    if n2c is None:
        n2c = default_n2_chunk(sys)
    X = np.zeros((4, sys.n1, sys.n_ao, sys.n_ao))
    rc = lib().oracle_int2_grad1_u12_ao(
        sys.n1, sys.n2, sys.n_ao, sys.n_nuc, sys.size_bh, _p(_f64(sys.r1)), _p(_f64(sys.r2)), _p(_f64(sys.wr2)),
        _p(_f64(sys.rn)), _p(_f64(sys.aos2)), _p(_f64(sys.c_bh)), _p(_i32(sys.m_bh), _ip), _p(_i32(sys.n_bh), _ip),
        _p(_i32(sys.o_bh), _ip), _p(X), n1c, n2c)
    if rc != 0:
        raise MemoryError(f"oracle_int2_grad1_u12_ao returned {rc}")
    return X


def int_2e_ao(sys, X):
    """int_2e_ao with numpy shape (n_ao,)*4: out[b, a, l, k] is Fortran
    int_2e_ao(k, l, a, b)."""
    X = _f64(X)
    out = np.zeros((sys.n_ao,) * 4)
    rc = lib().oracle_int_2e_ao(sys.n1, sys.n_ao, _p(_f64(sys.wr1)), _p(_f64(sys.aos1)), _p(X), _p(out))
    if rc != 0:
        raise MemoryError(f"oracle_int_2e_ao returned {rc}")
    return out


def int_2e_ao_cols(sys, X, c0, nc):
    """Columns [c0, c0 + nc) of int_2e_ao (Fortran column index a + n_ao b)
    without the n_ao^4 array: numpy (nc, n_ao^2), equal to
    int_2e_ao(sys, X).reshape(n_ao^2, n_ao^2)[c0:c0 + nc]."""
    X = _f64(X)
    nao2 = sys.n_ao ** 2
    out = np.zeros((nc, nao2))
    rc = lib().oracle_int_2e_ao_cols(sys.n1, sys.n_ao, _p(_f64(sys.wr1)), _p(_f64(sys.aos1)), _p(X), c0, nc, _p(out))
    if rc != 0:
        raise ValueError(f"oracle_int_2e_ao_cols returned {rc}")
    return out


def dd_2e(sys, X, D):
    """sum_{kl,ab} D_kl D_ab (kl|ab) over int_2e_ao without forming it; D
    (n_ao, n_ao) symmetric (any vector of n_ao^2 entries in the Fortran
    order k + n_ao l)."""
    X = _f64(X)
    d = _f64(np.asarray(D).reshape(-1))
    assert d.size == sys.n_ao ** 2
    r = np.zeros(1)
    rc = lib().oracle_dd_2e(sys.n1, sys.n_ao, _p(_f64(sys.wr1)), _p(_f64(sys.aos1)), _p(X), _p(d), _p(r))
    if rc != 0:
        raise ValueError(f"oracle_dd_2e returned {rc}")
    return float(r[0])


def cutc_int(sys, n1c=0, n2c=None):
    """(X[0:3], int_2e_ao) as cutc_int returns them."""
    X = int2_grad1_u12_ao(sys, n1c, n2c)
    return X[:3].copy(), int_2e_ao(sys, X)


def cutc_no(ne_a, ne_b, wr1, mos_l, mos_r, G):
    """Normal ordering (oracle/oracle_no.c).  Inputs in numpy C order of the
    Fortran layouts: wr1 (n_grid1,), mos_l / mos_r (n_mo, n_grid1) = Fortran
    (n_grid1, n_mo), G (n_mo, n_mo, 3, n_grid1) = Fortran
    int2_grad1_u12(n_grid1, 3, n_mo, n_mo), G[j, i, x, g].  Returns (no_2e,
    no_1e, no_0e): no_2e (n_mo,)*4 = Fortran (p, q, s, t) as [t, s, q, p],
    no_1e (n_mo, n_mo) = Fortran (p, q) as [q, p], no_0e a float."""
    G = _f64(G)
    nm, ng = G.shape[0], G.shape[3]
    assert G.shape == (nm, nm, 3, ng)
    no_2e = np.zeros((nm,) * 4)
    no_1e = np.zeros((nm, nm))
    no_0e = np.zeros(1)
    rc = lib().oracle_cutc_no(ng, nm, int(ne_a), int(ne_b), _p(_f64(wr1)), _p(_f64(mos_l)), _p(_f64(mos_r)), _p(G),
                              _p(no_2e), _p(no_1e), _p(no_0e))
    if rc != 0:
        raise ValueError(f"oracle_cutc_no returned {rc}")
    return no_2e, no_1e, float(no_0e[0])


def ao_to_mo(mo_l_coef, mo_r_coef, X):
    """AO -> MO transform (oracle/oracle_mo.c).  numpy C order of the Fortran
    layouts: mo_l_coef / mo_r_coef (n_mo, n_ao), X (3, n_grid1, n_ao, n_ao) =
    int2_grad1_u12_ao(n_ao, n_ao, n_grid1, 3).  Returns int2_grad1_u12 as
    (n_mo, n_mo, 3, n_grid1): [j, i, x, g] = sum_kl C_L[i, k] X[x, g, l, k] C_R[j, l]."""
    X = _f64(X)
    cl, cr = _f64(mo_l_coef), _f64(mo_r_coef)
    n1, na, nm = X.shape[1], X.shape[2], cl.shape[0]
    assert X.shape == (3, n1, na, na) and cl.shape == (nm, na) and cr.shape == (nm, na)
    out = np.zeros((nm, nm, 3, n1))
    rc = lib().oracle_ao_to_mo(n1, na, nm, _p(cl), _p(cr), _p(X), _p(out))
    if rc != 0:
        raise ValueError(f"oracle_ao_to_mo returned {rc}")
    return out


def aos(r, rn, ao_nucl, ao_power, ao_prim_num, ao_expo, ao_coef):
    """AO values and gradients (oracle/oracle_ao.c).  numpy C order of the
    Fortran layouts: r (n, 3), rn (n_nuc, 3), ao_nucl (n_ao,) 1-based,
    ao_power (3, n_ao) = Fortran (n_ao, 3), ao_prim_num (n_ao,), ao_expo /
    ao_coef (n_ao, prim_max) = Fortran (prim_max, n_ao).  Returns
    (4, n_ao, n) = Fortran aos_data(n, n_ao, 4)."""
    r, rn = _f64(r), _f64(rn)
    e, c = _f64(ao_expo), _f64(ao_coef)
    nu, pw, npr = _i32(ao_nucl), _i32(ao_power), _i32(ao_prim_num)
    n, n_ao = r.shape[0], nu.shape[0]
    pmax = e.shape[1] if e.ndim == 2 else 0
    out = np.zeros((4, n_ao, n))
    rc = lib().oracle_aos(n, rn.shape[0], n_ao, pmax, _p(r), _p(rn), _p(nu, _ip), _p(pw, _ip), _p(npr, _ip), _p(e),
                          _p(c), _p(out))
    if rc != 0:
        raise ValueError(f"oracle_aos returned {rc}")
    return out

