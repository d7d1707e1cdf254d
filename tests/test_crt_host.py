"""CPU checks of the modular (CRT) split-precision arithmetic.

cutc_crt_setup (host code in libcutcint.so) is called through ctypes; its
constants are checked with Python integers, and the device rebuild formula
(crt_value in cutc_amd/csrc/kernels/crt.hip) and the residue mapping are
restated in numpy float64 with the same operation order, so a wrong constant,
part split or rounding step fails here before any kernel runs."""
import ctypes
import math

import numpy as np
import pytest

from cutc_amd import cutc

NMAX = 16


KIND_I8, KIND_BF16, KIND_FP16 = 0, 1, 2


class Crt(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("bits", ctypes.c_int), ("m", ctypes.c_int * NMAX), ("y", ctypes.c_int * NMAX),
                ("inv_m", ctypes.c_double * NMAX), ("Mi", (ctypes.c_double * 3) * NMAX), ("M", ctypes.c_double * 3),
                ("Minv", ctypes.c_double), ("kind", ctypes.c_int)]


def setup(K, n, kind=KIND_I8):
    L = cutc.lib()
    L.cutc_crt_setup.restype = ctypes.c_int
    L.cutc_crt_setup.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Crt)]
    c = Crt()
    rc = L.cutc_crt_setup(K, n, kind, ctypes.byref(c))
    return rc, c


def parts(v):
    return int(v[0]) * 2 ** 80 + int(v[1]) * 2 ** 40 + int(v[2])


@pytest.mark.parametrize("kind", [KIND_I8, KIND_BF16, KIND_FP16])
@pytest.mark.parametrize("K,n", [(200000, 12), (60000, 12), (180000, 12), (1, 6), (131072, 16), (1000, 8)])
def test_constants(K, n, kind):
    rc, c = setup(K, n, kind)
    assert rc == 0 and c.kind == kind
    m = [c.m[i] for i in range(n)]
    # int8 residues lie in [-128, 127]; bf16 / fp16 residues within +-127, so
    # that 1024 products of two stay exact in an FP32 sum
    assert all(2 <= x <= (256 if kind == KIND_I8 else 255) for x in m)
    for i in range(n):
        for j in range(i):
            assert math.gcd(m[i], m[j]) == 1
    M = math.prod(m)
    assert parts(c.M) == M
    for i in range(n):
        Mi = M // m[i]
        assert parts(c.Mi[i]) == Mi
        assert (c.y[i] * Mi) % m[i] == 1
        assert all(0 <= c.Mi[i][t] < 2 ** 40 for t in range(3))
    b = c.bits
    # Cauchy-Schwarz: |V| < (2^b + sqrt(K)/2)^2 must stay below M/2 (checked
    # with exact integers: 4 (2^b + sqrt(K)/2)^2 = (2^(b+1) + sqrt(K))^2 <=
    # (2^(b+1) + isqrt(K) + 1)^2), and b is the largest such value up to 50
    def fits(bb):
        return (2 ** (bb + 1) + math.isqrt(K) + 1) ** 2 < 2 * M

    assert b <= 50 and fits(b)
    assert b == 50 or not (2 ** (b + 2) + math.isqrt(K)) ** 2 < 2 * M


def test_bad_arguments():
    assert setup(1000, 1)[0] != 0
    assert setup(1000, 17)[0] != 0
    assert setup(0, 12)[0] != 0
    assert setup(1000, 12, 3)[0] != 0
    assert setup(1000, 12, -1)[0] != 0


def residue(q, m, inv_m):
    """restatement of residue() in crt.hip (float64, same steps): q mod m in
    [-(m // 2), (m - 1) // 2]"""
    t = np.rint(q * inv_m)
    r = np.fma(-t, float(m), q) if hasattr(np, "fma") else q - t * m
    r = r.astype(np.int64)
    r = np.where(r > (m - 1) // 2, r - m, r)
    r = np.where(r < -(m // 2), r + m, r)
    return r


@pytest.mark.parametrize("m", [256, 255, 193, 191])
def test_residue_range_and_congruence(m):
    rng = np.random.default_rng(m)
    q = np.concatenate([rng.integers(-2 ** 51, 2 ** 51, 20000), np.arange(-3 * m, 3 * m),
                        [2 ** 51 - 1, -2 ** 51 + 1]]).astype(np.float64)
    r = residue(q, m, 1.0 / m)
    assert r.min() >= -(m // 2) and r.max() <= (m - 1) // 2
    assert np.array_equal((q.astype(np.int64) - r) % m, np.zeros_like(r))


def crt_value(c, res):
    """restatement of crt_value() in crt.hip; res: (n, ...) int residues"""
    sh = np.zeros(res.shape[1:])
    sm = np.zeros(res.shape[1:])
    sl = np.zeros(res.shape[1:])
    for i in range(c.n):
        u = (res[i] * c.y[i]) % c.m[i]  # numpy % is already >= 0 for m > 0
        du = u.astype(np.float64)
        sh = du * c.Mi[i][0] + sh
        sm = du * c.Mi[i][1] + sm
        sl = du * c.Mi[i][2] + sl
    two40, two80 = 2.0 ** 40, 2.0 ** 80
    s = (sh * two80 + sm * two40) + sl
    q = np.rint(s * c.Minv)
    vh, vm, vl = sh - q * c.M[0], sm - q * c.M[1], sl - q * c.M[2]
    return (vh * two80 + vm * two40) + vl


@pytest.mark.parametrize("kind", [KIND_I8, KIND_BF16])
@pytest.mark.parametrize("n", [8, 12, 16])
def test_rebuild_recovers_integers(n, kind):
    K = 200000
    rc, c = setup(K, n, kind)
    rng = np.random.default_rng(n)
    bound = (2 ** c.bits + math.isqrt(K) // 2 + 1) ** 2  # largest |V| the scaling allows
    vals = [int(x) for x in rng.integers(-2 ** 62, 2 ** 62, 2000)]
    vals = [v * (bound // 2 ** 62) for v in vals] + [0, 1, -1, bound, -bound, 12345, -2 ** 40]
    res = np.array([[((v % c.m[i]) + 128) % c.m[i] - 128 if c.m[i] == 256 else
                     (((v % c.m[i]) + c.m[i] // 2) % c.m[i]) - c.m[i] // 2 for v in vals] for i in range(n)])
    got = crt_value(c, res)
    want = np.array([float(v) for v in vals])
    err = np.abs(got - want) / np.maximum(np.abs(want), 1.0)
    assert err.max() <= 2.0 ** -50


def test_quantised_dot_products_match_exact_products():
    """Whole scheme on small vectors: rounding per row (2-norm exponents),
    residues, modular products, rebuild, scaling (the device steps, restated
    in numpy)."""
    K = 3000
    rc, c = setup(K, 12)
    rng = np.random.default_rng(1)
    A = rng.standard_normal((7, K)) * np.exp(rng.standard_normal((7, K)) * 3)
    B = rng.standard_normal((5, K))
    B[2] *= 1e-100  # a tiny row
    A[3, :] = 0.0  # a zero row
    A[4, :] = 0.0
    A[4, 17] = 3.0  # a row with a single entry: 2-norm = max

    def quant(Mat):
        ss = (Mat * Mat).sum(axis=1)
        E = np.where(ss > 0, np.frexp(np.sqrt(ss * (1 + 2.0 ** -29)))[1], 0)  # ||x|| < 2^E
        return np.rint(np.ldexp(Mat, (c.bits - E)[:, None])), E

    qa, ea = quant(A)
    qb, eb = quant(B)
    bound = 2.0 ** c.bits + np.sqrt(K) / 2
    assert (np.sqrt((qa * qa).sum(1)) < bound).all() and (np.sqrt((qb * qb).sum(1)) < bound).all()
    assert np.abs(qa).max() < 2.0 ** 52
    ra = np.stack([residue(qa, c.m[i], c.inv_m[i]) for i in range(c.n)])
    rb = np.stack([residue(qb, c.m[i], c.inv_m[i]) for i in range(c.n)])
    assert ra.min() >= -128 and ra.max() <= 127
    prod = np.stack([(ra[i] @ rb[i].T) % c.m[i] for i in range(c.n)])
    prod = np.where(prod > 127, prod - np.array([c.m[i] for i in range(c.n)])[:, None, None], prod)
    V = crt_value(c, prod)
    exact = np.array([[sum(int(x) * int(y) for x, y in zip(qa[i], qb[j])) for j in range(5)] for i in range(7)])
    np.testing.assert_allclose(V, exact.astype(np.float64), rtol=2.0 ** -50, atol=1.0)
    C = np.ldexp(V, (ea[:, None] + eb[None, :]) - 2 * c.bits)
    ref = A @ B.T
    # each rounding moves a row by at most sqrt(K)/2 quanta of 2^(e - bits)
    na = np.sqrt((A * A).sum(1))[:, None]
    nb = np.sqrt((B * B).sum(1))[None, :]
    tol = np.sqrt(K) * 2.0 ** -c.bits * (na * 2 * nb + 2 * na * nb) + 1e-300
    assert (np.abs(C - ref) <= tol).all()
    assert np.array_equal(C[3], np.zeros(5))
