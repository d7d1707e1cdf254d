"""Pins the FP64 CPU oracle (oracle/oracle.c) against independent checks.

The reference ships no golden vectors (its only known answer is the all-zero
demo of src/cutc_int_f.f90:41-55), so each oracle function is checked against
mathematics that fixes it: finite differences of an independently coded u,
closed forms, brute-force sums, BLAS semantics, invariances.
"""
import numpy as np
import pytest

from cutc_amd import synth
from oracle import oracle


def rbar(d):
    return d / (1.0 + d)


def u_value(sys, p1, p2):
    """u(r1, r2) = sum_A sum_p c (f1^m f2^n + f1^n f2^m) g^o with f = d/(1+d)
    (the Boys-Handy form, src/jast_bh.cu:38-43 documents its gradient)."""
    g = rbar(np.linalg.norm(p1 - p2))
    u = 0.0
    for a in range(sys.n_nuc):
        f1 = rbar(np.linalg.norm(p1 - sys.rn[a]))
        f2 = rbar(np.linalg.norm(p2 - sys.rn[a]))
        for p in range(sys.size_bh):
            c = sys.c_bh[a, p]
            if abs(c) < 1e-10:
                continue
            m, n, o = sys.m_bh[a, p], sys.n_bh[a, p], sys.o_bh[a, p]
            u += c * (f1 ** m * f2 ** n + f1 ** n * f2 ** m) * g ** o
    return u


def small_system(n1=12, n2=17, seed=5, terms=None):
    s = synth.triatomic(n1=n1, n2=n2, seed=seed)
    if terms is not None:
        rng = np.random.default_rng(seed)
        size = len(terms)
        s.c_bh = rng.uniform(-0.5, 0.5, (s.n_nuc, size))
        s.m_bh = np.array([[t[0] for t in terms]] * s.n_nuc, np.int32)
        s.n_bh = np.array([[t[1] for t in terms]] * s.n_nuc, np.int32)
        s.o_bh = np.array([[t[2] for t in terms]] * s.n_nuc, np.int32)
    return s


ALL_TERMS = [(0, 0, 1), (0, 0, 2), (2, 0, 0), (3, 0, 0), (2, 2, 0), (0, 0, 3), (1, 2, 1), (3, 1, 2), (6, 0, 7),
             (0, 6, 0), (4, 5, 3)]


def test_powd_int():
    for b in range(0, 10):
        for a in (0.0, 0.3, 0.77, 1.0, 0.999):
            assert oracle.powd_int(a, b) == pytest.approx(a ** b, rel=1e-15, abs=0)


@pytest.mark.parametrize("terms", [None, ALL_TERMS])
def test_jastrow_finite_difference(terms):
    s = small_system(terms=terms)
    G = oracle.jast_bh(s)
    h = 1e-5
    worst = 0.0
    for i1 in range(s.n1):
        for i2 in range(s.n2):
            p1, p2 = s.r1[i1].copy(), s.r2[i2]
            fd = np.zeros(3)
            for d in range(3):
                e = np.zeros(3)
                e[d] = h
                fd[d] = (u_value(s, p1 + e, p2) - u_value(s, p1 - e, p2)) / (2 * h)
            g = G[:3, i1, i2]
            worst = max(worst, np.max(np.abs(g - fd)) / max(1.0, np.max(np.abs(fd))))
            assert G[3, i1, i2] == pytest.approx(-0.5 * np.dot(g, g), rel=1e-14, abs=1e-300)
    assert worst < 1e-7


def test_jastrow_pure_ee_closed_form():
    """Only (0,0,o) terms with one c on every nucleus:
    grad_1 u = 2 n_nuc c o g^(o-1) (r1 - r2) / (r12 (1 + r12)^2)."""
    for o in (1, 2, 3, 5):
        s = small_system(n1=9, n2=11)
        s.c_bh = np.full((s.n_nuc, 1), 0.37)
        s.m_bh = np.zeros((s.n_nuc, 1), np.int32)
        s.n_bh = np.zeros((s.n_nuc, 1), np.int32)
        s.o_bh = np.full((s.n_nuc, 1), o, np.int32)
        G = oracle.jast_bh(s)
        d = s.r1[:, None, :] - s.r2[None, :, :]
        r12 = np.linalg.norm(d, axis=-1)
        g = r12 / (1 + r12)
        want = (2 * s.n_nuc * 0.37 * o * g ** (o - 1) / (r12 * (1 + r12) ** 2))[..., None] * d
        np.testing.assert_allclose(np.moveaxis(G[:3], 0, -1), want, rtol=1e-13, atol=0)


def test_jastrow_pure_en_closed_form():
    """Only (m,0,0) terms: grad_1 u = sum_A c_A m f1A^(m-1) (r1-R_A)/(d (1+d)^2)."""
    for m in (1, 2, 3, 6):
        s = small_system(n1=9, n2=11)
        rng = np.random.default_rng(m)
        s.c_bh = rng.uniform(-0.5, 0.5, (s.n_nuc, 1))
        s.m_bh = np.full((s.n_nuc, 1), m, np.int32)
        s.n_bh = np.zeros((s.n_nuc, 1), np.int32)
        s.o_bh = np.zeros((s.n_nuc, 1), np.int32)
        G = oracle.jast_bh(s)
        want = np.zeros((s.n1, 3))
        for a in range(s.n_nuc):
            d = s.r1 - s.rn[a]
            dn = np.linalg.norm(d, axis=-1)
            f1 = dn / (1 + dn)
            want += (s.c_bh[a, 0] * m * f1 ** (m - 1) / (dn * (1 + dn) ** 2))[:, None] * d
        np.testing.assert_allclose(np.moveaxis(G[:3], 0, -1), np.broadcast_to(want[:, None, :], (s.n1, s.n2, 3)),
                                   rtol=1e-13, atol=1e-300)


def test_jastrow_special_cases():
    s = small_system(n1=5, n2=6, terms=ALL_TERMS)
    # coincident r1 == r2 and r1 on a nucleus
    s.r2[0] = s.r1[0]
    s.r1[1] = s.rn[1]
    G = oracle.jast_bh(s)
    assert np.all(np.isfinite(G))
    # at r1 == r2 only the e-e part vanishes; compare against the same pair with
    # every o > 0 term removed
    s2 = small_system(n1=5, n2=6, terms=ALL_TERMS)
    s2.r2[0] = s2.r1[0]
    s2.r1[1] = s2.rn[1]
    s2.c_bh = np.where(s2.o_bh > 0, 0.0, s2.c_bh)
    G2 = oracle.jast_bh(s2)
    # terms with o > 0 contribute exactly 0 at g = 0 when o >= 2; for o == 1 the
    # e-n part carries g^1 = 0 and the e-e part carries grad g = 0
    np.testing.assert_allclose(G[:3, 0, 0], G2[:3, 0, 0], rtol=1e-14, atol=1e-300)
    # |c| < 1e-10 is skipped exactly
    s3 = small_system(n1=5, n2=6, terms=ALL_TERMS)
    s4 = small_system(n1=5, n2=6, terms=ALL_TERMS)
    s3.c_bh[:, 2] = 9e-11
    s4.c_bh[:, 2] = 0.0
    assert np.array_equal(oracle.jast_bh(s3), oracle.jast_bh(s4))


def test_jastrow_translation_invariance():
    s = small_system(terms=ALL_TERMS)
    G = oracle.jast_bh(s)
    shift = np.array([0.3, -1.7, 2.2])
    s.r1 = s.r1 + shift
    s.r2 = s.r2 + shift
    s.rn = s.rn + shift
    G2 = oracle.jast_bh(s)
    np.testing.assert_allclose(G2, G, rtol=1e-9, atol=1e-12 * np.abs(G).max())


def test_jastrow_zero_c_gives_zero():
    """The reference demo's known answer: c = 0, m = n = o = 1 -> G = 0."""
    s = small_system(n1=4, n2=5)
    s.c_bh[:] = 0.0
    s.m_bh[:] = 1
    s.n_bh[:] = 1
    s.o_bh[:] = 1
    assert not np.any(oracle.jast_bh(s))


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-0.5, 1.0), (2.0, 0.3)])
def test_dgemm_semantics(ta, tb, alpha, beta):
    rng = np.random.default_rng(3)
    m, n, k = 37, 29, 53
    A = rng.standard_normal((m, k) if ta == "N" else (k, m))
    B = rng.standard_normal((k, n) if tb == "N" else (n, k))
    C = rng.standard_normal((m, n))
    # column-major views: a Fortran (r, c) array is numpy (c, r) C-order
    Af = np.ascontiguousarray(A.T)
    Bf = np.ascontiguousarray(B.T)
    Cf = np.ascontiguousarray(C.T)
    oracle.dgemm(ta, tb, m, n, k, alpha, Af.ravel(), A.shape[0], Bf.ravel(), B.shape[0], beta, Cf.ravel(), m)
    opA = A if ta == "N" else A.T
    opB = B if tb == "N" else B.T
    np.testing.assert_allclose(Cf.T, alpha * opA @ opB + beta * C, rtol=1e-12, atol=1e-12)


def test_integrands_closed_forms():
    s = small_system(n1=10, n2=13)
    A = oracle.long_range(s)
    chi2 = s.aos2[0]
    np.testing.assert_allclose(A, s.wr2[None, None, :] * chi2[:, None, :] * chi2[None, :, :], rtol=1e-15, atol=0)
    np.testing.assert_allclose(A, A.transpose(1, 0, 2), rtol=1e-15, atol=0)
    H = oracle.herm(s)
    chi1 = s.aos1[0]
    np.testing.assert_allclose(H, s.wr1[None, None, :] * chi1[:, None, :] * chi1[None, :, :], rtol=1e-15, atol=0)
    NH = oracle.nonherm(s)
    d = s.aos1[1:]  # (3, n_ao, n1)
    # NH[i, j, m, r1] = w (chi_j d_m chi_i - chi_i d_m chi_j)
    want = s.wr1 * (chi1[None, :, None, :] * np.moveaxis(d, 0, 1)[:, None, :, :]
                    - chi1[:, None, None, :] * np.moveaxis(d, 0, 1)[None, :, :, :])
    np.testing.assert_allclose(NH, want, rtol=1e-14, atol=1e-300)
    np.testing.assert_array_equal(NH, -NH.transpose(1, 0, 2, 3))
    for i in range(s.n_ao):
        assert not np.any(NH[i, i])


def brute_X(s):
    G = oracle.jast_bh(s)
    chi = s.aos2[0]
    return np.einsum("q,iq,jq,mpq->mpij", s.wr2, chi, chi, G)


def brute_2e(s, X):
    chi = s.aos1[0]
    d = s.aos1[1:]
    w = s.wr1
    herm = np.einsum("p,pkl,ap,bp->abkl", w, X[3], chi, chi)
    # NH(a', b') with a' = i', b' = j' in M_np[i', j', k, l]:  chi_j' d chi_i' - chi_i' d chi_j'
    nh = np.einsum("p,mpkl,bp,map->abkl", w, X[:3], chi, d) - np.einsum("p,mpkl,ap,mbp->abkl", w, X[:3], chi, d)
    M = herm - 0.5 * nh
    return M + M.transpose(2, 3, 0, 1)


def test_stage1_bruteforce_and_chunking():
    s = small_system(n1=11, n2=23, terms=ALL_TERMS)
    X = oracle.int2_grad1_u12_ao(s)
    Xb = brute_X(s)
    np.testing.assert_allclose(X, Xb, rtol=1e-12, atol=1e-13 * np.abs(Xb).max())
    Xc = oracle.int2_grad1_u12_ao(s, n1c=4, n2c=7)
    np.testing.assert_allclose(Xc, X, rtol=1e-12, atol=1e-14 * np.abs(X).max())
    np.testing.assert_allclose(X, X.transpose(0, 1, 3, 2), rtol=1e-14, atol=1e-15 * np.abs(X).max())


def test_stage2_bruteforce_and_symmetry():
    s = small_system(n1=13, n2=19, terms=ALL_TERMS)
    X = oracle.int2_grad1_u12_ao(s)
    I = oracle.int_2e_ao(s, X)
    Ib = brute_2e(s, X)
    np.testing.assert_allclose(I, Ib, rtol=1e-11, atol=1e-13 * np.abs(Ib).max())
    Im = I.reshape(s.n_ao ** 2, s.n_ao ** 2)
    np.testing.assert_array_equal(Im, Im.T)


@pytest.mark.parametrize("c0,nc", [(0, 1), (0, 576), (37, 100), (575, 1), (200, 376)])
def test_int_2e_ao_column_block_matches_full(c0, nc):
    """oracle_int_2e_ao_cols (the large-shape checker) against the columns
    of the full oracle_int_2e_ao: a wrong offset, a missing M^T half or a
    dropped non-herm term changes whole columns."""
    s = small_system(n1=13, n2=19, terms=ALL_TERMS)
    X = oracle.int2_grad1_u12_ao(s)
    full = oracle.int_2e_ao(s, X).reshape(s.n_ao ** 2, s.n_ao ** 2)
    blk = oracle.int_2e_ao_cols(s, X, c0, nc)
    np.testing.assert_allclose(blk, full[c0:c0 + nc], rtol=1e-13, atol=1e-14 * np.abs(full).max())


def test_dd_contraction_matches_einsum():
    """oracle_dd_2e (sum D D int_2e_ao through the factors of M) against an
    einsum over the full int_2e_ao, for a symmetric PSD D and for a
    non-symmetric d (which the factor form handles as 2 d^T M d only when
    int_2e_ao = M + M^T is symmetric, as it is)."""
    s = small_system(n1=13, n2=19, terms=ALL_TERMS)
    X = oracle.int2_grad1_u12_ao(s)
    I = oracle.int_2e_ao(s, X)
    rng = np.random.default_rng(3)
    C = rng.standard_normal((s.n_ao, 10))
    for D in (C @ C.T, rng.standard_normal((s.n_ao, s.n_ao))):
        d = D.reshape(-1)  # numpy D[l, k] = Fortran D(k, l): index k + n_ao l
        want = float(d @ (I.reshape(s.n_ao ** 2, s.n_ao ** 2) @ d))
        got = oracle.dd_2e(s, X, D)
        assert abs(got - want) <= 1e-12 * abs(want), (got, want)


def test_linearity_in_weights():
    s = small_system(n1=8, n2=15, terms=ALL_TERMS)
    X, I = oracle.cutc_int(s)
    s.wr2 = 2.0 * s.wr2
    X2, I2 = oracle.cutc_int(s)
    np.testing.assert_allclose(X2, 2 * X, rtol=1e-14, atol=1e-300)
    np.testing.assert_allclose(I2, 2 * I, rtol=1e-13, atol=1e-13 * np.abs(I).max())
    s.wr1 = 3.0 * s.wr1
    X3, I3 = oracle.cutc_int(s)
    np.testing.assert_allclose(X3, X2, rtol=0, atol=0)
    np.testing.assert_allclose(I3, 3 * I2, rtol=1e-13, atol=1e-13 * np.abs(I2).max())


def test_trans_inplace():
    rng = np.random.default_rng(0)
    M = rng.standard_normal((31, 31))
    T = oracle.trans_inplace(M)
    np.testing.assert_array_equal(T, T.T)
    np.testing.assert_allclose(T, M + M.T, rtol=0, atol=1e-15)
    np.testing.assert_array_equal(np.diag(T), 2 * np.diag(M))


def test_zero_demo_known_answer():
    """src/cutc_int_f.f90:41-55: everything zero, m = n = o = 1 -> outputs 0."""
    s = small_system(n1=6, n2=7)
    for f in ("r1", "wr1", "r2", "wr2", "rn", "aos1", "aos2", "c_bh"):
        setattr(s, f, np.zeros_like(getattr(s, f)))
    s.m_bh[:] = 1
    s.n_bh[:] = 1
    s.o_bh[:] = 1
    X, I = oracle.cutc_int(s)
    assert not np.any(X) and not np.any(I)


def test_generator_deterministic():
    a = synth.triatomic(n1=50, n2=60)
    b = synth.triatomic(n1=50, n2=60)
    for f in ("r1", "wr1", "r2", "wr2", "aos1", "aos2", "c_bh"):
        assert np.array_equal(getattr(a, f), getattr(b, f))
    c = synth.triatomic(n1=50, n2=60, seed=7)
    assert not np.array_equal(a.r1, c.r1)
    # the grid-1 draws do not depend on the grid-2 size
    d = synth.triatomic(n1=50, n2=10)
    assert np.array_equal(a.r1, d.r1) and np.array_equal(a.r2[:10], d.r2)


def test_generator_rules():
    s = synth.cube20(n1=3000, n2=10)
    assert s.n_ao == 200 and s.n_nuc == 20 and s.size_bh == 6
    dmin = min(np.linalg.norm(s.rn[i] - s.rn[j]) for i in range(20) for j in range(i))
    assert dmin >= 1.8 and s.rn.min() >= 0 and s.rn.max() <= 12
    lw = np.log(s.wr1)
    assert abs(lw.mean() + 2) < 0.1 and abs(lw.std() - 1) < 0.1
    assert np.all(s.c_bh[:, 0] == 0.5) and np.all(np.abs(s.c_bh) <= 0.5)


def _basis(seed, shells, n_ao):
    lib = synth._synth()
    sn = np.array([x[0] for x in shells], np.int32)
    sl = np.array([x[1] for x in shells], np.int32)
    nuc = np.zeros(n_ao, np.int32)
    pw = np.zeros(3 * n_ao, np.int32)
    al = np.zeros(n_ao)
    nm = np.zeros(n_ao)
    assert lib.synth_basis(seed, len(shells), synth._ptr(sn, synth._ip), synth._ptr(sl, synth._ip), n_ao,
                           synth._ptr(nuc, synth._ip), synth._ptr(pw, synth._ip), synth._ptr(al),
                           synth._ptr(nm)) == n_ao
    return nuc, pw, al, nm


def _aos(P, basis):
    nuc, pw, al, nm = basis
    n_ao = len(al)
    P = np.ascontiguousarray(P)
    aos = np.zeros((4, n_ao, P.shape[0]))
    synth._synth().synth_aos(P.shape[0], synth._ptr(P), n_ao, synth._ptr(np.zeros(3)), synth._ptr(nuc, synth._ip),
                             synth._ptr(pw, synth._ip), synth._ptr(al), synth._ptr(nm), synth._ptr(aos))
    return aos


def test_generator_normalised_gaussians_and_gradients():
    shells = [(0, 0), (0, 1), (0, 2)]
    seed = next(sd for sd in range(1000) if _basis(sd, shells, 10)[2].max() < 2.5)
    basis = _basis(seed, shells, 10)
    x = np.arange(-6.0, 6.0 + 1e-9, 0.15)
    P = np.stack(np.meshgrid(x, x, x, indexing="ij"), -1).reshape(-1, 3)
    aos = _aos(P, basis)
    norms = (aos[0] ** 2).sum(axis=1) * 0.15 ** 3
    np.testing.assert_allclose(norms, 1.0, rtol=1e-6)
    # gradients against central differences of the values
    rng = np.random.default_rng(1)
    Q = rng.normal(size=(50, 3))
    h = 1e-6
    g = _aos(Q, basis)[1:]
    for d in range(3):
        e = np.zeros(3)
        e[d] = h
        fd = (_aos(Q + e, basis)[0] - _aos(Q - e, basis)[0]) / (2 * h)
        np.testing.assert_allclose(g[d], fd, rtol=1e-6, atol=1e-8)
