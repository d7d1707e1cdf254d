"""CPU check of the Jastrow monomial program (cutc_amd/csrc/jastrow_prog.c).

The host library function cutc_jast_program expands the Boys-Handy table into
monomials; here the pair kernel's evaluation of that program is restated in
numpy and compared with the oracle's direct term loop
(src/jast_bh.cu:198-254), so a dropped or misplaced monomial fails on the CPU
before any kernel runs."""
import ctypes

import numpy as np
import pytest

from cutc_amd import cutc, synth
from oracle import oracle


class JMono(ctypes.Structure):
    _fields_ = [("A", ctypes.c_int), ("slot", ctypes.c_int), ("i", ctypes.c_int), ("is_r", ctypes.c_int),
                ("coef", ctypes.c_double)]


class JProg(ctypes.Structure):
    _fields_ = [("mono", ctypes.POINTER(JMono)), ("cls", ctypes.POINTER(ctypes.c_int)), ("n_mono", ctypes.c_int),
                ("n_cls", ctypes.c_int), ("kmax", ctypes.c_int), ("ts", ctypes.c_int64)]


def program(s):
    L = cutc.lib()
    L.cutc_jast_program.restype = ctypes.c_int
    L.cutc_jast_program.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.POINTER(JProg)]
    L.cutc_jast_program_free.argtypes = [ctypes.POINTER(JProg)]
    c = np.ascontiguousarray(s.c_bh, np.float64)
    m, n, o = (np.ascontiguousarray(a, np.int32) for a in (s.m_bh, s.n_bh, s.o_bh))
    p = JProg()
    assert L.cutc_jast_program(s.n_nuc, s.size_bh, c.ctypes.data, m.ctypes.data, n.ctypes.data, o.ctypes.data,
                               ctypes.byref(p)) == 0
    mono = [(p.mono[t].A, p.mono[t].slot, p.mono[t].i, p.mono[t].is_r, p.mono[t].coef) for t in range(p.n_mono)]
    cls = [(p.cls[2 * t], p.cls[2 * t + 1]) for t in range(p.n_cls)]
    out = (mono, cls, p.kmax, p.ts)
    L.cutc_jast_program_free(ctypes.byref(p))
    return out


def rbar_grad(d):
    r2 = (d * d).sum(-1)
    ok = r2 > 1e-15
    r = np.sqrt(np.where(ok, r2, 1.0))
    f = np.where(ok, r / (1 + r), 0.0)
    gr = np.where(ok, 1.0 / (r * (1 + r) ** 2), 0.0)[..., None] * d
    return f, gr


def evaluate(s, prog):
    mono, cls, kmax, ts = prog
    n1, n2 = s.n1, s.n2
    tab = np.zeros((n1, ts))
    for (A, slot, i, is_r, coef) in mono:
        f, gr = rbar_grad(s.r1 - s.rn[A])
        v = coef * f ** i
        if is_r:
            tab[:, slot + 3] += v
        else:
            tab[:, slot:slot + 3] += v[:, None] * gr
    g, gg = rbar_grad(s.r1[:, None, :] - s.r2[None, :, :])
    f2 = np.stack([rbar_grad(s.r2 - s.rn[A])[0] for A in range(s.n_nuc)])  # (n_nuc, n2)
    u = np.zeros((n1, n2, 3))
    sc = np.zeros((n1, n2))
    for k in range(kmax + 1):
        gk = g ** k
        u += gk[..., None] * tab[:, None, 4 * k:4 * k + 3]
        sc += gk * tab[:, None, 4 * k + 3]
    base = 4 * (kmax + 1)
    for c, (j, k) in enumerate(cls):
        gk = g ** k
        for A in range(s.n_nuc):
            q = tab[:, base + 4 * (c * s.n_nuc + A): base + 4 * (c * s.n_nuc + A) + 4]
            w = gk * (f2[A] ** j)[None, :]
            u += w[..., None] * q[:, None, :3]
            sc += w * q[:, None, 3]
    u += sc[..., None] * gg
    return np.concatenate([np.moveaxis(u, -1, 0), (-0.5 * (u * u).sum(-1))[None]], 0)


TERMS = [(0, 0, 1), (0, 0, 2), (2, 0, 0), (3, 0, 0), (2, 2, 0), (0, 0, 3), (1, 2, 1), (3, 1, 2), (6, 0, 7),
         (0, 6, 0), (4, 5, 3), (1, 1, 0), (0, 1, 1)]


@pytest.mark.parametrize("terms", [None, TERMS])
def test_program_reproduces_term_loop(terms):
    s = synth.triatomic(n1=23, n2=31, seed=9)
    if terms is not None:
        rng = np.random.default_rng(2)
        s.c_bh = rng.uniform(-0.5, 0.5, (s.n_nuc, len(terms)))
        s.c_bh[1, 3] = 5e-11  # skipped exactly like the reference
        s.m_bh = np.array([[t[0] for t in terms]] * s.n_nuc, np.int32)
        s.n_bh = np.array([[t[1] for t in terms]] * s.n_nuc, np.int32)
        s.o_bh = np.array([[t[2] for t in terms]] * s.n_nuc, np.int32)
    s.r2[0] = s.r1[0]  # coincident pair
    s.r1[2] = s.rn[1]  # r1 on a nucleus
    prog = program(s)
    G = evaluate(s, prog)
    Go = oracle.jast_bh(s)
    np.testing.assert_allclose(G, Go, rtol=1e-12, atol=1e-14 * np.abs(Go).max())


def test_program_structure_bh6():
    """The six-term table gives kmax = 2 and a single (j, k) = (2, 0) class."""
    s = synth.cube20(n1=4, n2=4)
    mono, cls, kmax, ts = program(s)
    assert cls == [(2, 0)] and kmax == 2
    assert ts == 4 * 3 + 4 * 20


def test_program_nucleus_dependent_exponents():
    s = synth.triatomic(n1=11, n2=13, seed=4)
    rng = np.random.default_rng(5)
    s.m_bh = rng.integers(0, 5, s.m_bh.shape).astype(np.int32)
    s.n_bh = rng.integers(0, 5, s.n_bh.shape).astype(np.int32)
    s.o_bh = rng.integers(0, 4, s.o_bh.shape).astype(np.int32)
    G = evaluate(s, program(s))
    Go = oracle.jast_bh(s)
    np.testing.assert_allclose(G, Go, rtol=1e-12, atol=1e-14 * np.abs(Go).max())
