"""CPU checks of the AO -> MO transform oracle (oracle/oracle_mo.c).

CuTC has no such function; its definition is the change of basis
out(g, x, i, j) = sum_kl C_L(k, i) A(k, l, g, x) C_R(l, j).  Pinned by a
pure-Python loop on a tiny case, numpy einsum at other shapes (n_mo below,
equal to and above n_ao), the identity basis (a pure re-layout of A), the
exchange of the two coefficient matrices on a (k, l)-symmetric A (a transpose
of (i, j)) and empty sizes."""
import numpy as np
import pytest

from oracle import oracle


def inputs(n1, na, nm, seed, sym=False):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((3, n1, na, na))
    if sym:
        X = X + X.transpose(0, 1, 3, 2)
    return rng.standard_normal((nm, na)), rng.standard_normal((nm, na)), X


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_tiny_against_loops():
    cl, cr, X = inputs(2, 3, 2, 1)
    out = oracle.ao_to_mo(cl, cr, X)
    for g in range(2):
        for x in range(3):
            for i in range(2):
                for j in range(2):
                    s = 0.0
                    for k in range(3):
                        for l in range(3):
                            s += cl[i, k] * X[x, g, l, k] * cr[j, l]  # Fortran A(k, l, g, x)
                    assert abs(out[j, i, x, g] - s) <= 1e-13 * max(1.0, abs(s))


@pytest.mark.parametrize("n1,na,nm", [(7, 5, 5), (11, 9, 4), (5, 6, 10), (1, 1, 1), (4, 17, 13)])
def test_matches_einsum(n1, na, nm):
    cl, cr, X = inputs(n1, na, nm, n1 + na + nm)
    out = oracle.ao_to_mo(cl, cr, X)
    want = np.einsum("ik,xglk,jl->jixg", cl, X, cr)
    assert rel(out, want) < 1e-13


def test_identity_basis_is_a_relayout():
    _, _, X = inputs(6, 8, 8, 3)
    eye = np.eye(8)
    out = oracle.ao_to_mo(eye, eye, X)
    # out(g, x, i, j) = A(i, j, g, x): numpy out[j, i, x, g] = X[x, g, j, i]
    assert np.array_equal(out, X.transpose(2, 3, 0, 1))


def test_exchanging_coefficients_transposes_for_symmetric_input():
    cl, cr, X = inputs(5, 7, 6, 4, sym=True)
    a = oracle.ao_to_mo(cl, cr, X)
    b = oracle.ao_to_mo(cr, cl, X)
    assert rel(b, a.transpose(1, 0, 2, 3)) < 1e-13


def test_empty_sizes():
    cl, cr, X = inputs(0, 4, 3, 5)
    assert oracle.ao_to_mo(cl, cr, X).shape == (3, 3, 3, 0)
    cl, cr, X = inputs(3, 0, 2, 5)
    assert not oracle.ao_to_mo(cl, cr, X).any()
