"""CPU checks of the AO oracle (oracle/oracle_ao.c).

CuTC takes aos_data(n_grid, n_ao, 4) from its caller; the function is the
textbook contracted Cartesian Gaussian and its gradient.  Pinned by the
closed form of an s primitive, central finite differences of the values
(gradients), the synthetic-input generator's own AO code (cutc_amd/csrc/
synth.c, uncontracted AOs), linearity in the contraction, translation
invariance and argument checks."""
import numpy as np
import pytest

from cutc_amd import synth
from oracle import oracle


def contracted_basis(n_nuc, n_ao, pmax, seed):
    """random contracted Cartesian AOs, powers up to 3 per direction"""
    rng = np.random.default_rng(seed)
    nucl = rng.integers(1, n_nuc + 1, n_ao).astype(np.int32)
    power = rng.integers(0, 4, (3, n_ao)).astype(np.int32)
    nprim = rng.integers(1, pmax + 1, n_ao).astype(np.int32)
    expo = np.exp(rng.uniform(np.log(0.1), np.log(50.0), (n_ao, pmax)))
    coef = rng.standard_normal((n_ao, pmax))
    return dict(ao_nucl=nucl, ao_power=power, ao_prim_num=nprim, ao_expo=expo, ao_coef=coef)


def points(n, seed, scale=2.0):
    return np.random.default_rng(seed).standard_normal((n, 3)) * scale


def test_s_primitive_closed_form():
    rn = np.array([[0.3, -0.2, 0.5]])
    b = dict(ao_nucl=[1], ao_power=[[0], [0], [0]], ao_prim_num=[1], ao_expo=[[0.7]], ao_coef=[[1.3]])
    r = points(50, 1)
    out = oracle.aos(r, rn, **b)
    d = r - rn[0]
    v = 1.3 * np.exp(-0.7 * (d * d).sum(1))
    assert np.allclose(out[0, 0], v, rtol=1e-15, atol=0)
    for c in range(3):
        assert np.allclose(out[1 + c, 0], -2 * 0.7 * d[:, c] * v, rtol=1e-14, atol=1e-300)


def test_gradient_matches_finite_differences():
    rn = points(4, 2, 1.5)
    b = contracted_basis(4, 30, 3, 3)
    b["ao_expo"] = np.minimum(b["ao_expo"], 5.0)  # smooth enough for central differences
    r = points(40, 4)
    out = oracle.aos(r, rn, **b)
    h = 1e-5
    for c in range(3):
        e = np.zeros(3)
        e[c] = h
        fd = (oracle.aos(r + e, rn, **b)[0] - oracle.aos(r - e, rn, **b)[0]) / (2 * h)
        scale = np.abs(out[1 + c]).max()
        assert np.abs(fd - out[1 + c]).max() <= 1e-6 * scale


def test_matches_generator_for_uncontracted_aos():
    s = synth.make_system(5, 40, 300, 10, 17, name="aos")
    out = oracle.aos(s.r1, s.rn, **s.basis)
    assert np.abs(out - s.aos1).max() <= 1e-14 * np.abs(s.aos1).max()


def test_linear_in_the_contraction():
    rn = points(3, 5)
    b = contracted_basis(3, 12, 4, 6)
    r = points(30, 7)
    whole = oracle.aos(r, rn, **b)
    acc = np.zeros_like(whole)
    for p in range(4):
        bp = dict(b)
        bp["ao_prim_num"] = (b["ao_prim_num"] > p).astype(np.int32)
        bp["ao_expo"] = np.ascontiguousarray(b["ao_expo"][:, p:p + 1])
        bp["ao_coef"] = np.ascontiguousarray(b["ao_coef"][:, p:p + 1])
        acc += oracle.aos(r, rn, **bp)
    assert np.abs(acc - whole).max() <= 1e-13 * np.abs(whole).max()


def test_translation_invariance():
    rn = points(3, 8)
    b = contracted_basis(3, 10, 2, 9)
    r = points(25, 10)
    t = np.array([1.25, -0.5, 3.0])
    a, c = oracle.aos(r, rn, **b), oracle.aos(r + t, rn + t, **b)
    assert np.abs(a - c).max() <= 1e-12 * np.abs(a).max()


@pytest.mark.parametrize("field,value", [("ao_nucl", 0), ("ao_nucl", 4), ("ao_power", -1), ("ao_prim_num", 3),
                                         ("ao_prim_num", -1)])
def test_bad_tables(field, value):
    b = contracted_basis(3, 5, 2, 11)
    b[field] = b[field].copy()
    b[field].flat[2] = value
    with pytest.raises(ValueError):
        oracle.aos(points(4, 1), points(3, 2), **b)
