"""The committed fixture tests/golden/chain_tiny.npz (written once by
tests/golden/make_fixtures.py): a tiny seeded system (2 nuclei, n_ao = 8,
16 + 48 grid points, n_mo = 6) with the oracle's answers for the whole chain
aos -> cutc_int -> AO->MO -> cutc_no (open and closed shell).

CPU: the oracle still reproduces every array (<= 1e-14 relative Frobenius),
the generator still draws the same inputs, and the fixture has the symmetries
the reference's outputs have (int2_grad1_u12_ao symmetric in (k, l) up to
rounding, int_2e_ao exactly symmetric under (kl) <-> (ab) after the M + M^T
step, src/build_int_2e_ao.cu:118-124).

GPU: every host-array entry point of the library, called through the C ABI on
the fixture's inputs, matches the fixture to <= 1e-12 (FP64 throughout; only
the summation order differs)."""
import os

import numpy as np
import pytest

from cutc_amd import cutc, synth
from oracle import oracle

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "chain_tiny.npz")
ORACLE_TOL = 1e-14
GPU_TOL = 1e-12


@pytest.fixture(scope="module")
def g():
    with np.load(FIX) as d:
        return {k: d[k] for k in d.files}


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def basis(g):
    return {k: g[k] for k in ("ao_nucl", "ao_power", "ao_prim_num", "ao_expo", "ao_coef")}


def system(g):
    return synth.System(g["rn"], g["r1"], g["wr1"], g["r2"], g["wr2"], g["aos1"], g["aos2"], g["c_bh"], g["m_bh"],
                        g["n_bh"], g["o_bh"], name="golden-tiny", basis=basis(g))


# ------------------------------------------------------------------- CPU

def test_fixture_shapes(g):
    assert g["aos1"].shape == (4, 8, 16) and g["aos2"].shape == (4, 8, 48)
    assert g["int2_grad1_u12_ao"].shape == (3, 16, 8, 8) and g["int_2e_ao"].shape == (8,) * 4
    assert g["int2_grad1_u12"].shape == (6, 6, 3, 16) and g["no_2e"].shape == (6,) * 4
    for k in g:
        assert np.all(np.isfinite(g[k])), k


def test_fixture_symmetries(g):
    X, I = g["int2_grad1_u12_ao"], g["int_2e_ao"]
    assert rel(X, X.transpose(0, 1, 3, 2)) <= ORACLE_TOL  # two DGEMM sums of the same products
    M = I.reshape(64, 64)
    np.testing.assert_array_equal(M, M.T)
    assert np.count_nonzero(X) > 0.9 * X.size


def test_generator_reproduces_inputs(g):
    from tests.golden import make_fixtures as mf
    s = synth.make_system(mf.N_NUC, mf.N_AO, mf.N1, mf.N2, mf.SEED, nuclei=mf.NUCLEI, shells=mf.SHELLS)
    for k in ("rn", "r1", "wr1", "r2", "wr2"):
        np.testing.assert_array_equal(getattr(s, k), g[k])
    for k in ("c_bh", "m_bh", "n_bh", "o_bh"):
        np.testing.assert_array_equal(getattr(s, k), g[k])
    for k, v in s.basis.items():
        np.testing.assert_array_equal(v, g[k])


def test_oracle_reproduces_aos(g):
    b = basis(g)
    for r, want in ((g["r1"], g["aos1"]), (g["r2"], g["aos2"])):
        got = oracle.aos(r, g["rn"], b["ao_nucl"], b["ao_power"], b["ao_prim_num"], b["ao_expo"], b["ao_coef"])
        assert rel(got, want) <= ORACLE_TOL


def test_oracle_reproduces_cutc_int(g):
    X, I = oracle.cutc_int(system(g))
    assert rel(X, g["int2_grad1_u12_ao"]) <= ORACLE_TOL
    assert rel(I, g["int_2e_ao"]) <= ORACLE_TOL


def test_oracle_reproduces_mo_and_no(g):
    G = oracle.ao_to_mo(g["mo_l_coef"], g["mo_r_coef"], g["int2_grad1_u12_ao"])
    assert rel(G, g["int2_grad1_u12"]) <= ORACLE_TOL
    ne_a, ne_b = (int(v) for v in g["ne"])
    for (a, b), pre in (((ne_a, ne_b), ""), ((ne_a, ne_a), "cs_")):
        o2, o1, o0 = oracle.cutc_no(a, b, g["wr1"], g["mos_l"], g["mos_r"], g["int2_grad1_u12"])
        assert rel(o2, g[pre + "no_2e"]) <= ORACLE_TOL
        assert rel(o1, g[pre + "no_1e"]) <= ORACLE_TOL
        assert abs(o0 - g[pre + "no_0e"][0]) <= ORACLE_TOL * abs(g[pre + "no_0e"][0])


# ------------------------------------------------------------------- GPU

@pytest.mark.gpu
def test_gpu_aos_golden(g):
    for r, want in ((g["r1"], g["aos1"]), (g["r2"], g["aos2"])):
        got = cutc.cutc_aos(r, g["rn"], basis(g))
        assert rel(got, want) <= GPU_TOL


@pytest.mark.gpu
def test_gpu_cutc_int_golden(g):
    X, I = cutc.cutc_int(system(g))
    assert rel(X, g["int2_grad1_u12_ao"]) <= GPU_TOL
    assert rel(I, g["int_2e_ao"]) <= GPU_TOL


@pytest.mark.gpu
def test_gpu_split_golden(g):
    """The default split precision (modular int8) on the same inputs, to the
    split bar of 1e-10."""
    X, I, _ = cutc.cutc_int_ex(system(g), cutc.options(precision="int8"))
    assert rel(X, g["int2_grad1_u12_ao"]) <= 1e-10
    assert rel(I, g["int_2e_ao"]) <= 1e-10


@pytest.mark.gpu
def test_gpu_ao_to_mo_golden(g):
    G = cutc.cutc_ao_to_mo(g["mo_l_coef"], g["mo_r_coef"], g["int2_grad1_u12_ao"])
    assert rel(G, g["int2_grad1_u12"]) <= GPU_TOL


@pytest.mark.gpu
def test_gpu_cutc_no_golden(g):
    ne_a, ne_b = (int(v) for v in g["ne"])
    for (a, b), pre in (((ne_a, ne_b), ""), ((ne_a, ne_a), "cs_")):
        o2, o1, o0 = cutc.cutc_no(a, b, g["wr1"], g["mos_l"], g["mos_r"], g["int2_grad1_u12"])
        assert rel(o2, g[pre + "no_2e"]) <= GPU_TOL
        assert rel(o1, g[pre + "no_1e"]) <= GPU_TOL
        assert abs(o0 - g[pre + "no_0e"][0]) <= GPU_TOL * abs(g[pre + "no_0e"][0])
