"""The reference's special cases of the Boys-Handy Jastrow on the device,
through the exported wrapper cutc_int_bh and the whole path, against the
FP64 oracle (src/jast_bh.cu:143,169,189: squared distances <= 1e-15 give a
zero rbar and a zero gradient; :203-204: terms with |c| < 1e-10 are
skipped), with exponents 0, 6 and 7."""
import numpy as np
import pytest

from cutc_amd import cutc, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12


def relerr(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need the MI355X"
    return t


# exponents 0 (including an all-zero term, which contributes nothing), 6 and 7
TERMS = [(0, 0, 1), (0, 0, 2), (2, 0, 0), (3, 0, 0), (2, 2, 0), (0, 0, 3), (0, 0, 0), (6, 0, 7), (0, 6, 0),
         (7, 1, 6)]


def special_system(seed=21):
    """A triatomic system whose grids hold the degenerate configurations:
    r2 points equal to r1 points (r12 = 0), r1 and r2 points on nuclei
    (r1A = 0, r2A = 0), a pair 1e-9 apart (squared distance 3e-18, below
    the threshold) and one 3e-7 apart (above it)."""
    s = synth.triatomic(n1=40, n2=300, seed=seed)
    rng = np.random.default_rng(seed)
    s.c_bh = rng.uniform(-0.5, 0.5, (s.n_nuc, len(TERMS)))
    s.c_bh[:, 0] = 0.5
    s.m_bh = np.array([[t[0] for t in TERMS]] * s.n_nuc, np.int32)
    s.n_bh = np.array([[t[1] for t in TERMS]] * s.n_nuc, np.int32)
    s.o_bh = np.array([[t[2] for t in TERMS]] * s.n_nuc, np.int32)
    r1, r2 = s.r1.copy(), s.r2.copy()
    r2[0:5] = r1[0:5]                      # coincident r1 = r2
    r1[5] = s.rn[0]                        # r1 on nucleus 0
    r1[6] = s.rn[2]                        # r1 on nucleus 2
    r2[7] = s.rn[1]                        # r2 on nucleus 1
    r2[8] = r1[9] + np.array([1e-9, 1e-9, 1e-9])     # d^2 = 3e-18 <= 1e-15: zero
    r2[9] = r1[10] + np.array([3e-7, 0.0, 0.0])      # d^2 = 9e-14 > 1e-15: evaluated
    r1[11] = s.rn[0] + np.array([0.0, 2e-9, 0.0])    # r1A^2 = 4e-18: zero
    s.r1, s.r2 = np.ascontiguousarray(r1), np.ascontiguousarray(r2)
    return s


def jast_device(torch, s):
    ds = cutc.DeviceSystem.upload(s)
    G = torch.full((4, s.n1, s.n2), 7.0, dtype=torch.float64, device="cuda")
    cutc.cutc_int_bh(ds, G, 0, s.n1, s.n1, 0, s.n2, s.n2)
    torch.cuda.synchronize()
    return G.cpu().numpy()


def test_degenerate_distances(torch):
    s = special_system()
    g = jast_device(torch, s)
    want = oracle.jast_bh(s)
    print(f"special cases: relerr {relerr(g, want):.2e}", flush=True)
    assert relerr(g, want) <= TOL
    np.testing.assert_allclose(g, want, rtol=1e-11, atol=1e-13 * np.abs(want).max())
    assert np.all(np.isfinite(g))
    # the pure e-e part of a coincident pair vanishes: with only e-e terms
    # (m = n = 0) the whole gradient of those pairs is exactly 0
    ee = special_system()
    ee.m_bh[:] = 0
    ee.n_bh[:] = 0
    ee.o_bh[:] = np.maximum(ee.o_bh, 1)
    gee = jast_device(torch, ee)
    for i in range(5):
        assert np.all(gee[:, i, i] == 0.0)
    assert np.all(gee[:, 9, 8] == 0.0)
    assert np.any(gee[:, 10, 9] != 0.0)
    np.testing.assert_allclose(gee, oracle.jast_bh(ee), rtol=1e-11, atol=1e-13 * np.abs(gee).max())


def test_small_coefficients_skipped(torch):
    """|c| = 9e-11 < 1e-10 is skipped: bit-identical to c = 0.  |c| = 1.1e-10
    is kept (here on the (6, 0, 7) term, so that it changes G)."""
    def with_c(c7, c8):
        s = special_system()
        s.c_bh[:, 7] = c7
        s.c_bh[1, 8] = c8
        return s
    s0, s9, s11 = with_c(0.0, 0.0), with_c(9e-11, -9e-11), with_c(0.0, 0.0)
    s11.c_bh[0, 7] = 1.1e-10
    g0, g9, g11 = (jast_device(torch, s) for s in (s0, s9, s11))
    assert np.array_equal(g9, g0)
    assert not np.array_equal(g11, g0)
    for s, g in ((s0, g0), (s9, g9), (s11, g11)):
        assert relerr(g, oracle.jast_bh(s)) <= TOL


def test_whole_path_with_degenerate_points(torch):
    """cutc_int_device on the degenerate grids, and the reference's device
    stage functions on the same inputs, against the oracle."""
    s = special_system()
    ds = cutc.DeviceSystem.upload(s)
    X, I = cutc.alloc_outputs(ds)
    cutc.cutc_int_device(ds, X, I)
    torch.cuda.synchronize()
    Xo = oracle.int2_grad1_u12_ao(s)
    Io = oracle.int_2e_ao(s, Xo)
    ex, ei = relerr(X.cpu().numpy(), Xo), relerr(I.cpu().numpy(), Io)
    print(f"whole path, degenerate grids: X {ex:.2e} int_2e_ao {ei:.2e}", flush=True)
    assert ex <= TOL and ei <= TOL
    X2, I2 = cutc.alloc_outputs(ds)
    cutc.get_int2_grad1_u12_ao(ds, X2)
    cutc.get_int_2e_ao(ds, X2, I2)
    torch.cuda.synchronize()
    assert relerr(X2.cpu().numpy(), Xo) <= TOL and relerr(I2.cpu().numpy(), Io) <= TOL
