"""CPU checks of the normal-ordering oracle (oracle/oracle_no.c).

The reference (src/cutc_no.c and its no_*_kernel.cu) ships no known answers.
The oracle restates the kernels loop by loop; these tests pin it by
  * an independent tensor formulation of the same sums (numpy einsum over
    occupation weights, written from the kernels' index maps, not from the
    oracle's loops), closed and open shell,
  * the algebra every output obeys: degree 2 in int2_grad1_u12 and degree 1
    in each of wr1, mos_l_in_r and mos_r_in_r; additivity over disjoint grid
    subsets (every output is a sum over grid points); the index symmetry
    no_2e(p,q,s,t) = no_2e(q,p,t,s) left by N + N^T and the (q,s) swap;
    all-zero outputs without electrons,
  * argument checks.
The coefficients of the individual terms (the -2, 1/2, 1/4 factors of the
kernels) are pinned only by the restatement itself: parity of cutc_no is
partially pinned."""
import numpy as np
import pytest

from oracle import oracle


def inputs(ng, nm, seed):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((nm, nm, 3, ng))
    w = rng.random(ng) + 0.1
    L = rng.standard_normal((nm, ng))
    R = rng.standard_normal((nm, ng))
    return w, L, R, G


def tensor_form(na, nb, w, L, R, G):
    """no_2e (as [t, s, q, p]), no_1e ([q, p]), no_0e from the sums written as
    tensor contractions with orbital weights w1 and pair weights W."""
    Gm = G.transpose(3, 2, 1, 0)  # [g, x, i, j]
    Lm, Rm = L.T, R.T  # [g, e]
    ng, nm = Lm.shape
    ne = max(na, nb)
    w1 = np.array([1.0] * min(nb, ne) + [0.5] * (ne - nb))
    W = np.full((ne, ne), 0.5)
    W[:nb, :nb] = 1.0
    Ge = Gm[:, :, :ne, :ne]
    Le, Re = Lm[:, :ne], Rm[:, :ne]
    J = np.einsum("e,gxee->gx", w1, Ge)
    O = np.einsum("e,ge,ge->g", w1, Le, Re)
    M = np.einsum("ef,ge,gf,gxfe->gx", W, Le, Re, Ge)
    S = 2 * (J * J).sum(1) - np.einsum("ef,gxfe,gxef->g", W, Ge, Ge)
    A = w[:, None, None] * np.einsum("e,ge,gxpe->gxp", w1, Le, Gm[:, :, :, :ne])
    Rt = np.einsum("e,ge,gxep->gxp", w1, Re, Gm[:, :, :ne, :])
    C2 = np.einsum("e,gxpe,gxes->gps", w1, Gm[:, :, :, :ne], Gm[:, :, :ne, :])
    TE = np.einsum("ef,gf,gxef->gxe", W, Le, Ge)
    TF = np.einsum("ef,ge,gxef->gxf", W, Re, Ge)
    E = np.zeros((ng, 5, nm))
    E[:, 0] = w[:, None] * Lm
    E[:, 1] = -2 * np.einsum("gxp,gx->gp", A, J) + w[:, None] * np.einsum("gxpe,gxe->gp", Gm[:, :, :, :ne], TE)
    E[:, 2:] = A
    F = np.zeros((ng, 5, nm))
    F[:, 0] = (-2 * np.einsum("gxp,gx->gp", Rt, J) + Rm * S[:, None] +
               np.einsum("gxfp,gxf->gp", Gm[:, :, :ne, :], TF))
    F[:, 1] = Rm
    F[:, 2:] = Rt
    C1 = (np.einsum("gs,gxp->gxps", Rm, A) + np.einsum("gp,g,gxs->gxps", Lm, w, Rt) -
          (w * O)[:, None, None, None] * Gm - 2 * np.einsum("g,gp,gs,gx->gxps", w, Lm, Rm, J))
    D2 = np.einsum("g,gp,gs->gps", w, Lm, Rm)
    Dv = w[:, None] * (2 * O[:, None] * J - M)
    Dv3 = -w * O
    N = 0.5 * np.einsum("gxps,gxqt->psqt", C1, Gm) + 0.5 * np.einsum("gps,gqt->psqt", C2, D2)
    no2 = N + N.transpose(2, 3, 0, 1)  # (p, s, q, t)
    no2_pqst = no2.transpose(0, 2, 1, 3)  # (p, q, s, t)
    no1 = (2 * np.einsum("gxpq,gx->pq", Gm, Dv) + 2 * np.einsum("gpq,g->pq", C2, Dv3) +
           np.einsum("gcp,gcq->pq", E, F))
    e0 = np.einsum("gxp,gxp->", A[:, :, :nb], Rt[:, :, :nb]) + (w * (O * S - 2 * (J * M).sum(1))).sum()
    no0 = -2 * e0
    if na > nb:
        op = slice(nb, na)
        Gt = 0.5 * w[:, None, None] * np.einsum("ge,gxpe->gxp", Lm[:, op], Gm[:, :, :, op])
        Ht = 0.5 * np.einsum("ge,gxep->gxp", Rm[:, op], Gm[:, :, op, :])
        no1 = no1 + np.einsum("gxp,gxq->pq", Gt, Ht)
        xy = 0.0
        for f in range(na):
            es = slice(nb, na) if f < nb else slice(0, na)
            cx = 0.25 if f < nb else 0.5
            X = cx * w[:, None] * np.einsum("ge,gxe->gx", Lm[:, es], Gm[:, :, f, es])
            Y = np.einsum("ge,gxe->gx", Rm[:, es], Gm[:, :, es, f])
            xy += (X * Y).sum()
        no0 -= 2 * xy
    return no2_pqst.transpose(3, 2, 1, 0), no1.T, no0


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


@pytest.mark.parametrize("ng,nm,na,nb", [(70, 5, 2, 2), (133, 7, 3, 3), (90, 6, 4, 2), (64, 8, 5, 1), (50, 4, 1, 3),
                                         (40, 5, 0, 0), (33, 6, 6, 6)])
def test_oracle_matches_tensor_form(ng, nm, na, nb):
    w, L, R, G = inputs(ng, nm, ng + 10 * nm + na)
    o2, o1, o0 = oracle.cutc_no(na, nb, w, L, R, G)
    t2, t1, t0 = tensor_form(na, nb, w, L, R, G)
    assert rel(o2, t2) < 1e-12
    assert rel(o1, t1) < 1e-12
    assert abs(o0 - t0) <= 1e-12 * max(1.0, abs(t0))


@pytest.mark.parametrize("na,nb", [(3, 3), (4, 2)])
def test_homogeneity(na, nb):
    w, L, R, G = inputs(80, 6, 7)
    base = oracle.cutc_no(na, nb, w, L, R, G)
    cases = [((w, L, R, 1.7 * G), 1.7 ** 2), ((1.3 * w, L, R, G), 1.3), ((w, -0.6 * L, R, G), -0.6),
             ((w, L, 2.1 * R, G), 2.1)]
    for args, f in cases:
        out = oracle.cutc_no(na, nb, *args)
        for a, b in zip(out[:2], base[:2]):
            assert rel(a, f * b) < 1e-12
        assert abs(out[2] - f * base[2]) <= 1e-12 * abs(f * base[2])


@pytest.mark.parametrize("na,nb", [(3, 3), (5, 2)])
def test_additive_over_grid_points(na, nb):
    w, L, R, G = inputs(101, 6, 3)
    whole = oracle.cutc_no(na, nb, w, L, R, G)
    cut = 37
    a = oracle.cutc_no(na, nb, w[:cut], L[:, :cut], R[:, :cut], G[..., :cut])
    b = oracle.cutc_no(na, nb, w[cut:], L[:, cut:], R[:, cut:], G[..., cut:])
    for i in range(2):
        assert rel(a[i] + b[i], whole[i]) < 1e-12
    assert abs(a[2] + b[2] - whole[2]) <= 1e-12 * abs(whole[2])


def test_no_2e_index_symmetry():
    w, L, R, G = inputs(60, 5, 11)
    o2, _, _ = oracle.cutc_no(4, 2, w, L, R, G)
    # array [t, s, q, p] = no_2e(p, q, s, t); no_2e(q, p, t, s) is array [s, t, p, q]
    assert np.allclose(o2, o2.transpose(1, 0, 3, 2), rtol=0, atol=1e-12 * np.abs(o2).max())
    assert not np.allclose(o2, o2.transpose(0, 1, 3, 2))  # not symmetric in (p, q) alone


def test_no_electrons_give_zero():
    w, L, R, G = inputs(30, 4, 5)
    o2, o1, o0 = oracle.cutc_no(0, 0, w, L, R, G)
    assert not o2.any() and not o1.any() and o0 == 0.0


def test_bad_arguments():
    w, L, R, G = inputs(10, 3, 1)
    for na, nb in ((4, 1), (1, 4), (-1, 0)):
        with pytest.raises(ValueError):
            oracle.cutc_no(na, nb, w, L, R, G)
