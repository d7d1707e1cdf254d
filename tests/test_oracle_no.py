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
  * argument checks,
  * Wick's theorem: a brute-force normal ordering over spin orbitals of the
    three-electron operator built from int2_grad1_u12 gives all three outputs
    (closed shell, open shell, ne_a < ne_b) up to one convention constant each;
    this fixes the coefficient of every term (the -2, 1/2, 1/4 factors of the
    kernels), and a coefficient planted wrong in the numpy restatement fails
    it."""
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


def tensor_form(na, nb, w, L, R, G, plant=None):
    """no_2e (as [t, s, q, p]), no_1e ([q, p]), no_0e from the sums written as
    tensor contractions with orbital weights w1 and pair weights W.  plant:
    the name of one coefficient to scale by 1.1 (a planted mistake)."""
    k = lambda name: 1.1 if plant == name else 1.0
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
    E[:, 1] = -2 * k("E1") * np.einsum("gxp,gx->gp", A, J) + w[:, None] * np.einsum("gxpe,gxe->gp", Gm[:, :, :, :ne], TE)
    E[:, 2:] = A
    F = np.zeros((ng, 5, nm))
    F[:, 0] = (-2 * k("F0") * np.einsum("gxp,gx->gp", Rt, J) + Rm * S[:, None] +
               np.einsum("gxfp,gxf->gp", Gm[:, :, :ne, :], TF))
    F[:, 1] = Rm
    F[:, 2:] = Rt
    C1 = (np.einsum("gs,gxp->gxps", Rm, A) + np.einsum("gp,g,gxs->gxps", Lm, w, Rt) -
          (w * O)[:, None, None, None] * Gm - 2 * k("C1") * np.einsum("g,gp,gs,gx->gxps", w, Lm, Rm, J))
    D2 = np.einsum("g,gp,gs->gps", w, Lm, Rm)
    Dv = w[:, None] * (2 * O[:, None] * J - M)
    Dv3 = -w * O
    N = 0.5 * np.einsum("gxps,gxqt->psqt", C1, Gm) + 0.5 * k("N") * np.einsum("gps,gqt->psqt", C2, D2)
    no2 = N + N.transpose(2, 3, 0, 1)  # (p, s, q, t)
    no2_pqst = no2.transpose(0, 2, 1, 3)  # (p, q, s, t)
    no1 = (2 * np.einsum("gxpq,gx->pq", Gm, Dv) + 2 * k("no1") * np.einsum("gpq,g->pq", C2, Dv3) +
           np.einsum("gcp,gcq->pq", E, F))
    e0 = np.einsum("gxp,gxp->", A[:, :, :nb], Rt[:, :, :nb]) + (w * (O * S - 2 * k("e0") * (J * M).sum(1))).sum()
    no0 = -2 * e0
    if na > nb:
        op = slice(nb, na)
        Gt = 0.5 * k("Gt") * w[:, None, None] * np.einsum("ge,gxpe->gxp", Lm[:, op], Gm[:, :, :, op])
        Ht = 0.5 * np.einsum("ge,gxep->gxp", Rm[:, op], Gm[:, :, op, :])
        no1 = no1 + np.einsum("gxp,gxq->pq", Gt, Ht)
        xy = 0.0
        for f in range(na):
            es = slice(nb, na) if f < nb else slice(0, na)
            cx = (0.25 * k("X")) if f < nb else 0.5
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


# ---------------------------------------------------------------- Wick's theorem
#
# An independent derivation of all three outputs.  The three-electron operator
# whose normal-ordered parts cutc_no returns has the spatial integral
#   <p q r | L | s t u> = K(p s; q t, r u) + K(q t; p s, r u) + K(r u; p s, q t),
#   K(a b; q t, r u) = sum_g wr1(g) mos_l(g, a) mos_r(g, b)
#                      sum_x int2_grad1_u12(g, x, q, t) int2_grad1_u12(g, x, r, u)
# (one term per electron taken as the grid point; the other two are the
# integrated gradients).  Normal ordering of (1/6) sum L a+ a+ a+ a a a with
# respect to a determinant (bi-orthogonal: every contraction is a delta over
# occupied spin orbitals) leaves, with the ket antisymmetrised,
#   0e  E0 = 1/6 sum_{ijk occ} <ijk|L|ijk>_A
#   1e  F(P, Q) = 1/2 sum_{jk occ} <P jk|L|Q jk>_A
#   2e  V(P Q, S T) = sum_{k occ} <P Q k|L|S T k>_A
# computed here by brute force over spin orbitals.  The reference's outputs
# are these up to one constant each, and spin-averaged for open shells:
#   no_0e = -E0, no_1e = (F_aa + F_bb) / 2, no_2e = -(V_ab + V_ba) / 4
# (closed shell: F_aa = F_bb, V_ab = V_ba).  Orbitals e < ne_b are doubly
# occupied and ne_b <= e < ne_a singly (alpha); with ne_a < ne_b all ne_b are
# doubly occupied, as the reference's loops run.  Every coefficient inside each
# output (the -2, 1/2, 1/4 of the kernels) is fixed by this derivation, so a
# wrong or dropped term fails it.

def wick_parts(na, nb, w, L, R, G):
    import itertools
    Gf = G.transpose(3, 2, 1, 0)  # [g, x, i, j]
    K = np.einsum("g,ga,gb,gxqt,gxru->abqtru", w, L.T, R.T, Gf, Gf)
    L3 = (np.einsum("psqtru->pqrstu", K) + np.einsum("qtpsru->pqrstu", K) + np.einsum("rupsqt->pqrstu", K))
    nm = L3.shape[0]
    n = 2 * nm  # spin orbital P = 2 p + sigma
    Lso = np.zeros((n,) * 6)
    for sp in itertools.product(range(2), repeat=3):
        a, b, c = (slice(s, n, 2) for s in sp)
        Lso[a, b, c, a, b, c] = L3
    LA = np.zeros_like(Lso)
    for perm in itertools.permutations(range(3)):
        inv = sum(1 for i in range(3) for j in range(i + 1, 3) if perm[i] > perm[j])
        LA += (-1) ** inv * Lso.transpose(0, 1, 2, *(3 + p for p in perm))
    occ = [2 * e + s for e in range(nb) for s in range(2)] + [2 * e for e in range(nb, na)]
    E0 = sum(LA[i, j, k, i, j, k] for i in occ for j in occ for k in occ) / 6
    F = sum(LA[:, j, k, :, j, k] for j in occ for k in occ) / 2
    V = sum(LA[:, :, k, :, :, k] for k in occ)
    no1 = 0.5 * (F[0::2, 0::2] + F[1::2, 1::2])
    no2 = -0.25 * (V[0::2, 1::2, 0::2, 1::2] + V[1::2, 0::2, 1::2, 0::2])
    return no2, no1, -E0


@pytest.mark.parametrize("ng,nm,na,nb", [(7, 4, 2, 2), (9, 4, 3, 1), (6, 4, 4, 2), (8, 4, 1, 3), (5, 5, 3, 3),
                                         (5, 5, 4, 2), (6, 3, 3, 3)])
def test_oracle_matches_wick(ng, nm, na, nb):
    w, L, R, G = inputs(ng, nm, 100 + ng + nm + na)
    o2, o1, o0 = oracle.cutc_no(na, nb, w, L, R, G)
    t2, t1, t0 = wick_parts(na, nb, w, L, R, G)
    assert rel(o2.transpose(3, 2, 1, 0), t2) < 1e-12  # array [t, s, q, p] -> (p, q, s, t)
    assert rel(o1.T, t1) < 1e-12  # array [q, p] -> (p, q)
    assert abs(o0 - t0) <= 1e-12 * abs(t0)


@pytest.mark.parametrize("plant,na,nb", [("E1", 2, 2), ("F0", 2, 2), ("C1", 2, 2), ("N", 2, 2), ("no1", 2, 2),
                                         ("e0", 2, 2), ("Gt", 3, 1), ("X", 3, 1)])
def test_wick_catches_a_planted_coefficient(plant, na, nb):
    """One coefficient of the restatement scaled by 1.1 (the unplanted
    restatement equals the oracle, test_oracle_matches_tensor_form) moves at
    least one output away from the derivation."""
    w, L, R, G = inputs(7, 4, 31)
    t2, t1, t0 = wick_parts(na, nb, w, L, R, G)
    good = tensor_form(na, nb, w, L, R, G)
    bad = tensor_form(na, nb, w, L, R, G, plant=plant)
    assert rel(good[0].transpose(3, 2, 1, 0), t2) < 1e-12 and rel(good[1].T, t1) < 1e-12
    assert abs(good[2] - t0) <= 1e-12 * abs(t0)
    err = max(rel(bad[0].transpose(3, 2, 1, 0), t2), rel(bad[1].T, t1), abs(bad[2] - t0) / abs(t0))
    assert err > 1e-4, (plant, err)
