/*
 * FP64 CPU ORACLE for the normal-ordering path (cutc_no) -- TEST INFRASTRUCTURE.
 *
 * Only tests, the smoke check and the CPU-baseline leg of the benchmark load
 * this file (through liboracle.so); the HIP library never does.
 *
 * Restatement of src/cutc_no.c:14-778 and the kernels it launches, per grid
 * point in the kernels' own summation order, with host DGEMM for the vendor
 * calls.  Layouts (Fortran order, as the reference):
 *   wr1[g]                 grid-1 weights                 (n_grid1)
 *   mol[g + e ng]          mos_l_in_r                      (n_grid1, n_mo)
 *   mor[g + e ng]          mos_r_in_r                      (n_grid1, n_mo)
 *   G[g + x ng + i 3ng + j 3ng nm]  int2_grad1_u12         (n_grid1, 3, n_mo, n_mo)
 *   no_2e[p + q nm + s nm^2 + t nm^3], no_1e[p + q nm], no_0e[0]
 * Occupations: orbitals e < ne_b are doubly occupied; ne_b <= e < ne_a singly
 * (weight 1/2 in the open-shell kernels).  The reference runs its "_cs"
 * kernels when ne_a == ne_b and the "_os" kernels otherwise; the "_os"
 * kernels reduce to the "_cs" ones when their open loops are empty, so one
 * set of functions below follows the "_os" loops, and the open-shell extras
 * (src/cutc_no.c:584-694) run only when ne_a != ne_b, as in the reference.
 *
 * Citations per intermediate (src/):
 *   J   no_tmpJ_os_kernel.cu:4-45       O   no_tmpO_os_kernel.cu:4-35
 *   M   no_tmpM_os_kernel.cu:4-80       S   no_tmpS_os_kernel.cu:4-74
 *   A   no_2e_tmpA_os_kernel.cu:4-55    B   no_2e_tmpB_os_kernel.cu:4-57
 *   C2  no_tmpC2_os_kernel.cu:4-62      L   no_1e_tmpL_os_kernel.cu:4-55
 *   R   no_1e_tmpR_os_kernel.cu:4-55    E   no_1e_tmpE_os_kernel.cu:4-95
 *   F   no_1e_tmpF_os_kernel.cu:4-98    C1  no_2e_tmpC1_kernel.cu:4-79
 *   D2  no_2e_tmpD2_kernel.cu:4-43      D   no_1e_tmpD_kernel.cu:4-42
 *   U   no_0e_tmpU_kernel.cu:4-65 (32-point block sums, added on the host
 *       in block order, src/cutc_no.c:564-571)
 *   Gt  no_1e_tmpG_os_kernel.cu:4-49    Ht  no_1e_tmpH_os_kernel.cu:4-49
 *   X   no_0e_tmpX_os_kernel.cu:4-63    Y   no_0e_tmpY_os_kernel.cu:4-63
 *   assembly: src/cutc_no.c:419-468 (2e: two DGEMM, M + M^T, (q,s) swap
 *   of trans_pqst_psqt_inplace.cu:4-31), :481-525 (1e: two DGEMV, DGEMM),
 *   :534-571 (0e: DDOT + U), :584-694 (open-shell 1e DGEMM, 0e DDOT).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

void oracle_dgemm(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
                  const double *B, int64_t ldb, double beta, double *C, int64_t ldc);
void oracle_trans_inplace(double *data, int64_t size);

/* element (x, i, j) of int2_grad1_u12 at point g */
#define GX(x, i, j) G[g + (x) * ng + (int64_t)(i) * n1 + (int64_t)(j) * n2]

/* the per-point intermediates of one call */
typedef struct {
    double *J, *O, *M, *S, *A, *B, *C2, *L, *R, *E, *F, *C1, *D2, *D, *Ublk;
} no_tmp;

static void per_point(int64_t ng, int nm, int na, int nb, const double *w, const double *mol, const double *mor,
                      const double *G, no_tmp *t)
{
    const int64_t n1 = 3 * ng, n2 = n1 * nm, m2 = ng * nm;
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < ng; g++) {
        const double wg = w[g];
        double J[3], M[3], O = 0.0, S;
        /* J: diagonal elements, open orbitals at 1/2 */
        for (int x = 0; x < 3; x++) {
            double s = 0.0;
            for (int e = 0; e < nb; e++)
                s += GX(x, e, e);
            for (int e = nb; e < na; e++)
                s += 0.5 * GX(x, e, e);
            J[x] = s;
            t->J[g + x * ng] = s;
        }
        /* O */
        for (int e = 0; e < nb; e++)
            O += mol[g + e * ng] * mor[g + e * ng];
        for (int e = nb; e < na; e++)
            O += 0.5 * mol[g + e * ng] * mor[g + e * ng];
        t->O[g] = O;
        /* M: closed-closed, open-closed (both orders), open-open */
        for (int x = 0; x < 3; x++) {
            double s = 0.0;
            for (int ie = 0; ie < nb; ie++)
                for (int je = 0; je < nb; je++)
                    s += mol[g + ie * ng] * mor[g + je * ng] * GX(x, je, ie);
            for (int ie = nb; ie < na; ie++)
                for (int je = 0; je < nb; je++)
                    s += 0.5 * (mol[g + ie * ng] * mor[g + je * ng] * GX(x, je, ie) +
                                mol[g + je * ng] * mor[g + ie * ng] * GX(x, ie, je));
            for (int ie = nb; ie < na; ie++)
                for (int je = nb; je < na; je++)
                    s += 0.5 * mol[g + ie * ng] * mor[g + je * ng] * GX(x, je, ie);
            M[x] = s;
            t->M[g + x * ng] = s;
        }
        /* S */
        S = 2.0 * (J[0] * J[0] + J[1] * J[1] + J[2] * J[2]);
        for (int ie = 0; ie < nb; ie++)
            for (int je = 0; je < nb; je++)
                for (int x = 0; x < 3; x++)
                    S -= GX(x, je, ie) * GX(x, ie, je);
        for (int ie = nb; ie < na; ie++)
            for (int je = 0; je < nb; je++)
                for (int x = 0; x < 3; x++)
                    S -= GX(x, je, ie) * GX(x, ie, je);
        for (int ie = nb; ie < na; ie++)
            for (int je = nb; je < na; je++)
                for (int x = 0; x < 3; x++)
                    S -= 0.5 * GX(x, je, ie) * GX(x, ie, je);
        t->S[g] = S;
        /* A (= L, the same expression) and B (= w R) per orbital p */
        for (int p = 0; p < nm; p++) {
            for (int x = 0; x < 3; x++) {
                double a = 0.0, b = 0.0;
                for (int e = 0; e < nb; e++) {
                    a += wg * mol[g + e * ng] * GX(x, p, e);
                    b += mor[g + e * ng] * GX(x, e, p);
                }
                for (int e = nb; e < na; e++) {
                    a += 0.5 * wg * mol[g + e * ng] * GX(x, p, e);
                    b += 0.5 * mor[g + e * ng] * GX(x, e, p);
                }
                const int64_t ix = g + x * ng + (int64_t)p * n1;
                t->A[ix] = a;
                t->L[ix] = a;
                t->R[ix] = b;
                t->B[ix] = wg * b;
            }
        }
        /* C2(p, s) and D2(p, s) */
        for (int s = 0; s < nm; s++)
            for (int p = 0; p < nm; p++) {
                double c = 0.0;
                for (int e = 0; e < nb; e++)
                    for (int x = 0; x < 3; x++)
                        c += GX(x, p, e) * GX(x, e, s);
                for (int e = nb; e < na; e++)
                    for (int x = 0; x < 3; x++)
                        c += 0.5 * GX(x, p, e) * GX(x, e, s);
                t->C2[g + p * ng + s * m2] = c;
                t->D2[g + p * ng + s * m2] = wg * mol[g + p * ng] * mor[g + s * ng];
            }
        /* E and F (5 rows per orbital) */
        for (int p = 0; p < nm; p++) {
            const int64_t iL = g + (int64_t)p * n1, iE = g + (int64_t)p * 5 * ng;
            const double Lx = t->L[iL], Ly = t->L[iL + ng], Lz = t->L[iL + 2 * ng];
            const double Rx = t->R[iL], Ry = t->R[iL + ng], Rz = t->R[iL + 2 * ng];
            double e1 = -2.0 * (Lx * J[0] + Ly * J[1] + Lz * J[2]);
            for (int je = 0; je < nb; je++)
                for (int ie = 0; ie < nb; ie++)
                    for (int x = 0; x < 3; x++)
                        e1 += wg * mol[g + je * ng] * GX(x, p, ie) * GX(x, ie, je);
            for (int je = 0; je < nb; je++)
                for (int ie = nb; ie < na; ie++)
                    for (int x = 0; x < 3; x++)
                        e1 += 0.5 * wg * mol[g + je * ng] * GX(x, p, ie) * GX(x, ie, je) +
                              0.5 * wg * mol[g + ie * ng] * GX(x, p, je) * GX(x, je, ie);
            for (int je = nb; je < na; je++)
                for (int ie = nb; ie < na; ie++)
                    for (int x = 0; x < 3; x++)
                        e1 += 0.5 * wg * mol[g + je * ng] * GX(x, p, ie) * GX(x, ie, je);
            t->E[iE] = wg * mol[g + p * ng];
            t->E[iE + ng] = e1;
            t->E[iE + 2 * ng] = Lx;
            t->E[iE + 3 * ng] = Ly;
            t->E[iE + 4 * ng] = Lz;
            double f0 = -2.0 * (Rx * J[0] + Ry * J[1] + Rz * J[2]) + mor[g + p * ng] * S;
            for (int ie = 0; ie < nb; ie++)
                for (int je = 0; je < nb; je++)
                    for (int x = 0; x < 3; x++)
                        f0 += mor[g + ie * ng] * GX(x, je, p) * GX(x, ie, je);
            for (int ie = nb; ie < na; ie++)
                for (int je = 0; je < nb; je++)
                    for (int x = 0; x < 3; x++)
                        f0 += 0.5 * (mor[g + ie * ng] * GX(x, je, p) * GX(x, ie, je) +
                                     mor[g + je * ng] * GX(x, ie, p) * GX(x, je, ie));
            for (int ie = nb; ie < na; ie++)
                for (int je = nb; je < na; je++)
                    for (int x = 0; x < 3; x++)
                        f0 += 0.5 * mor[g + ie * ng] * GX(x, je, p) * GX(x, ie, je);
            t->F[iE] = f0;
            t->F[iE + ng] = mor[g + p * ng];
            t->F[iE + 2 * ng] = Rx;
            t->F[iE + 3 * ng] = Ry;
            t->F[iE + 4 * ng] = Rz;
        }
        /* C1(x, p, s): the weights, O and J enter multiplied by w */
        const double Ow = wg * O;
        for (int s = 0; s < nm; s++)
            for (int p = 0; p < nm; p++)
                for (int x = 0; x < 3; x++) {
                    const int64_t ix = g + x * ng + (int64_t)p * n1 + (int64_t)s * n2;
                    const double ml = mol[g + p * ng], mr = mor[g + s * ng];
                    t->C1[ix] = mr * t->A[g + x * ng + (int64_t)p * n1] + ml * t->B[g + x * ng + (int64_t)s * n1] -
                                Ow * G[ix] - 2.0 * ml * mr * (wg * J[x]);
                }
        /* D */
        for (int x = 0; x < 3; x++)
            t->D[g + x * ng] = wg * (2.0 * O * J[x] - M[x]);
        t->D[g + 3 * ng] = -wg * O;
    }
    /* U: 32-point block sums (the kernel's 32-thread blocks) */
    const int64_t nblk = (ng + 31) / 32;
    for (int64_t b = 0; b < nblk; b++) {
        double u[32];
        for (int k = 0; k < 32; k++) {
            const int64_t g = b * 32 + k;
            u[k] = 0.0;
            if (g < ng)
                u[k] = w[g] * (t->O[g] * t->S[g] - 2.0 * (t->J[g] * t->M[g] + t->J[g + ng] * t->M[g + ng] +
                                                         t->J[g + 2 * ng] * t->M[g + 2 * ng]));
        }
        for (int h = 16; h > 0; h >>= 1) /* the kernel's tree reduction */
            for (int k = 0; k < h; k++)
                u[k] += u[k + h];
        t->Ublk[b] = u[0];
    }
}

/* (p,q,s,t) -> (p,s,q,t) swap in place (trans_pqst_psqt_inplace.cu:4-31) */
static void swap_qs(int64_t nm, double *d)
{
    const int64_t n2 = nm * nm, n3 = n2 * nm;
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < nm; t++)
        for (int64_t p = 0; p < nm; p++)
            for (int64_t s = 0; s < nm; s++)
                for (int64_t q = 0; q < s; q++) {
                    const int64_t a = p + q * nm + s * n2 + t * n3, b = p + s * nm + q * n2 + t * n3;
                    const double v = d[a];
                    d[a] = d[b];
                    d[b] = v;
                }
}

static double dot(int64_t n, const double *a, const double *b)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; i++)
        s += a[i] * b[i];
    return s;
}

int oracle_cutc_no(int64_t ng, int nm, int na, int nb, const double *w, const double *mol, const double *mor,
                   const double *G, double *no_2e, double *no_1e, double *no_0e)
{
    if (ng < 0 || nm < 0 || na < 0 || nb < 0 || nb > nm || na > nm)
        return -2;
    const int64_t n1 = 3 * ng, nm2 = (int64_t)nm * nm;
    no_tmp t;
    const int64_t nblk = (ng + 31) / 32;
    double *buf = calloc((size_t)(3 * ng + ng + 3 * ng + ng + 4 * n1 * nm + 2 * ng * nm2 + 10 * ng * nm + n1 * nm2 +
                                  4 * ng + nblk + 1),
                         sizeof(double));
    if (!buf)
        return -3;
    double *q = buf;
    t.J = q, q += 3 * ng;
    t.O = q, q += ng;
    t.M = q, q += 3 * ng;
    t.S = q, q += ng;
    t.A = q, q += n1 * nm;
    t.B = q, q += n1 * nm;
    t.L = q, q += n1 * nm;
    t.R = q, q += n1 * nm;
    t.C2 = q, q += ng * nm2;
    t.D2 = q, q += ng * nm2;
    t.E = q, q += 5 * ng * nm;
    t.F = q, q += 5 * ng * nm;
    t.C1 = q, q += n1 * nm2;
    t.D = q, q += 4 * ng;
    t.Ublk = q;
    per_point(ng, nm, na, nb, w, mol, mor, G, &t);

    /* 2e (src/cutc_no.c:419-468) */
    oracle_dgemm('T', 'N', nm2, nm2, n1, 0.5, t.C1, n1, G, n1, 0.0, no_2e, nm2);
    oracle_dgemm('T', 'N', nm2, nm2, ng, 0.5, t.C2, ng, t.D2, ng, 1.0, no_2e, nm2);
    oracle_trans_inplace(no_2e, nm2);
    swap_qs(nm, no_2e);

    /* 1e (src/cutc_no.c:481-525): two DGEMV and a DGEMM */
    oracle_dgemm('T', 'N', nm2, 1, n1, 2.0, G, n1, t.D, n1, 0.0, no_1e, nm2);
    oracle_dgemm('T', 'N', nm2, 1, ng, 2.0, t.C2, ng, t.D + 3 * ng, ng, 1.0, no_1e, nm2);
    oracle_dgemm('T', 'N', nm, nm, 5 * ng, 1.0, t.E, 5 * ng, t.F, 5 * ng, 1.0, no_1e, nm);

    /* 0e (src/cutc_no.c:534-571) */
    double e0 = dot(n1 * nb, t.L, t.R);
    for (int64_t b = 0; b < nblk; b++)
        e0 += t.Ublk[b];
    e0 = -2.0 * e0;

    if (na != nb) { /* open-shell extras (src/cutc_no.c:584-694) */
        const int64_t n2 = n1 * nm;
        double *Gt = calloc((size_t)(2 * n1 * nm + 2 * n1 * (na > 0 ? na : 1)), sizeof(double));
        if (!Gt) {
            free(buf);
            return -3;
        }
        double *Ht = Gt + n1 * nm, *X = Ht + n1 * nm, *Y = X + n1 * (na > 0 ? na : 1);
#pragma omp parallel for schedule(static)
        for (int64_t g = 0; g < ng; g++) {
            for (int p = 0; p < nm; p++)
                for (int x = 0; x < 3; x++) {
                    double a = 0.0, b = 0.0;
                    for (int e = nb; e < na; e++) {
                        a += 0.5 * w[g] * mol[g + e * ng] * GX(x, p, e);
                        b += 0.5 * mor[g + e * ng] * GX(x, e, p);
                    }
                    Gt[g + x * ng + (int64_t)p * n1] = a;
                    Ht[g + x * ng + (int64_t)p * n1] = b;
                }
            for (int f = 0; f < na; f++)
                for (int x = 0; x < 3; x++) {
                    double a = 0.0, b = 0.0;
                    if (f < nb) {
                        for (int e = nb; e < na; e++) {
                            a += 0.25 * w[g] * mol[g + e * ng] * GX(x, f, e);
                            b += mor[g + e * ng] * GX(x, e, f);
                        }
                    } else {
                        for (int e = 0; e < na; e++) {
                            a += 0.5 * w[g] * mol[g + e * ng] * GX(x, f, e);
                            b += mor[g + e * ng] * GX(x, e, f);
                        }
                    }
                    X[g + x * ng + (int64_t)f * n1] = a;
                    Y[g + x * ng + (int64_t)f * n1] = b;
                }
        }
        (void)n2;
        oracle_dgemm('T', 'N', nm, nm, n1, 1.0, Gt, n1, Ht, n1, 1.0, no_1e, nm);
        e0 -= 2.0 * dot(n1 * na, X, Y);
        free(Gt);
    }
    no_0e[0] = e0;
    free(buf);
    return 0;
}
