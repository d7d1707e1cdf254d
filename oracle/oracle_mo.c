/*
 * FP64 CPU ORACLE for the AO -> MO transform of int2_grad1_u12 -- TEST
 * INFRASTRUCTURE.  Only tests, the smoke check and the CPU-baseline leg of
 * the benchmark load this file (through liboracle.so); the HIP library never
 * does.
 *
 * CuTC has no such function (its callers run it between cutc_int and
 * cutc_no).  Definition, per grid point g and component x:
 *   out(g, x, i, j) = sum_{k,l} C_L(k, i) A(k, l, g, x) C_R(l, j)
 * computed as the usual two DGEMMs of an AO -> MO change of basis,
 * T = A_gx C_R then C_L^T T, one grid point at a time.
 * Layouts (Fortran order):
 *   A[k + na l + na^2 (g + n1 x)]      int2_grad1_u12_ao (n_ao, n_ao, n_grid1, 3)
 *                                      (cutc_int's output, src/cutc_int.c:210-211)
 *   C_L[k + na i], C_R[l + na j]       (n_ao, n_mo)
 *   out[g + n1 x + 3 n1 i + 3 n1 nm j] int2_grad1_u12 (n_grid1, 3, n_mo, n_mo)
 *                                      (cutc_no's input, src/no_tmpJ_cs_kernel.cu:18)
 */
#include <stdint.h>
#include <stdlib.h>

void oracle_dgemm(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
                  const double *B, int64_t ldb, double beta, double *C, int64_t ldc);

int oracle_ao_to_mo(int64_t n1, int64_t na, int64_t nm, const double *CL, const double *CR, const double *A,
                    double *out)
{
    if (n1 < 0 || na < 0 || nm < 0)
        return -2;
    const int64_t R = 3 * n1;
    if (R == 0 || nm == 0)
        return 0;
    double *T = (double *)malloc(sizeof(double) * (size_t)((na > 0 ? na : 1) * nm));
    double *O = (double *)malloc(sizeof(double) * (size_t)(nm * nm));
    if (!T || !O) {
        free(T);
        free(O);
        return -3;
    }
    for (int64_t gx = 0; gx < R; gx++) {
        const double *Ag = A + na * na * gx;
        oracle_dgemm('N', 'N', na, nm, na, 1.0, Ag, na, CR, na, 0.0, T, na);
        oracle_dgemm('T', 'N', nm, nm, na, 1.0, CL, na, T, na, 0.0, O, nm);
        for (int64_t j = 0; j < nm; j++)
            for (int64_t i = 0; i < nm; i++)
                out[gx + R * (i + nm * j)] = O[i + nm * j];
    }
    free(T);
    free(O);
    return 0;
}
