/*
 * FP64 CPU ORACLE for the AOs on a grid (cutc_aos) -- TEST INFRASTRUCTURE.
 * Only tests, the smoke check and the CPU-baseline leg of the benchmark load
 * this file (through liboracle.so); the HIP library never does.
 *
 * CuTC takes aos_data(n_grid, n_ao, 4) from its caller (value, d/dx, d/dy,
 * d/dz; src/cutc_module.f90:28-29, src/short_range_integ_nonherm.cu:45-55).
 * Contracted Cartesian Gaussians:
 *   chi_i(r) = x^a y^b z^c sum_p c_p exp(-alpha_p |r - R_i|^2)
 * and the gradient from the product rule, d/dx (x^a e^{-alpha r^2}) =
 * (a x^{a-1} - 2 alpha x^{a+1}) e^{-alpha r^2}.  Layouts (Fortran order):
 * r(3, n), rn(3, n_nuc), ao_nucl(n_ao) 1-based, ao_power(n_ao, 3),
 * ao_prim_num(n_ao), ao_expo / ao_coef(prim_max, n_ao), out(n, n_ao, 4).
 */
#include <math.h>
#include <stdint.h>

static double ipw(double x, int e)
{
    double r = 1.0;
    for (int k = 0; k < e; k++)
        r *= x;
    return r;
}

int oracle_aos(int64_t n, int n_nuc, int n_ao, int prim_max, const double *r, const double *rn, const int *nucl,
               const int *power, const int *nprim, const double *expo, const double *coef, double *out)
{
    if (n < 0 || n_nuc < 0 || n_ao < 0 || prim_max < 0)
        return -2;
    for (int i = 0; i < n_ao; i++) {
        if (nucl[i] < 1 || nucl[i] > n_nuc || nprim[i] < 0 || nprim[i] > prim_max)
            return -2;
        for (int d = 0; d < 3; d++)
            if (power[i + n_ao * d] < 0)
                return -2;
    }
    const int64_t ld = n * n_ao;
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < n; g++) {
        for (int i = 0; i < n_ao; i++) {
            const double *R = rn + 3 * (nucl[i] - 1);
            const double d[3] = {r[3 * g] - R[0], r[3 * g + 1] - R[1], r[3 * g + 2] - R[2]};
            const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            double val = 0.0, grad[3] = {0.0, 0.0, 0.0};
            for (int p = 0; p < nprim[i]; p++) {
                const double al = expo[p + (int64_t)prim_max * i];
                const double e = coef[p + (int64_t)prim_max * i] * exp(-al * r2);
                double f[3], df[3];
                for (int c = 0; c < 3; c++) {
                    const int k = power[i + n_ao * c];
                    f[c] = ipw(d[c], k);
                    df[c] = (k ? k * ipw(d[c], k - 1) : 0.0) - 2.0 * al * ipw(d[c], k + 1);
                }
                val += e * f[0] * f[1] * f[2];
                grad[0] += e * df[0] * f[1] * f[2];
                grad[1] += e * f[0] * df[1] * f[2];
                grad[2] += e * f[0] * f[1] * df[2];
            }
            const int64_t o = g + n * i;
            out[o] = val;
            for (int c = 0; c < 3; c++)
                out[o + (c + 1) * ld] = grad[c];
        }
    }
    return 0;
}
