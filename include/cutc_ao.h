/*
 * Atomic orbitals on a grid: MI355X (gfx950) library, the producer of
 * cutc_int's aos_data1 / aos_data2 inputs.  Host code is C; the kernel is HIP.
 *
 * CuTC has no such function: its callers evaluate the AOs and pass
 * aos_data(n_grid, n_ao, 4) -- the value, then d/dx, d/dy, d/dz -- to
 * cutc_int (src/cutc_module.f90:28-29; components 1..3 are read as the
 * gradient by src/short_range_integ_nonherm.cu:45-55).  Here the AOs are
 * contracted Cartesian Gaussians, described by the per-AO arrays a quantum
 * chemistry code holds (Fortran order, as cutc_int's other arguments):
 *
 *   ao_nucl(n_ao)               centre of AO i: nucleus number, 1-based
 *   ao_power(n_ao, 3)           Cartesian powers (a, b, c) of AO i
 *   ao_prim_num(n_ao)           primitives of AO i (0..prim_max)
 *   ao_expo(prim_max, n_ao)     exponents alpha_p
 *   ao_coef(prim_max, n_ao)     contraction coefficients, normalisation
 *                               included
 *
 *   chi_i(r) = x^a y^b z^c sum_p c_p exp(-alpha_p |r - R|^2),
 *   (x, y, z) = r - R, R = rn(:, ao_nucl(i)).
 *
 * r(3, n_grid) and rn(3, n_nuc) are xyz-interleaved like cutc_int's r1, r2
 * and rn.  Output aos_data[g + n_grid (i + n_ao c)], c = 0 value, 1..3 the
 * gradient.
 */
#ifndef CUTC_AO_H
#define CUTC_AO_H

#include <stddef.h>
#include <stdint.h>

#include "cutc_int.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Host arrays in and out.  Returns 0, CUTC_ERR_ARG (-2) for negative sizes,
 * a nucleus number outside 1..n_nuc, a negative power or a primitive count
 * outside 0..prim_max, CUTC_ERR_NOMEM (-3), CUTC_ERR_HIP (-4). */
int cutc_aos(int n_grid, int n_nuc, int n_ao, int prim_max, double *h_r, double *h_rn, int *h_ao_nucl,
             int *h_ao_power, int *h_ao_prim_num, double *h_ao_expo, double *h_ao_coef, double *h_aos_data);

/* The same on device pointers, queued on a caller stream (NULL: the null
 * stream); no host synchronisation.  The tables are not read on the host:
 * an AO whose centre, powers or primitive count is out of range gets NaN in
 * its four rows (nothing is read through the bad entry).  Returns 0, or
 * CUTC_ERR_ARG for negative sizes, n_ao > 65535 or missing pointers. */
int cutc_aos_device(int n_grid, int n_nuc, int n_ao, int prim_max, const double *r, const double *rn,
                    const int *ao_nucl, const int *ao_power, const int *ao_prim_num, const double *ao_expo,
                    const double *ao_coef, double *aos_data, void *stream);

#ifdef __cplusplus
}
#endif
#endif
