/*
 * AOs on a grid (cutc_aos): host driver.  CuTC's callers evaluate
 * aos_data1 / aos_data2 on the host and pass them to cutc_int, which copies
 * them to the device (src/cutc_int.c:96-99, 1.67 GB for 20 nuclei,
 * n_ao = 200, 60k + 200k points).  cutc_aos_device writes them straight into
 * HBM from the basis description (kernels/ao.hip).
 */
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "cutc_ao.h"
#include "internal.h"
#include "kernels/launch.h"
#include "xfer.h"

static int hrc(hipError_t e)
{
    if (e == hipSuccess)
        return CUTC_OK;
    return e == hipErrorOutOfMemory ? CUTC_ERR_NOMEM : CUTC_ERR_HIP;
}

/* nucleus numbers in 1..n_nuc, powers >= 0, primitive counts in 0..prim_max */
static int check_tables(int n_nuc, int n_ao, int prim_max, const int *nucl, const int *power, const int *nprim)
{
    for (int i = 0; i < n_ao; i++) {
        if (nucl[i] < 1 || nucl[i] > n_nuc || nprim[i] < 0 || nprim[i] > prim_max)
            return CUTC_ERR_ARG;
        for (int d = 0; d < 3; d++)
            if (power[i + (int64_t)n_ao * d] < 0)
                return CUTC_ERR_ARG;
    }
    return CUTC_OK;
}

static int aos_path(int n_grid, int n_nuc, int n_ao, int prim_max, const double *r, const double *rn,
                    const int *nucl, const int *power, const int *nprim, const double *expo, const double *coef,
                    double *out, hipStream_t st)
{
    if (n_grid < 0 || n_nuc < 0 || n_ao < 0 || prim_max < 0 || n_ao > 65535)
        return CUTC_ERR_ARG;
    if (n_grid == 0 || n_ao == 0)
        return CUTC_OK;
    if (!r || !nucl || !power || !nprim || !out || n_nuc == 0 || !rn || (prim_max > 0 && (!expo || !coef)))
        return CUTC_ERR_ARG;
    return cutc_launch_aos(n_grid, r, rn, n_nuc, n_ao, nucl, power, nprim, prim_max, expo, coef, out, st);
}

int cutc_aos_device(int n_grid, int n_nuc, int n_ao, int prim_max, const double *r, const double *rn,
                    const int *ao_nucl, const int *ao_power, const int *ao_prim_num, const double *ao_expo,
                    const double *ao_coef, double *aos_data, void *stream)
{
    int rc = aos_path(n_grid, n_nuc, n_ao, prim_max, r, rn, ao_nucl, ao_power, ao_prim_num, ao_expo, ao_coef,
                      aos_data, (hipStream_t)stream);
    cutc_set_last_error(rc);
    return rc;
}

/* Host arrays: check the tables, upload, evaluate, download (through the
 * transfer pipeline), free. */
int cutc_aos(int n_grid, int n_nuc, int n_ao, int prim_max, double *h_r, double *h_rn, int *h_ao_nucl,
             int *h_ao_power, int *h_ao_prim_num, double *h_ao_expo, double *h_ao_coef, double *h_aos_data)
{
    int rc = CUTC_OK, locked = 0;
    void *d[8] = {NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL};
    cutc_xfer *xf = NULL;
    if (n_grid < 0 || n_nuc < 0 || n_ao < 0 || prim_max < 0 || n_ao > 65535) {
        rc = CUTC_ERR_ARG;
        goto out;
    }
    if (n_grid == 0 || n_ao == 0)
        goto out;
    if (!h_r || !h_rn || !h_ao_nucl || !h_ao_power || !h_ao_prim_num || !h_aos_data || n_nuc == 0 ||
        (prim_max > 0 && (!h_ao_expo || !h_ao_coef))) {
        rc = CUTC_ERR_ARG;
        goto out;
    }
    if ((rc = check_tables(n_nuc, n_ao, prim_max, h_ao_nucl, h_ao_power, h_ao_prim_num)) != CUTC_OK)
        goto out;
    {
        const size_t ng = (size_t)n_grid, na = (size_t)n_ao, pm = (size_t)prim_max;
        const size_t sz[8] = {8 * 3 * ng, 8 * 3 * (size_t)n_nuc, 4 * na, 4 * 3 * na, 4 * na, 8 * pm * na,
                              8 * pm * na, 8 * ng * na * 4};
        const void *hin[7] = {h_r, h_rn, h_ao_nucl, h_ao_power, h_ao_prim_num, h_ao_expo, h_ao_coef};
        /* the library lock from the first allocation to the last free */
        cutc_internal_lock();
        locked = 1;
        for (int i = 0; i < 8; i++)
            if (sz[i] > 0 && (rc = hrc(hipMalloc(&d[i], sz[i]))) != CUTC_OK)
                goto out;
        for (int i = 0; i < 7; i++)
            if (sz[i] > 0 && (rc = hrc(hipMemcpy(d[i], hin[i], sz[i], hipMemcpyHostToDevice))) != CUTC_OK)
                goto out;
        rc = cutc_launch_aos(n_grid, (const double *)d[0], (const double *)d[1], n_nuc, n_ao, (const int *)d[2],
                             (const int *)d[3], (const int *)d[4], prim_max, (const double *)d[5],
                             (const double *)d[6], (double *)d[7], NULL);
        if (rc == CUTC_OK) { /* the output through the transfer pipeline */
            rc = cutc_xfer_open(&xf, (size_t)64 << 20, 4, cutc_internal_copy_threads());
            if (rc == CUTC_OK)
                rc = cutc_xfer_d2h(xf, h_aos_data, d[7], sz[7], NULL);
            if (rc == CUTC_OK)
                rc = cutc_xfer_finish(xf, NULL);
        }
    }
out:
    cutc_xfer_close(xf);
    for (int i = 0; i < 8; i++)
        if (d[i])
            hipFree(d[i]);
    if (locked)
        cutc_internal_unlock();
    cutc_set_last_error(rc);
    return rc;
}
