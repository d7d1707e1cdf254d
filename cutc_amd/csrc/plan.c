/* Device-memory plan of the TC integral path (pure host function).
 *
 * Replaces the reference's chunk heuristic (src/build_int2_grad1_u12_ao.cu:
 * 63-99), whose 32-bit arithmetic overflows at n_ao = 200 and whose r2 chunk
 * does not match its own memory formula.  Here every size is 64-bit, all r2
 * points stay in one K sweep, and only r1 is chunked:
 *   stage 1: pairs table + Apk (n_pairs x ldk) + G chunk (4 x n1c x ldk)
 *            [+ f2 table for > 32 nuclei] [+ operand slices in split modes]
 *   stage 2: Hp and NHp for an r1 chunk (4 planes x roundup(n1c2,16) x n_pairs)
 * Chunks are balanced (equal sizes rounded up to a multiple of 64). */
#include <stdint.h>

#include "cutc_int.h"

static int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
static int64_t rup(int64_t a, int64_t b) { return cdiv(a, b) * b; }

/* Bytes of the split-mode operand copies per element of Apk and of G. */
static int64_t split_bytes_per_elem(int precision, int slices)
{
    if (precision == CUTC_PREC_FP64)
        return 0;
    return 2 * (int64_t)slices; /* 16-bit slices */
}

static int64_t balanced(int64_t n, int64_t cmax, int64_t align)
{
    if (cmax >= n)
        return n;
    int64_t nch = cdiv(n, cmax);
    int64_t c = rup(cdiv(n, nch), align);
    if (c > cmax)
        c = cmax / align * align;
    return c;
}

int cutc_make_plan(int64_t n1, int64_t n2, int n_ao, int n_nuc, int precision, int slices, size_t budget,
                   int64_t n1c_req, cutc_plan *pl)
{
    if (n1 < 0 || n2 < 0 || n_ao < 0 || n_nuc < 0 || !pl)
        return CUTC_ERR_ARG;
    const int64_t P = (int64_t)n_ao * (n_ao + 1) / 2;
    const int64_t ldk = rup(n2, 64);
    const int64_t sb = split_bytes_per_elem(precision, slices);
    pl->n_pairs = P;
    pl->ldk = ldk;

    /* stage 1 */
    int64_t fixed1 = 8 * P + 8 * P * ldk + sb * P * ldk + 4096;
    if (n_nuc > 32)
        fixed1 += 8 * (int64_t)n_nuc * ldk;
    int64_t per_r1 = (32 + 4 * sb) * ldk; /* G rows + their slices */
    int64_t n1c = 0;
    if (n1 > 0) {
        if ((int64_t)budget <= fixed1)
            return CUTC_ERR_NOMEM;
        int64_t cmax = per_r1 > 0 ? ((int64_t)budget - fixed1) / per_r1 : n1;
        if (cmax < 1)
            return CUTC_ERR_NOMEM;
        if (n1c_req > 0 && n1c_req < cmax)
            cmax = n1c_req;
        n1c = balanced(n1, cmax, cmax >= 64 ? 64 : 1);
    }
    pl->n1_chunk = n1c;
    pl->bytes_stage1 = (size_t)(fixed1 + per_r1 * n1c);

    /* stage 2 */
    int64_t fixed2 = 8 * P + 4096;
    int64_t per_r1b = 4 * 8 * P; /* Hp + 3 NHp planes */
    int64_t n1c2 = 0;
    if (n1 > 0 && P > 0) {
        if ((int64_t)budget <= fixed2 + 16 * per_r1b)
            return CUTC_ERR_NOMEM;
        int64_t cmax = ((int64_t)budget - fixed2) / per_r1b - 16;
        n1c2 = balanced(n1, cmax, cmax >= 64 ? 64 : 1);
    }
    pl->n1_chunk_2e = n1c2;
    pl->bytes_stage2 = (size_t)(fixed2 + per_r1b * rup(n1c2, 16));
    return CUTC_OK;
}
