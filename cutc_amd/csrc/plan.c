/* Device-memory plan of the TC integral path (pure host function).
 *
 * Replaces the reference's chunk heuristic (src/build_int2_grad1_u12_ao.cu:
 * 63-99), whose 32-bit arithmetic overflows at n_ao = 200 and whose r2 chunk
 * does not match its own memory formula.  Here every size is 64-bit, all r2
 * points stay in one K sweep, and only r1 is chunked:
 *   FP64 stage 1: pairs table + Apk (n_pairs x ldk) + G chunk (4 x n1c x ldk)
 *                 [+ f2 table for > 32 nuclei]
 *   INT8 stage 1: pairs + row maxima + pair residues (nmod x n_pairs x ldk
 *                 bytes, fragment-packed) + per r1 row: G (4 ldk doubles), its
 *                 residues (4 nmod ldk bytes) and the GEMM residues (nmod x 4
 *                 n_pairs bytes) + the block-sparse orderings and the
 *                 step lists of one K pass (sparse_bytes)
 *   stage 2:      Hp and NHp for an r1 chunk (4 planes x roundup(n1c2,16) x
 *                 n_pairs doubles) [INT8: + both operands' packed residues,
 *                 2 x nmod x 3 x n_pairs bytes per r1 row, + the GEMM residues
 *                 nmod x n_pairs^2 bytes, + the block-sparse orderings and
 *                 step lists]
 * Chunks are balanced (equal sizes rounded up to a multiple of 64).  The
 * modular precisions may also cut K (r2) into passes of k_chunk points whose
 * residue products are added modulo each modulus, when the pair planes of the
 * whole K leave room for too few r1 rows. */
#include <stdint.h>

#include "cutc_int.h"
#include "kernels/launch.h"

static int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
static int64_t rup(int64_t a, int64_t b) { return cdiv(a, b) * b; }

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

/* Block-sparse stage 1 (one K pass of a modular precision): the permuted r2,
 * wr2 and AO values, the sort keys and values, the natural-order pair
 * exponents, the occupancy flags, step lists and counts of every 256-pair
 * tile, each region rounded up to 256 bytes. */
static int64_t sparse_bytes(int64_t n2, int n_ao, int64_t P, int64_t ldk, int64_t kc, int kind)
{
    const int64_t ns = cutc_sort_len(n2);
    const int64_t tiles = CUTC_PK_ROWS_B(P) / CUTC_SP_ROWS, kt = ldk / CUTC_CRT_KSTEP(kind);
    return 8 * ((int64_t)n_ao + 4) * n2 + 12 * ns + 4 * P + 5 * tiles * (kc / CUTC_CRT_KSTEP(kind)) + 4 * tiles +
           (int64_t)cutc_cluster_bytes(P, kt) + 11 * 256;
}

int cutc_make_plan(int64_t n1, int64_t n2, int n_ao, int n_nuc, int precision, int slices, size_t budget,
                   int64_t n1c_req, cutc_plan *pl)
{
    if (n1 < 0 || n2 < 0 || n_ao < 0 || n_nuc < 0 || slices < 0 || !pl)
        return CUTC_ERR_ARG;
    if (slices == 0)
        slices = cutc_default_slices(precision);
    const int crt = precision == CUTC_PREC_INT8 || precision == CUTC_PREC_BF16_CRT || precision == CUTC_PREC_FP16_CRT;
    const int64_t rb = precision == CUTC_PREC_INT8 ? 1 : 2; /* bytes per residue-plane element */
    const int slice = precision == CUTC_PREC_BF16 || precision == CUTC_PREC_FP16;
    const int64_t nmod = (crt || slice) ? slices : 0;
    const int64_t P = (int64_t)n_ao * (n_ao + 1) / 2;
    const int64_t ldk = rup(n2, (crt || slice) ? CUTC_PK_KALIGN : 64);
    pl->n_pairs = P;
    pl->ldk = ldk;

    /* stage 1 */
    int64_t fixed1, per_r1;
    int64_t kc = ldk;
    if (crt) {
        /* the residue planes of one K pass of kc points; more passes when the
         * pair planes leave room for fewer than min(n1, 1024) r1 rows */
        const int64_t want = n1 < 1024 ? (n1 > 0 ? n1 : 1) : 1024;
        for (int64_t nk = 1; nk <= 64; nk++) {
            kc = rup(cdiv(ldk, nk), CUTC_PK_KALIGN);
            fixed1 = 16 * P + rb * nmod * CUTC_PK_ROWS_B(P) * kc + rb * nmod * 511 * kc + 4096;
            /* the block-sparse contraction's orderings, and the flags and lists of one pass */
            fixed1 += sparse_bytes(n2, n_ao, P, ldk, kc, precision == CUTC_PREC_INT8 ? CUTC_CRT_I8 : CUTC_CRT_BF16);
            per_r1 = 32 * ldk + 32 + 4 * rb * nmod * kc + nmod * 4 * P + 32 * cutc_jast_norm_blocks(ldk);
            if (n_nuc > 32)
                fixed1 += 8 * (int64_t)n_nuc * ldk;
            if ((int64_t)budget >= fixed1 + per_r1 * want || kc == CUTC_PK_KALIGN)
                break;
        }
        if (n_nuc > 32)
            fixed1 -= 8 * (int64_t)n_nuc * ldk; /* added once below */
    } else if (slice) { /* digit planes of 2-byte elements, FP64 results */
        fixed1 = 16 * P + 2 * nmod * CUTC_OZ_ROWS(P) * ldk + 2 * nmod * 127 * ldk + 4096;
        per_r1 = 32 * ldk + 32 + 8 * nmod * ldk + 32 * P;
    } else {
        fixed1 = 8 * P + 8 * P * ldk + 4096;
        per_r1 = 32 * ldk;
    }
    if (n_nuc > 32)
        fixed1 += 8 * (int64_t)n_nuc * ldk;
    int64_t n1c = 0;
    if (n1 > 0 && P > 0 && n2 > 0) {
        if ((int64_t)budget <= fixed1 + per_r1)
            return CUTC_ERR_NOMEM;
        int64_t cmax = ((int64_t)budget - fixed1) / per_r1;
        if (n1c_req > 0 && n1c_req < cmax)
            cmax = n1c_req;
        const int64_t align = crt ? 128 : 64; /* int8: 4 n1c fills whole 512-row tiles */
        n1c = balanced(n1, cmax, cmax >= align ? align : 1);
    } else if (n1 > 0) {
        n1c = n1;
    }
    pl->n1_chunk = n1c;
    pl->bytes_stage1 = (size_t)(fixed1 + per_r1 * n1c);
    pl->k_chunk = kc;

    /* stage 2 */
    int64_t fixed2 = 8 * P + 4096, per_r1b = 4 * 8 * P;
    if (crt) {
        /* + the block-sparse state: the point sort (up to 24 bytes per r1
         * row), the pair sort and H pair table, flags and step lists of every
         * (256-pair H tile, 32-point step) over the 3 K segments */
        const int64_t tiles = CUTC_PK_ROWS_A(P) / CUTC_SP_ROWS;
        const int64_t ks = rb == 1 ? 64 : 32; /* points per K step of the lists */
        fixed2 += 16 * P + 8 * CUTC_CRT_KSPLIT * P + nmod * P * P +
                  2 * rb * nmod * CUTC_PK_ROWS_A(P) * 3 * (CUTC_PK_KALIGN - 1) + 8 * P +
                  (int64_t)cutc_cluster_bytes(P, (CUTC_PK_KALIGN + 63 * ks) / ks) + 48 +
                  4 * tiles + 5 * tiles * 3 * CUTC_PK_KALIGN / 32 + 14 * 256;
        per_r1b += 2 * rb * nmod * 3 * CUTC_PK_ROWS_A(P) + 24 + (5 * 3 * tiles + 31) / 32 + (8 * P + 64 * ks - 1) / (64 * ks);
    } else if (slice) {
        fixed2 += 16 * P + 8 * CUTC_CRT_KSPLIT * P + 8 * P * P +
                  2 * 2 * nmod * CUTC_OZ_ROWS(P) * 3 * (CUTC_PK_KALIGN - 1);
        per_r1b += 2 * 2 * nmod * 3 * CUTC_OZ_ROWS(P);
    }
    int64_t n1c2 = 0;
    if (n1 > 0 && P > 0) {
        if ((int64_t)budget <= fixed2 + 144 * per_r1b)
            return CUTC_ERR_NOMEM;
        int64_t cmax = ((int64_t)budget - fixed2) / per_r1b - 144;
        n1c2 = balanced(n1, cmax, cmax >= 64 ? 64 : 1);
    }
    pl->n1_chunk_2e = n1c2;
    pl->bytes_stage2 = (size_t)(fixed2 + per_r1b * (rup(n1c2, 16) + 128));
    return CUTC_OK;
}
