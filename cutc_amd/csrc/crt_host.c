/*
 * Constants of the modular (CRT) split-precision contraction (host code).
 *
 * Scheme (Ozaki, Uchino & Imamura, "Ozaki Scheme II: a GEMM-oriented
 * emulation of floating-point matrix multiplication using an integer modular
 * technique"): each operand row x is scaled by 2^(bits - e), ||x||_2 < 2^e,
 * and rounded to integers q; the exact integer product V = sum_k a_k b_k is
 * obtained from its residues modulo n pairwise-coprime moduli m_i <= 256,
 * each residue product being an int8 x int8 -> int32 GEMM (or, with moduli
 * <= 255, a bf16 / fp16 GEMM with exact FP32 sums); V is rebuilt by
 * the Chinese remainder theorem and scaled back to FP64.  Correct
 * reconstruction needs |V| < M/2 with M = prod m_i.  Rounding moves a row by
 * at most sqrt(K)/2 in 2-norm, so ||q||_2 < 2^bits + sqrt(K)/2 and, by
 * Cauchy-Schwarz (the scheme's "fast mode" bound), |V| < (2^bits +
 * sqrt(K)/2)^2; `bits` is the largest value (at most 50, which keeps every
 * |q| < 2^51 exact in FP64) for which that square stays below M/2.
 *
 * CRT: with M_i = M / m_i and y_i = M_i^-1 mod m_i,
 *   V == S = sum_i u_i M_i  (mod M),  u_i = (c_i y_i) mod m_i in [0, m_i),
 * and V = S - M round(S / M).  S < n M < 2^128 is carried as three
 * 40-bit integer parts (hi 2^80 + mid 2^40 + lo) so that every partial sum is
 * exact in FP64.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "cutc_int.h"
#include "kernels/launch.h"

/* int8 residues: pairwise-coprime moduli <= 256 (largest first) */
static const int k_moduli[CUTC_CRT_MAX] = {256, 255, 253, 251, 247, 241, 239, 233,
                                           229, 227, 223, 217, 211, 199, 197, 193};
/* bf16 / fp16 residues (|r| <= 127, which keeps 1024 FP32 products exact):
 * the int8 set without 256 */
static const int k_moduli16[CUTC_CRT_MAX] = {255, 253, 251, 247, 241, 239, 233, 229,
                                             227, 223, 217, 211, 199, 197, 193, 191};

typedef unsigned __int128 u128;

static int inv_mod(int a, int m) /* a^-1 mod m, gcd(a, m) = 1 */
{
    int t = 0, nt = 1, r = m, nr = ((a % m) + m) % m;
    while (nr) {
        int q = r / nr, tmp;
        tmp = t - q * nt;
        t = nt;
        nt = tmp;
        tmp = r - q * nr;
        r = nr;
        nr = tmp;
    }
    return r == 1 ? (t < 0 ? t + m : t) : -1;
}

static void split3(u128 v, double out[3])
{
    const u128 mask = ((u128)1 << 40) - 1;
    out[2] = (double)(uint64_t)(v & mask);
    out[1] = (double)(uint64_t)((v >> 40) & mask);
    out[0] = (double)(uint64_t)(v >> 80);
}

int cutc_crt_setup(int64_t K, int n, int kind, cutc_crt *c)
{
    memset(c, 0, sizeof *c);
    if (n < 2 || n > CUTC_CRT_MAX || K < 1 || kind < CUTC_CRT_I8 || kind > CUTC_CRT_FP16)
        return CUTC_ERR_ARG;
    const int *mods = kind == CUTC_CRT_I8 ? k_moduli : k_moduli16;
    u128 M = 1;
    double log2M = 0.0;
    for (int i = 0; i < n; i++)
        log2M += log2((double)mods[i]);
    if (log2M > 127.0) /* M and the sums below must fit 128 bits */
        return CUTC_ERR_ARG;
    for (int i = 0; i < n; i++)
        M *= (u128)mods[i];
    /* largest bits with 2 log2(2^bits + sqrt(K)/2) < log2(M) - 1, with a
     * margin far above the rounding of log2M */
    const double half_rootk = 0.5 * sqrt((double)K);
    int bits = 50;
    while (bits >= 8 && 2.0 * log2(ldexp(1.0, bits) + half_rootk) >= log2M - 1.0 - 1e-9)
        bits--;
    if (bits < 8)
        return CUTC_ERR_ARG;
    c->n = n;
    c->bits = bits;
    c->kind = kind;
    for (int i = 0; i < n; i++) {
        const int m = mods[i];
        const u128 Mi = M / (u128)m;
        if (Mi >> 120) /* three 40-bit parts */
            return CUTC_ERR_ARG;
        c->m[i] = m;
        c->inv_m[i] = 1.0 / (double)m;
        c->y[i] = inv_mod((int)(Mi % (u128)m), m);
        if (c->y[i] < 0)
            return CUTC_ERR_ARG;
        split3(Mi, c->Mi[i]);
    }
    split3(M, c->M);
    c->Minv = 1.0 / ((double)c->M[0] * 1208925819614629174706176.0 + (double)c->M[1] * 1099511627776.0 +
                     (double)c->M[2]);
    return CUTC_OK;
}
