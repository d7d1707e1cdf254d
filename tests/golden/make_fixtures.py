"""Writes tests/golden/chain_tiny.npz: one tiny seeded system and the FP64
oracle's answers for every function on the path, so that a later change to the
oracle (or to a compiler, BLAS or thread count under it) that moves a result
shows up against data fixed once.

The reference ships no golden vectors (its only known answer is the all-zero
demo, src/cutc_int_f.f90:41-55 with src/jast_bh.cu:203), and it cannot run
here (CUDA + cuBLAS), so the expected outputs come from the oracle, which
tests/test_oracle*.py pin by closed forms, finite differences, symmetry and
brute force.

The chain follows what a caller of the reference does:
  aos_data1 / aos_data2 from the basis            (oracle_ao.c)
  cutc_int -> int2_grad1_u12_ao[0:3], int_2e_ao   (oracle.c; src/cutc_int.c:15-20)
  AO -> MO with left/right coefficients            (oracle_mo.c)
  cutc_no -> no_2e, no_1e, no_0e, open shell (ne_a != ne_b) and closed
  shell (ne_a == ne_b)                             (oracle_no.c; src/cutc_no.c:14-778)

Run from the repository root:  python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from cutc_amd import synth  # noqa: E402
from oracle import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "chain_tiny.npz")

# two nuclei, an s and a p shell on each: n_ao = 8
N_NUC, N_AO, N1, N2, SEED = 2, 8, 16, 48, 20260301
NUCLEI = [[0.0, 0.0, -0.7], [0.0, 0.0, 0.7]]
SHELLS = [(0, 0), (0, 1), (1, 0), (1, 1)]
N_MO, NE_A, NE_B = 6, 3, 2


def main():
    s = synth.make_system(N_NUC, N_AO, N1, N2, SEED, nuclei=NUCLEI, shells=SHELLS, name="golden-tiny")
    b = s.basis
    aos1 = oracle.aos(s.r1, s.rn, b["ao_nucl"], b["ao_power"], b["ao_prim_num"], b["ao_expo"], b["ao_coef"])
    aos2 = oracle.aos(s.r2, s.rn, b["ao_nucl"], b["ao_power"], b["ao_prim_num"], b["ao_expo"], b["ao_coef"])
    s.aos1, s.aos2 = aos1, aos2
    X, I = oracle.cutc_int(s)

    rng = np.random.default_rng(SEED)
    mo_l = rng.standard_normal((N_MO, N_AO)) / np.sqrt(N_AO)
    mo_r = mo_l + 0.1 * rng.standard_normal((N_MO, N_AO)) / np.sqrt(N_AO)
    G = oracle.ao_to_mo(mo_l, mo_r, X)
    mos_l = rng.standard_normal((N_MO, N1))
    mos_r = rng.standard_normal((N_MO, N1))
    no_2e, no_1e, no_0e = oracle.cutc_no(NE_A, NE_B, s.wr1, mos_l, mos_r, G)
    cs_2e, cs_1e, cs_0e = oracle.cutc_no(NE_A, NE_A, s.wr1, mos_l, mos_r, G)

    np.savez_compressed(
        OUT,
        # inputs (numpy C order of the Fortran layouts)
        rn=s.rn, r1=s.r1, wr1=s.wr1, r2=s.r2, wr2=s.wr2, c_bh=s.c_bh, m_bh=s.m_bh, n_bh=s.n_bh, o_bh=s.o_bh,
        ao_nucl=b["ao_nucl"], ao_power=b["ao_power"], ao_prim_num=b["ao_prim_num"], ao_expo=b["ao_expo"],
        ao_coef=b["ao_coef"], mo_l_coef=mo_l, mo_r_coef=mo_r, mos_l=mos_l, mos_r=mos_r,
        ne=np.array([NE_A, NE_B], np.int32),
        # expected outputs
        aos1=aos1, aos2=aos2, int2_grad1_u12_ao=X, int_2e_ao=I, int2_grad1_u12=G, no_2e=no_2e, no_1e=no_1e,
        no_0e=np.array([no_0e]), cs_no_2e=cs_2e, cs_no_1e=cs_1e, cs_no_0e=np.array([cs_0e]))
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes): {s.shape_str()}, n_mo={N_MO}")


if __name__ == "__main__":
    main()
