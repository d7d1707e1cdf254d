"""Split-precision (modular int8 / CRT) contractions against the FP64 oracle
and the FP64 GPU path.  Bar: relative Frobenius error <= 1e-10 on
int2_grad1_u12_ao and int_2e_ao with the default 12 moduli."""
import numpy as np
import pytest

from cutc_amd import cutc, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-10


def relerr(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


def run(torch, s, **opt):
    ds = cutc.DeviceSystem.upload(s)
    X, I = cutc.alloc_outputs(ds)
    st = cutc.cutc_int_device(ds, X, I, cutc.options(**opt))
    torch.cuda.synchronize()
    return X.cpu().numpy(), I.cpu().numpy(), st


@pytest.mark.parametrize("shape", [(1, 1, 1, 1), (2, 5, 3, 7), (3, 24, 37, 129), (7, 61, 77, 2049),
                                   (33, 40, 65, 333), (4, 30, 300, 5000)])
def test_int8_matches_oracle_small(torch, shape):
    n_nuc, n_ao, n1, n2 = shape
    print(f"int8 shape n_nuc={n_nuc} n_ao={n_ao} n1={n1} n2={n2}", flush=True)
    s = synth.make_system(n_nuc, n_ao, n1, n2, 300 + n_ao, name="int8")
    X, I, _ = run(torch, s, precision="int8")
    Xo = oracle.int2_grad1_u12_ao(s)
    Io = oracle.int_2e_ao(s, Xo)
    ex, ei = relerr(X, Xo), relerr(I, Io)
    print(f"  relerr X {ex:.3e} int_2e_ao {ei:.3e}", flush=True)
    assert ex <= TOL and ei <= TOL


def test_int8_triatomic_and_chunking(torch):
    s = synth.triatomic()
    X, I, st = run(torch, s, precision="int8")
    Xo = oracle.int2_grad1_u12_ao(s)
    Io = oracle.int_2e_ao(s, Xo)
    ex, ei = relerr(X, Xo), relerr(I, Io)
    print(f"triatomic int8: relerr X {ex:.3e} int_2e_ao {ei:.3e}", flush=True)
    assert ex <= TOL and ei <= TOL
    Xc, Ic, stc = run(torch, s, precision="int8", n1_chunk=640)
    assert stc["n1_chunks"] > 1
    assert relerr(Xc, X) <= 1e-13 and relerr(Ic, I) <= 1e-12


def test_int8_fewer_moduli_degrade_gracefully(torch):
    s = synth.triatomic(n1=200, n2=3000)
    Xo = oracle.int2_grad1_u12_ao(s)
    errs = []
    for n in (8, 10, 12, 14):
        X, I, _ = run(torch, s, precision="int8", slices=n)
        errs.append(relerr(X, Xo))
    print("moduli 8/10/12/14 relerr X:", [f"{e:.2e}" for e in errs], flush=True)
    assert errs[0] < 1e-4 and errs[2] <= TOL
    assert errs[0] > errs[2]


def test_int8_cube20_full_basis_subset(torch):
    """20 nuclei, n_ao = 200, all 200000 r2 points, 24 r1 points."""
    s = synth.cube20(n1=24, n2=200000)
    X, I, st = run(torch, s, precision="int8")
    Xf, If, _ = run(torch, s)
    Xo = oracle.int2_grad1_u12_ao(s)
    ex, exf = relerr(X, Xo), relerr(X, Xf)
    print(f"cube20 r1<24 int8: relerr X vs oracle {ex:.3e} vs FP64 GPU {exf:.3e}", flush=True)
    assert ex <= TOL
    Io = oracle.int_2e_ao(s, Xo)
    ei, eif = relerr(I, Io), relerr(I, If)
    print(f"cube20 r1<24 int8: relerr int_2e_ao vs oracle {ei:.3e} vs FP64 GPU {eif:.3e}", flush=True)
    assert ei <= TOL
