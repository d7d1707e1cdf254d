"""The 20-nucleus workload (n_ao = 200, all 200000 r2 points) on the GPU
against the FP64 oracle on 128 r1 points built in 3 r1 chunks (the
reference's chunk loop, src/build_int2_grad1_u12_ao.cu:113-331), through
cutc_int_device: FP64 (<= 1e-12), int8 with the library default of 11
moduli (the benchmark's timed build) and with 10, bf16 / fp16 modular with
10 (<= 1e-10), bf16 / fp16 digit slices with the stated bounds of the slice
scheme: 3 slices <= 1e-6, 4 slices <= 5e-9.  The n_ao^4 arrays are compared
on the device (torch as plumbing).  Errors are relative Frobenius norms per
X component and for int_2e_ao, and the relative error of
sum D D (kl|ab) for a seeded rank-10 PSD D."""
import numpy as np
import pytest

from cutc_amd import cutc, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12
TOL_SPLIT = 1e-10
TOL_SLICES = {3: 1e-6, 4: 5e-9}


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need the MI355X"
    return t


def rel_dev(torch, a, b):
    """relative Frobenius error of device tensor a against device tensor b"""
    nb = float(torch.linalg.vector_norm(b))
    return float(torch.linalg.vector_norm(a - b)) / (nb if nb > 0 else 1.0)


def density(n_ao, seed=20261205):
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((n_ao, 10))
    return C @ C.T


# ------------------------------------------------- 20 nuclei, n_ao = 200

@pytest.fixture(scope="module")
def cube20(torch):
    s = synth.cube20(n1=128, n2=200000)
    print(f"cube20 subset {s.shape_str()}: oracle", flush=True)
    Xo = oracle.int2_grad1_u12_ao(s)
    Io = oracle.int_2e_ao(s, Xo)
    D = density(s.n_ao)
    dd = float(D.reshape(-1) @ (Io.reshape(s.n_ao ** 2, -1) @ D.reshape(-1)))
    ds = cutc.DeviceSystem.upload(s)
    X, I = cutc.alloc_outputs(ds)
    ref = {"X": torch.from_numpy(Xo).cuda(), "I": torch.from_numpy(Io).cuda(), "dd": dd,
           "D": torch.from_numpy(D.reshape(-1)).cuda()}
    del Io, Xo
    yield s, ds, X, I, ref
    del X, I, ref
    cutc.release_workspace()
    torch.cuda.empty_cache()


def build_cube20(torch, cube20, precision, slices=None):
    s, ds, X, I, ref = cube20
    X.fill_(float("nan"))
    I.fill_(float("nan"))
    st = cutc.cutc_int_device(ds, X, I, cutc.options(precision, slices, n1_chunk=48))
    torch.cuda.synchronize()
    cutc.release_workspace()
    n = s.n_ao ** 2
    err = {f"X{m}": rel_dev(torch, X[m], ref["X"][m]) for m in range(4)}
    err["int_2e_ao"] = rel_dev(torch, I, ref["I"])
    dd = float(ref["D"] @ (I.reshape(n, n) @ ref["D"]))
    err["dd"] = abs(dd - ref["dd"]) / abs(ref["dd"])
    print(f"  {precision}x{slices}: " + " ".join(f"{k} {v:.2e}" for k, v in err.items())
          + f"  ({st['n1_chunks']} r1 chunks)", flush=True)
    return err, st


@pytest.mark.timeout(900)
def test_cube20_fp64_three_chunks(torch, cube20):
    err, st = build_cube20(torch, cube20, "fp64")
    assert st["n1_chunks"] >= 3
    assert max(err.values()) <= TOL_FP64, err


@pytest.mark.parametrize("precision,slices", [("int8", None), ("int8", 10), ("bf16crt", None), ("fp16crt", None)])
def test_cube20_modular(torch, cube20, precision, slices):
    """int8 at the library default (11 moduli, the benchmark's timed build)
    and at 10; bf16 / fp16 modular at their default of 10."""
    err, st = build_cube20(torch, cube20, precision, slices)
    assert st["n1_chunks"] >= 3
    assert max(err.values()) <= TOL_SPLIT, err


def test_cube20_fp16_modular_equals_bf16(torch, cube20):
    """moduli <= 255 give residues |r| <= 127, exact in bf16 and in fp16,
    and the FP32 sums are exact: the two formats give identical bits"""
    s, ds, X, I, ref = cube20
    build_cube20(torch, cube20, "bf16crt")
    Xb, Ib = X.clone(), I.clone()
    build_cube20(torch, cube20, "fp16crt")
    assert torch.equal(X, Xb) and torch.equal(I, Ib)
    del Xb, Ib


@pytest.mark.parametrize("precision", ["bf16", "fp16"])
@pytest.mark.parametrize("slices", [3, 4])
def test_cube20_slices(torch, cube20, precision, slices):
    err, _ = build_cube20(torch, cube20, precision, slices)
    assert max(err.values()) <= TOL_SLICES[slices], err
