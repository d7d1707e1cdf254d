"""Block-sparse modular contractions: the build that skips the all-zero
(256-pair tile, K step) tiles of the pair operand in stage 1 and of the H
operand in stage 2 equals the build that runs every step (options
dense=1), bit for bit, in int8, bf16 and fp16 modular arithmetic, while
running only a fraction of the tiles (stats k_density, k_density_2e).  The
orderings (r2 and each r1 chunk along a Morton curve, pair rows clustered by
their step supports) are applied in both builds, so X and int_2e_ao must
agree exactly; each is also checked against the FP64 build (<= 1e-10).

Systems: the 20-nucleus basis (n_ao = 200, 157 pair tiles) on 20000 r2
points, several tiles with empty or short lists and padded ones; a
3-nucleus system with one pair tile whose every step is occupied (the list
is the whole K); a system whose r2 count is not a multiple of the step; a
small basis (n_ao = 30, 2 x 2 tiles of 256 pairs) with 12000 r1 points, whose
stage-2 contraction runs its lists in K-split parts (4 for X^3 H, 12 for
X^{0..2} NH), each into its own residues, summed before the rebuild."""
import numpy as np
import pytest

from cutc_amd import cutc, synth

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need the MI355X"
    return t


def rel_dev(torch, a, b):
    nb = float(torch.linalg.vector_norm(b))
    return float(torch.linalg.vector_norm(a - b)) / (nb if nb > 0 else 1.0)


def build(torch, ds, precision, dense, n1_chunk=0):
    X, I = cutc.alloc_outputs(ds)
    st = cutc.cutc_int_device(ds, X, I, cutc.options(precision, dense=dense, n1_chunk=n1_chunk))
    torch.cuda.synchronize()
    return X, I, st


SYSTEMS = {
    "cube20": lambda: synth.cube20(n1=1024, n2=20000),
    "triatomic": lambda: synth.triatomic(n1=700, n2=3001),
    "odd": lambda: synth.make_system(7, 23, 515, 9973, 91, name="odd"),
    "splitk": lambda: synth.make_system(4, 30, 12000, 1500, 1201, name="splitk"),
}


@pytest.fixture(scope="module", params=list(SYSTEMS))
def system(torch, request):
    s = SYSTEMS[request.param]()
    ds = cutc.DeviceSystem.upload(s)
    X64, I64, _ = build(torch, ds, "fp64", 0)
    yield request.param, s, ds, X64, I64
    del X64, I64
    cutc.release_workspace()


@pytest.mark.parametrize("precision", ["int8", "bf16crt", "fp16crt"])
def test_sparse_equals_dense(torch, system, precision):
    name, s, ds, X64, I64 = system
    print(f"sparse {name} {s.shape_str()} {precision}", flush=True)
    Xs, Is, sts = build(torch, ds, precision, 0, n1_chunk=512)
    Xd, Id, std = build(torch, ds, precision, 1, n1_chunk=512)
    print(f"  k_density sparse {sts['k_density']:.3f} dense {std['k_density']:.3f}; "
          f"k_density_2e {sts['k_density_2e']:.3f}; gemm1 {sts['ms_gemm1']:.2f} / {std['ms_gemm1']:.2f} ms; "
          f"gemm2 {sts['ms_gemm2']:.2f} / {std['ms_gemm2']:.2f} ms", flush=True)
    assert torch.equal(Xs, Xd), "X differs between the list-driven and dense contraction"
    assert torch.equal(Is, Id), "int_2e_ao differs between the list-driven and dense contraction"
    assert std["k_density"] == 1.0 and std["k_density_2e"] == 0.0
    assert 0.0 < sts["k_density"] <= 1.0 and 0.0 < sts["k_density_2e"] <= 1.0
    if name == "cube20":
        assert sts["k_density"] < 0.5, sts["k_density"]
        assert sts["k_density_2e"] < 0.6, sts["k_density_2e"]
    ex, ei = rel_dev(torch, Xs, X64), rel_dev(torch, Is, I64)
    print(f"  vs FP64 build: X {ex:.2e} int_2e_ao {ei:.2e}", flush=True)
    assert ex <= TOL and ei <= TOL
    assert np.isfinite(float(torch.linalg.vector_norm(Is)))


@pytest.mark.parametrize("precision", ["int8", "bf16crt"])
def test_chunked_equals_one_chunk(torch, system, precision):
    """r1 chunks of 128 points (the Jastrow of each next chunk on the side
    stream, beside the current chunk's contraction) against the whole r1 set
    in one chunk: every G row's residues, exponents and rebuild are the
    same arithmetic, so X and int_2e_ao agree bit for bit."""
    name, s, ds, X64, I64 = system
    print(f"chunked {name} {s.shape_str()} {precision}", flush=True)
    Xc, Ic, stc = build(torch, ds, precision, 0, n1_chunk=128)
    X1, I1, st1 = build(torch, ds, precision, 0, n1_chunk=s.n1)
    print(f"  chunks {stc['n1_chunks']} / {st1['n1_chunks']}", flush=True)
    assert stc["n1_chunks"] == (s.n1 + 127) // 128 and st1["n1_chunks"] == 1
    assert torch.equal(Xc, X1), "X differs between chunked and one-chunk builds"
    assert torch.equal(Ic, I1), "int_2e_ao differs between chunked and one-chunk builds"
