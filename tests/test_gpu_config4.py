"""The 40-nucleus workload (n_ao = 300, all 300000 r2 points) streamed (X
never stored, cutc_int_streamed_device) on 16 r1 points, against the FP64
oracle and the FP64 build: FP64 (<= 1e-12 against the oracle's column
blocks and D D contraction), int8 and bf16 modular (<= 1e-10 over the whole
tensor against the FP64 build); the bf16 modular pair planes over all r2
points do not fit, so its stage 1 runs K in passes.  The oracle gives
three column blocks of int_2e_ao and the D D contraction without forming
the 64.8 GB array on the host."""
import numpy as np
import pytest

from cutc_amd import cutc, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12
TOL_SPLIT = 1e-10


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


# ------------------------------------------------ 40 nuclei, n_ao = 300

@pytest.fixture(scope="module")
def cube40(torch):
    s = synth.cube40(n1=16, n2=300000)
    print(f"cube40 subset {s.shape_str()}: oracle", flush=True)
    Xo = oracle.int2_grad1_u12_ao(s)
    n = s.n_ao ** 2
    blocks = [(0, 300), (44000, 300), (n - 300, 300)]
    cols = {c0: torch.from_numpy(oracle.int_2e_ao_cols(s, Xo, c0, nc)).cuda() for c0, nc in blocks}
    D = density(s.n_ao)
    dd = oracle.dd_2e(s, Xo, D)
    del Xo
    ds = cutc.DeviceSystem.upload(s)
    ref = {"cols": cols, "blocks": blocks, "dd": dd, "D": torch.from_numpy(D.reshape(-1)).cuda()}
    # the FP64 streamed build of the whole tensor, kept on the device as the
    # reference of the split builds once it has matched the oracle
    Ifp = torch.full((s.n_ao,) * 4, float("nan"), dtype=torch.float64, device="cuda")
    st = cutc.cutc_int_streamed_device(ds, Ifp, cutc.options("fp64"))
    torch.cuda.synchronize()
    cutc.release_workspace()
    print(f"  FP64 streamed build {st['ms_total'] / 1e3:.2f} s", flush=True)
    ref["fp64"] = Ifp
    yield s, ds, ref
    del Ifp, cols, ref
    cutc.release_workspace()
    torch.cuda.empty_cache()


def against_oracle(torch, s, ref, I):
    """relative Frobenius errors of three column blocks and the relative
    error of the D D contraction against the oracle"""
    n = s.n_ao ** 2
    M = I.reshape(n, n)
    errs = {}
    for c0, nc in ref["blocks"]:
        errs[f"cols{c0}"] = rel_dev(torch, M[c0:c0 + nc], ref["cols"][c0])
        assert torch.equal(M[c0:c0 + nc], M[:, c0:c0 + nc].t())  # (kl|ab) = (ab|kl), written once for both
    dd = float(ref["D"] @ (M @ ref["D"]))
    errs["dd"] = abs(dd - ref["dd"]) / abs(ref["dd"])
    return errs


@pytest.mark.timeout(900)
def test_cube40_streamed_fp64(torch, cube40):
    s, ds, ref = cube40
    errs = against_oracle(torch, s, ref, ref["fp64"])
    print("  cube40 streamed fp64 vs oracle: " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    assert bool(torch.isfinite(ref["fp64"]).all())
    assert max(errs.values()) <= TOL_FP64, errs


@pytest.mark.parametrize("precision", ["int8", "bf16crt"])
def test_cube40_streamed_split(torch, cube40, precision):
    """The split builds over the whole int_2e_ao against the FP64 build
    (<= 1e-10 relative Frobenius, the metric's bar), the D D contraction
    against the oracle (<= 1e-10), and each column block against the oracle
    with its error norm measured against the norm a block of its width has
    on average in the whole tensor, ||I|| sqrt(cols / n_ao^2) (<= 1e-10).
    A block's error relative to its own norm is printed too: blocks of
    small norm carry the same absolute error, so theirs is larger (with 16
    r1 points up to 1.9e-10 for int8 with 11 moduli, 6.6e-9 for bf16
    modular with 10)."""
    s, ds, ref = cube40
    if precision == "bf16crt":
        free = torch.cuda.mem_get_info()[0]
        p = cutc.make_plan(s.n1, s.n2, s.n_ao, s.n_nuc, free - (1 << 30), precision="bf16crt")
        print(f"  bf16crt plan: k_chunk {p.k_chunk} of ldk {p.ldk}", flush=True)
        assert p.k_chunk < p.ldk, "expected the K passes of the modular stage 1"
    I = torch.full_like(ref["fp64"], float("nan"))
    st = cutc.cutc_int_streamed_device(ds, I, cutc.options(precision))
    torch.cuda.synchronize()
    cutc.release_workspace()
    assert bool(torch.isfinite(I).all())
    e_full = rel_dev(torch, I, ref["fp64"])
    errs = against_oracle(torch, s, ref, I)
    n = s.n_ao ** 2
    scale = float(torch.linalg.vector_norm(ref["fp64"]))
    M = I.reshape(n, n)
    scaled = {c0: float(torch.linalg.vector_norm(M[c0:c0 + nc] - ref["cols"][c0])) / (scale * (nc / n) ** 0.5)
              for c0, nc in ref["blocks"]}
    print(f"  cube40 streamed {precision}: whole tensor vs FP64 {e_full:.2e}; vs oracle "
          + " ".join(f"{k} {v:.2e}" for k, v in errs.items()) + "; block errors / mean block norm "
          + " ".join(f"{v:.2e}" for v in scaled.values()) + f" ({st['ms_total'] / 1e3:.2f} s)", flush=True)
    del I, M
    assert e_full <= TOL_SPLIT and errs["dd"] <= TOL_SPLIT
    assert max(scaled.values()) <= TOL_SPLIT, scaled


def test_cube40_streamed_int8_two_chunks(torch, cube40):
    """Two r1 chunks of 8 points: the second reuses the first's r2 order and
    pair clustering (kept across the chunks of a streamed call); whole
    tensor within the metric's bar of the FP64 build."""
    s, ds, ref = cube40
    I = torch.full_like(ref["fp64"], float("nan"))
    st = cutc.cutc_int_streamed_device(ds, I, cutc.options("int8", n1_chunk=8))
    torch.cuda.synchronize()
    cutc.release_workspace()
    assert st["n1_chunks"] == 2
    assert bool(torch.isfinite(I).all())
    e_full = rel_dev(torch, I, ref["fp64"])
    print(f"  cube40 streamed int8 in 2 chunks: whole tensor vs FP64 {e_full:.2e} ({st['ms_total'] / 1e3:.2f} s)",
          flush=True)
    del I
    assert e_full <= TOL_SPLIT
