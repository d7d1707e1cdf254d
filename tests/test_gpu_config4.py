"""The 40-nucleus workload (n_ao = 300, all 300000 r2 points) streamed (X
never stored, cutc_int_streamed_device) on 16 r1 points, against the FP64
oracle: FP64 (<= 1e-12), int8 and bf16 modular (<= 1e-10); the bf16
modular pair planes over all r2 points do not fit, so its stage 1 runs K
in passes.  The oracle gives three column blocks of int_2e_ao and the
D D contraction without forming the 64.8 GB array on the host."""
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
    ds = cutc.DeviceSystem.upload(s)
    I = torch.empty((s.n_ao,) * 4, dtype=torch.float64, device="cuda")
    yield s, ds, I, {"cols": cols, "blocks": blocks, "dd": dd, "D": torch.from_numpy(D.reshape(-1)).cuda()}
    del I, cols
    cutc.release_workspace()
    torch.cuda.empty_cache()


@pytest.mark.timeout(900)
@pytest.mark.parametrize("precision", ["fp64", "int8", "bf16crt"])
def test_cube40_streamed(torch, cube40, precision):
    s, ds, I, ref = cube40
    n = s.n_ao ** 2
    if precision == "bf16crt":
        free = torch.cuda.mem_get_info()[0]
        p = cutc.make_plan(s.n1, s.n2, s.n_ao, s.n_nuc, free - (1 << 30), precision="bf16crt")
        print(f"  bf16crt plan: k_chunk {p.k_chunk} of ldk {p.ldk}", flush=True)
        assert p.k_chunk < p.ldk, "expected the K passes of the modular stage 1"
    I.fill_(float("nan"))
    st = cutc.cutc_int_streamed_device(ds, I, cutc.options(precision))
    torch.cuda.synchronize()
    cutc.release_workspace()
    M = I.reshape(n, n)
    errs = {}
    for c0, nc in ref["blocks"]:
        errs[f"cols{c0}"] = rel_dev(torch, M[c0:c0 + nc], ref["cols"][c0])
        assert torch.equal(M[c0:c0 + nc], M[:, c0:c0 + nc].t())  # (kl|ab) = (ab|kl), written once for both
    dd = float(ref["D"] @ (M @ ref["D"]))
    errs["dd"] = abs(dd - ref["dd"]) / abs(ref["dd"])
    print(f"  cube40 streamed {precision}: " + " ".join(f"{k} {v:.2e}" for k, v in errs.items())
          + f" ({st['ms_total'] / 1e3:.2f} s)", flush=True)
    assert bool(torch.isfinite(M[::997]).all())
    assert max(errs.values()) <= (TOL_FP64 if precision == "fp64" else TOL_SPLIT), errs
