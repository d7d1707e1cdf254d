"""AOs on the grid (cutc_aos) on the GPU against the FP64 oracle
(oracle/oracle_ao.c), and cutc_int fed with device-evaluated AOs against the
same build fed with the host AOs.  Bar: max abs error <= 1e-13 x the largest
value of each component (the device exp and the host libm exp may differ in
the last bit); build results within 1e-12 relative Frobenius error."""
import numpy as np
import pytest

from cutc_amd import cutc, synth
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


def contracted_basis(n_nuc, n_ao, pmax, seed):
    rng = np.random.default_rng(seed)
    return dict(ao_nucl=rng.integers(1, n_nuc + 1, n_ao).astype(np.int32),
                ao_power=rng.integers(0, 4, (3, n_ao)).astype(np.int32),
                ao_prim_num=rng.integers(1, pmax + 1, n_ao).astype(np.int32),
                ao_expo=np.exp(rng.uniform(np.log(0.1), np.log(50.0), (n_ao, pmax))),
                ao_coef=rng.standard_normal((n_ao, pmax)))


def close(got, want):
    for c in range(4):
        scale = np.abs(want[c]).max()
        err = np.abs(got[c] - want[c]).max()
        assert err <= 1e-13 * scale, (c, err, scale)


@pytest.mark.parametrize("n,n_nuc,n_ao,pmax", [(1000, 3, 24, 1), (777, 5, 61, 4), (4097, 20, 200, 3), (1, 1, 1, 1),
                                               (300, 2, 7, 6)])
def test_aos_device_matches_oracle(torch, n, n_nuc, n_ao, pmax):
    print(f"aos n={n} n_nuc={n_nuc} n_ao={n_ao} pmax={pmax}", flush=True)
    rng = np.random.default_rng(n + n_ao)
    r = rng.standard_normal((n, 3)) * 3.0
    rn = rng.standard_normal((n_nuc, 3)) * 2.0
    b = contracted_basis(n_nuc, n_ao, pmax, n_ao)
    want = oracle.aos(r, rn, **b)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    out = torch.full((4, n_ao, n), float("nan"), dtype=torch.float64, device="cuda")
    cutc.cutc_aos_device(t(r), t(rn), cutc.upload_basis(b), out)
    torch.cuda.synchronize()
    close(out.cpu().numpy(), want)


def test_aos_host_entry_and_errors(torch):
    print("aos host entry, errors", flush=True)
    s = synth.make_system(6, 50, 500, 10, 23, name="aos")
    got = cutc.cutc_aos(s.r1, s.rn, s.basis)
    close(got, s.aos1)  # the generator's own (uncontracted) AOs
    bad = dict(s.basis)
    bad["ao_nucl"] = s.basis["ao_nucl"].copy()
    bad["ao_nucl"][3] = 7  # only 6 nuclei
    with pytest.raises(cutc.CutcError) as e:
        cutc.cutc_aos(s.r1, s.rn, bad)
    assert e.value.code == cutc.CUTC_ERR_ARG
    # the device entry point does not read the tables on the host: the bad
    # AO gets NaN rows, the others their values
    out = torch.zeros((4, 50, 500), dtype=torch.float64, device="cuda")
    cutc.cutc_aos_device(*(torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (s.r1, s.rn)),
                         cutc.upload_basis(bad), out)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    assert np.isnan(o[:, 3]).all() and not np.isnan(np.delete(o, 3, axis=1)).any()
    close(np.delete(o, 3, axis=1), np.delete(s.aos1, 3, axis=1))


def test_cutc_int_on_device_aos(torch):
    """no aos_data transfer: the AOs are evaluated in HBM and the build reads
    them there"""
    s = synth.make_system(8, 60, 700, 5000, 41, name="devaos")
    print(f"cutc_int on device AOs: {s.shape_str()}", flush=True)
    res = []
    for ds in (cutc.DeviceSystem.upload(s), cutc.DeviceSystem.upload_with_device_aos(s)):
        X, I = cutc.alloc_outputs(ds)
        cutc.cutc_int_device(ds, X, I)
        torch.cuda.synchronize()
        res.append((X.cpu().numpy(), I.cpu().numpy()))
    (X0, I0), (X1, I1) = res
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    print(f"  relerr X {rel(X1, X0):.2e} int_2e_ao {rel(I1, I0):.2e}", flush=True)
    assert rel(X1, X0) <= 1e-12 and rel(I1, I0) <= 1e-12
