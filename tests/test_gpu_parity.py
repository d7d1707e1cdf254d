"""Parity of the HIP path (libcutcint.so, through its C ABI) with the FP64 CPU
oracle.  Bar: relative Frobenius error <= 1e-12 on int2_grad1_u12_ao and
int_2e_ao (the tolerance of the FP64 path), elementwise checks on the
Jastrow block and the integrands."""
import os
import subprocess

import numpy as np
import pytest

from cutc_amd import cutc, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12


def relerr(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need the MI355X"
    return t


ALL_TERMS = [(0, 0, 1), (0, 0, 2), (2, 0, 0), (3, 0, 0), (2, 2, 0), (0, 0, 3), (1, 2, 1), (3, 1, 2), (6, 0, 7),
             (0, 6, 0), (4, 5, 3)]


def with_terms(s, terms, seed=1):
    rng = np.random.default_rng(seed)
    s.c_bh = rng.uniform(-0.5, 0.5, (s.n_nuc, len(terms)))
    s.c_bh[:, 0] = 0.5
    s.m_bh = np.array([[t[0] for t in terms]] * s.n_nuc, np.int32)
    s.n_bh = np.array([[t[1] for t in terms]] * s.n_nuc, np.int32)
    s.o_bh = np.array([[t[2] for t in terms]] * s.n_nuc, np.int32)
    return s


def run_device(torch, s, **opt):
    ds = cutc.DeviceSystem.upload(s)
    X, I = cutc.alloc_outputs(ds)
    stats = cutc.cutc_int_device(ds, X, I, cutc.options(**opt))
    torch.cuda.synchronize()
    return X.cpu().numpy(), I.cpu().numpy(), stats


@pytest.mark.parametrize("n_nuc,n1,n2,terms", [(3, 21, 1000, None), (3, 19, 777, ALL_TERMS), (20, 17, 513, None),
                                               (40, 9, 300, None), (40, 5, 129, ALL_TERMS)])
def test_jastrow_block_reference_layout(torch, n_nuc, n1, n2, terms):
    if n_nuc == 3:
        s = synth.triatomic(n1=n1 + 5, n2=n2 + 7, seed=11)
    else:
        s = synth.make_system(n_nuc, 2 * n_nuc, n1 + 5, n2 + 7, 13, name="jast")
    if terms:
        s = with_terms(s, terms)
    print(f"jastrow n_nuc={n_nuc} n1={n1} n2={n2} terms={'all' if terms else 'bh6'}", flush=True)
    ds = cutc.DeviceSystem.upload(s)
    ii0, jj0, n1_tot, n2_tot = 3, 5, n1 + 2, n2 + 4
    G = torch.full((4, n1_tot, n2_tot), 7.0, dtype=torch.float64, device="cuda")
    cutc.cutc_int_bh(ds, G, ii0, n1, n1_tot, jj0, n2, n2_tot)
    g = G.cpu().numpy()
    want = oracle.jast_bh(s, ii0, n1, jj0, n2)
    got = g[:, :n1, :n2]
    assert relerr(got, want) <= TOL
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-13 * np.abs(want).max())
    # untouched outside the block
    assert np.all(g[:, n1:, :] == 7.0) and np.all(g[:, :n1, n2:] == 7.0)


def test_integrand_wrappers_reference_layout(torch):
    s = synth.triatomic(n1=131, n2=517, seed=3)
    ds = cutc.DeviceSystem.upload(s)
    n_ao = s.n_ao
    # long range over an r2 window, leading dimension larger than the window
    jj0, n2e, n2t = 17, 300, 320
    out = torch.zeros((n_ao, n_ao, n2t), dtype=torch.float64, device="cuda")
    cutc.int_long_range(ds, out, jj0, n2e, n2t)
    chi = s.aos2[0][:, jj0:jj0 + n2e]
    want = s.wr2[jj0:jj0 + n2e] * chi[:, None, :] * chi[None, :, :]
    np.testing.assert_allclose(out.cpu().numpy()[:, :, :n2e], want, rtol=1e-15, atol=0)
    H = torch.zeros((n_ao, n_ao, s.n1), dtype=torch.float64, device="cuda")
    cutc.int_short_range_herm(ds, H)
    np.testing.assert_array_equal(H.cpu().numpy(), oracle.herm(s))
    NH = torch.zeros((n_ao, n_ao, 3, s.n1), dtype=torch.float64, device="cuda")
    cutc.int_short_range_nonherm(ds, NH)
    np.testing.assert_allclose(NH.cpu().numpy(), oracle.nonherm(s), rtol=1e-15, atol=1e-300)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 200, 577])
def test_trans_inplace(torch, n):
    rng = np.random.default_rng(n)
    M = rng.standard_normal((n, n))
    T = torch.from_numpy(M.copy()).cuda()
    cutc.trans_inplace(T)
    np.testing.assert_array_equal(T.cpu().numpy(), oracle.trans_inplace(M))


SHAPES = [  # n_nuc, n_ao, n1, n2
    (1, 1, 1, 1), (2, 5, 3, 7), (3, 24, 37, 129), (5, 17, 100, 300), (3, 24, 130, 1000), (7, 61, 77, 2049),
    (33, 40, 65, 333),
]


@pytest.mark.parametrize("shape", SHAPES)
def test_path_matches_oracle_small_shapes(torch, shape):
    n_nuc, n_ao, n1, n2 = shape
    print(f"path shape n_nuc={n_nuc} n_ao={n_ao} n1={n1} n2={n2}", flush=True)
    s = synth.make_system(n_nuc, n_ao, n1, n2, 100 + n_ao, name="small")
    if n_ao > 20:
        s = with_terms(s, ALL_TERMS)
    X, I, _ = run_device(torch, s)
    Xo = oracle.int2_grad1_u12_ao(s)
    Io = oracle.int_2e_ao(s, Xo)
    assert relerr(X, Xo) <= TOL, relerr(X, Xo)
    assert relerr(I, Io) <= TOL, relerr(I, Io)
    # the GPU's own X feeds the oracle's stage 2 as well
    assert relerr(I, oracle.int_2e_ao(s, X)) <= TOL


def test_triatomic_full_host_path(torch):
    """The 3-nucleus workload (n_ao = 24, n1 = 4000, n2 = 16000) through the
    reference entry point cutc_int on host arrays."""
    s = synth.triatomic()
    X3, I = cutc.cutc_int(s)
    Xo, Io = oracle.cutc_int(s)
    ex, ei = relerr(X3, Xo), relerr(I, Io)
    print(f"triatomic: relerr X {ex:.3e} int_2e_ao {ei:.3e}", flush=True)
    assert ex <= TOL and ei <= TOL
    # scalar sum_{kl,ab} D_kl D_ab (kl|ab) for a rank-10 PSD density
    rng = np.random.default_rng(7)
    C = rng.standard_normal((s.n_ao, 10))
    D = C @ C.T
    dd = np.einsum("kl,ab,balk->", D, D, I)
    ddo = np.einsum("kl,ab,balk->", D, D, Io)
    assert abs(dd - ddo) <= TOL * abs(ddo)


REFERENCE_LINES = [  # src/cutc_int.c, src/build_int2_grad1_u12_ao.cu, src/build_int_2e_ao.cu, in their order
    " Computing TC-Integrals With CuTC",  # cutc_int.c
    " free  memory = ", " total memory = ", "n_grid1_pass = ", "n_grid1_rest = ", "n1_pass = ",
    "n_grid2_pass = ", "n_grid2_rest = ", "n2_pass = ",  # build_int2_grad1_u12_ao.cu:60-61,81-95
    "Ellapsed time for DGEMM to build int2_grad1_u12_ao = ", "Ellapsed time for int_long_range kernel = ",
    "Ellapsed time for cutc_int_bh kernel = ",  # build_int2_grad1_u12_ao.cu:342-344
    "Ellapsed time for get_int2_grad1_u12_ao = ",  # cutc_int.c:153
    "Ellapsed time for DGEMM to build int_2e_ao = ", "Ellapsed time for int_short_range_herm kernel = ",
    "Ellapsed time for int_short_range_nonherm kernel = ", "Ellapsed time for addT = ",  # build_int_2e_ao.cu:139-142
    "Ellapsed time for get_int_2e_ao = ",  # cutc_int.c:198
    "Ellapsed time for Device <-> Host transf = ", "Ellapsed (total) time on GPU for cutc_int = ",  # cutc_int.c:230,236
]


def test_reference_stdout_lines(torch, capfd):
    """cutc_int prints the reference's progress and stage-time lines, with
    the same labels and in the same order, each with a number"""
    s = synth.triatomic(n1=300, n2=2000)
    capfd.readouterr()
    cutc.cutc_int(s)
    out = capfd.readouterr().out.splitlines()
    pos = 0
    for label in REFERENCE_LINES:
        while pos < len(out) and not out[pos].startswith(label):
            pos += 1
        assert pos < len(out), f"missing or out of order: {label!r}\n" + "\n".join(out)
        rest = out[pos][len(label):].split()
        if rest:
            float(rest[0])
        pos += 1
    # the four values of the chunk plan for one r1 chunk of all points
    txt = "\n".join(out)
    assert "n_grid1_pass = 300" in txt and "n1_pass = 1" in txt and "n_grid2_pass = 2000" in txt


def test_chunking_and_reference_stage_functions(torch):
    s = with_terms(synth.make_system(6, 33, 300, 1500, 77, name="chunk"), ALL_TERMS)
    X0, I0, st0 = run_device(torch, s)
    X1, I1, st1 = run_device(torch, s, n1_chunk=64)
    assert st1["n1_chunks"] > 1 and st0["n1_chunks"] == 1
    assert relerr(X1, X0) <= 1e-14 and relerr(I1, I0) <= 1e-13
    # get_int2_grad1_u12_ao + get_int_2e_ao (reference device entry points)
    ds = cutc.DeviceSystem.upload(s)
    X, I = cutc.alloc_outputs(ds)
    cutc.get_int2_grad1_u12_ao(ds, X)
    cutc.get_int_2e_ao(ds, X, I)
    torch.cuda.synchronize()
    Xg, Ig = X.cpu().numpy(), I.cpu().numpy()
    assert relerr(Xg, X0) <= 1e-14 and relerr(Ig, I0) <= 1e-14
    # and against the oracle directly
    Xo = oracle.int2_grad1_u12_ao(s)
    Io = oracle.int_2e_ao(s, Xo)
    assert relerr(Xg, Xo) <= TOL and relerr(Ig, Io) <= TOL


@pytest.mark.parametrize("n1_chunk,precision", [(700, "fp64"), (0, "fp64"), (500, "int8")])
def test_host_entry_pipelined_download(torch, n1_chunk, precision):
    """cutc_int_ex on host arrays downloads each r1 chunk of X while the build
    runs (pinned staging ring, copy stream, host copy threads): the host
    arrays equal the device results, and the blocking copies (host_copy=1)
    give the same bytes.  n1_chunk 700: 5 chunks of 20 MB per component; 0:
    one chunk whose 86 MB components cross the 64 MiB staging pieces."""
    s = synth.make_system(8, 60, 3000, 4000, 31, name="xfer")
    print(f"pipelined download n1_chunk={n1_chunk} {precision} {s.shape_str()}", flush=True)
    X0, I0, st0 = run_device(torch, s, n1_chunk=n1_chunk, precision=precision)
    Xp, Ip, stp = cutc.cutc_int_ex(s, cutc.options(precision=precision, n1_chunk=n1_chunk))
    Xb, Ib, _ = cutc.cutc_int_ex(s, cutc.options(precision=precision, n1_chunk=n1_chunk, host_copy=1))
    assert stp["n1_chunks"] == st0["n1_chunks"]
    assert (stp["n1_chunks"] > 1) == (n1_chunk > 0)
    assert np.array_equal(Xp, Xb) and np.array_equal(Ip, Ib)
    assert relerr(Xp, X0[:3]) <= 1e-15 and relerr(Ip, I0) <= 1e-14
    # empty grid 2: X written without r1 chunks, downloaded after the build
    e2 = s.subset_r2(0)
    Xz, Iz, _ = cutc.cutc_int_ex(e2, cutc.options(precision=precision))
    assert Xz.shape == Xp.shape and not np.any(Xz) and not np.any(Iz)


def test_empty_and_error_paths(torch):
    s = synth.triatomic(n1=5, n2=9)
    # n1 = 0: int_2e_ao is 0
    e1 = s.subset_r1(0)
    X, I, _ = run_device(torch, e1)
    assert X.size == 0 and not np.any(I)
    # n2 = 0: X and int_2e_ao are 0
    e2 = s.subset_r2(0)
    X, I, _ = run_device(torch, e2)
    assert not np.any(X) and not np.any(I)
    # launch geometry beyond the device: -1 (src/cutc_int.c:79-86)
    with pytest.raises(cutc.CutcError) as e:
        cutc.cutc_int(s, geometry=(1, 1, 1, 64, 64, 1))
    assert e.value.code == cutc.CUTC_ERR_GEOMETRY
    # negative exponent: argument error, no launch
    bad = synth.triatomic(n1=5, n2=9)
    bad.o_bh[0, 0] = -1
    with pytest.raises(cutc.CutcError) as e:
        cutc.cutc_int(bad)
    assert e.value.code == cutc.CUTC_ERR_ARG
    # working set beyond the memory limit
    with pytest.raises(cutc.CutcError) as e:
        run_device(torch, s, mem_limit=1024)
    assert e.value.code == cutc.CUTC_ERR_NOMEM


def test_split_k_step2_matches_oracle(torch):
    """A small pair count with a long r1 range: the FP64 step-2 contraction
    runs split along K (8 slices here), reduced in order."""
    s = synth.make_system(4, 30, 12000, 1500, 1201, name="splitk")
    print(f"split-K step 2: {s.shape_str()}", flush=True)
    X, I, st = run_device(torch, s)
    Xo = oracle.int2_grad1_u12_ao(s)
    Io = oracle.int_2e_ao(s, Xo)
    e = relerr(I, Io)
    print(f"  int_2e_ao vs oracle {e:.2e}", flush=True)
    assert e <= 1e-12 and relerr(X, Xo) <= 1e-12


def test_cube20_full_basis_subset(torch):
    """20 nuclei, n_ao = 200, all n2 = 200000 points, a 24-point r1 subset."""
    s = synth.cube20(n1=24, n2=200000)
    X, I, st = run_device(torch, s)
    Xo = oracle.int2_grad1_u12_ao(s)
    ex = relerr(X, Xo)
    print(f"cube20 r1<24: relerr X {ex:.3e}", flush=True)
    assert ex <= TOL
    Io = oracle.int_2e_ao(s, Xo)
    ei = relerr(I, Io)
    print(f"cube20 r1<24: relerr int_2e_ao {ei:.3e}", flush=True)
    assert ei <= TOL


@pytest.mark.skipif(not os.path.exists("/opt/rocm/llvm/bin/amdflang"), reason="amdflang not present")
def test_fortran_caller_zero_demo(torch, tmp_path):
    """The reference demo's only known answer, through the Fortran binding."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    from cutc_amd import _lib
    libdir = os.path.dirname(_lib.lib_path("libcutcint.so"))
    exe = tmp_path / "demo"
    subprocess.run(["/opt/rocm/llvm/bin/amdflang", "-J", str(tmp_path), os.path.join(root, "include", "cutc_module.f90"),
                    os.path.join(root, "tests", "fortran", "cutc_demo.f90"), "-o", str(exe), f"-L{libdir}",
                    "-lcutcint", f"-Wl,-rpath,{libdir}"], check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    last = r.stdout.strip().splitlines()[-1].split()
    assert last[0] == "status" and int(last[1]) == 0
    assert float(last[3]) == 0.0


@pytest.mark.skipif(not os.path.exists("/opt/rocm/llvm/bin/amdflang"), reason="amdflang not present")
def test_fortran_chain(torch, tmp_path):
    """AOs -> cutc_int -> cutc_ao_to_mo -> cutc_no from Fortran through
    include/cutc_module.f90 (tests/fortran/cutc_chain.f90), against the
    oracle's chain on the same inputs (the device AOs differ from the host
    ones in the last bits, hence 1e-11)."""
    print("fortran chain", flush=True)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    from cutc_amd import _lib
    libdir = os.path.dirname(_lib.lib_path("libcutcint.so"))
    exe = tmp_path / "chain"
    subprocess.run(["/opt/rocm/llvm/bin/amdflang", "-J", str(tmp_path), os.path.join(root, "include", "cutc_module.f90"),
                    os.path.join(root, "tests", "fortran", "cutc_chain.f90"), "-o", str(exe), f"-L{libdir}",
                    "-lcutcint", f"-Wl,-rpath,{libdir}"], check=True, capture_output=True)
    s = synth.triatomic(n1=300, n2=2000, seed=41)
    b = s.basis
    n_mo, na, nb = 12, 4, 3
    rng = np.random.default_rng(3)
    cl = rng.standard_normal((n_mo, s.n_ao)) / np.sqrt(s.n_ao)  # numpy (n_mo, n_ao) = Fortran cl(n_ao, n_mo)
    cr = cl + 0.05 * rng.standard_normal((n_mo, s.n_ao)) / np.sqrt(s.n_ao)
    pmax = b["ao_expo"].shape[1]
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([s.n_nuc, s.n_ao, s.n1, s.n2, s.size_bh, pmax, n_mo, na, nb], np.int32).tofile(f)
        for a in (s.r1, s.wr1, s.r2, s.wr2, s.rn):
            np.ascontiguousarray(a, np.float64).tofile(f)
        for k in ("ao_nucl", "ao_power", "ao_prim_num"):
            np.ascontiguousarray(b[k], np.int32).tofile(f)
        for k in ("ao_expo", "ao_coef"):
            np.ascontiguousarray(b[k], np.float64).tofile(f)
        np.ascontiguousarray(s.c_bh, np.float64).tofile(f)
        for a in (s.m_bh, s.n_bh, s.o_bh):
            np.ascontiguousarray(a, np.int32).tofile(f)
        cl.tofile(f)
        cr.tofile(f)
    r = subprocess.run([str(exe), str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(fout, np.uint8)
    st = raw[:20].view(np.int32)
    assert not st.any(), st
    d = raw[20:].view(np.float64)
    na4, n1 = s.n_ao ** 4, s.n1
    sizes = [na4, n1 * 3 * n_mo * n_mo, n_mo ** 4, n_mo * n_mo, 1]
    parts = np.split(d, np.cumsum(sizes)[:-1])
    Xo, Io = oracle.cutc_int(s)
    Go = oracle.ao_to_mo(cl, cr, Xo)
    chi = oracle.aos(s.r1, s.rn, **b)[0]
    want_no = oracle.cutc_no(na, nb, s.wr1, cl @ chi, cr @ chi, Go)
    errs = [relerr(parts[0], Io.reshape(-1)), relerr(parts[1], Go.reshape(-1)),
            relerr(parts[2], want_no[0].reshape(-1)), relerr(parts[3], want_no[1].reshape(-1)),
            abs(parts[4][0] - want_no[2]) / abs(want_no[2])]
    print("  relerr int_2e_ao %.2e G %.2e no_2e %.2e no_1e %.2e no_0e %.2e" % tuple(errs), flush=True)
    assert max(errs) <= 1e-11

