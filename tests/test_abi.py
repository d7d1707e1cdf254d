"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
public header declares (no compute calls without a GPU), and its memory plan
behaves (pure host function)."""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from cutc_amd import _lib, cutc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("cutc_int.h", "cutc_no.h", "cutc_mo.h", "cutc_ao.h")]
HEADER = HEADERS[0]
GB = 1 << 30


def header_functions():
    names = []
    for h in HEADERS:
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*([A-Za-z_][A-Za-z_0-9]*)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_header_symbol():
    path = _lib.lib_path("libcutcint.so")
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    declared = header_functions()
    assert len(declared) >= 16
    missing = [n for n in declared if n not in exported]
    assert not missing, missing
    assert set(cutc.HEADER_SYMBOLS) <= set(declared)


def test_library_loads_and_binds():
    L = cutc.lib()
    for name in cutc.HEADER_SYMBOLS:
        assert hasattr(L, name)
    assert L.cutc_error_string(-1).decode().startswith("launch geometry")
    o = cutc.options()
    assert o.precision == 0 and o.slices == 0


def test_default_options_resolve_for_every_precision():
    """cutc_default_options leaves slices at 0, which every precision
    resolves to its own default: a C or Fortran caller that takes the
    defaults and sets only the precision gets a valid call.  Checked
    through the pure planner (no GPU): slices 0 plans exactly as the
    precision's default count."""
    L = cutc.lib()
    o = cutc.Options()
    L.cutc_default_options(ctypes.byref(o))
    assert o.precision == cutc.PREC_FP64 and o.slices == 0 and o.n1_chunk == 0 and o.mem_limit == 0
    for name, prec in cutc.PRECISIONS.items():
        p0, pd = cutc.Plan(), cutc.Plan()
        assert L.cutc_make_plan(60000, 200000, 200, 20, prec, o.slices, 200 * GB, 0, ctypes.byref(p0)) == 0, name
        assert L.cutc_make_plan(60000, 200000, 200, 20, prec, cutc.DEFAULT_SLICES[prec], 200 * GB, 0,
                                ctypes.byref(pd)) == 0, name
        assert (p0.n1_chunk, p0.bytes_stage1, p0.bytes_stage2, p0.k_chunk) == \
            (pd.n1_chunk, pd.bytes_stage1, pd.bytes_stage2, pd.k_chunk), name
    assert cutc.DEFAULT_SLICES[cutc.PREC_INT8] == 11
    p = cutc.Plan()
    assert L.cutc_make_plan(10, 10, 4, 1, cutc.PREC_INT8, -1, GB, 0, ctypes.byref(p)) == cutc.CUTC_ERR_ARG


def test_reference_entry_points_keep_their_signatures():
    """Argument lists of the reference symbols (src/cutc_int.c:15-20,
    include/cutc_int.h:4-15) as declared in include/cutc_int.h."""
    src = re.sub(r"\s+", " ", open(HEADER).read())
    assert ("int cutc_int(int nxBlocks, int nyBlocks, int nzBlocks, int blockxSize, int blockySize, int blockzSize, "
            "int n_grid1, int n_grid2, int n_ao, int n_nuc, int size_bh, double *h_r1, double *h_wr1, double *h_r2, "
            "double *h_wr2, double *h_rn, double *h_aos_data1, double *h_aos_data2, double *h_c_bh, int *h_m_bh, "
            "int *h_n_bh, int *h_o_bh, double *h_int2_grad1_u12_ao, double *h_int_2e_ao);") in src
    assert ("void get_int2_grad1_u12_ao(cutc_dim3 dimGrid, cutc_dim3 dimBlock, int n_grid1, int n_grid2, int n_ao, "
            "int n_nuc, int size_bh, double *r1, double *r2, double *wr2, double *rn, double *aos_data2, "
            "double *c_bh, int *m_bh, int *n_bh, int *o_bh, double *int2_grad1_u12_ao);") in src
    assert ("void get_int_2e_ao(int n_grid1, int n_ao, double *wr1, double *aos_data1, double *int2_grad1_u12_ao, "
            "double *tc_int_2e_ao);") in src
    assert ctypes.sizeof(cutc.Dim3) == 12


def test_plan_fits_budget_and_covers_r1():
    # the 20-nucleus workload on one GPU: n_ao = 200, n1 = 60k, n2 = 200k
    for budget in (64 * GB, 120 * GB, 200 * GB):
        p = cutc.make_plan(60000, 200000, 200, 20, budget)
        assert p.n_pairs == 20100
        assert p.ldk % 64 == 0 and p.ldk >= 200000
        assert 1 <= p.n1_chunk <= 60000
        assert p.bytes_stage1 <= budget and p.bytes_stage2 <= budget
        assert p.n1_chunk % 64 == 0 or p.n1_chunk == 60000
        # 64-bit sizes: the Apk workspace alone is 32 GB
        assert p.bytes_stage1 > 20100 * 200000 * 8


def test_plan_small_and_edge_cases():
    p = cutc.make_plan(4000, 16000, 24, 3, 8 * GB)
    assert p.n1_chunk == 4000 and p.n1_chunk_2e == 4000
    p = cutc.make_plan(0, 0, 0, 0, GB)
    assert p.n1_chunk == 0
    p = cutc.make_plan(1, 1, 1, 1, GB)
    assert p.n1_chunk == 1 and p.n_pairs == 1
    with pytest.raises(cutc.CutcError) as e:
        cutc.make_plan(60000, 200000, 200, 20, 1 * GB)
    assert e.value.code == cutc.CUTC_ERR_NOMEM
    with pytest.raises(cutc.CutcError) as e:
        cutc.make_plan(-1, 10, 10, 1, GB)
    assert e.value.code == cutc.CUTC_ERR_ARG
    # a requested chunk is honoured when it fits
    p = cutc.make_plan(10000, 5000, 30, 4, 8 * GB, n1_chunk=1000)
    assert p.n1_chunk <= 1000


def test_plan_large_streamed_shape_is_64bit():
    # 40 nuclei, n_ao = 300, n1 = 80k, n2 = 300k: workspaces beyond 2^31 elements
    p = cutc.make_plan(80000, 300000, 300, 40, 200 * GB)
    assert p.n_pairs == 45150
    assert p.bytes_stage1 > 2 ** 31 * 8
    assert p.bytes_stage1 <= 200 * GB


@pytest.mark.skipif(not os.path.exists("/opt/rocm/llvm/bin/amdflang"), reason="amdflang not present")
def test_fortran_binding_compiles_and_links(tmp_path):
    """include/cutc_module.f90 (the binding with the C definition's 6 launch
    integers) compiles, and a Fortran caller links against libcutcint.so."""
    fc = "/opt/rocm/llvm/bin/amdflang"
    mod = os.path.join(ROOT, "include", "cutc_module.f90")
    demo = os.path.join(ROOT, "tests", "fortran", "cutc_demo.f90")
    libdir = os.path.dirname(_lib.lib_path("libcutcint.so"))
    exe = tmp_path / "demo"
    r = subprocess.run([fc, "-J", str(tmp_path), mod, demo, "-o", str(exe), f"-L{libdir}", "-lcutcint",
                        f"-Wl,-rpath,{libdir}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert exe.exists()


def test_plan_modular_k_passes():
    """The 16-bit modular planes of all 300k r2 points (10 moduli x 45150
    pairs x 2 bytes = 271 GB) do not fit beside the rest: the planner cuts K
    into passes (multiples of 768 points) and keeps r1 chunks of >= 1024."""
    GB = 1 << 30
    one = cutc.make_plan(80000, 300000, 300, 40, 1 << 44, precision="bf16crt")
    assert one.k_chunk == one.ldk
    p = cutc.make_plan(80000, 300000, 300, 40, 240 * GB, precision="bf16crt")
    assert p.k_chunk < p.ldk and p.k_chunk % 768 == 0
    assert p.n1_chunk >= 1024 and p.bytes_stage1 <= 240 * GB
    # int8 at the same budget needs one pass; FP64 never cuts K
    assert cutc.make_plan(80000, 300000, 300, 40, 240 * GB, precision="int8").k_chunk == one.ldk
    f = cutc.make_plan(80000, 300000, 300, 40, 240 * GB)
    assert f.k_chunk == f.ldk


def test_argument_errors_before_any_device_call():
    """The entry points reject bad sizes and occupations with -2 before
    touching the device (runs without a GPU)."""
    L = cutc.lib()
    assert L.cutc_ao_to_mo(-1, 2, 2, None, None, None, None) == cutc.CUTC_ERR_ARG
    assert L.cutc_ao_to_mo(2, 2, 2, None, None, None, None) == cutc.CUTC_ERR_ARG  # missing arrays
    assert L.cutc_aos(-1, 1, 1, 1, None, None, None, None, None, None, None, None) == cutc.CUTC_ERR_ARG
    assert L.cutc_aos(4, 1, 70000, 1, None, None, None, None, None, None, None, None) == cutc.CUTC_ERR_ARG
    nucl = np.array([2], np.int32)  # nucleus 2 of 1: checked on the host
    power = np.zeros(3, np.int32)
    nprim = np.ones(1, np.int32)
    r = np.zeros(3)
    e = np.ones(1)
    assert L.cutc_aos(1, 1, 1, 1, r.ctypes.data, r.ctypes.data, nucl.ctypes.data, power.ctypes.data,
                      nprim.ctypes.data, e.ctypes.data, e.ctypes.data, r.ctypes.data) == cutc.CUTC_ERR_ARG
    assert L.cutc_no(10, 4, 5, 2, None, None, None, None, None, None, None) == cutc.CUTC_ERR_ARG
    assert L.cutc_no(10, 4, 1, -1, None, None, None, None, None, None, None) == cutc.CUTC_ERR_ARG
    assert L.cutc_get_last_error() == cutc.CUTC_ERR_ARG


def test_batch_branch_count_checked_before_any_device_call():
    """cutc_int_batch_device refuses more than 4 concurrent branches (or a
    negative count) with -2 before touching the device"""
    L = cutc.lib()
    st = cutc.BatchStats()
    for streams in (5, -1):
        o = cutc.options("fp64", streams=streams)
        assert L.cutc_int_batch_device(ctypes.byref(o), 0, None, None, ctypes.byref(st)) == cutc.CUTC_ERR_ARG


def test_one_hip_runtime_per_process():
    """The wrapper loads torch before the library, so a process that uses both
    maps a single libamdhip64 (torch asks for "libamdhip64.so", the library
    for "libamdhip64.so.7"; the other order maps two runtimes and torch's then
    sees no device)."""
    pytest.importorskip("torch")
    cutc.lib()
    with open("/proc/self/maps") as f:
        mapped = {line.split()[-1] for line in f if "libamdhip64" in line}
    assert len(mapped) == 1, mapped
