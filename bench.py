"""Benchmark of the TC integral build on one MI355X: split precision vs FP64.

A step is one full build on inputs already resident in HBM: Jastrow pair
blocks -> integrands -> step-1 contraction X_kl(r1) -> step-2 contraction ->
symmetrised int_2e_ao (cutc_int_device through the C ABI), on the workload
20 nuclei in a 12-bohr cube, n_ao = 200, n1 = 60,000, n2 = 200,000 (seeded
synthetic inputs).

Default run: one FP64 build (the plain path, after a small warm-up call),
then W warm-up and K timed builds with the split-precision contraction
(--precision, default int8 = modular int8 MFMA with the library's default
of 11 moduli), then the split results are compared with the FP64 ones:
relative Frobenius and max-abs error of every X component and of int_2e_ao,
and of sum_{kl,ab} D_kl D_ab (kl|ab) for a seeded rank-10 PSD density D.
Then the BF16 leg: one warm-up and one timed build with the contractions as
bf16 modular products (10 moduli, the library default), compared with the
same FP64 build (--no-bf16 skips it).

--workload batch64 times 64 independent systems of uneven sizes: 64
sequential FP64 calls against one batched hipGraph launch in FP64, bf16
slices and int8 (errors against the sequential FP64 results).

--workload cube40 runs the large streamed case instead (40 nuclei, n_ao = 300,
n1 = 80,000, n2 = 300,000, cutc_int_streamed_device: X is never stored, each
r1 chunk goes straight into the int_2e_ao accumulator); errors are reported
for int_2e_ao and the D D contraction.

value = effective TFLOP/s of the timed (split) build = W_ref / seconds, with
the reference formulation's work W_ref = 8 n_ao^2 n1 n2 + 8 n_ao^4 n1 (full
AO pairs, the metric's definition).  The build executes only the packed
k <= l pairs, and the modular stage-1 contraction only the (256-pair tile,
K step) tiles with a non-zero residue (stats k_density): roofline
fractions are computed on the executed work.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import statistics
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "int_2e_ao build s + eff TFLOP/s: split-BF16 vs FP64 on GPU at ≤1e-10 rel err"
PEAK = {"fp64": 78.6, "int8": 5000.0, "bf16": 2500.0, "fp16": 2500.0, "bf16crt": 2500.0,
        "fp16crt": 2500.0}  # dense MFMA, MI355X (int8: TOPS)


BATCH_PRECISIONS = ("fp64", "bf16", "bf16crt", "int8")
# HBM-side bytes per stage-1 contraction launch on the default workload (n1 = 60000,
# n2 = 200000, the library's default moduli), from two rocprofv3 --pmc passes (FETCH_SIZE,
# WRITE_SIZE) scaled by a 1 GiB copy kernel measured in the same run
# (profiles/cube20_traffic.txt); not measured for other shapes.
MEASURED_TRAFFIC = {"int8": 132.0e9, "fp64": 17.83e12,
                    # the transform's GEMM launch (n_ao = n_mo = 200, n_grid1 = 60000) and the AO kernel
                    # (260000 points, n_ao = 200), profiles/ao_to_mo_nao200_kernel_stats.txt and
                    # profiles/aos_cube20_kernel_stats.txt
                    "mo": 2 * 42588146 * 1024 + 56250000 * 1024, "ao": 2 * 3556 * 1024 + 1625000 * 1024}


def w_ref(n_ao, n1, n2):
    return 8.0 * n_ao ** 2 * n1 * n2 + 8.0 * n_ao ** 4 * n1


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed builds (default: 5 for cube20, 1 for the 20x larger cube40)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="int8", choices=["fp64", "int8", "bf16", "fp16", "bf16crt", "fp16crt"],
                    help="int8 / bf16crt / fp16crt: modular (CRT) products; bf16 / fp16: digit slices")
    ap.add_argument("--slices", type=int, default=None,
                    help="int8: moduli (default: the library's, 11); bf16crt / fp16crt: moduli (default 10); "
                         "bf16 / fp16: digit slices per operand (default 4)")
    ap.add_argument("--no-bf16", action="store_true", help="skip the BF16 leg (bf16 modular build beside the "
                                                           "timed split build)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison build")
    ap.add_argument("--workload", default="cube20", choices=["cube20", "cube40", "batch64", "no", "mo", "ao"],
                    help="cube20: 20 nuclei, n_ao=200, X stored (the metric's workload); cube40: 40 nuclei, "
                         "n_ao=300, n1=80000, n2=300000, streamed (X never stored); batch64: 64 independent "
                         "systems, sequential FP64 calls vs one batched graph launch; no: normal ordering "
                         "(cutc_no), n_mo=200, n_grid1=60000, 30 doubly occupied orbitals; mo: AO -> MO "
                         "transform of int2_grad1_u12 (cutc_ao_to_mo), n_ao=n_mo=200, n_grid1=60000; ao: AO values and "
                         "gradients on both grids of the cube20 workload (cutc_aos_device), 3 primitives per AO")
    ap.add_argument("--batch-scale", type=float, default=1.0, help="batch64: grid sizes times this factor")
    ap.add_argument("--batch-streams", type=int, default=0,
                    help="batch64: concurrent graph branches (0: the library default, 4 for the modular forms and 1 for "
                         "the others); FP64 is also timed on one branch")
    ap.add_argument("--variants", default="",
                    help="further split builds compared with the same FP64 build, e.g. bf16:3,bf16:4,fp16:4")
    ap.add_argument("--variant-warmup", type=int, default=1, help="untimed builds before each variant's timed one")
    ap.add_argument("--host-path", action="store_true",
                    help="also time one cutc_int_ex call on host arrays (uploads, build, downloads of X[0..2] and "
                         "int_2e_ao): the PCIe-inclusive rate, reported beside value")
    ap.add_argument("--n1", type=int, default=None)
    ap.add_argument("--n2", type=int, default=None)
    ap.add_argument("--cpu-sample", type=int, default=32, help="r1 points of the CPU-baseline sample (0: skip)")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="first run one device-to-device copy of 1 GiB, a plain copy kernel of known size that "
                         "calibrates FETCH_SIZE / WRITE_SIZE counters read by rocprofv3 --pmc in the same run")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 1 if args.workload == "cube40" else 5
    return args


def log(args, *a):
    if not args.quiet:
        print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(synth, sysfull, npts, runs=3):
    """The FP64 oracle (C + OpenMP + OpenBLAS) timed on this host, median of
    `runs`: the 3-nucleus workload in full (n_ao = 24, n1 = 4000, n2 =
    16000), and the first npts r1 points of the benchmarked system (all its
    r2 points), whose rate is `value`; the full system's time at that rate
    is reported as a linear extrapolation in n1."""
    from oracle import oracle
    oracle.lib()

    def timed(s):
        ts = []
        for _ in range(runs):
            t0 = time.perf_counter()
            X = oracle.int2_grad1_u12_ao(s)
            I = oracle.int_2e_ao(s, X)
            ts.append(time.perf_counter() - t0)
            del X, I
        med = statistics.median(ts)
        w = w_ref(s.n_ao, s.n1, s.n2)
        return {"workload": s.shape_str(), "seconds": ts, "seconds_median": med, "w_ref_flops": w,
                "tflops_median": w / med / 1e12}

    c0 = timed(synth.triatomic())
    sub = timed(sysfull.subset_r1(npts))
    w_full = w_ref(sysfull.n_ao, sysfull.n1, sysfull.n2)
    rate = sub["tflops_median"]
    return {"value": rate, "unit": "TFLOP/s", "cores": oracle.threads(), "kind": "port",
            "cpu_model": cpu_model(), "runs": runs,
            "sample": f"FP64 oracle ({oracle.blas_name()} DGEMM, {oracle.threads()} threads) on the first {npts} of "
                      f"{sysfull.n1} r1 points with all {sysfull.n2} r2 points, n_ao={sysfull.n_ao}: median "
                      f"{sub['seconds_median']:.2f} s of {runs}; value = W_ref(sample) / s",
            "seconds": sub["seconds_median"], "config1_sample": sub, "config0_full": c0,
            "config1_full_s_extrapolated": w_full / (rate * 1e12),
            "extrapolation_note": "linear in n1 from the sample's rate (the oracle's work is linear in n1)"}


def build(cutc, ds, X, I, opt, stream):
    if X is None:
        return cutc.cutc_int_streamed_device(ds, I, opt, stream)
    return cutc.cutc_int_device(ds, X, I, opt, stream)


def timed_builds(cutc, torch, ds, X, I, opt, n, stream, per_step=None):
    """n builds back to back between HIP events on the build stream; returns
    the mean ms per build (the whole timed region / n) and the statistics of
    every build.  per_step, a list, receives each build's own event time."""
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(n + 1)]
    stats = []
    torch.cuda.synchronize()
    ev[0].record(stream)
    for i in range(n):
        stats.append(build(cutc, ds, X, I, opt, stream))
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    if per_step is not None:
        per_step.extend(ev[i].elapsed_time(ev[i + 1]) for i in range(n))
    return ev[0].elapsed_time(ev[n]) / n, stats


def density(n_ao, seed=20261205):
    import numpy as np
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((n_ao, 10))
    return C @ C.T


def batch_bench(args, torch, cutc, synth):
    """64 independent systems (n_nuc U{3..12}, n_ao U{30..150}, n1 U{5k..30k},
    n2 U{20k..100k}): 64 sequential FP64 streamed calls against one batched
    graph launch (cutc_int_batch_device) in FP64, bf16 slices and int8."""
    t0 = time.perf_counter()
    sys_ = synth.batch_systems(64, scale=args.batch_scale)
    log(args, f"generated 64 systems in {time.perf_counter() - t0:.1f} s")
    dss = [cutc.DeviceSystem.upload(s) for s in sys_]
    W = sum(w_ref(s.n_ao, s.n1, s.n2) for s in sys_)
    stream = torch.cuda.current_stream()
    Iseq = [torch.empty((s.n_ao,) * 4, dtype=torch.float64, device="cuda") for s in sys_]
    Ib = [torch.empty_like(I) for I in Iseq]
    # warm-up: load the kernels of every path on a small system
    small = synth.make_system(3, 24, 300, 2000, 1, name="warm")
    dsm = cutc.DeviceSystem.upload(small)
    Ism = torch.empty((24,) * 4, dtype=torch.float64, device="cuda")
    for prec in BATCH_PRECISIONS:
        cutc.cutc_int_streamed_device(dsm, Ism, cutc.options(prec), stream)
        cutc.cutc_int_batch_device([dsm], [Ism], cutc.options(prec), stream)
    # the system with the largest workspace once, so that the timed calls find the library's workspace grown
    big = max(range(len(sys_)), key=lambda i: sys_[i].n_ao ** 2 * sys_[i].n2)
    cutc.cutc_int_streamed_device(dss[big], Iseq[big], cutc.options("fp64"), stream)
    # sequential FP64 calls
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for ds, I in zip(dss, Iseq):
        cutc.cutc_int_streamed_device(ds, I, cutc.options("fp64"), stream)
    torch.cuda.synchronize()
    seq_s = time.perf_counter() - t0
    log(args, f"64 sequential fp64 calls: {seq_s:.2f} s")
    cutc.release_workspace()
    out = {"metric": "batch64: int_2e_ao build s + eff TFLOP/s, 64 systems, one graph launch vs 64 FP64 calls",
           "unit": "TFLOP/s", "n_gpus": 1, "steps": 1, "warmup": 1, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "data": "synthetic",
           "config": {"workload": f"batch64: 64 systems, n_nuc U{{3..12}}, n_ao U{{30..150}}, "
                                  f"n1 U{{5k..30k}} x {args.batch_scale}, n2 U{{20k..100k}} x {args.batch_scale}",
                      "n_ao_range": [min(s.n_ao for s in sys_), max(s.n_ao for s in sys_)]},
           "w_ref_flops": W, "sequential_fp64_s": seq_s, "sequential_fp64_eff_tflops": W / seq_s / 1e12,
           "batch": {}}
    runs = [("fp64", 1)] + [(p, args.batch_streams) for p in BATCH_PRECISIONS]
    for prec, nstreams in runs:
        opt = cutc.options(prec, streams=nstreams)
        first = None
        if prec in ("int8", "bf16crt"):
            # the modular batches spread over several graph branches, whose
            # workspace the first call allocates: time the call after it
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st0 = cutc.cutc_int_batch_device(dss, Ib, opt, stream)
            torch.cuda.synchronize()
            first = {"call_wall_s": time.perf_counter() - t0, "setup_ms": st0["ms_setup"],
                     "graph_launch_s": st0["ms_launch"] / 1e3}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = cutc.cutc_int_batch_device(dss, Ib, opt, stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        errs = []
        for I, Ir in zip(Ib, Iseq):
            d = float(torch.linalg.norm(I - Ir) / torch.linalg.norm(Ir))
            errs.append(d)
        cutc.release_workspace()
        res = {"graph_launch_s": st["ms_launch"] / 1e3, "call_wall_s": wall, "setup_ms": st["ms_setup"],
               "autotune_ms": st["ms_tune"], "graph_nodes": st["n_nodes"], "systems_split_k": st["n_split"],
               "branches": st["n_branches"],
               "eff_tflops": W / (st["ms_launch"] / 1e3) / 1e12, "speedup_vs_sequential_fp64": seq_s / wall,
               "max_rel_frob_vs_sequential_fp64": max(errs)}
        if first is not None:
            res["first_call"] = first
        log(args, f"batch {prec}: {res}")
        label = f"{prec}x{opt.slices}" if prec != "fp64" else prec
        out["batch"][label + ("_1branch" if nstreams == 1 and prec == "fp64" else "")] = res
    best = out["batch"]["fp64"]
    out.update({"value": best["eff_tflops"], "ms_per_step": best["graph_launch_s"] * 1e3, "dtype": "f64"})
    print(json.dumps(out), flush=True)


class stdout_to_stderr:
    """Send what C code prints to stdout (the reference entry points print
    their progress lines) to stderr, so that stdout carries only the JSON
    line."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        os.dup2(2, 1)

    def __exit__(self, *exc):
        os.dup2(self.saved, 1)
        os.close(self.saved)


def no_bench(args, torch, cutc):
    """Normal ordering (cutc_no_device) on synthetic MO-basis inputs resident in
    HBM: int2_grad1_u12 (n_grid1, 3, n_mo, n_mo), orbitals on the grid, weights.
    The work is dominated by the 2e contractions, W = 2 n_mo^4 (3 + 1) n_grid1."""
    nm = 200
    ng = args.n1 or 60000
    na = nb = 30
    gen = torch.Generator(device="cuda")
    gen.manual_seed(20261206)
    G = torch.randn((nm, nm, 3, ng), dtype=torch.float64, device="cuda", generator=gen)
    w = torch.rand(ng, dtype=torch.float64, device="cuda", generator=gen)
    L = torch.randn((nm, ng), dtype=torch.float64, device="cuda", generator=gen)
    R = L + 0.1 * torch.randn((nm, ng), dtype=torch.float64, device="cuda", generator=gen)
    o2 = torch.empty((nm,) * 4, dtype=torch.float64, device="cuda")
    o1 = torch.empty((nm, nm), dtype=torch.float64, device="cuda")
    o0 = torch.empty(1, dtype=torch.float64, device="cuda")
    W = 2.0 * nm ** 4 * 4 * ng
    stream = torch.cuda.current_stream()
    for i in range(args.warmup):
        st = cutc.cutc_no_device(na, nb, w, L, R, G, o2, o1, o0, stream=stream)
        log(args, f"warmup {i}: {st['ms_total'] / 1e3:.2f} s")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    sts = []
    torch.cuda.synchronize()
    ev[0].record(stream)
    for i in range(args.steps):
        sts.append(cutc.cutc_no_device(na, nb, w, L, R, G, o2, o1, o0, stream=stream))
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    step_ms = [ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps)]
    ms = ev[0].elapsed_time(ev[args.steps]) / args.steps
    last = sts[-1]
    log(args, f"cutc_no: {ms / 1e3:.2f} s {last}")
    g2 = last["ms_gemm2e"] / 1e3
    achieved = last["flops_gemm2e"] / g2 / 1e12
    out = {"metric": "cutc_no build s + eff TFLOP/s (2e contraction work 2 n_mo^4 (3+1) n_grid1)",
           "value": W / (ms / 1e3) / 1e12, "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "ms_per_step_median": statistics.median(step_ms),
           "step_ms": step_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": f"normal ordering: n_mo={nm}, n_grid1={ng}, ne_a=ne_b={nb} (closed shell)",
                      "chunk": last["chunk"], "n_chunks": last["n_chunks"]},
           "build_s": ms / 1e3, "w_flops": W,
           "stage_ms": {k: last[k] for k in ("ms_point", "ms_pair", "ms_gemm2e", "ms_1e", "ms_symm")},
           "roofline": {"kernel": "gemm_f64_kernel<0,0,7> (2e contractions, 4 launches per chunk)", "bound": "mfma",
                        "achieved": achieved, "peak": PEAK["fp64"], "unit": "TFLOP/s",
                        "frac": achieved / PEAK["fp64"], "traffic": None}}
    del o2, o1, o0
    if args.host_path:
        # the reference's entry point on host arrays: uploads, build, downloads
        hG, hw, hL, hR = (t.cpu().numpy() for t in (G, w, L, R))
        del G, L, R, w
        cutc.release_workspace()
        torch.cuda.empty_cache()
        hp = {}
        for mode in ("first_call", "repeated"):
            with stdout_to_stderr():
                t0 = time.perf_counter()
                r = cutc.cutc_no(na, nb, hw, hL, hR, hG)
                wall = time.perf_counter() - t0
            del r
            hp[mode] = {"wall_s": wall, "eff_tflops": W / wall / 1e12}
            log(args, f"cutc_no on host arrays ({mode}): {wall:.2f} s")
        hp["note"] = ("cutc_no (host arrays, the reference's entry point): int2_grad1_u12 (57.6 GB) and the other "
                      "inputs uploaded, no_2e (12.8 GB), no_1e, no_0e downloaded to pageable host memory")
        out["host_path"] = hp
        del hG
    else:
        del G, L, R, w
    cutc.release_workspace()
    torch.cuda.empty_cache()
    out["cpu_baseline"] = None
    if args.cpu_sample > 0:
        import numpy as np
        from oracle import oracle
        rng = np.random.default_rng(1)
        n = args.cpu_sample
        Gs = rng.standard_normal((nm, nm, 3, n))
        ws = rng.random(n)
        Ls = rng.standard_normal((nm, n))
        Rs = Ls + 0.1 * rng.standard_normal((nm, n))
        oracle.lib()
        t0 = time.perf_counter()
        oracle.cutc_no(na, nb, ws, Ls, Rs, Gs)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": 2.0 * nm ** 4 * 4 * n / dt / 1e12, "unit": "TFLOP/s",
                               "cores": oracle.threads(), "kind": "port",
                               "sample": f"FP64 oracle ({oracle.blas_name()} DGEMM), n_mo={nm}, {n} of {ng} grid "
                                         f"points: {dt:.2f} s; value = W(sample) / s", "seconds": dt}
    print(json.dumps(out), flush=True)


def mo_bench(args, torch, cutc):
    """AO -> MO transform (cutc_ao_to_mo_device) of a resident
    int2_grad1_u12_ao (n_ao, n_ao, n_grid1, 3) into int2_grad1_u12 (n_grid1,
    3, n_mo, n_mo): two FP64 GEMMs, W = 6 n_grid1 n_ao n_mo (n_ao + n_mo)."""
    na = nm = 200
    n1 = args.n1 or 60000
    gen = torch.Generator(device="cuda")
    gen.manual_seed(20261207)
    X = torch.randn((3, n1, na, na), dtype=torch.float64, device="cuda", generator=gen)
    cl = torch.randn((nm, na), dtype=torch.float64, device="cuda", generator=gen) / na ** 0.5
    cr = torch.randn((nm, na), dtype=torch.float64, device="cuda", generator=gen) / na ** 0.5
    out = torch.empty((nm, nm, 3, n1), dtype=torch.float64, device="cuda")
    W = 6.0 * n1 * na * nm * (na + nm)
    stream = torch.cuda.current_stream()
    for i in range(args.warmup):
        st = cutc.cutc_ao_to_mo_device(cl, cr, X, out, stream=stream)
        log(args, f"warmup {i}: {st['ms_total']:.1f} ms")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    sts = []
    torch.cuda.synchronize()
    ev[0].record(stream)
    for i in range(args.steps):
        sts.append(cutc.cutc_ao_to_mo_device(cl, cr, X, out, stream=stream))
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    step_ms = [ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps)]
    ms = ev[0].elapsed_time(ev[args.steps]) / args.steps
    last = sts[-1]
    log(args, f"ao_to_mo: {ms:.1f} ms {last}")
    gemm_ms = last["ms_left"] + last["ms_right"]
    achieved = last["flops"] / (gemm_ms / 1e3) / 1e12
    # HBM bytes the two GEMMs must move at least: A read, Z written and read, out written
    hbm = 8.0 * (3 * n1) * (na * na + 2 * na * nm + nm * nm)
    out_js = {"metric": "cutc_ao_to_mo ms + eff TFLOP/s (W = 6 n_grid1 n_ao n_mo (n_ao + n_mo))",
              "value": W / (ms / 1e3) / 1e12, "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps,
              "warmup": args.warmup, "ms_per_step": ms, "ms_per_step_median": statistics.median(step_ms),
              "step_ms": step_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
              "data": "synthetic",
              "config": {"workload": f"AO->MO transform: n_ao={na}, n_mo={nm}, n_grid1={n1}",
                         "mo_chunk": last["mo_chunk"], "n_chunks": last["n_chunks"]},
              "w_flops": W, "min_hbm_bytes": hbm, "hbm_GBps_at_min_bytes": hbm / (ms / 1e3) / 1e9,
              "stage_ms": {"ms_left": last["ms_left"], "ms_right": last["ms_right"]},
              "roofline": {"kernel": "gemm_f64_kernel<3,3,7> (2 launches; even n_ao: 16-byte pair loads)", "bound": "mfma", "achieved": achieved,
                           "peak": PEAK["fp64"], "unit": "TFLOP/s", "frac": achieved / PEAK["fp64"],
                           "traffic": MEASURED_TRAFFIC["mo"] if (n1, na, nm) == (60000, 200, 200) else None,
                           "traffic_note": "HBM-side bytes per launch (L2 fills x2 + write-backs), "
                                           "profiles/ao_to_mo_nao200_kernel_stats.txt"}}
    del out
    if args.host_path:
        # the host-array entry point: uploads, transform, download
        hX, hcl, hcr = X.cpu().numpy(), cl.cpu().numpy(), cr.cpu().numpy()
        del X
        cutc.release_workspace()
        torch.cuda.empty_cache()
        hp = {}
        for mode in ("first_call", "repeated"):
            with stdout_to_stderr():
                t0 = time.perf_counter()
                r = cutc.cutc_ao_to_mo(hcl, hcr, hX)
                wall = time.perf_counter() - t0
            del r
            hp[mode] = {"wall_s": wall}
            log(args, f"cutc_ao_to_mo on host arrays ({mode}): {wall:.2f} s")
        hp["note"] = ("cutc_ao_to_mo on host arrays: 57.6 GB uploaded, 57.6 GB downloaded (pageable host memory) "
                      "through the transfer pipeline")
        out_js["host_path"] = hp
        del hX
    else:
        del X
    del cl, cr
    cutc.release_workspace()
    torch.cuda.empty_cache()
    out_js["cpu_baseline"] = None
    if args.cpu_sample > 0:
        import numpy as np
        from oracle import oracle
        rng = np.random.default_rng(1)
        n = 64 * args.cpu_sample
        Xs = rng.standard_normal((3, n, na, na))
        cls, crs = rng.standard_normal((nm, na)), rng.standard_normal((nm, na))
        oracle.lib()
        t0 = time.perf_counter()
        oracle.ao_to_mo(cls, crs, Xs)
        dt = time.perf_counter() - t0
        out_js["cpu_baseline"] = {"value": 6.0 * n * na * nm * (na + nm) / dt / 1e12, "unit": "TFLOP/s",
                                  "cores": oracle.threads(), "kind": "port",
                                  "sample": f"FP64 oracle ({oracle.blas_name()} DGEMM per grid point), n_ao=n_mo="
                                            f"{na}, {n} of {n1} grid points: {dt:.2f} s; value = W(sample) / s",
                                  "seconds": dt}
    print(json.dumps(out_js), flush=True)


def ao_bench(args, torch, cutc, synth):
    """AO values and gradients (cutc_aos_device) on both grids of the
    20-nucleus workload (60000 + 200000 points, n_ao = 200), each AO of the
    synthetic basis contracted over 3 primitives.  HBM-bound: 32 bytes
    written per (point, AO)."""
    import numpy as np
    s = synth.cube20(n1=args.n1 or 60000, n2=200000)
    b = dict(s.basis)
    f = np.array([0.5, 1.0, 2.0])
    b["ao_expo"] = np.ascontiguousarray(b["ao_expo"] * f)
    b["ao_coef"] = np.ascontiguousarray(b["ao_coef"] * np.array([0.3, 0.5, 0.2]))
    b["ao_prim_num"] = np.full(s.n_ao, 3, np.int32)
    db = cutc.upload_basis(b)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    r = t(np.concatenate([s.r1, s.r2]))
    rn = t(s.rn)
    n, n_ao = r.shape[0], s.n_ao
    out = torch.empty((4, n_ao, n), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        cutc.cutc_aos_device(r, rn, db, out, stream=stream)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    ev[0].record(stream)
    for i in range(args.steps):
        cutc.cutc_aos_device(r, rn, db, out, stream=stream)
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    step_ms = [ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps)]
    ms = ev[0].elapsed_time(ev[args.steps]) / args.steps
    nbytes = 32.0 * n * n_ao
    # what the device evaluation replaces: the upload of the same aos_data from pageable host memory
    host = np.concatenate([s.aos1, s.aos2], axis=2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev = torch.from_numpy(host).to("cuda")
    torch.cuda.synchronize()
    h2d = time.perf_counter() - t0
    del dev, host
    out_js = {"metric": "cutc_aos_device ms + GB/s of aos_data written", "value": nbytes / (ms / 1e3) / 1e9,
              "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
              "ms_per_step_median": statistics.median(step_ms), "step_ms": step_ms, "higher_is_better": True,
              "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
              "config": {"workload": f"AOs of the cube20 workload: n_ao={n_ao} (3 primitives each), "
                                     f"{s.n1} + {s.n2} grid points"},
              "roofline": {"kernel": "aos_kernel", "bound": "hbm", "achieved": nbytes / (ms / 1e3) / 1e9,
                           "peak": 8000.0, "unit": "GB/s", "frac": nbytes / (ms / 1e3) / 1e9 / 8000.0,
                           "traffic": MEASURED_TRAFFIC["ao"] if n == 260000 else None, "bytes_per_launch": nbytes,
                           "traffic_note": "HBM-side bytes per launch (L2 fills x2 + write-backs), "
                                           "profiles/aos_cube20_kernel_stats.txt"},
              "h2d_pageable_ms_same_bytes": h2d * 1e3}
    log(args, f"aos: {ms:.2f} ms, {out_js['value']:.0f} GB/s; upload of the same bytes {h2d * 1e3:.0f} ms")
    if args.host_path:
        rr = np.concatenate([s.r1, s.r2])
        hp = {}
        for mode in ("first_call", "repeated"):
            with stdout_to_stderr():
                t0 = time.perf_counter()
                r_ = cutc.cutc_aos(rr, s.rn, b)
                wall = time.perf_counter() - t0
            del r_
            hp[mode] = {"wall_s": wall}
            log(args, f"cutc_aos on host arrays ({mode}): {wall * 1e3:.0f} ms")
        hp["note"] = "cutc_aos on host arrays: the 1.664 GB of aos_data downloaded through the transfer pipeline"
        out_js["host_path"] = hp
    del out, r
    torch.cuda.empty_cache()
    out_js["cpu_baseline"] = None
    if args.cpu_sample > 0:
        from oracle import oracle
        oracle.lib()
        t0 = time.perf_counter()
        for rg in (s.r1, s.r2):
            oracle.aos(rg, s.rn, **b)
        dt = time.perf_counter() - t0
        out_js["cpu_baseline"] = {"value": nbytes / dt / 1e9, "unit": "GB/s", "cores": oracle.threads(),
                                  "kind": "port", "sample": f"FP64 oracle (C, OpenMP), the whole workload ({n} points, "
                                                            f"n_ao={n_ao}): {dt:.2f} s", "seconds": dt}
    print(json.dumps(out_js), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (no multi-GPU path in this build)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    if args.traffic_probe:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log(args, "traffic probe: copied 1 GiB device to device")

    if args.workload == "batch64":
        from cutc_amd import cutc, synth
        batch_bench(args, torch, cutc, synth)
        return
    if args.workload == "no":
        from cutc_amd import cutc
        no_bench(args, torch, cutc)
        return
    if args.workload == "mo":
        from cutc_amd import cutc
        mo_bench(args, torch, cutc)
        return
    if args.workload == "ao":
        from cutc_amd import cutc, synth
        ao_bench(args, torch, cutc, synth)
        return

    from cutc_amd import cutc, synth

    t0 = time.perf_counter()
    streamed = args.workload == "cube40"
    if streamed:
        s = synth.cube40(n1=args.n1 or 80000, n2=args.n2 or 300000)
    else:
        s = synth.cube20(n1=args.n1 or 60000, n2=args.n2 or 200000)
    log(args, f"generated {s.shape_str()} in {time.perf_counter() - t0:.1f} s")
    ds = cutc.DeviceSystem.upload(s)
    if streamed:
        X = None
        I = torch.empty((s.n_ao,) * 4, dtype=torch.float64, device="cuda")
    else:
        X, I = cutc.alloc_outputs(ds)
    stream = torch.cuda.current_stream()
    n_ao, P = s.n_ao, s.n_ao * (s.n_ao + 1) // 2
    W = w_ref(n_ao, s.n1, s.n2)
    Dh = torch.from_numpy(density(n_ao)).reshape(-1)
    D = Dh.cuda()

    def dd(Imat):  # sum D_kl D_ab (kl|ab)
        return float(D @ (Imat.reshape(n_ao * n_ao, n_ao * n_ao) @ D))

    def errors_vs_fp64():
        """errors of the split build in X, I against the FP64 build, every tensor"""
        err = {}
        # the FP64 results are kept on the host (the split builds get the
        # whole device): compare in slabs
        num = den = mx = rmx = 0.0
        for a0 in range(0, n_ao, 10):
            b = Iref[a0:a0 + 10].to(I.device)
            d = I[a0:a0 + 10] - b
            num += float((d * d).sum())
            den += float((b * b).sum())
            mx = max(mx, float(d.abs().max()))
            rmx = max(rmx, float(b.abs().max()))
            del b, d
        err["int_2e_ao"] = {"rel_frob": (num / den) ** 0.5 if den > 0 else 0.0, "max_abs": mx, "ref_max_abs": rmx}
        for m in range(0 if streamed else 4):
            num = den = mx = rmx = 0.0
            step = 2000
            for r0 in range(0, s.n1, step):
                a = X[m, r0:r0 + step]
                b = Xref_host[m, r0:r0 + step].to(a.device)
                d = a - b
                num += float((d * d).sum())
                den += float((b * b).sum())
                mx = max(mx, float(d.abs().max()))
                rmx = max(rmx, float(b.abs().max()))
            err[f"X{m}"] = {"rel_frob": (num / den) ** 0.5 if den > 0 else 0.0, "max_abs": mx, "ref_max_abs": rmx}
        dd_split = dd(I)
        err["DD_contraction"] = {"value": dd_split, "fp64_value": dd_ref,
                                 "rel": abs(dd_split - dd_ref) / abs(dd_ref)}
        err["max_rel_frob"] = max(v["rel_frob"] for v in err.values() if isinstance(v, dict) and "rel_frob" in v)
        return err

    out = {"metric": METRIC}
    fp64 = None
    Xref_host = None
    if args.precision == "fp64" or not args.no_fp64:
        small = s.subset_r1(min(256, s.n1))
        dsm = cutc.DeviceSystem.upload(small)
        Xs, Is = cutc.alloc_outputs(dsm)
        build(cutc, dsm, None if streamed else Xs, Is, cutc.options("fp64"), stream)  # loads the kernels
        del Xs, Is, dsm
        nrep = args.steps if args.precision == "fp64" else 1
        fp64_step_ms = []
        ms, st = timed_builds(cutc, torch, ds, X, I, cutc.options("fp64"), nrep, stream, fp64_step_ms)
        log(args, f"fp64 build: {ms / 1e3:.2f} s  {st[-1]}")
        fp64 = {"build_s": ms / 1e3, "eff_tflops": W / (ms / 1e3) / 1e12, "stats": st[-1]}
        if args.precision != "fp64":
            cutc.release_workspace()  # the device workspace of one precision is not reused by the other
            # the FP64 results go to host memory and their DD contraction is
            # taken there, so that the split builds plan with the same device
            # memory as without the FP64 build
            Iref = I.cpu()
            dd_ref = float(Dh @ (Iref.reshape(n_ao * n_ao, n_ao * n_ao) @ Dh))
            Xref_host = None if streamed else X.cpu()
            torch.cuda.empty_cache()

    if args.precision == "fp64":
        ms, last = fp64["build_s"] * 1e3, fp64["stats"]
        step_ms = fp64_step_ms
        prec_label, opt = "fp64", cutc.options("fp64")
    else:
        opt = cutc.options(args.precision, args.slices)
        prec_label = f"{args.precision}x{opt.slices}"
        for i in range(args.warmup):
            st = build(cutc, ds, X, I, opt, stream)
            log(args, f"warmup {i}: {st['ms_total'] / 1e3:.2f} s")
        step_ms = []
        ms, sts = timed_builds(cutc, torch, ds, X, I, opt, args.steps, stream, step_ms)
        last = sts[-1]
        log(args, f"{prec_label} build: {ms / 1e3:.2f} s  {last}")
    cutc.release_workspace()  # room for the error analysis below

    # roofline of the dominant kernel (the step-1 contraction), on the
    # algorithmic work (packed pairs, unpadded n2); launch times are the
    # tracer's HIP events on the build's stream
    exec1 = 2.0 * P * 4 * s.n1 * s.n2
    default_shape = (not streamed and (s.n1, s.n2) == (60000, 200000)
                     and (args.precision == "fp64" or (args.precision == "int8" and opt.slices == 11)))
    traffic = MEASURED_TRAFFIC[args.precision] if default_shape else None
    if args.precision == "fp64":
        launches = last["n_gemm1"]
        achieved = exec1 / (last["ms_gemm1"] / 1e3) / 1e12
        roof = {"kernel": "gemm_f64_kernel (step-1 contraction)", "bound": "mfma", "achieved": achieved,
                "peak": PEAK["fp64"], "unit": "TFLOP/s", "frac": achieved / PEAK["fp64"], "traffic": traffic,
                "launches": launches, "avg_launch_ms": last["ms_gemm1"] / max(1, launches),
                "work": f"2 x {P} pairs x 4 n1 x n2 FP64 flops",
                "traffic_note": "HBM-side bytes per launch (L2 fills x2 + write-backs), profiles/cube20_traffic.txt"}
    elif args.precision in ("bf16", "fp16"):
        launches = last["n_gemm1"]
        npairs = opt.slices * (opt.slices + 1) // 2
        sl_flops = 2.0 * npairs * P * 4 * s.n1 * s.n2
        achieved = sl_flops / (last["ms_gemm1"] / 1e3) / 1e12
        roof = {"kernel": f"oz_gemm_kernel<{opt.slices}> (step-1 contraction, one launch per r1 chunk)",
                "bound": "mfma", "achieved": achieved, "peak": PEAK[args.precision], "unit": "TFLOP/s",
                "frac": achieved / PEAK[args.precision], "traffic": None, "launches": launches,
                "avg_launch_ms": last["ms_gemm1"] / max(1, launches),
                "work": f"2 x {npairs} slice products x {P} pairs x 4 n1 x n2 {args.precision} flops",
                "fp64_equiv_tflops": exec1 / (last["ms_gemm1"] / 1e3) / 1e12}
    else:
        # int8: one launch per modulus and r1 chunk (chunks of >= 1024 tiles,
        # as on this workload); 16-bit: one launch per r1 chunk, all moduli
        i8 = args.precision == "int8"
        launches = last["n_gemm1"] * (opt.slices if i8 else 1)
        # the list-driven contraction runs k_density of the (256-pair tile,
        # K step) tiles: the executed work is the dense work times that
        dens = last["k_density"] if last["k_density"] > 0 else 1.0
        ops = 2.0 * opt.slices * P * 4 * s.n1 * s.n2 * dens
        achieved = ops / (last["ms_gemm1"] / 1e3) / 1e12
        roof = {"kernel": ("crt_gemm_pk_kernel (step-1 contraction, one launch per modulus and r1 chunk)" if i8
                           else "crt16_gemm_kernel (step-1 contraction, one launch per r1 chunk, all moduli)"),
                "bound": "mfma", "achieved": achieved, "peak": PEAK[args.precision],
                "unit": "TOPS (int8)" if i8 else "TFLOP/s", "frac": achieved / PEAK[args.precision],
                "traffic": traffic, "launches": launches, "avg_launch_ms": last["ms_gemm1"] / max(1, launches),
                "k_density": dens,
                "work": f"2 x {opt.slices} moduli x {P} pairs x 4 n1 x n2 x k_density "
                        f"{'int8 ops' if i8 else 'flops'} (the listed tiles)",
                "fp64_equiv_tflops": exec1 / (last["ms_gemm1"] / 1e3) / 1e12,
                "traffic_note": "HBM-side bytes per launch (L2 fills x2 + write-backs), profiles/cube20_traffic.txt"}

    out.update({
        "value": W / (ms / 1e3) / 1e12,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "ms_per_step_median": statistics.median(step_ms),
        "step_ms": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else f"{prec_label}->f64",
        "data": "synthetic",
        "config": {"workload": (f"cube40: 40 nuclei, n_ao={n_ao}, n1={s.n1}, n2={s.n2}, size_bh=6, streamed "
                                "(X never stored)") if streamed else
                               f"cube20: 20 nuclei, n_ao={n_ao}, n1={s.n1}, n2={s.n2}, size_bh=6",
                   "precision": prec_label, "n1_chunk": last["n1_chunk"]},
        "build_s": ms / 1e3,
        "w_ref_flops": W,
        "w_exec_fp64_equiv": last["flops_exec"],
        "stage_ms": {k: last[k] for k in ("ms_jastrow", "ms_integrands", "ms_split", "ms_gemm1", "ms_gemm2",
                                          "ms_symm")},
        "roofline": roof,
    })
    if fp64 is not None and args.precision != "fp64":
        out["fp64_build_s"] = fp64["build_s"]
        out["fp64_eff_tflops"] = fp64["eff_tflops"]
        out["speedup_vs_fp64"] = fp64["build_s"] / (ms / 1e3)
        fp64_g1 = fp64["stats"]["ms_gemm1"] / 1e3
        out["fp64_roofline"] = {"kernel": "gemm_f64_kernel<0,0,*>", "achieved": exec1 / fp64_g1 / 1e12,
                                "peak": PEAK["fp64"], "frac": exec1 / fp64_g1 / 1e12 / PEAK["fp64"]}
        err = errors_vs_fp64()
        out["errors_vs_fp64"] = err
        out["meets_1e-10"] = err["max_rel_frob"] <= 1e-10
        if not args.no_bf16 and args.precision not in ("bf16crt", "fp16crt"):
            # the BF16 leg: the contractions as bf16 modular products at the library default
            bopt = cutc.options("bf16crt")
            try:
                for _ in range(args.variant_warmup):
                    build(cutc, ds, X, I, bopt, stream)
                bms, bst = timed_builds(cutc, torch, ds, X, I, bopt, 1, stream)
            except cutc.CutcError as exc:
                cutc.release_workspace()
                out["bf16_split"] = {"precision": "bf16crt", "error": str(exc)}
            else:
                cutc.release_workspace()
                b = bst[-1]
                berr = errors_vs_fp64()
                bdens = b["k_density"] if b["k_density"] > 0 else 1.0
                bops = 2.0 * bopt.slices * P * 4 * s.n1 * s.n2 * bdens
                blaunch = b["n_gemm1"]  # one launch per r1 chunk, all moduli
                out["bf16_split"] = {
                    "precision": f"bf16crt x{bopt.slices} moduli (bf16 MFMA, FP32 accumulation kept exact, CRT "
                                 "rebuild in FP64)",
                    "build_s": bms / 1e3, "eff_tflops": W / (bms / 1e3) / 1e12,
                    "speedup_vs_fp64": fp64["build_s"] / (bms / 1e3), "errors_vs_fp64": berr,
                    "meets_1e-10": berr["max_rel_frob"] <= 1e-10,
                    "stage_ms": {k: b[k] for k in ("ms_jastrow", "ms_integrands", "ms_split", "ms_gemm1",
                                                   "ms_gemm2", "ms_symm")},
                    "roofline": {"kernel": "crt16_gemm_kernel (step-1 contraction, one launch per r1 chunk, "
                                           "all moduli)", "bound": "mfma",
                                 "achieved": bops / (b["ms_gemm1"] / 1e3) / 1e12, "peak": PEAK["bf16"],
                                 "unit": "TFLOP/s", "frac": bops / (b["ms_gemm1"] / 1e3) / 1e12 / PEAK["bf16"],
                                 "launches": blaunch, "avg_launch_ms": b["ms_gemm1"] / max(1, blaunch),
                                 "k_density": bdens,
                                 "work": f"2 x {bopt.slices} moduli x {P} pairs x 4 n1 x n2 x k_density bf16 "
                                         "flops (the listed tiles)"}}
                log(args, f"bf16 leg: {bms / 1e3:.2f} s, {fp64['build_s'] / (bms / 1e3):.2f}x FP64, "
                          f"max rel frob {berr['max_rel_frob']:.2e}")
        # further split variants against the same FP64 build (one warm-up build each)
        variants = []
        for v in filter(None, args.variants.split(",")):
            vp, vs = v.split(":")
            vopt = cutc.options(vp, int(vs))
            try:
                for _ in range(args.variant_warmup):
                    build(cutc, ds, X, I, vopt, stream)
                vms, vst = timed_builds(cutc, torch, ds, X, I, vopt, 1, stream)
            except cutc.CutcError as exc:  # e.g. a working set that does not fit
                cutc.release_workspace()
                variants.append({"precision": f"{vp}x{vs}", "error": str(exc)})
                log(args, f"variant {vp}x{vs}: {exc}")
                continue
            cutc.release_workspace()
            verr = errors_vs_fp64()
            variants.append({"precision": f"{vp}x{vs}", "build_s": vms / 1e3, "eff_tflops": W / (vms / 1e3) / 1e12,
                             "speedup_vs_fp64": fp64["build_s"] / (vms / 1e3), "stage_ms": {
                                 k: vst[-1][k] for k in ("ms_jastrow", "ms_split", "ms_gemm1", "ms_gemm2")},
                             "errors_vs_fp64": verr})
            log(args, f"variant {vp}x{vs}: {vms / 1e3:.2f} s, max rel frob {verr['max_rel_frob']:.2e}")
        if variants:
            out["variants"] = variants
    del X, I, ds
    torch.cuda.empty_cache()
    if args.host_path and not streamed:
        Xref_host = Iref = None
        hp = {}
        cutc.release_workspace()
        # first call allocates the device workspace; the second (pipelined) and
        # third (blocking) reuse it
        for mode, hc in (("pipelined_first_call", 0), ("pipelined", 0), ("blocking", 1)):
            hopt = cutc.options(precision=opt.precision, slices=opt.slices, host_copy=hc)
            t0 = time.perf_counter()
            Xh, Ih, hst = cutc.cutc_int_ex(s, hopt)
            wall = time.perf_counter() - t0
            del Xh, Ih
            hp[mode] = {"wall_s": wall, "eff_tflops": W / wall / 1e12, "build_ms": hst["ms_total"],
                        "n1_chunks": hst["n1_chunks"], "ms_gemm1": hst["ms_gemm1"], "ms_gemm2": hst["ms_gemm2"],
                        "ms_split": hst["ms_split"], "ms_host_in": hst["ms_host_in"],
                        "ms_host_out": hst["ms_host_out"]}
            log(args, f"host path ({mode}): {wall:.2f} s, build {hst['ms_total'] / 1e3:.2f} s, "
                      f"{hst['n1_chunks']} chunks, allocation + upload {hst['ms_host_in'] / 1e3:.2f} s, "
                      f"after the build {hst['ms_host_out'] / 1e3:.2f} s")
        cutc.release_workspace()
        hp["note"] = ("cutc_int_ex on host arrays: inputs uploaded, X[0..2] and int_2e_ao downloaded to pageable "
                      "(freshly allocated) host memory; pipelined: each r1 chunk of X is downloaded while the build "
                      "runs; blocking: hipMemcpy after the build, as the reference")
        out["host_path"] = hp
    out["cpu_baseline"] = cpu_baseline(synth, s, args.cpu_sample) if args.cpu_sample > 0 else None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
