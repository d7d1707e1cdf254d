"""Benchmark of the TC integral build on one MI355X.

A step is one full build on inputs already resident in HBM: Jastrow pair
blocks -> integrands -> step-1 contraction X_kl(r1) -> step-2 contraction ->
symmetrised int_2e_ao (cutc_int_device through the C ABI).  Workload: 20
nuclei in a 12-bohr cube, n_ao = 200, n1 = 60,000, n2 = 200,000 (seeded
synthetic inputs).

value = effective TFLOP/s = W_ref / build seconds, with the reference
formulation's work W_ref = 8 n_ao^2 n1 n2 + 8 n_ao^4 n1 (full AO pairs).  The
build executes only the packed k <= l pairs; the roofline fraction is taken on
the executed work W_exec, never on W_ref.

Prints one JSON line (rank 0, N = 1).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "int_2e_ao build s + eff TFLOP/s: split-BF16 vs FP64 on GPU at \u22641e-10 rel err"
PEAK = {"fp64": 78.6, "bf16": 2516.0, "fp16": 2516.0}  # dense MFMA TFLOP/s, MI355X


def w_ref(n_ao, n1, n2):
    return 8.0 * n_ao ** 2 * n1 * n2 + 8.0 * n_ao ** 4 * n1


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "bf16", "fp16"])
    ap.add_argument("--slices", type=int, default=4)
    ap.add_argument("--n1", type=int, default=60000)
    ap.add_argument("--n2", type=int, default=200000)
    ap.add_argument("--cpu-sample", type=int, default=32, help="r1 points of the CPU-baseline sample (0: skip)")
    ap.add_argument("--quiet", action="store_true")
    return ap.parse_args()


def log(args, *a):
    if not args.quiet:
        print(*a, file=sys.stderr, flush=True)


def cpu_baseline(args, sysfull):
    """The FP64 oracle (C + OpenMP + OpenBLAS) on the first cpu_sample r1
    points of the same system, timed on this host."""
    from oracle import oracle
    s = sysfull.subset_r1(args.cpu_sample)
    oracle.lib()
    t0 = time.perf_counter()
    X = oracle.int2_grad1_u12_ao(s)
    I = oracle.int_2e_ao(s, X)
    dt = time.perf_counter() - t0
    del X, I
    w = w_ref(s.n_ao, s.n1, s.n2)
    return {"value": w / dt / 1e12, "unit": "TFLOP/s", "cores": oracle.threads(), "kind": "port",
            "sample": f"oracle ({oracle.blas_name()} DGEMM) on the first {s.n1} of {sysfull.n1} r1 points, all "
                      f"{s.n2} r2 points, n_ao={s.n_ao}: {dt:.2f} s; value = W_ref(sample)/s",
            "seconds": dt}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (no multi-GPU path in this build)",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from cutc_amd import cutc, synth

    t0 = time.perf_counter()
    s = synth.cube20(n1=args.n1, n2=args.n2)
    log(args, f"generated {s.shape_str()} in {time.perf_counter() - t0:.1f} s")
    ds = cutc.DeviceSystem.upload(s)
    X, I = cutc.alloc_outputs(ds)
    torch.cuda.synchronize()
    opt = cutc.options(args.precision, args.slices)
    stream = torch.cuda.current_stream()

    for i in range(args.warmup):
        st = cutc.cutc_int_device(ds, X, I, opt, stream)
        log(args, f"warmup {i}: {st['ms_total'] / 1e3:.2f} s")
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    stats = []
    ev0.record(stream)
    for i in range(args.steps):
        stats.append(cutc.cutc_int_device(ds, X, I, opt, stream))
        log(args, f"step {i}: {stats[-1]['ms_total'] / 1e3:.2f} s  {stats[-1]}")
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / args.steps

    W = w_ref(s.n_ao, s.n1, s.n2)
    last = stats[-1]
    # dominant kernel: the step-1 contraction; per-launch time from the library's
    # stream-ordered hipEvents around every launch
    g1_ms = last["ms_gemm1"] / max(1, last["n_gemm1"])
    P = s.n_ao * (s.n_ao + 1) // 2
    n1c = last["n1_chunk"]
    chunk_flops = 2.0 * P * 4 * n1c * s.n2  # executed FP64 work of one full-size launch
    # average over launches with their own sizes
    g1_flops_total = 2.0 * P * 4 * s.n1 * s.n2
    achieved = g1_flops_total / (last["ms_gemm1"] / 1e3) / 1e12
    peak = PEAK[args.precision] if args.precision == "fp64" else PEAK["fp64"]
    roof = {"kernel": "gemm_f64_kernel (step-1 contraction)" if args.precision == "fp64" else "split step-1",
            "bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
            "traffic": None, "launches": last["n_gemm1"], "avg_launch_ms": g1_ms,
            "flops_per_full_launch": chunk_flops}
    out = {
        "metric": METRIC,
        "value": W / (ms / 1e3) / 1e12,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else f"{args.precision}x{args.slices}->f64",
        "data": "synthetic",
        "config": {"workload": f"cube20: 20 nuclei, n_ao={s.n_ao}, n1={s.n1}, n2={s.n2}, size_bh=6",
                   "precision": args.precision, "slices": args.slices if args.precision != "fp64" else None,
                   "n1_chunk": n1c},
        "build_s": ms / 1e3,
        "w_ref_flops": W,
        "w_exec_flops": last["flops_exec"],
        "exec_tflops": last["flops_exec"] / (ms / 1e3) / 1e12,
        "stage_ms": {k: last[k] for k in ("ms_jastrow", "ms_integrands", "ms_gemm1", "ms_gemm2", "ms_symm")},
        "roofline": roof,
        "cpu_baseline": None,
    }
    del X, I, ds
    torch.cuda.empty_cache()
    if args.cpu_sample > 0:
        out["cpu_baseline"] = cpu_baseline(args, s)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
