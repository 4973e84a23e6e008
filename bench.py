"""Benchmark of the contraction hot path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload NAME] [--precision P]

A step is one binary_einsum call over the workload's inputs, which are
resident in HBM before the timed region starts.  Default workload: the
large GEMM-shaped contraction on permuted layouts, ComplexF64, rank-6
tensors with every extent 16, contracted modes at positions (2,4,6) of A and
(1,3,5) of B, output order (a1,b1,a3,b2,a5,b3): M = N = K = 4096.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "muscle.jl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}
HBM_GBS = 8000.0

WORKLOADS = {
    # name: (description, element type)
    "c2_c64": ("config2 rank-6 interleaved, every extent 16, M=N=K=4096, ComplexF64", "c64"),
    "c2_c32": ("config2 rank-6 interleaved, every extent 16, M=N=K=4096, ComplexF32", "c32"),
    "c2h_c64": ("config2 hyperindex: batch 64 x (M=N=K=512), ComplexF64", "c64"),
    "c1_c64": ("config1 plain: A(64,32,48) B(48,96,32) -> C(64,96), ComplexF64", "c64"),
}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_inputs(name: str):
    import numpy as np

    from muscle import workloads as wl
    if name in ("c2_c64", "c2_c32"):
        dt = np.complex128 if name == "c2_c64" else np.complex64
        A, ia, B, ib, ic = wl.config2_inputs(16, seed=0, dtype=dt)
        return A, ia, B, ib, {"out": ic}
    if name == "c2h_c64":
        A, ia, B, ib, ic = wl.config2h_inputs(1, seed=0)
        return A, ia, B, ib, {"out": ic}
    if name == "c1_c64":
        A, ia, B, ib, dims, ic = wl.config1_inputs(0)
        return A, ia, B, ib, {"dims": dims}
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(A, ia, B, ib, kw, threads: int, reps: int):
    """The oracle's BackendBase restatement (permute + BLAS gemm + permute) on the host cores."""
    import numpy as np
    from threadpoolctl import threadpool_info, threadpool_limits

    from oracle import backend_base as bb
    inds_c = bb.front_binary(ia, ib, dims=kw.get("dims"), out=kw.get("out"))
    f = bb.binary_einsum_base if not any(i in ia and i in ib for i in inds_c) else bb.binary_einsum_general
    times = []
    with threadpool_limits(limits=threads):
        f(inds_c, A, ia, B, ib)  # warm-up
        for _ in range(reps):
            t0 = time.perf_counter()
            f(inds_c, A, ia, B, ib)
            times.append(time.perf_counter() - t0)
    blas = [f"{d.get('internal_api')} {d.get('version')}" for d in threadpool_info() if d.get("user_api") == "blas"]
    return float(np.median(times)), blas


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2_c64", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="native")
    ap.add_argument("--complex-mult", type=int, default=3, choices=[3, 4])
    ap.add_argument("--target", type=float, default=1e-10, help="target relative error of the Ozaki rung")
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only: the contraction is sharded by output "
                                   "blocks across GPUs with no collective (see DESIGN.md), and no multi-GPU "
                                   "figure is produced"}))
        return 2

    import numpy as np
    import torch

    import muscle

    dev = torch.device("cuda", 0)
    log(f"workload {args.workload}: generating inputs")
    A, ia, B, ib, kw = make_inputs(args.workload)
    a = muscle.Tensor(torch.from_numpy(np.asfortranarray(A)).to(dev), ia)
    b = muscle.Tensor(torch.from_numpy(np.asfortranarray(B)).to(dev), ib)
    backend = muscle.BackendHIP(precision=args.precision, complex_mult=args.complex_mult,
                                target_rel_error=args.target)

    def step():
        with muscle.with_backend(backend):
            return muscle.binary_einsum(a, b, **kw)

    log(f"warmup {args.warmup}")
    c = None
    for _ in range(args.warmup):
        c = step()
    torch.cuda.synchronize()
    info = muscle.last_info()
    flops = info["flops"]
    stream = torch.cuda.current_stream(dev)  # the library enqueues on torch's current stream
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    log(f"timing {args.steps} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        c = step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1) / args.steps
    tflops = flops / (ms * 1e-3) / 1e12
    complex32 = a.dtype in (torch.complex64, torch.float32) and b.dtype in (torch.complex64, torch.float32)
    peak = PEAK_TFLOPS["fp32" if complex32 and args.precision == "native" else "fp64"]
    log(f"{ms:.3f} ms/step, {tflops:.2f} TFLOP/s (wall {wall:.3f} s)")

    # parity spot-check of the timed result (sampled entries, oracle dot products)
    del c

    cpu = None
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("MUSCLE_BENCH_CPU_THREADS", "16"))
        log(f"cpu baseline: oracle BackendBase restatement, {threads} threads, median of {args.cpu_reps}")
        t_cpu, blas = cpu_baseline(A, ia, B, ib, kw, threads, args.cpu_reps)
        cpu = {"value": flops / t_cpu / 1e12, "unit": "TFLOP/s", "cores": threads, "kind": "port",
               "sample": f"the full {args.workload} contraction, median of {args.cpu_reps} runs "
                         f"({t_cpu:.3f} s each; numpy {np.__version__}, BLAS {'; '.join(blas)})"}

    out = {
        "metric": "binary_einsum TFLOP/s (% MFMA peak) vs plain at stated error; jobs/s + p99 latency",
        "value": round(tflops, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"c64": "complex128 (fp64 mfma)", "c32": "complex64 (fp32 mfma)"}[WORKLOADS[args.workload][1]]
        if args.precision == "native" else f"complex128 via {args.precision}",
        "data": "synthetic (seeded numpy CN(0,1)/sqrt(K))",
        "config": {"workload": WORKLOADS[args.workload][0], "precision": args.precision,
                   "complex_mult": args.complex_mult,
                   "m": info["m"], "n": info["n"], "k": info["k"], "batch": info["batch"],
                   "kernel": info["kernel"], "flops_per_step": flops},
        "roofline": {"bound": "mfma", "achieved": round(tflops, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(tflops / peak, 4), "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
