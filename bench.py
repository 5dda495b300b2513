#!/usr/bin/env python3
"""Benchmark of the FP64 partial-pivot LU hot path (getrf + getrs) on one MI355X.

Workload: plain FP64 LU at scale, n = 16384, A iid N(0,1), 16 right-hand sides iid N(0,1)
(numpy PCG64 seeded; the same bytes feed the GPU and the CPU baseline).  One step = the
LinearSolve `init`+`solve!` work of LUFactorization: copy A into the resident factor buffer,
getrf, getrs with 16 right-hand sides, inputs already resident in HBM.

Prints one JSON line.  `value` = TFLOP/s of one step (luflop(n) + 2 n^2 nrhs flops, the
reference's own flop convention, lib/LinearSolveAutotune/src/algorithms.jl:147-159).
"""
import argparse
import ctypes
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0


def load_package():
    name = "linearsolve_jl_amd"
    if name in sys.modules:
        return sys.modules[name]
    d = os.path.join(ROOT, "linearsolve.jl_amd")
    spec = importlib.util.spec_from_file_location(name, os.path.join(d, "__init__.py"),
                                                  submodule_search_locations=[d])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def luflop(n):
    return n + n * (n - 1) / 2 + (n - 1) * n * (2 * n - 1) / 3


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary, if present."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    ent = d.get(kernel)
    return None if ent is None else ent.get("hbm_bytes_per_launch")


def cpu_baseline(A, B, n, nrhs):
    """LAPACK dgetrf + dgetrs (OpenBLAS, the library the reference's LUFactorization calls) on the
    host cores, on a bounded sample of the same workload."""
    import scipy
    import scipy.linalg.lapack as lapack
    try:
        from threadpoolctl import threadpool_info
        cores = max((i.get("num_threads", 1) for i in threadpool_info() if i.get("internal_api") == "openblas"),
                    default=os.cpu_count())
    except Exception:
        cores = os.cpu_count()
    # size the sample so that it stays within ~30 s: time n/4 first
    m = n // 4
    t0 = time.perf_counter()
    lapack.dgetrf(np.asfortranarray(A[:m, :m]))
    t_small = time.perf_counter() - t0
    ns = n if t_small * 64 < 25.0 else n // 2
    As = A if ns == n else np.asfortranarray(A[:ns, :ns])
    t0 = time.perf_counter()
    lu, piv, info = lapack.dgetrf(As, overwrite_a=False)
    x, _ = lapack.dgetrs(lu, piv, B[:ns, :nrhs])
    t = time.perf_counter() - t0
    flops = luflop(ns) + 2.0 * ns * ns * nrhs
    return dict(value=flops / t / 1e12, unit="TFLOP/s", cores=int(cores), kind="reference",
                sample=f"one dgetrf+dgetrs ({nrhs} RHS) at n={ns} of the same matrix, OpenBLAS "
                       f"{getattr(scipy, '__version__', '?')} via scipy (LAPACK, the library "
                       f"LUFactorization calls; julia is not installed), {t:.2f} s"), (piv + 1 if ns == n else None)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--nrhs", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the dense LU would need a 2D block-cyclic layout with per-column pivot all-reduces)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    lsx = load_package()
    from linearsolve_jl_amd import _abi

    n, nrhs = args.n, args.nrhs
    log(f"generating A ({n}x{n}) and B ({n}x{nrhs})")
    rng = np.random.default_rng(np.random.SeedSequence([2, n]))
    G = rng.standard_normal((n, n))          # C-order buffer == column-major A^T ... A := G.T
    Gb = rng.standard_normal((nrhs, n))
    A = G.T                                  # column-major view, no copy
    B = Gb.T
    dA = torch.from_numpy(G).cuda()          # device bytes == column-major A
    dB = torch.from_numpy(Gb).cuda()
    dX = torch.empty((nrhs, n), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()

    ctx = lsx.Context.get(0)
    L = _abi.lib()
    h = ctypes.c_void_p()
    ctx.check(L.lsx_lu_create(ctx.handle, n, _abi.LSX_F64, ctypes.byref(h)), "lsx_lu_create")

    def step():
        info = ctx.check(L.lsx_lu_factor(h, dA.data_ptr(), n), "lsx_lu_factor")
        ctx.check(L.lsx_lu_solve(h, nrhs, dB.data_ptr(), n, dX.data_ptr(), n), "lsx_lu_solve")
        return info

    for i in range(args.warmup):
        step()
        log(f"warmup {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    ctx.check(L.lsx_lu_profile(h, 1, None), "lsx_lu_profile")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        info = step()
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    prof = _abi.Profile()
    ctx.check(L.lsx_lu_profile(h, 0, ctypes.byref(prof)), "lsx_lu_profile")
    log(f"timed {args.steps} steps: {t / args.steps * 1e3:.2f} ms/step, info={info}")

    # correctness of the last step: backward error and pivots
    X = dX.cpu().numpy().T
    r = B - A @ X
    nA = np.abs(A).sum(axis=1).max()
    be = max(np.abs(r[:, j]).max() / (nA * np.abs(X[:, j]).max() + np.abs(B[:, j]).max()) for j in range(nrhs))
    ipiv = np.empty(n, dtype=np.int64)
    ctx.check(L.lsx_lu_get_factors(h, None, n, ipiv.ctypes.data), "lsx_lu_get_factors")
    log(f"backward error {be:.3e}")

    flops = luflop(n) + 2.0 * n * n * nrhs
    ms = t / args.steps * 1e3
    value = flops / (ms / 1e3) / 1e12
    breakdown = {}
    for i, cls in enumerate(_abi.PROFILE_CLASSES):
        if prof.launches[i]:
            breakdown[cls] = dict(launches_per_step=prof.launches[i] / args.steps,
                                  ms_per_step=prof.ms[i] / args.steps)
    g = _abi.PROFILE_CLASSES.index("gemm")
    gemm_tflops = prof.flops[g] / (prof.ms[g] / 1e3) / 1e12 if prof.ms[g] > 0 else 0.0
    roofline = dict(bound="mfma", kernel="k_gemm_f64 (trailing update A22 -= L21*U12)",
                    achieved=round(gemm_tflops, 3), peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=round(gemm_tflops / FP64_PEAK_TFLOPS, 4),
                    flops_per_launch=prof.flops[g] / max(prof.launches[g], 1),
                    avg_launch_us=prof.ms[g] / max(prof.launches[g], 1) * 1e3,
                    traffic=traffic_from_profiles("k_gemm_f64"))

    result = {
        "metric": "solves/s, TFLOP/s at <=1e-13 bwd err: mixed-IR vs FP64 LU; batched p50/p99 ms",
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (A, B iid N(0,1) from numpy PCG64 seed [2, n])",
        "config": {"workload": f"plain FP64 partial-pivot LU getrf+getrs, n={n}, {nrhs} RHS",
                   "n": n, "nrhs": nrhs, "parallelism": "single GPU"},
        "solves_per_s": round(1e3 / ms, 3),
        "backward_error": be,
        "roofline": roofline,
        "kernel_breakdown": breakdown,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (LAPACK on host cores)")
        cb, lapack_ipiv = cpu_baseline(A, B, n, nrhs)
        result["cpu_baseline"] = cb
        if lapack_ipiv is not None:
            result["pivots_bit_exact_vs_lapack"] = bool(np.array_equal(ipiv, lapack_ipiv))
    L.lsx_lu_destroy(h)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
