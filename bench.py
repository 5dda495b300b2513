#!/usr/bin/env python3
"""Benchmark of the FP64 partial-pivot LU hot path (getrf + getrs) on one MI355X.

Workload: plain FP64 LU at scale, n = 16384, A iid N(0,1), 16 right-hand sides iid N(0,1)
(numpy PCG64 seeded; the same bytes feed the GPU and the CPU baseline).  One step = the
LinearSolve `init`+`solve!` work of LUFactorization: copy A into the resident factor buffer,
getrf, getrs with 16 right-hand sides, inputs already resident in HBM.

Prints one JSON line.  `value` = TFLOP/s of one step (luflop(n) + 2 n^2 nrhs flops, the
reference's own flop convention, lib/LinearSolveAutotune/src/algorithms.jl:147-159).

Beside it, in the same line (each can be switched off with --sections):
  mixed    n = 16384, A = Q1 diag(s) Q2^T with s log-spaced from 1 to 1/cond, cond in
           {1e2, 1e4, 1e6}, one right-hand side: FP32 and BF16 factorization + FP64 refinement
           (IR, then GMRES-IR where IR stalls) to backward error <= 1e-13, against the plain FP64
           LU of the same system;
  gmres    SimpleGMRES(30), Jacobi, on the 1024 x 1024 convection-diffusion grid (Pe = 5), 300
           Arnoldi steps: us per step and algorithmic HBM GB/s; the CPU restatement on 256 x 256;
  batched  20,000 systems A_i = I - g_i J_i, n_i uniform in [8, 512], one right-hand side each,
           one persistent launch, against a loop of plain per-system GPU LU solves (timed on a
           sample of the same systems);
  serving  Poisson arrivals of the 70 % dense / 30 % SimpleGMRES mix, p50 / p99 latency and the
           served rate.
"""
import argparse
import ctypes
import importlib.util
import json
import os
import sys
import time

import numpy as np

# The serving lanes (batched, 2 dense, 2 GMRES) and the LU's lookahead streams each want a
# hardware queue of their own: streams that share one run one after the other, so a dense
# lane's short kernels would wait behind a whole GMRES launch.  HIP's default is 4 per process
# (the sections before serving create and drop further streams, CU-masked ones among them);
# it has to be set before the runtime starts.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

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
                sample=f"one dgetrf+dgetrs ({nrhs} RHS) at n={ns} of the same matrix, scipy "
                       f"{getattr(scipy, '__version__', '?')}'s OpenBLAS (LAPACK, the library "
                       f"LUFactorization calls; julia is not installed), {t:.2f} s"), (piv + 1 if ns == n else None)


def backward_error_t(A, X, B):
    """max over columns of ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf), in FP64 on the device."""
    import torch
    R = B - A @ X
    nA = A.abs().sum(dim=1).max()
    num = R.abs().amax(dim=0)
    den = nA * X.abs().amax(dim=0) + B.abs().amax(dim=0)
    return float((num / den).max())


def timed(fn, reps):
    import torch
    torch.cuda.synchronize()
    best = float("inf")
    out = None
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best, out


def run_mixed(lsx, _abi, ctx, L, n, conds, reps, cpu=True):
    """Mixed-precision LU + FP64 refinement against plain FP64 LU on conditioned systems."""
    import torch
    P = lsx.problems
    hs = {}
    for name, dt in (("fp64", _abi.LSX_F64), ("fp32", _abi.LSX_F32), ("bf16", _abi.LSX_BF16)):
        h = ctypes.c_void_p()
        ctx.check(L.lsx_lu_create(ctx.handle, n, dt, ctypes.byref(h)), "lsx_lu_create")
        hs[name] = h
    g = torch.Generator(device="cuda")
    g.manual_seed(31)
    b = torch.randn(n, 1, dtype=torch.float64, device="cuda", generator=g)
    x = torch.empty_like(b)
    out = {}
    try:
        for cond in conds:
            log(f"mixed: generating cond={cond:.0e} system (QR of two Gaussian matrices on the GPU)")
            A = P.conditioned(n, cond, device="cuda")   # (n, n) tensor, row-major
            Acm = A.t().contiguous()                      # row-major A^T = column-major A
            torch.cuda.synchronize()
            ent = {}

            def f64():
                ctx.check(L.lsx_lu_factor(hs["fp64"], Acm.data_ptr(), n), "lsx_lu_factor")
                ctx.check(L.lsx_lu_solve(hs["fp64"], 1, b.data_ptr(), n, x.data_ptr(), n), "lsx_lu_solve")

            t64, _ = timed(f64, reps)
            ent["fp64"] = dict(ms=round(t64 * 1e3, 3), backward_error=backward_error_t(A, x, b))
            if cpu and cond == conds[0]:
                # the reference's CPU path on the same system: LAPACK dgetrf + dgetrs, all BLAS threads
                import scipy.linalg.lapack as lapack
                Ah = np.asfortranarray(Acm.cpu().numpy().T)
                bh = b.cpu().numpy()[:, 0].copy()
                t0 = time.perf_counter()
                lu, piv, info = lapack.dgetrf(Ah, overwrite_a=True)
                lapack.dgetrs(lu, piv, bh)
                tc = time.perf_counter() - t0
                del Ah, lu
                ent["cpu_lapack_fp64"] = dict(ms=round(tc * 1e3, 1), threads=int(_blas_cores()), kind="reference",
                                              sample="one dgetrf + dgetrs of this system (scipy's OpenBLAS)")
                log(f"  cond={cond:.0e} CPU LAPACK FP64: {tc * 1e3:.0f} ms")
            for name in ("fp32", "bf16"):
                st = _abi.RefineStats()
                # the workload's refinement: GMRES-IR at cond 1e6 for FP32 factors; BF16 factors take
                # the automatic sequence (IR, GMRES-IR when IR stalls, FP32 factors when that stalls)
                rmode = 1 if (name == "fp32" and cond >= 1e6) else 2

                def mixed():
                    ctx.check(L.lsx_lu_factor(hs[name], Acm.data_ptr(), n), "lsx_lu_factor")
                    ctx.check(L.lsx_lu_refine(hs[name], 1, b.data_ptr(), n, x.data_ptr(), n, 1e-13, 100, rmode,
                                              ctypes.byref(st)), "lsx_lu_refine")

                tm, _ = timed(mixed, reps)
                bd = None
                if cond == conds[0]:  # one extra, profiled run: where the time goes
                    ctx.check(L.lsx_lu_profile(hs[name], 1, None), "lsx_lu_profile")
                    mixed()
                    torch.cuda.synchronize()
                    prof = _abi.Profile()
                    ctx.check(L.lsx_lu_profile(hs[name], 0, ctypes.byref(prof)), "lsx_lu_profile")
                    bd = {c: round(prof.ms[i], 3) for i, c in enumerate(_abi.PROFILE_CLASSES) if prof.launches[i]}
                ent[name] = dict(ms=round(tm * 1e3, 3), speedup_vs_fp64=round(t64 / tm, 3),
                                 backward_error=backward_error_t(A, x, b),
                                 mode=("IR", "GMRES-IR")[st.mode_used], iterations=int(st.iterations),
                                 converged=bool(st.converged), escalated_to_fp32=bool(st.escalated))
                if bd is not None:
                    ent[name]["kernel_ms_breakdown"] = bd
                log(f"  cond={cond:.0e} {name}: {tm * 1e3:.1f} ms vs fp64 {t64 * 1e3:.1f} ms "
                    f"({t64 / tm:.2f}x), bwd {ent[name]['backward_error']:.2e}")
            out[f"cond={cond:.0e}"] = ent
            del A, Acm
            torch.cuda.empty_cache()
    finally:
        for h in hs.values():
            L.lsx_lu_destroy(h)
    return out


def run_batched(lsx, _abi, ctx, L, count, reps, loop_sample):
    """20,000 small systems in one persistent launch vs a loop of per-system plain GPU LU solves."""
    import torch
    P = lsx.problems
    g = P.rng_for(4, count, 4)
    sizes = g.integers(8, 512 + 1, size=count).astype(np.int64)
    gam = g.uniform(1e-3, 1e-1, size=count)
    nn = sizes * sizes
    aoff = np.concatenate([[0], np.cumsum(nn)[:-1]])
    voff = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    total, totv = int(nn.sum()), int(sizes.sum())
    log(f"batched: {count} systems, {total * 8 / 2**30:.1f} GiB of matrices, generated on the GPU")
    tg = torch.Generator(device="cuda")
    tg.manual_seed(44)
    A = torch.randn(total, dtype=torch.float64, device="cuda", generator=tg)
    log("batched: J generated")
    scale = torch.repeat_interleave(torch.from_numpy(-gam / np.sqrt(sizes)).cuda(), torch.from_numpy(nn).cuda())
    A.mul_(scale)
    del scale
    didx = np.concatenate([aoff[i] + np.arange(sizes[i]) * (sizes[i] + 1) for i in range(count)])
    A[torch.from_numpy(didx).cuda()] += 1.0
    b = torch.randn(totv, dtype=torch.float64, device="cuda", generator=tg)
    x = torch.empty_like(b)
    torch.cuda.synchronize()
    log("batched: systems ready, creating the batch handle")
    # autotuned size-binned panel widths per factor type (tile_table.json)
    tables = {dt: lsx.load_tile_table(dtype=dt) for dt in (_abi.LSX_F64, _abi.LSX_F32)}
    info = np.zeros(count, dtype=np.int64)
    rs = np.random.default_rng(7)
    pick = rs.choice(count, size=min(200, count), replace=False)

    def sampled_backward_error():
        worst = 0.0
        for i in pick:
            n = int(sizes[i])
            Ai = A[aoff[i]:aoff[i] + n * n].view(n, n).t().cpu().numpy()
            bi = b[voff[i]:voff[i] + n].cpu().numpy()
            xi = x[voff[i]:voff[i] + n].cpu().numpy()
            r = bi - Ai @ xi
            worst = max(worst, np.abs(r).max() / (np.abs(Ai).sum(axis=1).max() * np.abs(xi).max() + np.abs(bi).max()))
        return worst

    def run(dtype):
        hb = ctypes.c_void_p()
        ctx.check(L.lsx_batch_create(ctx.handle, count, sizes.ctypes.data, dtype, ctypes.byref(hb)), "lsx_batch_create")
        table = tables[dtype]
        if table:
            up = np.array([u for u, _ in table], dtype=np.int32)
            bw = np.array([w for _, w in table], dtype=np.int32)
            ctx.check(L.lsx_batch_set_tiles(hb, len(table), up.ctypes.data, bw.ctypes.data), "lsx_batch_set_tiles")

        def batched():
            return ctx.check(L.lsx_batch_factor_solve(hb, A.data_ptr(), b.data_ptr(), x.data_ptr(),
                                                      info.ctypes.data, None), "lsx_batch_factor_solve")

        batched()
        t, nf = timed(batched, reps)
        iters = np.zeros(count, dtype=np.int32)
        berr = np.zeros(count, dtype=np.float64)
        ctx.check(L.lsx_batch_stats(hb, iters.ctypes.data, berr.ctypes.data), "lsx_batch_stats")
        L.lsx_batch_destroy(hb)
        return t, nf, iters, berr

    t32, nf32, it32, be32 = run(_abi.LSX_F32)
    worst32 = sampled_backward_error()
    log(f"batched: FP32 + FP64 refinement {t32 * 1e3:.1f} ms, refinement steps max {int(it32.max())}, "
        f"sampled backward error {worst32:.2e}")
    tb, nfail, _, _ = run(_abi.LSX_F64)
    log(f"batched: {reps} timed FP64 launches done, {tb * 1e3:.1f} ms")
    worst = sampled_backward_error()
    log(f"batched: sampled backward error {worst:.2e}")
    # per-system loop of the plain FP64 GPU LU (handles made beforehand, outside the timing)
    loop_idx = rs.choice(count, size=min(loop_sample, count), replace=False)
    handles = {}
    for i in loop_idx:
        n = int(sizes[i])
        if n not in handles:
            h = ctypes.c_void_p()
            ctx.check(L.lsx_lu_create(ctx.handle, n, _abi.LSX_F64, ctypes.byref(h)), "lsx_lu_create")
            handles[n] = h
    torch.cuda.synchronize()
    log(f"batched: {len(handles)} per-size LU handles made, timing the per-system loop")
    t0 = time.perf_counter()
    for i in loop_idx:
        n = int(sizes[i])
        h = handles[n]
        ctx.check(L.lsx_lu_factor(h, A.data_ptr() + 8 * int(aoff[i]), n), "lsx_lu_factor")
        ctx.check(L.lsx_lu_solve(h, 1, b.data_ptr() + 8 * int(voff[i]), n, x.data_ptr() + 8 * int(voff[i]), n),
                  "lsx_lu_solve")
    torch.cuda.synchronize()
    tl = (time.perf_counter() - t0) / len(loop_idx)
    for h in handles.values():
        L.lsx_lu_destroy(h)
    flops = float(sum(luflop(int(n)) + 2.0 * int(n) * int(n) for n in sizes))
    res = dict(systems=count, sizes="n uniform in [8, 512]", ms=round(tb * 1e3, 3),
               solves_per_s=round(count / tb, 1), tflops=round(flops / tb / 1e12, 3), failed=int(nfail),
               backward_error_max_sampled=worst,
               per_system_loop=dict(solves_per_s=round(1.0 / tl, 1), sample=len(loop_idx),
                                    kind="lsx_lu_factor + lsx_lu_solve per system, handles reused"),
               speedup_vs_loop=round((count / tb) * tl, 2),
               tile_table=tables[_abi.LSX_F64] or "library default",
               fp32_refined=dict(ms=round(t32 * 1e3, 3), solves_per_s=round(count / t32, 1), failed=int(nf32),
                                 tile_table=tables[_abi.LSX_F32] or "library default",
                                 refinement_steps_max=int(it32.max()), refinement_steps_mean=round(float(it32.mean()), 2),
                                 backward_error_max=float(be32.max()), backward_error_max_sampled=worst32,
                                 speedup_vs_loop=round((count / t32) * tl, 2)))
    log(f"batched: {tb * 1e3:.1f} ms, {count / tb:.0f} solves/s, loop {1 / tl:.0f} solves/s, "
        f"speedup {res['speedup_vs_loop']}x, bwd {worst:.2e}")
    del A, b, x
    torch.cuda.empty_cache()
    return res


def gmres_step_bytes(N, nnz, iters, memory):
    """HBM bytes of `iters` Arnoldi steps of restarted GMRES(memory) with a diagonal left
    preconditioner, step k of a cycle.  Returns (convention, minimum):
      convention  (128 + 40 k) N, the per-step figure of the unfused algorithm (SpMV ~80 N,
                  Jacobi 24 N, norm 8 N, next basis vector 16 N, MGS 40 k N: dot and axpy each
                  read the basis vector and q, and the axpy writes q);
      minimum     what the fused kernel must move: the CSR (12 nnz + 4 N), the gathered basis
                  vector (8 N), the diagonal (8 N), each of the k basis vectors once (8 k N: q stays
                  on chip), the next basis vector written (8 N)."""
    conv = mini = 0.0
    for i in range(iters):
        k = i % memory + 1
        conv += (128.0 + 40.0 * k) * N
        mini += 12.0 * nnz + 4.0 * N + 8.0 * N + 8.0 * N + 8.0 * k * N + 8.0 * N
    return conv, mini


def _blas_cores():
    try:
        from threadpoolctl import threadpool_info
        return max((i.get("num_threads", 1) for i in threadpool_info() if i.get("internal_api") == "openblas"),
                   default=os.cpu_count())
    except Exception:
        return os.cpu_count()


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "?"


def run_cpu_baselines(lsx, _abi, ctx, L, budget_s=12.0):
    """The reference's CPU path on the host, per workload, with all BLAS threads and with one:
    LAPACK dgetrf/dgetrs from scipy's OpenBLAS (the library LUFactorization calls; julia is not
    installed) and, for SimpleGMRES, the oracle's C restatement of src/simplegmres.jl (one
    thread).  Each figure is a bounded sample of its workload; the GPU figure beside it is the
    same sample on the device."""
    import torch
    import scipy.linalg.lapack as lapack
    from threadpoolctl import threadpool_limits
    P = lsx.problems
    cores = int(_blas_cores())
    out = dict(cpu=_cpu_model(), blas_threads_all=cores)

    def lapack_solve(A, b):
        lu, piv, info = lapack.dgetrf(A)
        x, _ = lapack.dgetrs(lu, piv, b)
        return lu, piv, x

    # dense LU, n = 1024 (A, b, b' iid U(0,1)): factor + solve, then a re-solve with the factors
    n = 1024
    A, b, b2 = P.dense_uniform(n)
    A = np.asfortranarray(A)
    ent = {}
    for label, th in (("all_cores", cores), ("one_thread", 1)):
        with threadpool_limits(limits=th, user_api="blas"):
            lapack_solve(A, b)
            reps, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < 1.0:
                lu, piv, x = lapack_solve(A, b)
                reps += 1
            tf = (time.perf_counter() - t0) / reps
            reps, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < 0.5:
                lapack.dgetrs(lu, piv, b2)
                reps += 1
            tr = (time.perf_counter() - t0) / reps
        ent[label] = dict(threads=th, factor_solve_ms=round(tf * 1e3, 3), resolve_ms=round(tr * 1e3, 4))
    h = ctypes.c_void_p()
    ctx.check(L.lsx_lu_create(ctx.handle, n, _abi.LSX_F64, ctypes.byref(h)), "lsx_lu_create")
    dA = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    db, db2 = torch.from_numpy(b).cuda(), torch.from_numpy(b2).cuda()
    dx = torch.empty_like(db)
    torch.cuda.synchronize()

    def gpu_fs():
        ctx.check(L.lsx_lu_factor(h, dA.data_ptr(), n), "lsx_lu_factor")
        ctx.check(L.lsx_lu_solve(h, 1, db.data_ptr(), n, dx.data_ptr(), n), "lsx_lu_solve")

    def gpu_rs():
        ctx.check(L.lsx_lu_solve(h, 1, db2.data_ptr(), n, dx.data_ptr(), n), "lsx_lu_solve")

    tg, _ = timed(gpu_fs, 20)
    tgr, _ = timed(gpu_rs, 50)
    ipiv = np.empty(n, dtype=np.int64)
    ctx.check(L.lsx_lu_get_factors(h, None, n, ipiv.ctypes.data), "lsx_lu_get_factors")
    L.lsx_lu_destroy(h)
    ent["gpu"] = dict(factor_solve_ms=round(tg * 1e3, 3), resolve_ms=round(tgr * 1e3, 4),
                      pivots_bit_exact_vs_lapack=bool(np.array_equal(ipiv, piv + 1)))
    out["dense_lu_n1024_uniform"] = ent
    log(f"cpu baselines: n=1024 factor+solve CPU {ent['all_cores']['factor_solve_ms']} ms "
        f"({cores} threads) / {ent['one_thread']['factor_solve_ms']} ms (1 thread), GPU {ent['gpu']['factor_solve_ms']} ms")

    # plain FP64 LU at scale, one thread: dgetrf + dgetrs (16 RHS) at n = 8192
    n = 8192
    Ag, Bg = P.dense_gaussian(n, 16)
    Ag = np.asfortranarray(Ag)
    Bg = np.asfortranarray(Bg)
    with threadpool_limits(limits=1, user_api="blas"):
        t0 = time.perf_counter()
        lu, piv, x = lapack_solve(Ag, Bg)
        t1 = time.perf_counter() - t0
    out["dense_lu_n8192_one_thread"] = dict(threads=1, seconds=round(t1, 3),
                                            tflops=round((luflop(n) + 2.0 * n * n * 16) / t1 / 1e12, 4))
    del Ag, Bg, lu, x
    log(f"cpu baselines: n=8192 one thread {t1:.1f} s")

    # batched: a per-system LAPACK loop over a sample of the 20,000-system distribution
    sizes, Ap, bp = P.batched_systems(count=1000, seed=9)
    offs = np.concatenate([[0], np.cumsum(sizes * sizes)[:-1]])
    voffs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    blocks = [(np.asfortranarray(Ap[o:o + k * k].reshape(k, k, order="F")), bp[v:v + k])
              for o, v, k in zip(offs, voffs, (int(k) for k in sizes))]
    ent = {}
    for label, th in (("all_cores", cores), ("one_thread", 1)):
        with threadpool_limits(limits=th, user_api="blas"):
            t0 = time.perf_counter()
            done = 0
            for Ai, bi in blocks:
                lapack_solve(Ai, bi)
                done += 1
                if time.perf_counter() - t0 > budget_s:
                    break
            t = time.perf_counter() - t0
        ent[label] = dict(threads=th, systems=done, solves_per_s=round(done / t, 1))
    out["batched_lu_loop"] = dict(sample="1000 systems of the batched distribution (n uniform in [8, 512], "
                                         "A = I - g J), one dgetrf + dgetrs per system", **ent)
    log(f"cpu baselines: batched loop {ent['all_cores']['solves_per_s']} / {ent['one_thread']['solves_per_s']} solves/s")
    del blocks, Ap

    return out


def run_gmres(lsx, _abi, ctx, L, m, pe, memory, maxiters, cpu_m):
    """One SimpleGMRES(memory) solve with the Jacobi preconditioner on the m x m convection-
    diffusion grid (inputs resident in HBM), and the same solver's CPU restatement (the oracle,
    one thread) on a smaller grid, both for `maxiters` Arnoldi steps at most."""
    import torch
    P = lsx.problems
    out = {}
    for side in (m, cpu_m):
        rp, ci, v = P.convection_diffusion(side, pe)
        N = side * side
        b = P.rng_for(12, side).standard_normal(N)
        A = ctypes.c_void_p()
        ctx.check(L.lsx_csr_create(ctx.handle, N, len(v), rp.ctypes.data, ci.ctypes.data, v.ctypes.data,
                                   ctypes.byref(A)), "lsx_csr_create")
        db = torch.from_numpy(b).cuda()
        dx = torch.empty_like(db)
        dpl = torch.full((N,), 4.0, dtype=torch.float64, device="cuda")
        st = _abi.GmresStats()
        tol = float(np.sqrt(np.finfo(np.float64).eps))

        def solve():
            ctx.check(L.lsx_gmres_solve(A, db.data_ptr(), dx.data_ptr(), memory, 1, tol, 1e-8, maxiters,
                                        dpl.data_ptr(), None, ctypes.byref(st)), "lsx_gmres_solve")

        t, _ = timed(solve, 2)
        L.lsx_csr_destroy(A)
        it = int(st.iters)
        conv, mini = gmres_step_bytes(N, len(v), it, memory)
        ent = dict(grid=f"{side}x{side}", N=N, nnz=len(v), iterations=it, retcode=int(st.retcode),
                   resid=float(st.resid), ms=round(t * 1e3, 3), us_per_step=round(t / max(it, 1) * 1e6, 2),
                   bytes_convention=conv, bytes_minimum=mini,
                   achieved_GBs_convention=round(conv / t / 1e9, 1), achieved_GBs_minimum=round(mini / t / 1e9, 1),
                   hbm_frac_minimum=round(mini / t / 1e9 / HBM_PEAK_GBS, 4))
        if side == cpu_m:
            sys.path.insert(0, ROOT)
            from oracle import oracle as O  # the CPU restatement, timed as the baseline only
            t0 = time.perf_counter()
            r = O.gmres(rp, ci, v, b, memory=memory, restart=True, abstol=tol, reltol=1e-8, maxiters=maxiters,
                        pl=np.full(N, 4.0))
            tc = time.perf_counter() - t0
            ent["cpu_oracle"] = dict(ms=round(tc * 1e3, 3), iterations=int(r["iters"]), cores=1, kind="port",
                                     us_per_step=round(tc / max(int(r["iters"]), 1) * 1e6, 2))
        out[f"m={side}"] = ent
        log(f"gmres: grid {side}^2, {it} steps in {t * 1e3:.1f} ms ({ent['us_per_step']} us/step), "
            f"{ent['achieved_GBs_minimum']} GB/s (fused minimum bytes)" + (f"; CPU oracle {ent['cpu_oracle']['ms']:.0f} ms"
                                                         if "cpu_oracle" in ent else ""))
        del db, dx, dpl
    out["batched"] = run_gmres_batched(lsx, _abi, ctx, L)
    return dict(memory=memory, maxiters=maxiters, pe=pe, precond="Jacobi (diagonal 4)", runs=out)


def run_gmres_batched(lsx, _abi, ctx, L, nblocks=4096, side=16, loop_sample=64):
    """SimpleGMRES(blocksize) on a block-diagonal matrix of `nblocks` convection-diffusion
    blocks (side x side grids, Pe cycling 0..9), one launch, against a loop of per-block
    SimpleGMRES solves (a sample of the same blocks, handles made beforehand)."""
    import torch
    import scipy.sparse as sp
    P = lsx.problems
    bs = side * side
    mats = [P.convection_diffusion(side, float(k % 10)) for k in range(10)]
    blocks = [sp.csr_matrix((v, ci, rp), shape=(bs, bs)) for rp, ci, v in mats]
    S = sp.block_diag([blocks[k % 10] for k in range(nblocks)], format="csr")
    S.sort_indices()
    N = S.shape[0]
    b = P.rng_for(13, nblocks).standard_normal(N)
    tol = float(np.sqrt(np.finfo(np.float64).eps))
    # the int32 copies must stay referenced while the library reads them
    rp32, ci32, vals = S.indptr.astype(np.int32), S.indices.astype(np.int32), np.ascontiguousarray(S.data)
    A = ctypes.c_void_p()
    ctx.check(L.lsx_csr_create(ctx.handle, N, S.nnz, rp32.ctypes.data, ci32.ctypes.data, vals.ctypes.data,
                               ctypes.byref(A)), "lsx_csr_create")
    db = torch.from_numpy(b).cuda()
    dx = torch.empty_like(db)
    dpl = torch.full((N,), 4.0, dtype=torch.float64, device="cuda")
    st = _abi.GmresStats()

    def batched():
        ctx.check(L.lsx_gmres_solve_batched(A, bs, db.data_ptr(), dx.data_ptr(), 20, 1, tol, 1e-8, 1000,
                                            dpl.data_ptr(), None, ctypes.byref(st)), "lsx_gmres_solve_batched")

    tb, _ = timed(batched, 3)
    L.lsx_csr_destroy(A)
    # the loop: one handle per distinct block, solves timed
    hs = []
    for B in blocks:
        h = ctypes.c_void_p()
        brp, bci, bv = B.indptr.astype(np.int32), B.indices.astype(np.int32), np.ascontiguousarray(B.data)
        ctx.check(L.lsx_csr_create(ctx.handle, bs, B.nnz, brp.ctypes.data, bci.ctypes.data, bv.ctypes.data,
                                   ctypes.byref(h)), "lsx_csr_create")
        hs.append(h)
    eps = tol + 1e-8 * float(np.linalg.norm(b / 4.0))
    st1 = _abi.GmresStats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(loop_sample):
        sl = slice(k * bs, (k + 1) * bs)
        ctx.check(L.lsx_gmres_solve(hs[k % 10], db[sl].data_ptr(), dx[sl].data_ptr(), 20, 1, eps, 0.0, 1000,
                                    dpl[sl].data_ptr(), None, ctypes.byref(st1)), "lsx_gmres_solve")
    torch.cuda.synchronize()
    tl = (time.perf_counter() - t0) / loop_sample
    for h in hs:
        L.lsx_csr_destroy(h)
    res = dict(blocks=nblocks, blocksize=bs, N=N, ms=round(tb * 1e3, 3), solves_per_s=round(nblocks / tb, 1),
               iterations_slowest_block=int(st.iters), retcode=int(st.retcode),
               per_block_loop=dict(solves_per_s=round(1.0 / tl, 1), sample=loop_sample),
               speedup_vs_loop=round(nblocks / tb * tl, 2))
    log(f"gmres batched: {nblocks} blocks of {bs} in {tb * 1e3:.1f} ms ({res['solves_per_s']} solves/s), "
        f"loop {res['per_block_loop']['solves_per_s']} solves/s -> {res['speedup_vs_loop']}x")
    del db, dx, dpl
    return res


def _submit_synth(srv, req, gmres_maxiters):
    """gmres_maxiters 0: SimpleGMRES runs to its stopping rule with the reference's default
    maxiters = length(b) (src/common.jl:764)."""
    if req["kind"] == "dense":
        return srv.submit_dense_synth(req["n"], req["seed"])
    return srv.submit_gmres_synth(req["m"], req["pe"], req["seed"], memory=30, reltol=1e-8,
                                  maxiters=gmres_maxiters or None)


def serving_loop_baseline(lsx, trace, gmres_maxiters, budget_gib, max_seconds=30.0):
    """The per-problem loop the serving path is compared with: the same requests (the trace's
    first ones, in order, for at most max_seconds), each submitted only after the previous one
    completed, so nothing is batched or overlapped; same kernels and device-side inputs."""
    srv = lsx.Server(hbm_budget=int(budget_gib) << 30, nstreams=4)
    for req in trace[:3]:  # warm-up: lane handles and code objects
        t = _submit_synth(srv, req, gmres_maxiters)
        while srv.poll(t) is None:
            time.sleep(0.0001)
    t0 = time.perf_counter()
    done = 0
    for req in trace:
        t = _submit_synth(srv, req, gmres_maxiters)
        while srv.poll(t) is None:
            time.sleep(0.0001)
        done += 1
        if time.perf_counter() - t0 > max_seconds:
            break
    el = time.perf_counter() - t0
    srv.close()
    return dict(requests=done, seconds=round(el, 3), req_per_s=round(done / el, 1))


def run_serving(lsx, seconds, rate, gmres_maxiters, budget_gib, cpu=True):
    """Poisson arrivals of the serving mix on the synthetic (device-generated) request path."""
    P = lsx.problems
    trace = P.serving_trace(rate=rate, seconds=seconds)
    srv = lsx.Server(hbm_budget=int(budget_gib) << 30, nstreams=4)
    pending = []
    c0 = srv.clock()
    t0 = time.perf_counter()
    for req in trace:
        while time.perf_counter() - t0 < req["t"]:
            pass
        pending.append((_submit_synth(srv, req, gmres_maxiters), req["kind"]))
    log(f"serving: {len(trace)} requests submitted in {time.perf_counter() - t0:.1f} s; draining")
    last = time.perf_counter()
    while srv.stats()["completed"] < len(trace):
        time.sleep(0.02)
        if time.perf_counter() - last > 10.0:
            last = time.perf_counter()
            st = srv.stats()
            log(f"serving: {st['completed']} of {len(trace)} done, {st['waiting']} waiting")
    srv.drain()
    t_end = srv.clock()
    lat = {"dense": [], "gmres": []}
    conv = errors = iters = 0
    steps = []  # (grid side, Arnoldi steps) of each served SimpleGMRES request
    for (t, kind), req in zip(pending, trace):
        r = srv.poll(t)
        if r["status"] != 0:  # no valid result (a wait gave up, bad input): not served
            errors += 1
            continue
        lat[kind].append(r["t_done"] - r["t_submit"])
        if kind == "gmres":
            iters += int(r["iters"])
            steps.append((req["m"], int(r["iters"])))
            if r["retcode"] == lsx.ReturnCode.Success:
                conv += 1
    st = srv.stats()
    srv.close()
    allv = np.array(lat["dense"] + lat["gmres"])
    q = lambda v, p: round(float(np.percentile(v, p)) * 1e3, 3) if len(v) else None
    served = len(allv) / (t_end - c0)
    log(f"serving: served {served:.2f} req/s in {t_end - c0:.1f} s; timing the one-at-a-time loop on the same trace")
    loop = serving_loop_baseline(lsx, trace, gmres_maxiters, budget_gib)
    res = dict(offered_req_per_s=rate, arrival_seconds=seconds, requests=len(trace), errors=errors,
               makespan_s=round(t_end - c0, 3), served_req_per_s=round(served, 2),
               p50_ms=q(allv, 50), p99_ms=q(allv, 99),
               dense=dict(count=len(lat["dense"]), p50_ms=q(lat["dense"], 50), p99_ms=q(lat["dense"], 99)),
               gmres=dict(count=len(lat["gmres"]), p50_ms=q(lat["gmres"], 50), p99_ms=q(lat["gmres"], 99),
                          converged=conv, arnoldi_steps=iters,
                          maxiters=gmres_maxiters or "length(b) (reference default)"),
               batched_launches=st["batched_launches"], batched_requests=st["batched_requests"],
               peak_hbm_gib=round(st["peak_bytes"] / 2**30, 2),
               loop_baseline=loop, speedup_vs_loop=round(served / loop["req_per_s"], 2))
    log(f"serving: served {res['served_req_per_s']} req/s, p50 {res['p50_ms']} ms, p99 {res['p99_ms']} ms; "
        f"loop {loop['req_per_s']} req/s -> {res['speedup_vs_loop']}x")
    if cpu:
        res["cpu_replay"] = serving_cpu_replay(lsx, trace, steps)
    return res


def serving_cpu_replay(lsx, trace, steps):
    """The serving requests one at a time on the host: every dense request by LAPACK dgetrf +
    dgetrs (all BLAS threads, then one), timed; every SimpleGMRES request at the oracle's cost
    per Arnoldi step (the C restatement of src/simplegmres.jl, one thread; 60 steps timed per grid
    size, the cost scaled by N between the timed sizes) times the steps the served solve took
    (the same algorithm, so the same count within one)."""
    import scipy.linalg.lapack as lapack
    from threadpoolctl import threadpool_limits
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    P = lsx.problems
    cores = int(_blas_cores())
    dense = [req for req in trace if req["kind"] == "dense"]
    blocks = [P.dense_block(req["n"], req["seed"]) for req in dense]
    out = {}
    for label, th in (("all_cores", cores), ("one_thread", 1)):
        with threadpool_limits(limits=th, user_api="blas"):
            t0 = time.perf_counter()
            for Ai, bi in blocks:
                lu, piv, info = lapack.dgetrf(Ai)
                lapack.dgetrs(lu, piv, bi)
            out[label] = time.perf_counter() - t0
    per_step = {}
    for m in (64, 256, 1024):
        rp, ci, v = P.convection_diffusion(m, 5.0)
        N = m * m
        b = P.rng_for(13, m).standard_normal(N)
        t0 = time.perf_counter()
        O.gmres(rp, ci, v, b, memory=30, restart=True, reltol=1e-30, abstol=0.0, maxiters=60, pl=np.full(N, 4.0))
        per_step[m] = (time.perf_counter() - t0) / 60
    ms_ = np.array(sorted(per_step))
    cost = np.array([per_step[m] / (m * m) for m in ms_])  # seconds per step per row
    gmres_s = sum(it * np.interp(m, ms_, cost) * m * m for m, it in steps)
    nreq = len(dense) + len(steps)
    res = dict(requests=nreq, gmres_seconds_estimated=round(gmres_s, 1),
               gmres_step_ms={int(m): round(per_step[m] * 1e3, 3) for m in ms_})
    for label in ("all_cores", "one_thread"):
        res[label] = dict(dense_seconds=round(out[label], 3),
                          req_per_s=round(nreq / (out[label] + gmres_s), 3))
    res["threads_all"] = cores
    log(f"serving: CPU replay {res['all_cores']['req_per_s']} req/s (dense {out['all_cores']:.2f} s measured, "
        f"SimpleGMRES {gmres_s:.0f} s from the oracle's step cost)")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--nrhs", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sections", default="mixed,gmres,batched,serving,cpu",
                    help="comma list of extra sections beside the FP64 LU line (mixed, gmres, batched, serving, "
                         "cpu; '' = none)")
    ap.add_argument("--mixed-n", type=int, default=16384)
    ap.add_argument("--batch-count", type=int, default=20000)
    ap.add_argument("--serve-seconds", type=float, default=0.05,
                    help="arrival window of the serving trace (every request is served to completion)")
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-gmres-maxiters", type=int, default=0,
                    help="0: length(b), the reference's default; SimpleGMRES runs to reltol 1e-8")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy three times (HBM byte-counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read_per_launch": src.numel() * 8, "copy_bytes_written_per_launch": src.numel() * 8,
                          "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the dense LU would need a 2D block-cyclic layout with per-column pivot all-reduces)",
              file=sys.stderr)
        sys.exit(2)
    sections = [x for x in args.sections.split(",") if x]

    import torch
    lsx = load_package()
    from linearsolve_jl_amd import _abi
    import linearsolve_jl_amd.problems  # noqa: F401  (lsx.problems)

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
    t0 = time.perf_counter()
    for i in range(args.steps):
        info = step()
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    log(f"timed {args.steps} steps: {t / args.steps * 1e3:.2f} ms/step, info={info}")
    # the per-kernel breakdown and the dominant kernel's launch time, from further steps with a
    # HIP event pair around every launch on its own stream (the events add gaps between kernels,
    # so these steps are not the timed ones)
    psteps = max(1, min(args.steps, 3))
    ctx.check(L.lsx_lu_profile(h, 1, None), "lsx_lu_profile")
    for i in range(psteps):
        step()
    torch.cuda.synchronize()
    prof = _abi.Profile()
    ctx.check(L.lsx_lu_profile(h, 0, ctypes.byref(prof)), "lsx_lu_profile")
    log(f"profiled {psteps} more steps")

    # correctness of the last step: backward error and pivots
    X = dX.cpu().numpy().T
    r = B - A @ X
    nA = np.abs(A).sum(axis=1).max()
    be = max(np.abs(r[:, j]).max() / (nA * np.abs(X[:, j]).max() + np.abs(B[:, j]).max()) for j in range(nrhs))
    ipiv = np.empty(n, dtype=np.int64)
    ctx.check(L.lsx_lu_get_factors(h, None, n, ipiv.ctypes.data), "lsx_lu_get_factors")
    log(f"backward error {be:.3e}")
    # one right-hand side with the resident factors: forward and adjoint (A^T x = b) solves
    db1 = dB[0].clone()
    dx1 = torch.empty_like(db1)
    solve_ms = {}
    for name, fn in (("forward", L.lsx_lu_solve), ("adjoint", L.lsx_lu_solve_adjoint)):
        fn(h, 1, db1.data_ptr(), n, dx1.data_ptr(), n)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            ctx.check(fn(h, 1, db1.data_ptr(), n, dx1.data_ptr(), n), name)
        torch.cuda.synchronize()
        solve_ms[name] = round((time.perf_counter() - t0) / 5 * 1e3, 3)
    log(f"one-RHS solves: forward {solve_ms['forward']} ms, adjoint {solve_ms['adjoint']} ms")
    del db1, dx1
    # the same step with A and B in (pageable) host memory: the PCIe-inclusive rate (not `value`)
    Xh = np.empty_like(Gb)
    t0 = time.perf_counter()
    ctx.check(L.lsx_lu_factor(h, G.ctypes.data, n), "lsx_lu_factor")
    ctx.check(L.lsx_lu_solve(h, nrhs, Gb.ctypes.data, n, Xh.ctypes.data, n), "lsx_lu_solve")
    host_ms = (time.perf_counter() - t0) * 1e3
    log(f"step from host buffers: {host_ms:.2f} ms")
    del Xh
    L.lsx_lu_destroy(h)
    del dA, dB, dX
    torch.cuda.empty_cache()

    flops = luflop(n) + 2.0 * n * n * nrhs
    ms = t / args.steps * 1e3
    value = flops / (ms / 1e3) / 1e12
    breakdown = {}
    for i, cls in enumerate(_abi.PROFILE_CLASSES):
        if prof.launches[i]:
            breakdown[cls] = dict(launches_per_step=prof.launches[i] / psteps,
                                  ms_per_step=prof.ms[i] / psteps)
    g = _abi.PROFILE_CLASSES.index("gemm")
    gemm_tflops = prof.flops[g] / (prof.ms[g] / 1e3) / 1e12 if prof.ms[g] > 0 else 0.0
    kname = "k_gemm<double, 0, 128, 128, 512>"
    roofline = dict(bound="mfma", kernel=f"{kname} (trailing update A22 -= L21*U12, 128x128 tiles on eight waves, K=128)",
                    achieved=round(gemm_tflops, 3), peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=round(gemm_tflops / FP64_PEAK_TFLOPS, 4),
                    flops_per_launch=prof.flops[g] / max(prof.launches[g], 1),
                    algorithmic_bytes_per_launch=prof.bytes[g] / max(prof.launches[g], 1),
                    avg_launch_us=prof.ms[g] / max(prof.launches[g], 1) * 1e3,
                    launches_per_step=prof.launches[g] / psteps,
                    traffic=traffic_from_profiles(kname),
                    traffic_source="profiles/pmc_traffic.json (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)")

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
        "host_buffers_ms_per_step": round(host_ms, 3),
        "one_rhs_solve_ms": solve_ms,
        "roofline": roofline,
        "kernel_breakdown": breakdown,
        "profiled_steps": psteps,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (LAPACK on host cores)")
        cb, lapack_ipiv = cpu_baseline(A, B, n, nrhs)
        result["cpu_baseline"] = cb
        if lapack_ipiv is not None:
            result["pivots_bit_exact_vs_lapack"] = bool(np.array_equal(ipiv, lapack_ipiv))
    del G, Gb, A, B
    if "mixed" in sections:
        result["mixed"] = dict(n=args.mixed_n, nrhs=1, tol=1e-13,
                               runs=run_mixed(lsx, _abi, ctx, L, args.mixed_n, (1e2, 1e4, 1e6), 2,
                                              cpu=not args.no_cpu_baseline))
    if "gmres" in sections:
        result["gmres"] = run_gmres(lsx, _abi, ctx, L, 1024, 5.0, 30, 300, 256)
    if "batched" in sections:
        result["batched"] = run_batched(lsx, _abi, ctx, L, args.batch_count, 3, 300)
    if "serving" in sections:
        result["serving"] = run_serving(lsx, args.serve_seconds, args.serve_rate, args.serve_gmres_maxiters, 64,
                                        cpu=not args.no_cpu_baseline)
    if "cpu" in sections and not args.no_cpu_baseline:
        result["cpu_baselines"] = run_cpu_baselines(lsx, _abi, ctx, L)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
