"""Autotuner of the batched LU's size-binned tile tables, one per factor type.

For every factor type (FP64; FP32 with FP64 refinement) and size bin it builds a sample batch of
that bin's sizes (A = I - g J, the batched workload's distribution, generated on the GPU), times
the persistent batched launch with each candidate panel width, keeps the fastest, and writes the
tables to tile_table.json beside this file; _BatchHandle loads the table of its factor type when
the file exists (lsx_batch_set_tiles).  The panel width fixes the blocking of a system: widths 8
and 16 hold the panel's L21 in LDS (FP64) for trailing-update tiles dealt to the waves in turn;
width 32 keeps a row block of L21 in each wave's registers and sweeps its columns, with half the
trailing-matrix traffic of width 16.

    python linearsolve.jl_amd/autotune.py [out.json]   # default: tile_table.json beside this file
"""
import json
import os
import sys
import time

import numpy as np

DEFAULT_BINS = (32, 64, 96, 128, 192, 256, 320, 384, 448, 512)
WIDTHS = (8, 16, 32)
DTYPES = (("bins", "LSX_F64"), ("bins_fp32", "LSX_F32"))


def _package():
    name = "linearsolve_jl_amd"
    if name in sys.modules:
        return sys.modules[name]
    import importlib.util
    d = os.path.dirname(os.path.abspath(__file__))
    spec = importlib.util.spec_from_file_location(name, os.path.join(d, "__init__.py"),
                                                  submodule_search_locations=[d])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def sample_batch(sizes, seed):
    """A_i = I - g_i J_i (J iid N(0,1)/sqrt(n_i), g_i ~ U[1e-3, 1e-1]) and b iid N(0,1), packed
    column-major on the GPU."""
    import torch
    sizes = np.asarray(sizes, dtype=np.int64)
    g = np.random.default_rng(seed)
    gam = g.uniform(1e-3, 1e-1, size=len(sizes))
    nn = sizes * sizes
    aoff = np.concatenate([[0], np.cumsum(nn)[:-1]])
    tg = torch.Generator(device="cuda")
    tg.manual_seed(int(seed))
    A = torch.randn(int(nn.sum()), dtype=torch.float64, device="cuda", generator=tg)
    A.mul_(torch.repeat_interleave(torch.from_numpy(-gam / np.sqrt(sizes)).cuda(), torch.from_numpy(nn).cuda()))
    didx = np.concatenate([aoff[i] + np.arange(sizes[i]) * (sizes[i] + 1) for i in range(len(sizes))])
    A[torch.from_numpy(didx).cuda()] += 1.0
    b = torch.randn(int(sizes.sum()), dtype=torch.float64, device="cuda", generator=tg)
    return A, b


def tune(bins=DEFAULT_BINS, widths=WIDTHS, per_bin=1024, reps=3, seed=0, path=None, write=True, dtypes=DTYPES):
    """Returns {key: [(upper, width), ...]} per factor type and writes it (with the timings) to
    path (tile_table.json)."""
    import torch
    lsx = _package()
    from linearsolve_jl_amd import _abi
    ctx = lsx.Context.get(0)
    path = path or lsx.TILE_TABLE
    rng = np.random.default_rng(seed)
    samples, lo = [], 1
    for up in bins:
        samples.append((up, rng.integers(lo, up + 1, size=per_bin)))
        lo = up + 1
    out = {"systems_per_bin": per_bin, "device": torch.cuda.get_device_name(0)}
    tables = {}
    for key, dname in dtypes:
        dtype = getattr(_abi, dname)
        table, timings = [], {}
        for up, sizes in samples:
            A, b = sample_batch(sizes, seed + up)
            x = torch.empty_like(b)
            h = lsx._BatchHandle(ctx, sizes, dtype, tiles=None)
            times = {}
            for w in widths:
                h.set_tiles([(up, w)])
                h.factor_solve(A, b, x)
                best = float("inf")
                for _ in range(reps):
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    h.factor_solve(A, b, x)
                    torch.cuda.synchronize()
                    best = min(best, time.perf_counter() - t0)
                times[w] = best
            win = min(times, key=times.get)
            table.append((int(up), int(win)))
            timings[str(up)] = {str(w): round(t * 1e3, 4) for w, t in times.items()}
            print(f"{dname} bin n <= {up}: " + ", ".join(f"bw {w}: {t * 1e3:.3f} ms" for w, t in times.items())
                  + f" -> {win}", flush=True)
            del h, A, b, x
        tables[key] = table
        out[key] = table
        out["ms_per_launch" + key[4:]] = timings
    if write:
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
    return tables


if __name__ == "__main__":
    tune(path=sys.argv[1] if len(sys.argv) > 1 else None)
