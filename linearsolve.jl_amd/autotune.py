"""Autotuner of the batched LU's size-binned tile table.

For every size bin it builds a sample batch of that bin's sizes (A = I - g J, the batched
workload's distribution, generated on the GPU), times the persistent batched launch with each
candidate panel width, keeps the fastest, and writes the table to tile_table.json beside this
file; _BatchHandle loads that file when it exists (lsx_batch_set_tiles).

    python linearsolve.jl_amd/autotune.py [out.json]   # default: tile_table.json beside this file
"""
import json
import os
import sys
import time

import numpy as np

DEFAULT_BINS = (32, 64, 128, 256, 384, 512)
WIDTHS = (8, 16)


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


def tune(bins=DEFAULT_BINS, widths=WIDTHS, per_bin=1024, reps=3, seed=0, path=None, write=True):
    """Returns [(upper, width), ...] and writes it (with the timings) to path (tile_table.json)."""
    import torch
    lsx = _package()
    from linearsolve_jl_amd import _abi
    ctx = lsx.Context.get(0)
    path = path or lsx.TILE_TABLE
    rng = np.random.default_rng(seed)
    table, timings = [], {}
    lo = 1
    for up in bins:
        sizes = rng.integers(lo, up + 1, size=per_bin)
        A, b = sample_batch(sizes, seed + up)
        x = torch.empty_like(b)
        h = lsx._BatchHandle(ctx, sizes, _abi.LSX_F64, tiles=None)
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
        print(f"bin n <= {up}: " + ", ".join(f"bw {w}: {t * 1e3:.3f} ms" for w, t in times.items())
              + f" -> {win}", flush=True)
        del h, A, b, x
        lo = up + 1
    if write:
        with open(path, "w") as f:
            json.dump({"bins": table, "ms_per_launch": timings, "systems_per_bin": per_bin,
                       "device": torch.cuda.get_device_name(0)}, f, indent=1)
    return table


if __name__ == "__main__":
    tune(path=sys.argv[1] if len(sys.argv) > 1 else None)
