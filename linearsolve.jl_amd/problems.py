"""Seeded synthetic inputs of the benchmark workloads (numpy PCG64 + SeedSequence, so the same
bytes feed the GPU path and the CPU baseline)."""
import numpy as np


def rng_for(*key):
    return np.random.default_rng(np.random.SeedSequence([int(k) for k in key]))


def dense_uniform(n, seed=1):
    """Plain dense LU, CPU reference case: A, b and a second b' iid U(0,1)."""
    g = rng_for(1, n, seed)
    return g.random((n, n)), g.random(n), g.random(n)


def dense_gaussian(n, nrhs=16, seed=2):
    """Plain FP64 LU at scale: A iid N(0,1) (column-major), nrhs right-hand sides iid N(0,1)."""
    g = rng_for(2, n, seed)
    A = g.standard_normal((n, n)).T  # F-ordered view: the generated buffer is column-major A
    B = g.standard_normal((nrhs, n)).T
    return A, B


def conditioned(n, cond, seed=3, device=None):
    """A = Q1 diag(s) Q2^T, Q from the QR of Gaussian matrices (signs fixed so that R has a
    positive diagonal), s log-spaced from 1 down to 1/cond.  device='cuda' builds it with torch
    on the GPU (input generation only, off the measured path)."""
    g = rng_for(3, n, int(np.log10(cond) * 100), seed)
    G1 = g.standard_normal((n, n))
    G2 = g.standard_normal((n, n))
    s = np.logspace(0.0, -np.log10(cond), n)
    if device is None:
        Q1, R1 = np.linalg.qr(G1)
        Q1 *= np.sign(np.diag(R1))
        Q2, R2 = np.linalg.qr(G2)
        Q2 *= np.sign(np.diag(R2))
        return (Q1 * s) @ Q2.T
    import torch
    t1 = torch.from_numpy(G1).to(device)
    t2 = torch.from_numpy(G2).to(device)
    del G1, G2
    Q1, R1 = torch.linalg.qr(t1)
    Q1 *= torch.sign(torch.diagonal(R1))
    del R1, t1
    Q2, R2 = torch.linalg.qr(t2)
    Q2 *= torch.sign(torch.diagonal(R2))
    del R2, t2
    ts = torch.from_numpy(s).to(device)
    return (Q1 * ts) @ Q2.T


def batched_systems(count=20000, nmin=8, nmax=512, seed=4):
    """Batched small dense LU: n_i ~ U{nmin..nmax}, A_i = I - g_i J_i, J iid N(0,1)/sqrt(n_i),
    g_i ~ U[1e-3, 1e-1], one right-hand side iid N(0,1).  Returns (sizes, packed A, packed b)
    with every A_i column-major and the blocks back to back."""
    g = rng_for(4, count, seed)
    sizes = g.integers(nmin, nmax + 1, size=count).astype(np.int64)
    gam = g.uniform(1e-3, 1e-1, size=count)
    total = int((sizes * sizes).sum())
    A = np.empty(total)
    off = 0
    for i, n in enumerate(sizes):
        n = int(n)
        J = g.standard_normal((n, n)) / np.sqrt(n)
        blk = -gam[i] * J
        blk[np.diag_indices(n)] += 1.0
        A[off:off + n * n] = blk.ravel(order="F")
        off += n * n
    b = g.standard_normal(int(sizes.sum()))
    return sizes, A, b


def dense_block(n, seed):
    """One system of the batched / serving distribution: A = I - g J, b iid N(0,1)."""
    g = rng_for(5, n, seed)
    gam = g.uniform(1e-3, 1e-1)
    A = -gam * g.standard_normal((n, n)) / np.sqrt(n)
    A[np.diag_indices(n)] += 1.0
    return np.asfortranarray(A), g.standard_normal(n)


def convection_diffusion(m, pe):
    """-Laplace(u) + Pe (u_x + u_y) on the unit square, Dirichlet, h = 1/(m+1), central
    differences, scaled by h^2: diagonal 4, west/south -1 - Pe h/2, east/north -1 + Pe h/2.
    Returns CSR (rowptr int32, colind int32 sorted, vals) of order N = m^2 (row = i + m j)."""
    N = m * m
    h = 1.0 / (m + 1)
    lo, hi = -1.0 - pe * h / 2.0, -1.0 + pe * h / 2.0
    idx = np.arange(N).reshape(m, m)  # idx[j, i] = i + m j
    rows, cols, vals = [], [], []

    def add(mask_r, dr, v):
        r = idx[mask_r]
        rows.append(r.ravel())
        cols.append((r + dr).ravel())
        vals.append(np.full(r.size, v))

    full = np.ones((m, m), dtype=bool)
    south = full.copy()
    south[0, :] = False
    west = full.copy()
    west[:, 0] = False
    east = full.copy()
    east[:, m - 1] = False
    north = full.copy()
    north[m - 1, :] = False
    add(south, -m, lo)
    add(west, -1, lo)
    add(full, 0, 4.0)
    add(east, 1, hi)
    add(north, m, hi)
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = np.concatenate(vals)
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    rowptr = np.zeros(N + 1, dtype=np.int64)
    np.add.at(rowptr, r + 1, 1)
    rowptr = np.cumsum(rowptr).astype(np.int32)
    return rowptr, c.astype(np.int32), v


def serving_trace(rate=2000.0, seconds=60.0, seed=6, dense_frac=0.7, nmin=16, nmax=4096, mmin=64, mmax=1024):
    """Poisson arrivals; 70% dense systems with n log-uniform in [nmin, nmax], 30% SimpleGMRES
    on convection-diffusion grids with side ~ U{mmin..mmax} and Pe ~ U[0, 10]."""
    g = rng_for(6, int(rate), int(seconds), seed)
    t = 0.0
    out = []
    while True:
        t += g.exponential(1.0 / rate)
        if t >= seconds:
            break
        if g.random() < dense_frac:
            n = int(round(np.exp(g.uniform(np.log(nmin), np.log(nmax)))))
            out.append(dict(t=t, kind="dense", n=n, seed=int(g.integers(1 << 30))))
        else:
            out.append(dict(t=t, kind="gmres", m=int(g.integers(mmin, mmax + 1)), pe=float(g.uniform(0.0, 10.0)),
                            seed=int(g.integers(1 << 30))))
    return out
