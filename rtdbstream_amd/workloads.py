"""Seeded synthetic point streams for the sliding-DBSCAN workloads.

All generators use numpy's PCG64 (np.random.default_rng(seed)), return FP32
point-major arrays of shape (n, dim), and are deterministic for a seed.

  blobs2d   - 2-D: 70% from 20 isotropic Gaussian blobs (sigma 1, centres
              U[0,100]^2), 30% uniform noise in [0,100]^2, shuffled.
              (W 10,000, S 1,000, 20 slides, Eps 0.8, MinPts 5, seed 0)
  walks3d   - 3-D: 500 random walks (start U(100x100x10), step N(0, 0.05^2)
              per axis, reflecting walls), interleaved by timestep with the
              walks permuted within a step; 15% uniform noise in the box at
              uniformly random stream positions.
              (W 100,000, S 10,000, 50 slides, Eps 0.5, MinPts 10, seed 1)
  skewed3d  - 3-D: 200 anisotropic Gaussian clusters (per-axis sigma
              U[0.2,2], centres U[0,1000]^3; weights ~ rank^-SKEW_ZIPF),
              10% uniform noise in the box, shuffled.
              (W 1,000,000, S 50,000, 40 slides, Eps 1.0, MinPts 20, seed 2)
  multi     - the 64-stream batch: stream i (seed 1000+i) alternates blobs2d
              and walks3d; W = round10(logU[5e3, 2e5]), S = W/10,
              Eps ~ U[0.3, 1.2], MinPts in {5, 10, 20}.
  served    - 256 streams from all three generators, W = round10(logU[2e3, 5e5]).
"""
import math

import numpy as np

WORKLOADS = {
    "blobs2d": dict(dim=2, window=10_000, stride=1_000, slides=20, eps=0.8, min_pts=5, seed=0),
    "walks3d": dict(dim=3, window=100_000, stride=10_000, slides=50, eps=0.5, min_pts=10, seed=1),
    "skewed3d": dict(dim=3, window=1_000_000, stride=50_000, slides=40, eps=1.0, min_pts=20, seed=2),
}

# Zipf exponent of the skewed3d cluster weights. The per-axis sigma range
# alone (down to 0.2) already puts ~300x the mean occupancy of non-empty
# Eps-cells into the densest cell (cell_skew of the first window: 307), above
# the ~100x aimed for, so the weights are left equal.
SKEW_ZIPF = 0.0


def blobs2d(n, seed=0, n_blobs=20, sigma=1.0, box=100.0, frac_blob=0.7):
    rng = np.random.default_rng(seed)
    nb = int(round(n * frac_blob))
    centres = rng.uniform(0.0, box, size=(n_blobs, 2))
    which = rng.integers(0, n_blobs, size=nb)
    blob = centres[which] + rng.normal(0.0, sigma, size=(nb, 2))
    noise = rng.uniform(0.0, box, size=(n - nb, 2))
    pts = np.concatenate([blob, noise]).astype(np.float32)
    return pts[rng.permutation(n)]


def _reflect(x, lo, hi):
    span = hi - lo
    y = np.mod(x - lo, 2.0 * span)
    return lo + np.where(y > span, 2.0 * span - y, y)


def walks3d(n, seed=1, n_walks=500, step=0.05, box=(100.0, 100.0, 10.0), frac_noise=0.15):
    rng = np.random.default_rng(seed)
    box = np.asarray(box, dtype=np.float64)
    n_noise = int(round(n * frac_noise))
    n_walk = n - n_noise
    steps = -(-n_walk // n_walks)  # timesteps needed
    start = rng.uniform(0.0, 1.0, size=(n_walks, 3)) * box
    inc = rng.normal(0.0, step, size=(steps, n_walks, 3))
    pos = start[None, :, :] + np.cumsum(inc, axis=0)
    for a in range(3):
        pos[..., a] = _reflect(pos[..., a], 0.0, box[a])
    # interleave by timestep, walks permuted within each step
    order = np.argsort(rng.random((steps, n_walks)), axis=1)
    walk_pts = np.take_along_axis(pos, order[:, :, None], axis=1).reshape(-1, 3)[:n_walk]
    noise = rng.uniform(0.0, 1.0, size=(n_noise, 3)) * box
    is_noise = np.zeros(n, dtype=bool)
    is_noise[rng.choice(n, size=n_noise, replace=False)] = True
    out = np.empty((n, 3), dtype=np.float64)
    out[is_noise] = noise
    out[~is_noise] = walk_pts
    return out.astype(np.float32)


def skewed3d(n, seed=2, n_clusters=200, box=1000.0, frac_noise=0.10, zipf=SKEW_ZIPF):
    rng = np.random.default_rng(seed)
    n_noise = int(round(n * frac_noise))
    nc = n - n_noise
    w = 1.0 / np.arange(1, n_clusters + 1, dtype=np.float64) ** zipf
    w = w[rng.permutation(n_clusters)]
    w /= w.sum()
    counts = rng.multinomial(nc, w)
    centres = rng.uniform(0.0, box, size=(n_clusters, 3))
    sig = rng.uniform(0.2, 2.0, size=(n_clusters, 3))
    which = np.repeat(np.arange(n_clusters), counts)
    pts = centres[which] + rng.normal(0.0, 1.0, size=(nc, 3)) * sig[which]
    noise = rng.uniform(0.0, box, size=(n_noise, 3))
    allp = np.concatenate([pts, noise]).astype(np.float32)
    return allp[rng.permutation(n)]


def generate(name, n=None, seed=None):
    cfg = WORKLOADS[name]
    if n is None:
        n = cfg["window"] + cfg["slides"] * cfg["stride"]
    if seed is None:
        seed = cfg["seed"]
    return {"blobs2d": blobs2d, "walks3d": walks3d, "skewed3d": skewed3d}[name](n, seed=seed)


def cell_skew(points, eps):
    """max / mean occupancy of the non-empty cells of side eps."""
    p = np.asarray(points, dtype=np.float64)
    ix = np.floor((p - p.min(axis=0)) / eps).astype(np.int64)
    _, cnt = np.unique(ix, axis=0, return_counts=True)
    return float(cnt.max()) / float(cnt.mean())


def _round10(x):
    return max(10, int(round(x / 10.0)) * 10)


def multi_stream_specs(n_streams=64, slides=20, wmin=5_000, wmax=200_000, base_seed=1000):
    """Per-stream parameters of the 64-stream batch (deterministic)."""
    specs = []
    for i in range(n_streams):
        rng = np.random.default_rng(base_seed + i)
        W = _round10(math.exp(rng.uniform(math.log(wmin), math.log(wmax))))
        eps = float(np.float32(rng.uniform(0.3, 1.2)))
        min_pts = int(rng.choice([5, 10, 20]))
        gen = "blobs2d" if i % 2 == 0 else "walks3d"
        specs.append(dict(index=i, gen=gen, dim=2 if gen == "blobs2d" else 3, window=W, stride=W // 10,
                          eps=eps, min_pts=min_pts, slides=slides, seed=base_seed + i))
    return specs


def served_stream_specs(n_streams=256, slides=20, wmin=2_000, wmax=500_000, base_seed=5000):
    """Per-stream parameters of the served workload (deterministic)."""
    gens = ["blobs2d", "walks3d", "skewed3d"]
    specs = []
    for i in range(n_streams):
        rng = np.random.default_rng(base_seed + i)
        W = _round10(math.exp(rng.uniform(math.log(wmin), math.log(wmax))))
        gen = gens[i % 3]
        base = WORKLOADS[gen]
        specs.append(dict(index=i, gen=gen, dim=base["dim"], window=W, stride=W // 10, eps=base["eps"],
                          min_pts=base["min_pts"], slides=slides, seed=base_seed + i))
    return specs


def stream_points(spec):
    n = spec["window"] + spec["slides"] * spec["stride"]
    return {"blobs2d": blobs2d, "walks3d": walks3d, "skewed3d": skewed3d}[spec["gen"]](n, seed=spec["seed"])
