"""Seeded synthetic workloads of the matrix, isochrone and serving
configurations, shared by the benchmark and the full-size parity tests.

Graphs are jittered road grids (GridSpec): the 224 x 224 grid of the plain CPU
configuration and the 1414 x 1414 grid with 2 % long motorway chords (about
2M NBG nodes, an edge-based CCH of about 6.7M nodes) of the others.  CCH
weights are seconds, so isochrone limits of 5-60 minutes are minutes * 60.
"""
import numpy as np

GRAPH_SEED = 20261203

# name: (grid rows, chord fraction, S, T)
WORKLOADS = {
    "matrix10k": (1414, 0.02, 10000, 10000),
    "matrix1k": (1414, 0.02, 1000, 1000),
    "config1": (224, 0.0, 100, 100),
    "mid": (448, 0.02, 2000, 2000),
}

ISO_MINUTES = (5, 10, 15, 30, 45, 60)


def grid_spec(rows, chords, seed=GRAPH_SEED):
    from .graph import GridSpec
    return GridSpec(rows=rows, cols=rows, seed=seed, chord_frac=chords)


def make_queries_tagged(g, count, seed, hot_frac=0.2, n_hot=5, sigma_frac=0.02):
    """80 % uniform ranks + 20 % around 5 Gaussian hot-spots (sigma = 2 % of
    the extent), snapped to the nearest node; returned shuffled, with a flag
    per entry that is set for the hot-spot draws."""
    rng = np.random.default_rng(seed)
    n_hot_pts = int(round(hot_frac * count))
    uni = rng.integers(0, g.n, count - n_hot_pts)
    if n_hot_pts == 0 or g.x is None:
        q = np.concatenate([uni, rng.integers(0, g.n, n_hot_pts)])
        hot = np.zeros(count, dtype=bool)
    else:
        from scipy.spatial import cKDTree
        xy = np.c_[g.x, g.y].astype(np.float64)
        lo, hi = xy.min(0), xy.max(0)
        ext = hi - lo
        centers = lo + rng.random((n_hot, 2)) * ext
        which = rng.integers(0, n_hot, n_hot_pts)
        pts = centers[which] + rng.normal(0.0, 1.0, (n_hot_pts, 2)) * (sigma_frac * ext)
        _, idx = cKDTree(xy).query(pts)
        q = np.concatenate([uni, idx])
        hot = np.r_[np.zeros(count - n_hot_pts, dtype=bool), np.ones(n_hot_pts, dtype=bool)]
    perm = rng.permutation(count)
    return q[perm].astype(np.uint32), hot[perm]


def make_queries(g, count, seed, **kw):
    return make_queries_tagged(g, count, seed, **kw)[0]


def isochrone_lanes(g, count, seed):
    """Origins uniform random, limits drawn uniformly from 5-60 minutes (in
    seconds, uneven per lane)."""
    rng = np.random.default_rng(seed)
    org = rng.integers(0, g.n, count).astype(np.uint32)
    lim = (rng.choice(ISO_MINUTES, count) * 60).astype(np.uint32)
    return org, lim


def serving_requests(g, count, seed, s_range=(10, 2000), matrix_frac=0.7):
    """The mixed stream: 70 % matrices with S and T each log-uniform in
    [10, 2000] (uniform random nodes), 30 % isochrones from a uniform random
    node with a limit of 5-60 minutes (seconds)."""
    from .serving import IsochroneRequest, MatrixRequest
    rng = np.random.default_rng(seed)
    lo, hi = np.log(s_range[0]), np.log(s_range[1])
    out = []
    for _ in range(count):
        if rng.random() < matrix_frac:
            S, T = (int(round(np.exp(rng.uniform(lo, hi)))) for _ in range(2))
            out.append(MatrixRequest(rng.integers(0, g.n, S).astype(np.uint32),
                                     rng.integers(0, g.n, T).astype(np.uint32)))
        else:
            out.append(IsochroneRequest(int(rng.integers(0, g.n)), int(rng.choice(ISO_MINUTES)) * 60))
    return out
