"""Synthetic inputs of the benchmark workloads (meshes, sparse patterns, value sets).

All generators are deterministic functions of their sizes and a seed (numpy PCG64), so a
run here and a run on another machine see identical inputs.

  icosphere(s)             icosahedron with s midpoint subdivisions (s = 5: 10,242 vertices)
  vertex_adjacency(V, F)   symmetric vertex-adjacency pattern with the diagonal, CSC
  grid_mesh(nx, ny, ...)   regular grid on [0,1]^2 split into triangles with alternating
                           (checkerboard) diagonals, optional jitter and height field
  cloth(side, seed)        rest sheet + folded, noisy deformed positions
  load_mtx(path)           MatrixMarket coordinate real general -> CSC
Dense n x 3 position arrays are returned row-per-vertex; the C ABI expects them
column-major, see to_colmajor().
"""
import numpy as np


def to_colmajor(X):
    return np.ascontiguousarray(np.asarray(X, dtype=np.float64).T.reshape(-1))


def icosphere(subdivisions):
    t = (1.0 + 5.0 ** 0.5) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    V = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in V]
    for _ in range(subdivisions):
        mid = {}
        F2 = []

        def midpoint(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in mid:
                m = V[a] + V[b]
                V.append(m / np.linalg.norm(m))
                mid[key] = len(V) - 1
            return mid[key]

        for (a, b, c) in F:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            F2 += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = F2
    return np.array(V), np.array(F, dtype=np.int32)


def vertex_adjacency(nv, F):
    """CSC (indptr, indices) of the symmetric vertex adjacency with the diagonal."""
    F = np.asarray(F)
    r = np.concatenate([F[:, 0], F[:, 1], F[:, 2], F[:, 1], F[:, 2], F[:, 0], np.arange(nv)])
    c = np.concatenate([F[:, 1], F[:, 2], F[:, 0], F[:, 0], F[:, 1], F[:, 2], np.arange(nv)])
    key = np.unique(c.astype(np.int64) * nv + r)
    cols, rows = key // nv, key % nv
    indptr = np.zeros(nv + 1, np.int64)
    np.add.at(indptr, cols + 1, 1)
    indptr = np.cumsum(indptr)
    return indptr.astype(np.int32), rows.astype(np.int32)


def grid_faces(nx, ny):
    """Triangles of an nx x ny vertex grid (vertex i + j nx), diagonal by parity of i + j."""
    i, j = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), indexing="ij")
    i, j = i.reshape(-1), j.reshape(-1)
    v00 = i + j * nx
    v10 = v00 + 1
    v01 = v00 + nx
    v11 = v01 + 1
    even = (i + j) % 2 == 0
    f1 = np.where(even[:, None], np.stack([v00, v10, v11], 1), np.stack([v00, v10, v01], 1))
    f2 = np.where(even[:, None], np.stack([v00, v11, v01], 1), np.stack([v10, v11, v01], 1))
    F = np.empty((2 * len(v00), 3), np.int32)
    F[0::2] = f1
    F[1::2] = f2
    return F


def grid_mesh(nx, ny, jitter=0.0, height=0.0, seed=0):
    """Grid on [0,1]^2 with xy += N(0, (jitter h)^2), z = height sin(4 pi x) sin(4 pi y)."""
    h = 1.0 / (nx - 1)
    x, y = np.meshgrid(np.linspace(0, 1, nx), np.linspace(0, 1, ny), indexing="xy")
    x, y = x.reshape(-1), y.reshape(-1)
    if jitter:
        rng = np.random.default_rng(seed)
        x = x + rng.normal(0.0, jitter * h, x.shape)
        y = y + rng.normal(0.0, jitter * h, y.shape)
    z = height * np.sin(4 * np.pi * x) * np.sin(4 * np.pi * y) if height else np.zeros_like(x)
    return np.stack([x, y, z], 1), grid_faces(nx, ny)


def cloth(side, seed, fold=0.2, noise=1e-3):
    """(rest, F, X): flat side x side sheet on [0,1]^2; X = rest + fold sin(pi x) in z + N(0, noise^2)."""
    rest, F = grid_mesh(side, side)
    rng = np.random.default_rng(seed)
    X = rest.copy()
    X[:, 2] += fold * np.sin(np.pi * rest[:, 0])
    X += rng.normal(0.0, noise, X.shape)
    return rest, F, X


def wheel(n, seed, fold=0.1, noise=1e-3):
    """(rest, F, X): a flat disc of two rings of n vertices (radii 0.5 and 1) around a center vertex
    shared by n triangles, so the center's Hessian entries sum n element contributions;
    X = rest + fold x^2 in z + N(0, noise^2)."""
    t = 2.0 * np.pi * np.arange(n) / n
    ring = lambda r: np.stack([r * np.cos(t), r * np.sin(t), np.zeros(n)], 1)  # noqa: E731
    rest = np.concatenate([np.zeros((1, 3)), ring(0.5), ring(1.0)])
    F = np.empty((3 * n, 3), np.int32)
    for i in range(n):
        a, b = 1 + i, 1 + (i + 1) % n
        F[3 * i] = (0, a, b)
        F[3 * i + 1] = (a, a + n, b + n)
        F[3 * i + 2] = (a, b + n, b)
    rng = np.random.default_rng(seed)
    X = rest.copy()
    X[:, 2] += fold * rest[:, 0] ** 2
    X += rng.normal(0.0, noise, X.shape)
    return rest, F, X


def random_values(n, seed, lo=-1.0, hi=1.0):
    return np.random.default_rng(seed).uniform(lo, hi, n)


def load_mtx(path):
    """MatrixMarket 'coordinate real general' -> (n, indptr, indices, values) in CSC."""
    with open(path) as f:
        header = f.readline()
        if "coordinate" not in header:
            raise ValueError("only coordinate MatrixMarket files are supported")
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        nr, nc, nnz = (int(t) for t in line.split())
        data = np.loadtxt(f, ndmin=2)
    r = data[:, 0].astype(np.int64) - 1
    c = data[:, 1].astype(np.int64) - 1
    v = data[:, 2]
    order = np.lexsort((r, c))
    r, c, v = r[order], c[order], v[order]
    indptr = np.zeros(nc + 1, np.int64)
    np.add.at(indptr, c + 1, 1)
    return nc, np.cumsum(indptr).astype(np.int32), r.astype(np.int32), v


# workload definitions (sizes of the benchmark configurations)
C1_SUBDIV, C1_SEED = 5, 1
C2_SIDE, C2_SEED = 512, 2
C3_SIDE, C3_SEED = 256, 3
K_STRETCH, K_SHEAR, K_BEND = 1e3, 1e2, 1e-2


def c1_matrix(subdiv=C1_SUBDIV, seed=C1_SEED):
    V, F = icosphere(subdiv)
    indptr, indices = vertex_adjacency(len(V), F)
    return len(V), indptr, indices, random_values(len(indices), seed)


def c2_mesh(side=C2_SIDE, seed=C2_SEED):
    return grid_mesh(side, side, jitter=0.1, height=0.05, seed=seed)


def c3_cloth(side=C3_SIDE, seed=C3_SEED):
    return cloth(side, seed)


C4_COUNT, C4_MIN_SIDE, C4_MAX_SIDE, C4_SEED = 256, 32, 256, 4


def c4_sides(count=C4_COUNT, lo=C4_MIN_SIDE, hi=C4_MAX_SIDE, seed=C4_SEED):
    """Side lengths drawn log-uniformly from [lo, hi] (rounded)."""
    rng = np.random.default_rng(seed)
    return [int(round(float(np.exp(rng.uniform(np.log(lo), np.log(hi)))))) for _ in range(count)]


def c4_cloths(count=C4_COUNT, lo=C4_MIN_SIDE, hi=C4_MAX_SIDE, seed=C4_SEED):
    """Batched cloths: one (rest, F, X) per instance, each with its own deformation seed."""
    return [cloth(s, seed * 100003 + k) for k, s in enumerate(c4_sides(count, lo, hi, seed))]


def interior_edge_count(F):
    """Edges shared by exactly two faces (the bending stencils of a cloth)."""
    F = np.asarray(F, np.int64)
    e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
    e.sort(axis=1)
    _, counts = np.unique(e[:, 0] * (int(F.max()) + 1) + e[:, 1], return_counts=True)
    return int((counts == 2).sum())


def bw_algorithmic_counts(num_vertices, F):
    """Inputs, per-element constants and connectivity indices of one cloth Hessian evaluation:
    3 coordinates per vertex; 5 rest constants per face (area, 2x2 inverse) and 10 bending
    constants per interior edge; 3 vertex indices per face and 4 per interior edge."""
    ne = interior_edge_count(F)
    return {"n_in": 3 * num_vertices, "n_const": 5 * len(F) + 10 * ne, "n_conn": 3 * len(F) + 4 * ne,
            "interior_edges": ne}


C5_COUNT, C5_MIN_N, C5_MAX_N, C5_SEED = 64, 10_000, 1_000_000, 5
C5_BAND, C5_BAND_FRACTION = 64, 0.75


def c5_sizes(count=C5_COUNT, lo=C5_MIN_N, hi=C5_MAX_N, seed=C5_SEED):
    """Matrix sizes drawn log-uniformly from [lo, hi]."""
    rng = np.random.default_rng(seed)
    return [int(round(float(np.exp(rng.uniform(np.log(lo), np.log(hi)))))) for _ in range(count)]


def powerlaw_counts(n, rng, lo=4, hi=64, alpha=2.0):
    """n integers on [lo, hi] with P(k) proportional to k^-alpha (inverse-CDF sampling)."""
    k = np.arange(lo, hi + 1, dtype=np.float64)
    p = k ** -alpha
    cdf = np.cumsum(p / p.sum())
    return (lo + np.searchsorted(cdf, rng.random(n), side="right")).clip(lo, hi).astype(np.int64)


def c5_pattern(index, count=C5_COUNT, seed=C5_SEED):
    """(n, indptr, indices, values) of synthetic pattern `index`: nnz per row from the power law
    on [4, 64] (alpha 2), three quarters of a row's entries inside the band |i - j| <= 32
    (band width 64), the rest uniform over the row; symmetrised, with the diagonal; values
    U(-1, 1).  CSC, rows sorted and unique per column."""
    n = c5_sizes(count, seed=seed)[index]
    rng = np.random.default_rng(seed * 7919 + index)
    r = powerlaw_counts(n, rng)
    rows = np.repeat(np.arange(n, dtype=np.int64), r)
    band = rng.random(len(rows)) < C5_BAND_FRACTION
    half = C5_BAND // 2
    off = rng.integers(-half, half + 1, len(rows))
    cols = np.where(band, np.clip(rows + off, 0, n - 1), rng.integers(0, n, len(rows)))
    diag = np.arange(n, dtype=np.int64)
    R = np.concatenate([rows, cols, diag])
    C = np.concatenate([cols, rows, diag])
    key = np.unique(C * n + R)
    cc, rr = key // n, key % n
    indptr = np.zeros(n + 1, np.int64)
    np.add.at(indptr, cc + 1, 1)
    indptr = np.cumsum(indptr)
    vals = np.random.default_rng(seed * 104729 + index).uniform(-1.0, 1.0, len(rr))
    return n, indptr.astype(np.int32), rr.astype(np.int32), vals
