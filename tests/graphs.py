"""Small graph builders for tests (symmetric CSR, sorted rows, no self-loops)."""
import numpy as np

MASK64 = (1 << 64) - 1


def splitmix64_np(x):
    """splitmix64 output function over a uint64 numpy array."""
    z = (x.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15))
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def from_edges(n, u, v, w=None, uniform=None):
    """Symmetrise, drop self-loops, merge duplicates (weights: first kept), sort rows."""
    u = np.asarray(u, dtype=np.int64)
    v = np.asarray(v, dtype=np.int64)
    if w is None:
        w = np.ones(len(u), dtype=np.int64)
        uniform = True if uniform is None else uniform
    w = np.asarray(w, dtype=np.int64)
    keep = u != v
    u, v, w = u[keep], v[keep], w[keep]
    a = np.minimum(u, v)
    b = np.maximum(u, v)
    key = a * n + b
    key, idx = np.unique(key, return_index=True)
    a, b, w = key // n, key % n, w[idx]
    src = np.concatenate([a, b])
    dst = np.concatenate([b, a])
    ww = np.concatenate([w, w])
    order = np.lexsort((dst, src))
    src, dst, ww = src[order], dst[order], ww[order]
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(rowptr, src + 1, 1)
    rowptr = np.cumsum(rowptr)
    return {"rowptr": rowptr, "col": dst.astype(np.int32), "ew": ww.astype(np.int32),
            "vw": np.ones(n, dtype=np.int32), "uniform": bool(uniform)}


def grid2d(nx, ny):
    ids = np.arange(nx * ny).reshape(ny, nx)
    u = np.concatenate([ids[:, :-1].ravel(), ids[:-1, :].ravel()])
    v = np.concatenate([ids[:, 1:].ravel(), ids[1:, :].ravel()])
    return from_edges(nx * ny, u, v)


def stencil3d(nx, ny, nz, seed=0):
    """7-point stencil, id (z*ny+y)*nx+x, w = 1 + splitmix64(seed ^ (min*N+max)) % 10."""
    N = nx * ny * nz
    ids = np.arange(N).reshape(nz, ny, nx)
    u = np.concatenate([ids[:, :, :-1].ravel(), ids[:, :-1, :].ravel(), ids[:-1, :, :].ravel()])
    v = np.concatenate([ids[:, :, 1:].ravel(), ids[:, 1:, :].ravel(), ids[1:, :, :].ravel()])
    key = (np.minimum(u, v) * N + np.maximum(u, v)).astype(np.uint64)
    w = 1 + (splitmix64_np(np.uint64(seed) ^ key) % np.uint64(10)).astype(np.int64)
    return from_edges(N, u, v, w, uniform=False)


def rmat(scale, ef, seed=0, a=0.57, b=0.19, c=0.19):
    rs = np.random.default_rng(seed)
    n = 1 << scale
    m = ef * n
    u = np.zeros(m, dtype=np.int64)
    v = np.zeros(m, dtype=np.int64)
    for _ in range(scale):
        r = rs.random(m)
        bu = (r >= a + b).astype(np.int64)
        bv = (((r >= a) & (r < a + b)) | (r >= a + b + c)).astype(np.int64)
        u = (u << 1) | bu
        v = (v << 1) | bv
    return from_edges(n, u, v)


def random_graph(n, m, seed=0, wmax=1):
    rs = np.random.default_rng(seed)
    u = rs.integers(0, n, m)
    v = rs.integers(0, n, m)
    w = rs.integers(1, wmax + 1, m) if wmax > 1 else None
    return from_edges(n, u, v, w, uniform=(wmax == 1))


def with_weights(g, seed=0, wmax=4, vwmax=1):
    rs = np.random.default_rng(seed)
    rp, col = g["rowptr"], g["col"]
    n = len(rp) - 1
    src = np.repeat(np.arange(n), np.diff(rp))
    a = np.minimum(src, col).astype(np.uint64)
    b = np.maximum(src, col).astype(np.uint64)
    w = 1 + (splitmix64_np(np.uint64(seed) ^ (a * np.uint64(n) + b)) % np.uint64(wmax)).astype(np.int32)
    h = dict(g)
    h["ew"] = w.astype(np.int32)
    h["uniform"] = False
    if vwmax > 1:
        h["vw"] = rs.integers(1, vwmax + 1, n).astype(np.int32)
    return h


def star(n):
    return from_edges(n, np.zeros(n - 1, dtype=np.int64), np.arange(1, n))


def path(n):
    return from_edges(n, np.arange(n - 1), np.arange(1, n))


def clique(n):
    u, v = np.triu_indices(n, 1)
    return from_edges(n, u, v)


def heavy_bipartite(hubs, leaves, seed=0, density=0.5, hub_w=1000, leaf_pairs=False):
    """Hubs of vertex weight hub_w joined to unit-weight leaves at random, edge
    weights in [1, 5]. Weighted HEC (heuristics.hpp:211-248) maps every vertex
    to itself here once leaves only see hubs, so coarsening reaches a fixed
    point and runs to the 200-level limit (contract.hpp:508-527). With
    leaf_pairs, leaves 2i and 2i+1 are also joined, so the first level merges
    them and the fixed point starts one level later."""
    rng = np.random.default_rng(seed)
    hu, lv = np.nonzero(rng.random((hubs, leaves)) < density)
    u = list(hu)
    v = list(hubs + lv)
    if leaf_pairs:
        for i in range(0, leaves - 1, 2):
            u.append(hubs + i)
            v.append(hubs + i + 1)
    g = from_edges(hubs + leaves, np.array(u), np.array(v), w=rng.integers(1, 6, len(u)))
    g["vw"] = np.ones(hubs + leaves, dtype=np.int32)
    g["vw"][:hubs] = hub_w
    return g
