"""TEST INFRASTRUCTURE -- ctypes wrapper over the CPU oracle (oracle/_build/libjet_oracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module. It restates the reference's hot path on the CPU (see
oracle/jet_oracle.cpp for the reference file:line each function follows).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libjet_oracle.so")
_OMP_PATH = os.path.join(_HERE, "_build", "libjet_oracle_omp.so")
_lib = None
_omp = None

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_mutants = {}


def mutant(idx):
    """The oracle built with planted mistake ORC_MUTANT=idx (mutation tests only)."""
    if idx not in _mutants:
        path = os.path.join(_HERE, "_build", f"libjet_oracle_mut{idx}.so")
        if not os.path.exists(path):
            build()
        _mutants[idx] = _configure(ctypes.CDLL(path))
    return _mutants[idx]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _configure(ctypes.CDLL(_LIB_PATH))
    return _lib


def omp_lib(threads=0):
    """The same restatement built with OpenMP (identical results, loops over
    independent items on `threads` threads; 0 keeps the current count): the
    multi-core CPU baseline. Returns (library, threads in effect)."""
    global _omp
    if _omp is None:
        if not os.path.exists(_OMP_PATH):
            build()
        _omp = _configure(ctypes.CDLL(_OMP_PATH))
    return _omp, _omp.orc_threads(int(threads))


def _configure(L):
    if True:
        c = ctypes
        L.orc_mix64.restype = c.c_uint64
        L.orc_mix64.argtypes = [c.c_uint64]
        L.orc_rng.restype = c.c_uint64
        L.orc_rng.argtypes = [c.c_uint64, c.c_uint32, c.c_uint32, c.c_uint32, c.c_uint32]
        L.orc_murmur3_32.restype = c.c_uint32
        L.orc_murmur3_32.argtypes = [c.c_char_p, c.c_int64, c.c_uint32]
        L.orc_pod_hash_i32.restype = c.c_uint32
        L.orc_pod_hash_i32.argtypes = [c.c_int32]
        L.orc_xorshift32.restype = c.c_uint32
        L.orc_xorshift32.argtypes = [c.c_uint32]
        L.orc_gain_bucket.restype = c.c_int32
        L.orc_gain_bucket.argtypes = [c.c_int64, c.c_int64]
        L.orc_hec_hn.restype = None
        L.orc_hec_hn.argtypes = [c.c_int32, i64p, i32p, i32p, i32p, c.c_int32, c.c_uint64, c.c_int32, i32p]
        L.orc_map_construct.restype = c.c_int32
        L.orc_map_construct.argtypes = [c.c_int32, i32p, i32p]
        L.orc_coarsen_map.restype = c.c_int32
        L.orc_coarsen_map.argtypes = [c.c_int32, c.c_int32, c.c_int32, c.c_uint64, c.c_int32, i64p, i32p, i32p,
                                      i32p, i32p]
        L.orc_contract.restype = c.c_int64
        L.orc_contract.argtypes = [c.c_int32, i64p, i32p, i32p, i32p, i32p, c.c_int32, i64p, i32p, i32p, i32p]
        L.orc_total_cut.restype = c.c_int64
        L.orc_total_cut.argtypes = [c.c_int32, i64p, i32p, i32p, i32p]
        L.orc_conn_build.restype = c.c_int64
        L.orc_conn_build.argtypes = [c.c_int32, i64p, i32p, i32p, i32p, c.c_int32, i64p, i32p, i64p, i32p]
        L.orc_refine_step.restype = c.c_int32
        L.orc_refine_step.argtypes = [c.c_int32, c.c_int32, i64p, i32p, i32p, i32p, c.c_int32, c.c_double,
                                      c.c_double, i32p, i32p, i32p]
        L.orc_refine_level.restype = c.c_int
        L.orc_refine_level.argtypes = [c.c_int32, i64p, i32p, i32p, i32p, c.c_int32, c.c_double, c.c_double,
                                       c.c_int32, c.c_int32, i32p, i64p]
        L.orc_partition.restype = c.c_int
        L.orc_partition.argtypes = [c.c_int32, c.c_int32, c.c_double, c.c_double, c.c_int32, c.c_int32, i64p,
                                    i32p, i32p, i32p, c.c_int32, c.c_uint64, i32p, c.POINTER(c.c_int64), i64p,
                                    c.POINTER(c.c_int32), c.c_int32]
        L.jet_host_initpart.restype = c.c_int
        L.jet_host_initpart.argtypes = [c.c_int32, i64p, i32p, i32p, i32p, c.c_int32, c.c_double, c.c_uint64, i32p]
        L.orc_refine_level_trace.restype = c.c_int32
        L.orc_refine_level_trace.argtypes = [c.c_int32, i64p, i32p, i32p, i32p, c.c_int32, c.c_double, c.c_double,
                                             c.c_int32, c.c_int32, i32p, i64p, i64p, c.c_int32]
        L.orc_set_slot_order_ties.restype = None
        L.orc_set_slot_order_ties.argtypes = [c.c_int32]
        L.orc_threads.restype = c.c_int32
        L.orc_threads.argtypes = [c.c_int32]
    return L


def _g(g):
    rowptr = np.ascontiguousarray(g["rowptr"], dtype=np.int64)
    col = np.ascontiguousarray(g["col"], dtype=np.int32)
    ew = np.ascontiguousarray(g["ew"], dtype=np.int32)
    return len(rowptr) - 1, rowptr, col, ew


def _vw(g):
    n = len(g["rowptr"]) - 1
    vw = g.get("vw")
    if vw is None:
        return np.ones(n, dtype=np.int32)
    return np.ascontiguousarray(vw, dtype=np.int32)


def mix64(x):
    return lib().orc_mix64(x)


def rng(seed, level, site, pas, idx):
    return lib().orc_rng(seed, level, site, pas, idx)


def murmur3_32(data: bytes, seed: int):
    return lib().orc_murmur3_32(data, len(data), seed)


def pod_hash_i32(key):
    return lib().orc_pod_hash_i32(key)


def xorshift32(x):
    return lib().orc_xorshift32(x)


def gain_bucket(gain, vw):
    return lib().orc_gain_bucket(int(gain), int(vw))


def hec_hn(g, uniform, seed, level):
    n, rp, col, ew = _g(g)
    out = np.empty(n, dtype=np.int32)
    lib().orc_hec_hn(n, rp, col, ew, _vw(g), int(uniform), seed, level, out)
    return out


def map_construct(hn):
    hn = np.ascontiguousarray(hn, dtype=np.int32)
    out = np.empty(len(hn), dtype=np.int32)
    nc = lib().orc_map_construct(len(hn), hn, out)
    return out, nc


def coarsen_map(g, alg, uniform, seed, level=0):
    n, rp, col, ew = _g(g)
    out = np.empty(n, dtype=np.int32)
    nc = lib().orc_coarsen_map(alg, level, int(uniform), seed, n, rp, col, ew, _vw(g), out)
    return out, nc


def contract(g, mp, nc):
    n, rp, col, ew = _g(g)
    m = len(col)
    orp = np.empty(nc + 1, dtype=np.int64)
    ocol = np.empty(max(m, 1), dtype=np.int32)
    oew = np.empty(max(m, 1), dtype=np.int32)
    ovw = np.empty(max(nc, 1), dtype=np.int32)
    mc = lib().orc_contract(n, rp, col, ew, _vw(g), np.ascontiguousarray(mp, dtype=np.int32), nc, orp, ocol, oew,
                            ovw)
    return {"rowptr": orp, "col": ocol[:mc].copy(), "ew": oew[:mc].copy(), "vw": ovw[:nc].copy()}


def total_cut(g, part):
    n, rp, col, ew = _g(g)
    return lib().orc_total_cut(n, rp, col, ew, np.ascontiguousarray(part, dtype=np.int32))


def conn_build(g, part, k):
    n, rp, col, ew = _g(g)
    cap = int(np.minimum(np.diff(rp), k).sum())
    off = np.empty(n + 1, dtype=np.int64)
    prt = np.empty(max(cap, 1), dtype=np.int32)
    val = np.empty(max(cap, 1), dtype=np.int64)
    ts = np.empty(n, dtype=np.int32)
    w = lib().orc_conn_build(n, rp, col, ew, np.ascontiguousarray(part, dtype=np.int32), k, off, prt, val, ts)
    return off, prt[:w].copy(), val[:w].copy(), ts


def refine_step(kind, g, part, k, imb, filter_ratio=0.75):
    n, rp, col, ew = _g(g)
    moves = np.empty(max(n, 1), dtype=np.int32)
    dest = np.empty(max(n, 1), dtype=np.int32)
    cnt = lib().orc_refine_step(kind, n, rp, col, ew, _vw(g), k, imb, filter_ratio,
                                np.ascontiguousarray(part, dtype=np.int32), moves, dest)
    return moves[:cnt].copy(), dest[:cnt].copy()


def refine_level(g, part, k, imb, tol=0.999, ultra=False, uniform=False):
    n, rp, col, ew = _g(g)
    p = np.array(part, dtype=np.int32)
    rec = np.zeros(8, dtype=np.int64)
    lib().orc_refine_level(n, rp, col, ew, _vw(g), k, imb, tol, int(ultra), int(uniform), p, rec)
    return p, rec


def refine_level_trace(g, part, k, imb, tol=0.999, ultra=False, uniform=False, L=None, max_iters=4096):
    """refine_level plus one record per iteration: (kind 0 lp / 1 weak / 2
    strong, moves, cut2, total_imb, take 0 none / 1 balance / 2 cut, count).
    L: a library from mutant(), or None for the oracle itself."""
    n, rp, col, ew = _g(g)
    p = np.array(part, dtype=np.int32)
    rec = np.zeros(8, dtype=np.int64)
    tr = np.zeros((max_iters, 6), dtype=np.int64)
    it = (L or lib()).orc_refine_level_trace(n, rp, col, ew, _vw(g), k, imb, tol, int(ultra), int(uniform), p, rec,
                                             tr, max_iters)
    return p, rec, [tuple(int(x) for x in row) for row in tr[:min(it, max_iters)]]


def set_slot_order_ties(on):
    """Tie-break in jet_lp / rebalance_weak: False = lowest part id (default),
    True = first maximum in the Serial backend's slot order."""
    lib().orc_set_slot_order_ties(int(bool(on)))


def partition(g, k, imb=1.03, alg=0, seed=0, tol=0.999, ultra=False, uniform=None, repeat_shortcut=False, L=None):
    """Full multilevel partition; returns (part, cut, level records). With
    repeat_shortcut, the exact shortcuts for a weighted-HEC coarsening fixed
    point are taken (shared repeated levels, no re-refinement of an unchanged
    repeated level); the default is the literal level loop. L: another build
    of the oracle (omp_lib()), or None for the serial one."""
    n, rp, col, ew = _g(g)
    if uniform is None:
        uniform = bool(g.get("uniform", False))
    part = np.empty(n, dtype=np.int32)
    cut = ctypes.c_int64(0)
    levels = np.zeros(8 * 256, dtype=np.int64)
    nl = ctypes.c_int32(0)
    rc = (L or lib()).orc_partition(alg, k, imb, tol, int(ultra), n, rp, col, ew, _vw(g), int(uniform), seed, part,
                             ctypes.byref(cut), levels, ctypes.byref(nl), int(repeat_shortcut))
    if rc != 0:
        raise RuntimeError(f"oracle partition failed: {rc}")
    return part, cut.value, levels[: 8 * nl.value].reshape(-1, 8)


def initpart(g, k, imb, seed=0):
    n, rp, col, ew = _g(g)
    part = np.empty(n, dtype=np.int32)
    rc = lib().jet_host_initpart(n, rp, col, ew, _vw(g), k, imb, seed, part)
    if rc != 0:
        raise RuntimeError("initpart failed")
    return part
