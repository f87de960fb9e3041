"""Python binding of the MI355X Jet partitioner (ctypes over libjet_amd.so).

Mirrors the reference's entry points (header/jet.h:46-107): ``partition``,
``partition_host``, ``partition_serial`` and their ``_big`` variants, taking a
CSR graph and a ``Config`` with the fields of ``config_t``
(header/jet_config.h:43-52). Every call runs the HIP path; there is no CPU
fallback, and loading fails loudly when the native library is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# JET_AMD_LIB selects another build of the library (for comparing builds).
LIB_PATH = os.environ.get("JET_AMD_LIB") or os.path.join(os.path.dirname(_HERE), "libjet_amd.so")

JET_OK, JET_EINVAL, JET_ENOMEM, JET_EINIT, JET_EDEVICE, JET_EUNSUPPORTED = 0, -1, -2, -3, -4, -5
JET_EAGAIN = 1
PTR_HOST, PTR_DEVICE = 0, 1
MAX_LEVELS = 256


class JetError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"jet error {code}: {msg}")
        self.code = code


class Config(ctypes.Structure):
    """config_t (header/jet_config.h:43-52)."""
    _fields_ = [("coarsening_alg", ctypes.c_int32), ("num_iter", ctypes.c_int32),
                ("max_imb_ratio", ctypes.c_double), ("num_parts", ctypes.c_int32),
                ("refine_tolerance", ctypes.c_double), ("dump_coarse", ctypes.c_int32),
                ("verbose", ctypes.c_int32), ("ultra_settings", ctypes.c_int32)]

    def __init__(self, num_parts=2, max_imb_ratio=1.03, coarsening_alg=0, num_iter=1, refine_tolerance=0.999,
                 ultra_settings=False, verbose=False):
        super().__init__(coarsening_alg, num_iter, max_imb_ratio, num_parts, refine_tolerance, 0, int(verbose),
                         int(ultra_settings))


class LevelStats(ctypes.Structure):
    _fields_ = [("edge_cut", ctypes.c_int64), ("imb", ctypes.c_double), ("num_edges", ctypes.c_int64),
                ("num_vertices", ctypes.c_int64), ("total_ref_time", ctypes.c_double),
                ("total_iterations", ctypes.c_int32), ("lp_iterations", ctypes.c_int32)]


class Stats(ctypes.Structure):
    """experiment_data (header/experiment_data.hpp:50-243)."""
    _fields_ = [("total_s", ctypes.c_double), ("coarsen_s", ctypes.c_double), ("map_s", ctypes.c_double),
                ("build_s", ctypes.c_double), ("init_partition_s", ctypes.c_double), ("refine_s", ctypes.c_double),
                ("num_levels", ctypes.c_int32), ("attempts", ctypes.c_int32), ("finest_imb_ratio", ctypes.c_double),
                ("finest_edge_cut", ctypes.c_int64), ("largest_part", ctypes.c_int64),
                ("smallest_part", ctypes.c_int64), ("levels", LevelStats * MAX_LEVELS)]


class HostGraph(ctypes.Structure):
    """jet_host_graph_t (library-allocated CSR in host memory)."""
    _fields_ = [("n", ctypes.c_int64), ("m", ctypes.c_int64), ("row_offsets", ctypes.POINTER(ctypes.c_int64)),
                ("col", ctypes.POINTER(ctypes.c_int32)), ("ew", ctypes.POINTER(ctypes.c_int32)),
                ("uniform_ew", ctypes.c_int32)]


class Quality(ctypes.Structure):
    """jet_quality_t (src/part_stat.hpp, app/part_eval.cpp)."""
    _fields_ = [("edge_cut", ctypes.c_int64), ("comm_size", ctypes.c_int64), ("max_part_cut", ctypes.c_int64),
                ("largest_part", ctypes.c_int64), ("smallest_part", ctypes.c_int64), ("max_imb", ctypes.c_double),
                ("min_imb", ctypes.c_double)]


class JobTimes(ctypes.Structure):
    """jet_job_times_t: seconds since the context was created."""
    _fields_ = [("submit_s", ctypes.c_double), ("start_s", ctypes.c_double), ("end_s", ctypes.c_double)]


_lib = None
VP = ctypes.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"native library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        c = ctypes
        i32, i64, u64, dbl = c.c_int32, c.c_int64, c.c_uint64, c.c_double
        L.jet_last_error.restype = c.c_char_p
        L.jet_device_count.restype = c.c_int
        sigs = {
            "jet_config_default": [c.POINTER(Config)],
            "jet_partition_csr": [c.POINTER(Config), i64, VP, VP, VP, VP, i32, i32, u64, VP, c.POINTER(i64),
                                  c.POINTER(Stats)],
            "jet_partition_csr32": [c.POINTER(Config), i64, VP, VP, VP, VP, i32, i32, u64, VP, c.POINTER(i64),
                                    c.POINTER(Stats)],
            "jet_partition_csr_big": [c.POINTER(Config), i64, VP, VP, VP, VP, i32, i32, u64, VP, c.POINTER(i64),
                                      c.POINTER(Stats)],
            "jet_partition_csr_biggest": [c.POINTER(Config), i64, VP, VP, VP, VP, i32, i32, u64, VP,
                                          c.POINTER(i64), c.POINTER(Stats)],
            "jet_coarsen_map": [i32, i32, i32, u64, i64, VP, VP, VP, VP, i32, VP, c.POINTER(i32)],
            "jet_hec_heavy_neighbours": [i32, i32, u64, i64, VP, VP, VP, VP, i32, VP],
            "jet_map_construct": [i64, VP, i32, VP, c.POINTER(i32)],
            "jet_contract": [i64, VP, VP, VP, VP, VP, i32, i32, VP, VP, VP, VP, c.POINTER(i64)],
            "jet_conn_tables": [i64, VP, VP, VP, VP, i32, i32, VP, VP, VP, VP],
            "jet_refine_step": [i32, i64, VP, VP, VP, VP, VP, i32, dbl, dbl, i32, VP, VP, c.POINTER(i32)],
            "jet_refine_level": [i64, VP, VP, VP, VP, i32, dbl, dbl, i32, i32, i32, VP, VP],
            "jet_refine_level_trace": [i64, VP, VP, VP, VP, i32, dbl, dbl, i32, i32, i32, VP, VP, VP, i32, VP],
            "jet_project": [i64, VP, VP, i32, VP],
            "jet_edge_cut": [i64, VP, VP, VP, VP, i32, c.POINTER(i64)],
            "jet_edge_cut_noalloc": [i64, VP, VP, VP, VP, c.POINTER(i64)],
            "jet_part_sizes": [i64, VP, VP, i32, i32, VP],
            "jet_gain_bucket_batch": [i64, VP, VP, VP],
            "jet_generate_grid2d": [i64, i64, VP, VP],
            "jet_generate_stencil3d": [i64, i64, i64, u64, i32, VP, VP, VP],
            "jet_generate_rmat": [i32, i32, dbl, dbl, dbl, u64, VP, VP, c.POINTER(i64)],
            "jet_ctx_create": [i32, i32, i64, c.POINTER(VP)],
            "jet_ctx_destroy": [VP],
            "jet_submit": [VP, c.POINTER(Config), i64, VP, VP, VP, VP, i32, i32, u64, VP, c.POINTER(i64)],
            "jet_poll": [VP, i64, i32, c.POINTER(i64), c.POINTER(Stats), c.POINTER(JobTimes)],
            "jet_partition_batch": [VP, i32, c.POINTER(Config), VP, VP, VP, VP, VP, VP, VP, VP, i32, VP, VP, VP],
            "jet_load_metis": [c.c_char_p, c.POINTER(HostGraph)],
            "jet_free_graph": [c.POINTER(HostGraph)],
            "jet_load_config": [c.c_char_p, c.POINTER(Config)],
            "jet_write_partition": [c.c_char_p, i64, VP],
            "jet_read_partition": [c.c_char_p, i64, VP],
            "jet_partition_quality": [i64, VP, VP, VP, VP, VP, i32, i32, c.POINTER(Quality)],
            "jet_cut_heatmap": [i64, VP, VP, VP, VP, i32, i32, VP],
            "jet_load_hierarchy": [c.c_char_p, i32, c.POINTER(VP)],
            "jet_hierarchy_levels": [VP],
            "jet_hierarchy_vertices": [VP, i32],
            "jet_hierarchy_edges": [VP, i32],
            "jet_free_hierarchy": [VP],
            "jet_load_coarse_part": [c.c_char_p, i64, VP],
            "jet_uncoarsen_hierarchy": [c.POINTER(Config), VP, VP, VP, c.POINTER(i64), c.POINTER(Stats)],
            "jet_tuning_set": [c.c_char_p, i64],
            "jet_tuning_get": [c.c_char_p, c.POINTER(i64)],
            "jet_tuning_load": [c.c_char_p],
            "jet_prof_enable": [c.c_char_p],
            "jet_prof_read": [c.c_char_p, c.POINTER(i64), c.POINTER(dbl), c.POINTER(dbl)],
        }
        for name, args in sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = None if name in ("jet_config_default", "jet_free_graph", "jet_free_hierarchy") else c.c_int
        L.jet_hierarchy_levels.restype = i32
        L.jet_hierarchy_vertices.restype = i64
        L.jet_hierarchy_edges.restype = i64
        _lib = L
    return _lib


def _check(rc):
    if rc != JET_OK:
        raise JetError(rc, lib().jet_last_error().decode())


def _p(a):
    """Pointer of a numpy array, or an int device address, or None."""
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data


def _c(a, dt):
    return None if a is None else np.ascontiguousarray(a, dtype=dt)


def device_count():
    return lib().jet_device_count()


def _graph(g):
    return (_c(g["rowptr"], np.int64), _c(g["col"], np.int32), _c(g.get("ew"), np.int32),
            _c(g.get("vw"), np.int32))


def partition_csr(g, cfg, seed=0, uniform=None, stats=False):
    """jet_partition_csr on host arrays; returns (part, cut[, Stats])."""
    rp, col, ew, vw = _graph(g)
    n = len(rp) - 1
    if uniform is None:
        uniform = bool(g.get("uniform", False))
    part = np.empty(n, dtype=np.int32)
    cut = ctypes.c_int64(0)
    st = Stats()
    _check(lib().jet_partition_csr(ctypes.byref(cfg), n, _p(rp), _p(col), _p(ew), _p(vw), int(uniform), PTR_HOST,
                                   seed, _p(part), ctypes.byref(cut), ctypes.byref(st)))
    return (part, cut.value, st) if stats else (part, cut.value)


def partition_device(n, rowptr_ptr, col_ptr, ew_ptr, vw_ptr, part_ptr, cfg, seed=0, uniform=False):
    """jet_partition_csr on device pointers (ints); returns (cut, Stats)."""
    cut = ctypes.c_int64(0)
    st = Stats()
    _check(lib().jet_partition_csr(ctypes.byref(cfg), n, rowptr_ptr, col_ptr, ew_ptr, vw_ptr, int(uniform),
                                   PTR_DEVICE, seed, part_ptr, ctypes.byref(cut), ctypes.byref(st)))
    return cut.value, st


# Reference entry-point names (header/jet.h:46-107). They return
# (part, edge_cut, stats) for a CSR dict {rowptr, col, ew, vw}.
def _partition_named(g, config, uniform_ew, seed, big):
    # converted arrays are bound to names so that they outlive the call
    rp = np.ascontiguousarray(g["rowptr"])
    n = len(rp) - 1
    part = np.empty(n, dtype=np.int32)
    cut = ctypes.c_int64(0)
    st = Stats()
    if big == "biggest":
        rp64, col, ew, vw = (_c(rp, np.int64), _c(g["col"], np.int64), _c(g.get("ew"), np.int64),
                             _c(g.get("vw"), np.int64))
        rc = lib().jet_partition_csr_biggest(ctypes.byref(config), n, _p(rp64), _p(col), _p(ew), _p(vw),
                                             int(uniform_ew), PTR_HOST, seed, _p(part), ctypes.byref(cut),
                                             ctypes.byref(st))
    elif big:
        rp64, col, ew, vw = (_c(rp, np.int64), _c(g["col"], np.int32), _c(g.get("ew"), np.int64),
                             _c(g.get("vw"), np.int64))
        rc = lib().jet_partition_csr_big(ctypes.byref(config), n, _p(rp64), _p(col), _p(ew), _p(vw),
                                         int(uniform_ew), PTR_HOST, seed, _p(part), ctypes.byref(cut),
                                         ctypes.byref(st))
    else:
        rp32, col, ew, vw = (_c(rp, np.int32), _c(g["col"], np.int32), _c(g.get("ew"), np.int32),
                             _c(g.get("vw"), np.int32))
        rc = lib().jet_partition_csr32(ctypes.byref(config), n, _p(rp32), _p(col), _p(ew), _p(vw), int(uniform_ew),
                                       PTR_HOST, seed, _p(part), ctypes.byref(cut), ctypes.byref(st))
    _check(rc)
    return part, cut.value, st


def partition(g, config, uniform_ew=False, seed=0):
    return _partition_named(g, config, uniform_ew, seed, False)


partition_host = partition
partition_serial = partition


def partition_big(g, config, uniform_ew=False, seed=0):
    return _partition_named(g, config, uniform_ew, seed, True)


partition_host_big = partition_big
partition_serial_big = partition_big


def partition_biggest(g, config, uniform_ew=False, seed=0):
    """64-bit ordinals (header/jet.h:88-107); n must be below 2^31 - 2."""
    return _partition_named(g, config, uniform_ew, seed, "biggest")


partition_host_biggest = partition_biggest
partition_serial_biggest = partition_biggest


# ---- stage entry points
def coarsen_map(g, alg, uniform, seed, level=0):
    rp, col, ew, vw = _graph(g)
    n = len(rp) - 1
    out = np.empty(n, dtype=np.int32)
    nc = ctypes.c_int32(0)
    _check(lib().jet_coarsen_map(alg, level, int(uniform), seed, n, _p(rp), _p(col), _p(ew), _p(vw), PTR_HOST,
                                 _p(out), ctypes.byref(nc)))
    return out, nc.value


def hec_hn(g, uniform, seed, level):
    rp, col, ew, vw = _graph(g)
    n = len(rp) - 1
    out = np.empty(n, dtype=np.int32)
    _check(lib().jet_hec_heavy_neighbours(level, int(uniform), seed, n, _p(rp), _p(col), _p(ew), _p(vw), PTR_HOST,
                                          _p(out)))
    return out


def map_construct(hn):
    hn = _c(hn, np.int32)
    out = np.empty(len(hn), dtype=np.int32)
    nc = ctypes.c_int32(0)
    _check(lib().jet_map_construct(len(hn), _p(hn), PTR_HOST, _p(out), ctypes.byref(nc)))
    return out, nc.value


def contract(g, mp, nc):
    rp, col, ew, vw = _graph(g)
    n = len(rp) - 1
    m = len(col)
    orp = np.empty(nc + 1, dtype=np.int64)
    ocol = np.empty(max(m, 1), dtype=np.int32)
    oew = np.empty(max(m, 1), dtype=np.int32)
    ovw = np.empty(max(nc, 1), dtype=np.int32)
    mc = ctypes.c_int64(0)
    mp = _c(mp, np.int32)
    _check(lib().jet_contract(n, _p(rp), _p(col), _p(ew), _p(vw), _p(mp), nc, PTR_HOST, _p(orp),
                              _p(ocol), _p(oew), _p(ovw), ctypes.byref(mc)))
    return {"rowptr": orp, "col": ocol[:mc.value].copy(), "ew": oew[:mc.value].copy(), "vw": ovw[:nc].copy()}


def conn_tables(g, part, k):
    """Raw tables: (off[n+1], slot_part, slot_val, size[n])."""
    rp, col, ew, _ = _graph(g)
    n = len(rp) - 1
    cap = int(np.minimum(np.diff(rp), k).sum())
    off = np.empty(n + 1, dtype=np.int64)
    sp = np.empty(max(cap, 1), dtype=np.int32)
    sv = np.empty(max(cap, 1), dtype=np.int32)
    sz = np.empty(n, dtype=np.int32)
    part = _c(part, np.int32)
    _check(lib().jet_conn_tables(n, _p(rp), _p(col), _p(ew), _p(part), k, PTR_HOST, _p(off), _p(sp),
                                 _p(sv), _p(sz)))
    return off, sp[:cap], sv[:cap], sz


def refine_step(kind, g, part, k, imb, filter_ratio=0.75):
    rp, col, ew, vw = _graph(g)
    n = len(rp) - 1
    mv = np.empty(max(n, 1), dtype=np.int32)
    de = np.empty(max(n, 1), dtype=np.int32)
    cnt = ctypes.c_int32(0)
    part = _c(part, np.int32)
    _check(lib().jet_refine_step(kind, n, _p(rp), _p(col), _p(ew), _p(vw), _p(part), k, imb,
                                 filter_ratio, PTR_HOST, _p(mv), _p(de), ctypes.byref(cnt)))
    return mv[:cnt.value].copy(), de[:cnt.value].copy()


def refine_level(g, part, k, imb, tol=0.999, ultra=False, uniform=False):
    rp, col, ew, vw = _graph(g)
    n = len(rp) - 1
    p = np.array(part, dtype=np.int32)
    rec = np.zeros(8, dtype=np.int64)
    _check(lib().jet_refine_level(n, _p(rp), _p(col), _p(ew), _p(vw), k, imb, tol, int(ultra), int(uniform),
                                  PTR_HOST, _p(p), _p(rec)))
    return p, rec


def refine_level_trace(g, part, k, imb, tol=0.999, ultra=False, uniform=False, max_iters=4096):
    """refine_level on the multi-kernel path plus one record per iteration:
    (kind 0 lp / 1 weak / 2 strong, -1, cut2, total_imb, take 0 none / 1
    balance / 2 cut, count)."""
    rp, col, ew, vw = _graph(g)
    n = len(rp) - 1
    p = np.array(part, dtype=np.int32)
    rec = np.zeros(8, dtype=np.int64)
    tr = np.zeros((max_iters, 6), dtype=np.int64)
    it = ctypes.c_int32(0)
    _check(lib().jet_refine_level_trace(n, _p(rp), _p(col), _p(ew), _p(vw), k, imb, tol, int(ultra), int(uniform),
                                        PTR_HOST, _p(p), _p(rec), _p(tr), max_iters, ctypes.byref(it)))
    return p, rec, [tuple(int(x) for x in row) for row in tr[:min(it.value, max_iters)]]


def project(mp, coarse):
    mp = _c(mp, np.int32)
    out = np.empty(len(mp), dtype=np.int32)
    coarse = _c(coarse, np.int32)
    _check(lib().jet_project(len(mp), _p(mp), _p(coarse), PTR_HOST, _p(out)))
    return out


def edge_cut2(g, part):
    rp, col, ew, _ = _graph(g)
    out = ctypes.c_int64(0)
    part = _c(part, np.int32)
    _check(lib().jet_edge_cut(len(rp) - 1, _p(rp), _p(col), _p(ew), _p(part), PTR_HOST,
                              ctypes.byref(out)))
    return out.value


def tuning_set(name, value):
    """Sets a result-neutral tuning parameter (jet_tuning_set)."""
    _check(lib().jet_tuning_set(name.encode(), int(value)))


def tuning_get(name):
    v = ctypes.c_int64(0)
    _check(lib().jet_tuning_get(name.encode(), ctypes.byref(v)))
    return v.value


# kernel families of the profiler (include/jet_capi.h)
PROF_FAMILIES = ("lp_select", "afterburner", "update", "weak_select")


def prof_enable(spec):
    """Times ("<family>") or counts ("<family>:count", "all:count") a kernel
    family from now on; "none" stops (jet_prof_enable)."""
    _check(lib().jet_prof_enable(spec.encode()))


def prof_read(family):
    """(units, total milliseconds, algorithmic bytes) of an enabled family."""
    u, ms, b = ctypes.c_int64(0), ctypes.c_double(0), ctypes.c_double(0)
    _check(lib().jet_prof_read(family.encode(), ctypes.byref(u), ctypes.byref(ms), ctypes.byref(b)))
    return u.value, ms.value, b.value


# the result-neutral tuning parameters (include/jet_capi.h)
TUNING_NAMES = ("wg_max_n", "batch_wg_max_n", "wg_big_deg", "small_level_n", "ab_lanes_small", "upd_lanes", "upd_big_slices",
                "lp_one_pass_n", "ct_wave_max", "ct_block_max", "redraw_lds_max",
                "sel_max_blocks", "sel_copy_wgs")


def tuning_all():
    """{name: value} of every tuning parameter."""
    return {n: tuning_get(n) for n in TUNING_NAMES}


def tuning_load(path):
    """Applies a tuning table file (jet_tuning_load)."""
    _check(lib().jet_tuning_load(os.fsencode(path)))


class tuned:
    """Context manager: sets tuning parameters, restores them on exit."""

    def __init__(self, **kv):
        self.kv = kv
        self.old = {}

    def __enter__(self):
        for k, v in self.kv.items():
            self.old[k] = tuning_get(k)
            tuning_set(k, v)
        return self

    def __exit__(self, *exc):
        for k, v in self.old.items():
            tuning_set(k, v)


def edge_cut2_noalloc(n, rowptr_ptr, col_ptr, ew_ptr, part_ptr):
    """jet_edge_cut_noalloc on device pointers (ints; ew_ptr may be None)."""
    out = ctypes.c_int64(0)
    _check(lib().jet_edge_cut_noalloc(n, rowptr_ptr, col_ptr, ew_ptr, part_ptr, ctypes.byref(out)))
    return out.value


def edge_cut2_device(n, rowptr_ptr, col_ptr, ew_ptr, part_ptr):
    """jet_edge_cut on device pointers (ints; ew_ptr may be None)."""
    out = ctypes.c_int64(0)
    _check(lib().jet_edge_cut(n, rowptr_ptr, col_ptr, ew_ptr, part_ptr, PTR_DEVICE, ctypes.byref(out)))
    return out.value


def part_sizes(vw, part, k):
    out = np.empty(k, dtype=np.int64)
    vw, part = _c(vw, np.int32), _c(part, np.int32)
    _check(lib().jet_part_sizes(len(part), _p(vw), _p(part), k, PTR_HOST, _p(out)))
    return out


def gain_bucket_batch(gains, vws):
    g = _c(gains, np.int64)
    w = _c(vws, np.int64)
    out = np.empty(len(g), dtype=np.int32)
    _check(lib().jet_gain_bucket_batch(len(g), _p(g), _p(w), _p(out)))
    return out


# ---- generators (host)
def grid2d(nx, ny):
    n = nx * ny
    m = 2 * (nx * (ny - 1) + ny * (nx - 1))
    rp = np.empty(n + 1, dtype=np.int64)
    col = np.empty(max(m, 1), dtype=np.int32)
    _check(lib().jet_generate_grid2d(nx, ny, _p(rp), _p(col)))
    return {"rowptr": rp, "col": col[:m], "ew": np.ones(m, dtype=np.int32), "vw": np.ones(n, dtype=np.int32),
            "uniform": True}


def stencil3d(nx, ny, nz, seed=0, wmax=10):
    n = nx * ny * nz
    m = 2 * (nx * ny * (nz - 1) + nx * nz * (ny - 1) + ny * nz * (nx - 1))
    rp = np.empty(n + 1, dtype=np.int64)
    col = np.empty(max(m, 1), dtype=np.int32)
    ew = np.empty(max(m, 1), dtype=np.int32)
    _check(lib().jet_generate_stencil3d(nx, ny, nz, seed, wmax, _p(rp), _p(col), _p(ew)))
    return {"rowptr": rp, "col": col[:m], "ew": ew[:m], "vw": np.ones(n, dtype=np.int32), "uniform": wmax <= 1}


def rmat(scale, ef, seed=0, a=0.57, b=0.19, c=0.19):
    n = 1 << scale
    rp = np.empty(n + 1, dtype=np.int64)
    col = np.empty(2 * ef * n, dtype=np.int32)
    m = ctypes.c_int64(0)
    _check(lib().jet_generate_rmat(scale, ef, a, b, c, seed, _p(rp), _p(col), ctypes.byref(m)))
    col = col[:m.value].copy()
    return {"rowptr": rp, "col": col, "ew": np.ones(m.value, dtype=np.int32), "vw": np.ones(n, dtype=np.int32),
            "uniform": True}


# ---------------------------------------------------------------- batch / serve
class Context:
    """jet_ctx_t: `workers` GPU workers (own HIP stream and device memory each)
    admitting queued jobs under a device-memory budget (bytes; 0 = 80% of
    free memory). Use as a context manager or call close()."""

    def __init__(self, workers=4, budget=0, device=0):
        h = ctypes.c_void_p()
        _check(lib().jet_ctx_create(device, workers, int(budget), ctypes.byref(h)))
        self._h = h
        self._keep = {}

    def close(self):
        if self._h:
            lib().jet_ctx_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, g, cfg, seed=0, uniform=None, part_out=None):
        """Queues a job on host arrays; returns its ticket. The result lands in
        part_out (allocated when None) and is returned by poll()."""
        rp, col, ew, vw = _graph(g)
        n = len(rp) - 1
        if uniform is None:
            uniform = bool(g.get("uniform", False))
        part = np.empty(n, dtype=np.int32) if part_out is None else part_out
        t = ctypes.c_int64(0)
        c = Config()
        ctypes.pointer(c)[0] = cfg
        _check(lib().jet_submit(self._h, ctypes.byref(c), n, _p(rp), _p(col), _p(ew), _p(vw), int(uniform), PTR_HOST,
                                seed, _p(part), ctypes.byref(t)))
        self._keep[t.value] = (rp, col, ew, vw, part, c)
        return t.value

    def submit_device(self, n, rowptr_ptr, col_ptr, ew_ptr, vw_ptr, part_ptr, cfg, seed=0, uniform=False):
        """Queues a job on device pointers (ints); returns its ticket."""
        t = ctypes.c_int64(0)
        _check(lib().jet_submit(self._h, ctypes.byref(cfg), n, rowptr_ptr, col_ptr, ew_ptr, vw_ptr, int(uniform),
                                PTR_DEVICE, seed, part_ptr, ctypes.byref(t)))
        self._keep[t.value] = (cfg,)
        return t.value

    def poll(self, ticket, wait_ms=-1):
        """None while the job runs (wait_ms >= 0 and not finished), else
        (part or None, cut, Stats, JobTimes); raises JetError on failure."""
        cut = ctypes.c_int64(0)
        st = Stats()
        tm = JobTimes()
        rc = lib().jet_poll(self._h, ticket, int(wait_ms), ctypes.byref(cut), ctypes.byref(st), ctypes.byref(tm))
        if rc == JET_EAGAIN:
            return None
        kept = self._keep.pop(ticket, None)
        _check(rc)
        part = kept[4] if kept is not None and len(kept) > 4 else None
        return part, cut.value, st, tm

    def batch(self, graphs, cfgs, seeds=None, stats=False):
        """jet_partition_batch over a list of host CSR dicts; returns
        (list of parts, cuts[, list of Stats])."""
        count = len(graphs)
        ns = [len(g["rowptr"]) - 1 for g in graphs]
        ms = [len(g["col"]) for g in graphs]
        voff = np.zeros(count + 1, dtype=np.int64)
        voff[1:] = np.cumsum(ns)
        eoff = np.zeros(count + 1, dtype=np.int64)
        eoff[1:] = np.cumsum(ms)
        rp = np.concatenate([np.asarray(g["rowptr"], dtype=np.int64) for g in graphs])
        col = np.concatenate([np.asarray(g["col"], dtype=np.int32) for g in graphs])
        ew = np.concatenate([np.asarray(g["ew"], dtype=np.int32) if g.get("ew") is not None
                             else np.ones(len(g["col"]), dtype=np.int32) for g in graphs])
        vw = np.concatenate([np.asarray(g["vw"], dtype=np.int32) if g.get("vw") is not None
                             else np.ones(len(g["rowptr"]) - 1, dtype=np.int32) for g in graphs])
        uni = np.array([int(bool(g.get("uniform", False))) for g in graphs], dtype=np.int32)
        carr = (Config * count)(*cfgs)
        sd = np.array(seeds if seeds is not None else [0] * count, dtype=np.uint64)
        part = np.empty(int(voff[-1]), dtype=np.int32)
        cuts = np.empty(count, dtype=np.int64)
        st = (Stats * count)()
        _check(lib().jet_partition_batch(self._h, count, carr, _p(sd), _p(voff), _p(eoff), _p(rp), _p(col), _p(ew),
                                         _p(vw), _p(uni), PTR_HOST, _p(part), _p(cuts), st))
        parts = [part[voff[i]:voff[i + 1]] for i in range(count)]
        return (parts, cuts, list(st)) if stats else (parts, cuts)


# ---------------------------------------------------------------- files, quality
def load_metis(path):
    """app/io.hpp METIS graph file -> CSR dict {rowptr, col, ew, uniform}."""
    hg = HostGraph()
    _check(lib().jet_load_metis(os.fsencode(path), ctypes.byref(hg)))
    try:
        n, m = hg.n, hg.m
        g = {"rowptr": np.ctypeslib.as_array(hg.row_offsets, shape=(n + 1,)).copy(),
             "col": np.ctypeslib.as_array(hg.col, shape=(max(m, 1),))[:m].copy(),
             "uniform": bool(hg.uniform_ew)}
        g["ew"] = (np.ctypeslib.as_array(hg.ew, shape=(max(m, 1),))[:m].copy() if hg.ew
                   else np.ones(m, dtype=np.int32))
        return g
    finally:
        lib().jet_free_graph(ctypes.byref(hg))


def load_config(path):
    """app/io.hpp config file -> Config."""
    cfg = Config()
    _check(lib().jet_load_config(os.fsencode(path), ctypes.byref(cfg)))
    return cfg


def write_partition(path, part):
    p = _c(part, np.int32)
    _check(lib().jet_write_partition(os.fsencode(path), len(p), _p(p)))


def read_partition(path, n):
    out = np.empty(n, dtype=np.int32)
    _check(lib().jet_read_partition(os.fsencode(path), n, _p(out)))
    return out


def partition_quality(g, part, k):
    """jet_partition_quality on host arrays -> Quality."""
    rp, col, ew, vw = _graph(g)
    q = Quality()
    part = _c(part, np.int32)
    _check(lib().jet_partition_quality(len(rp) - 1, _p(rp), _p(col), _p(ew), _p(vw), _p(part), k,
                                       PTR_HOST, ctypes.byref(q)))
    return q


def cut_heatmap(g, part, k):
    """jet_cut_heatmap on host arrays -> int64 [k, k]."""
    rp, col, ew, _ = _graph(g)
    part = _c(part, np.int32)
    out = np.empty(k * k, dtype=np.int64)
    _check(lib().jet_cut_heatmap(len(rp) - 1, _p(rp), _p(col), _p(ew), _p(part), k, PTR_HOST, _p(out)))
    return out.reshape(k, k)


# ---------------------------------------------------------------- dump / replay
DUMP_32, DUMP_BIG, DUMP_BIGGEST = 0, 1, 2


class Hierarchy:
    """A coarse hierarchy file (dump_coarse's coarse_graphs.out) loaded into
    host memory by jet_load_hierarchy; `kind` gives the widths it was written
    with (DUMP_32, DUMP_BIG, DUMP_BIGGEST)."""

    def __init__(self, path, kind=DUMP_32):
        self._h = ctypes.c_void_p()
        _check(lib().jet_load_hierarchy(os.fsencode(path), kind, ctypes.byref(self._h)))

    @property
    def levels(self):
        return lib().jet_hierarchy_levels(self._h)

    def vertices(self, level):
        return lib().jet_hierarchy_vertices(self._h, level)

    def edges(self, level):
        return lib().jet_hierarchy_edges(self._h, level)

    def close(self):
        if self._h:
            lib().jet_free_hierarchy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def uncoarsen(self, cfg, coarse_part, stats=False):
        """jet_uncoarsen_hierarchy: refine from the coarsest level's partition
        to the finest; returns (part, cut) or (part, cut, Stats)."""
        cp = _c(coarse_part, np.int32)
        part = np.empty(self.vertices(0), dtype=np.int32)
        cut = ctypes.c_int64()
        st = Stats()
        _check(lib().jet_uncoarsen_hierarchy(ctypes.byref(cfg), self._h, _p(cp), _p(part), ctypes.byref(cut),
                                             ctypes.byref(st)))
        return (part, cut.value, st) if stats else (part, cut.value)


def load_coarse_part(path, n):
    out = np.empty(n, dtype=np.int32)
    _check(lib().jet_load_coarse_part(os.fsencode(path), n, _p(out)))
    return out
