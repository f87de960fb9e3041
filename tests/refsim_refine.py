"""Python restatement of one level of Jet refinement (src/jet_refiner.hpp),
written from the reference for the tests (small inputs only).

It is independent of the oracle's C++ in two ways: it follows the reference's
kernels in their own order and with their own scratch arrays (minibucket
arrays, the two cookie-cutter scans, the `max_vwgt` re-draw counter), and it
models each connectivity table by its CONTENT instead of its slots. Row i is a
dict part -> connectivity plus the table size; the slot mechanics of the
reference fix exactly which parts are present:

* a build (init_conn_data :958-1048, build_row_cdata_large :648-739) inserts
  every adjacent part, then shrinks the size to used + max(used/4, 3) when that
  is below min(deg, k);
* update_small's subtraction (:785-806) frees a slot (HASH_RECLAIM) exactly
  when the connectivity reaches zero in a table of size < k; in a table of size
  k the part stays with connectivity 0;
* update_small's addition (:810-884) inserts an absent part into a free slot,
  and appends past the table (growing its size by one) only when all `size`
  slots hold parts.

Serial-backend order is used wherever the reference races: atomic positions in
minibuckets follow increasing vertex id, and the strong re-draws (:538-541) run
in move-list order. Ties between parts of equal connectivity go to the lowest
part id (the build's documented rule); the reference keeps the first maximum
in slot order, which a content model cannot express.

`refine_level` returns the best partition, the level record and a trace of
every iteration: (kind, moves, cut, total_imb, take, count).
"""
import math

import numpy as np

NULL_PART = -1
NO_MOVE = -3
MAX_BUCKETS = 50
MID_BUCKET = 25
MAX_SECTIONS = 128
GAIN_MIN = -(2**63)


def gain_bucket(gx, vwgt):
    """jet_refiner.hpp:326-356 in float32 arithmetic."""
    gain = np.float32(gx) / np.float32(vwgt)
    if gain > 0:
        return 0
    if gain == 0:
        return 1
    t = MID_BUCKET
    gain = np.float32(abs(gain))
    if gain < 1:
        while gain < 1:
            gain = np.float32(gain * np.float32(1.5))
            t -= 1
        if t < 2:
            t = 2
    else:
        while gain > 1:
            gain = np.float32(gain / np.float32(1.5))
            t += 1
        if t > MAX_BUCKETS:
            t = MAX_BUCKETS - 1
    return t


def sections_for(n, k):
    """jet_refiner.hpp:423-428 (and :556-561)."""
    sections = MAX_SECTIONS
    size = (n + sections * k) // (sections * k)
    if size < 4096:
        size = 4096
        sections = (n + size * k) // (size * k)
    return sections


class Level:
    def __init__(self, g, k, imb):
        self.rp = [int(x) for x in g["rowptr"]]
        self.col = [int(x) for x in g["col"]]
        self.ew = [int(x) for x in g["ew"]]
        self.n = len(self.rp) - 1
        self.vw = [int(x) for x in g["vw"]] if g.get("vw") is not None else [1] * self.n
        self.k = k
        self.imb = imb

    def nbrs(self, i):
        for j in range(self.rp[i], self.rp[i + 1]):
            yield self.col[j], self.ew[j]


class State:
    """conn_data (:116-123) plus the scratch that persists across iterations."""

    def __init__(self, L, part):
        self.L = L
        n, k = L.n, L.k
        self.tab = [None] * n
        self.size = [0] * n
        for i in range(n):
            self.build_row(part, i)
        self.dest_cache = [NULL_PART] * n
        self.lock = [0] * n
        self.gain_persistent = [0] * n
        self.dest = [0] * n

    def build_row(self, part, i):
        L = self.L
        d = {}
        for v, w in L.nbrs(i):
            d[part[v]] = d.get(part[v], 0) + w
        cap = min(L.rp[i + 1] - L.rp[i], L.k)
        used = len(d)
        size = used + max(used // 4, 3)
        self.tab[i] = d
        self.size[i] = size if size < cap else cap

    def lookup(self, i, p):  # :891-901
        return self.tab[i].get(p, 0)


def jet_lp(S, part, fr):
    """:200-323."""
    L = S.L
    n = L.n
    for i in range(n):
        best = S.dest_cache[i]
        if best != NULL_PART:
            S.dest[i] = best
            continue
        p = part[i]
        others = {q: c for q, c in S.tab[i].items() if q != p and c > 0}
        best = NO_MOVE
        b_conn = 0
        if others:
            b_conn = max(others.values())
            best = min(q for q, c in others.items() if c == b_conn)
        p_conn = S.tab[i].get(p, 0)
        S.gain_persistent[i] = 0
        if best != NO_MOVE:
            if b_conn >= p_conn or (p_conn - b_conn) < math.floor(fr * p_conn):
                S.gain_persistent[i] = b_conn - p_conn
            else:
                best = NO_MOVE
        S.dest_cache[i] = best
        S.dest[i] = best
    pregain = [0] * n
    cand = []
    for i in range(n):
        if S.dest[i] != NO_MOVE and S.lock[i] == 0:
            cand.append(i)
            pregain[i] = S.gain_persistent[i]
        else:
            pregain[i] = GAIN_MIN
            S.lock[i] = 0
    for i in cand:
        best, p, ig = S.dest[i], part[i], pregain[i]
        change = 0
        for v, w in L.nbrs(i):
            vg = pregain[v]
            if vg > ig or (vg == ig and v < i):
                vp = S.dest[v]
                if vp == p:
                    change -= w
                elif vp == best:
                    change += w
                vp = part[v]
                if vp == p:
                    change += w
                elif vp == best:
                    change -= w
        if ig + change >= 0:
            S.lock[i] = 1
    return [i for i in cand if S.lock[i]]


def get_evictions(S, part, sizes, bid, save_atomic, bucket_sizes, width, size_max, adjust):
    """:359-412. Returns (moves, bucket_offsets, evict_adjust)."""
    L = S.L
    offsets = [0] * len(bucket_sizes)
    run = 0
    for b, x in enumerate(bucket_sizes):
        offsets[b] = run
        run += x
    evict_adjust = [0] * L.k
    moves = []
    for i in range(L.n):
        b = bid[i]
        if b == -1:
            continue
        p = part[i]
        score = save_atomic[i] + offsets[b] - offsets[p * width]
        limit = sizes[p] - size_max
        if score < limit:
            if adjust and score + L.vw[i] >= limit:
                evict_adjust[p] = score + L.vw[i]
            moves.append(i)
    return moves, offsets, evict_adjust


def rebalance_weak(S, part, sizes, opt, size_max):
    """:551-645 with get_evictions<false>."""
    L = S.L
    n, k = L.n, L.k
    sections = sections_for(n, k)
    t_mb = MAX_BUCKETS * k * sections
    # a gain bucket of exactly 50 is not capped (:349-351); for the last part
    # it indexes up to `sections` minibuckets past t_mb, which are kept here
    # (zero-initialised and scanned with the rest)
    bucket_sizes = [0] * (t_mb + sections + 1)
    max_dest = int(size_max * 0.99)
    if max_dest < size_max - 100:
        max_dest = size_max - 100
    undersized = [p for p in range(k) if sizes[p] < max_dest]
    save_gains = [0] * n
    for i in range(n):
        p = part[i]
        if sizes[p] > size_max and L.vw[i] < 1.5 * (sizes[p] - opt):
            best, gain, p_gain = p, 0, 0
            recv = {q: c for q, c in S.tab[i].items() if sizes[q] < max_dest and c > 0}
            if recv:
                gain = max(recv.values())
                best = min(q for q, c in recv.items() if c == gain)
            p_gain = S.tab[i].get(p, 0)
            if gain > 0:
                S.dest[i] = best
                save_gains[i] = gain - p_gain
            elif undersized:
                S.dest[i] = undersized[i % len(undersized)]
                save_gains[i] = -p_gain
            else:
                S.dest[i] = p  # no undersized part: the reference divides by zero
        else:
            S.dest[i] = p
    bid = [-1] * n
    vscore = [0] * n
    for i in range(n):
        p = part[i]
        if p != S.dest[i]:
            g_id = (MAX_BUCKETS * p + gain_bucket(save_gains[i], L.vw[i])) * sections + (i % sections)
            bid[i] = g_id
            vscore[i] = bucket_sizes[g_id]  # atomic_fetch_add in increasing i
            bucket_sizes[g_id] += L.vw[i]
    moves, _, _ = get_evictions(S, part, sizes, bid, vscore, bucket_sizes, MAX_BUCKETS * sections, size_max, False)
    return moves


def rebalance_strong(S, part, sizes, opt, size_max):
    """:417-546 with get_evictions<true>."""
    L = S.L
    n, k = L.n, L.k
    sections = sections_for(n, k)
    t_mb = MAX_BUCKETS * k * sections
    bucket_sizes = [0] * (t_mb + 1)
    max_dest = max(opt + 1, int(size_max * 0.99))
    caps = [max_dest - sizes[p] for p in range(k) if sizes[p] < max_dest]
    max_vwgt = max(caps) if caps else None  # Kokkos::Max of nothing: the lowest value
    bid = [-1] * n
    save_atomic = [0] * n
    for i in range(n):
        p = part[i]
        vw_ok = max_vwgt is not None and L.vw[i] <= 2 * max_vwgt
        if sizes[p] > size_max and vw_ok and L.vw[i] < 2 * (sizes[p] - opt):
            tk = tg = 0
            p_gain = 0
            for q, c in S.tab[i].items():
                if q == p:
                    p_gain = c
                elif sizes[q] < max_dest:
                    tg += c
                    tk += 1
            if tk == 0:
                tk = 1
            gain = tg // tk - p_gain
            gt = gain_bucket(gain, min(L.vw[i], sizes[p] - size_max))
            if gt < MAX_BUCKETS:
                g_id = (MAX_BUCKETS * p + gt) * sections + (i % sections) + 1
                bid[i] = g_id
                save_atomic[i] = bucket_sizes[g_id]
                bucket_sizes[g_id] += L.vw[i]
    width = MAX_BUCKETS * sections
    moves, offsets, evict_adjust = get_evictions(S, part, sizes, bid, save_atomic, bucket_sizes, width, size_max,
                                                 True)
    # cookie cutter (:491-518): two team scans over the parts
    update = 0
    for p in range(k):
        add = evict_adjust[p]
        begin = MAX_BUCKETS * p * sections
        if add == 0:
            add = offsets[begin + width] - offsets[begin]
        evict_adjust[p] = offsets[begin] - update
        update += add
    counter = update  # max_vwgt() = the total after the last part
    evict_start = [0] * (k + 1)
    update = 0
    for p in range(k):
        if max_dest > sizes[p]:
            update += max_dest - sizes[p]
        evict_start[p + 1] = update
    for v in moves:  # adjust scores (:519-525)
        save_atomic[v] = save_atomic[v] + offsets[bid[v]] - evict_adjust[part[v]]
    for v in moves:  # select destination parts (:526-544), one vertex after the other
        p = 0
        placed = False
        while p < k:
            while p <= k and evict_start[p] <= save_atomic[v]:
                p += 1
            p -= 1
            if p < k and L.vw[v] // 2 <= evict_start[p + 1] - save_atomic[v]:
                S.dest[v] = p
                placed = True
                break
            elif p < k:
                save_atomic[v] = counter
                counter += L.vw[v]
        if not placed:
            S.dest[v] = part[v]
    return moves


def update_small(S, part, moves):
    """:781-888, subtractions for every move, then additions (Serial order)."""
    L = S.L
    for i in moves:
        p = S.dest[i]  # the old part
        for v, w in L.nbrs(i):
            x = S.tab[v][p]
            S.tab[v][p] = x - w
            if S.size[v] < L.k and x == w:
                del S.tab[v][p]  # HASH_RECLAIM
    for i in moves:
        best = part[i]
        for v, w in L.nbrs(i):
            S.dest_cache[v] = NULL_PART
            if best not in S.tab[v]:
                if len(S.tab[v]) == S.size[v]:
                    S.size[v] += 1  # no free slot: dense append past the table
                S.tab[v][best] = 0
            S.tab[v][best] += w


def update_large(S, part, moves):
    """:743-777."""
    L = S.L
    mark = set()
    for i in moves:
        for v, _ in L.nbrs(i):
            mark.add(v)
    for v in sorted(mark):
        S.build_row(part, v)
        S.dest_cache[v] = NULL_PART


def perform_moves(S, part, moves, sizes, opt):
    """:905-955. Returns (cut change, new total_imb)."""
    L = S.L
    cc1 = sum(S.lookup(i, S.dest[i]) - S.lookup(i, part[i]) for i in moves)
    for i in moves:
        p, best = part[i], S.dest[i]
        S.dest_cache[i] = NULL_PART
        sizes[p] -= L.vw[i]
        sizes[best] += L.vw[i]
        part[i] = best
        S.dest[i] = p
    if len(moves) > L.n // 10:
        update_large(S, part, moves)
    else:
        update_small(S, part, moves)
    cc2 = sum(S.lookup(i, part[i]) - S.lookup(i, S.dest[i]) for i in moves)
    mx = max(sizes)
    return cc1 + cc2, (mx - opt if mx > opt else 0)


def total_cut2(L, part):
    return sum(w for i in range(L.n) for v, w in L.nbrs(i) if part[v] != part[i])


def refine_level(g, part0, k, imb, tol=0.999, ultra=False, uniform=False):
    """jet_refine (:1050-1142) with fresh refine_data (the coarsest level)."""
    L = Level(g, k, imb)
    best_part = [int(x) for x in part0]
    best_cut = total_cut2(L, best_part)
    best_sizes = [0] * k
    for i in range(L.n):
        best_sizes[best_part[i]] += L.vw[i]
    total = sum(L.vw)
    opt = (total + k - 1) // k
    size_max = int(opt * imb)
    mx = max(best_sizes)
    best_imb = mx - opt if mx > opt else 0
    cut, cimb, sizes = best_cut, best_imb, list(best_sizes)
    imb_max = size_max - opt
    part = list(best_part)
    S = State(L, part)
    if ultra:
        temps = []
        t = 0.85
        while t > 0:
            temps.append(t)
            t -= 0.05
    elif uniform:
        temps = [0.25]
    else:
        temps = [0.75]
    iters = lp_iters = 0
    balance_counter = 0
    trace = []
    for fr in temps:
        count = 0
        while True:
            c0 = count
            count += 1
            if c0 > 11:
                break
            iters += 1
            if cimb <= imb_max:
                moves = jet_lp(S, part, fr)
                kind = 0
                balance_counter = 0
                lp_iters += 1
            else:
                if balance_counter < 2:
                    moves = rebalance_weak(S, part, sizes, opt, size_max)
                    kind = 1
                else:
                    moves = rebalance_strong(S, part, sizes, opt, size_max)
                    kind = 2
                balance_counter += 1
            cc, cimb = perform_moves(S, part, moves, sizes, opt)
            cut -= cc
            take = 0
            if best_imb > imb_max and cimb < best_imb:
                take = 1
                count = 0
            elif cut < best_cut and (cimb <= imb_max or cimb <= best_imb):
                if cut < tol * best_cut:
                    count = 0
                take = 2
            if take:
                best_cut, best_imb, best_sizes = cut, cimb, list(sizes)
                best_part = list(part)
            trace.append((kind, len(moves), cut, cimb, take, count))
    rec = [best_cut, best_imb, L.n, len(L.col), iters, lp_iters, opt, size_max]
    return np.array(best_part, dtype=np.int32), rec, trace
