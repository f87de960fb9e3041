"""Literal Python transliterations of the reference's order-dependent kernels,
executed under the schedule the oracle fixes (tests only, small inputs).

These execute the reference lambdas' CAS/load/store steps one by one, as
opposed to the oracle's derived closed forms, so they pin those derivations.
"""
import oracle

MAX = 2**31 - 1


def map_construct(hn):
    """heuristics.hpp:89-169 under: all first CASes, then second steps in
    decreasing u (swap=1) / increasing u (swap=0)."""
    n = len(hn)
    vcmap = [MAX] * n
    perm = list(range(n))
    swap = 1
    while perm:
        acting = [u for u in perm if ((1 if u < hn[u] else 0) ^ swap)]
        ok = {}
        for u in acting:
            if vcmap[u] == MAX:
                vcmap[u] = MAX - 1
                ok[u] = True
            else:
                ok[u] = False
        for u in sorted(acting, reverse=bool(swap)):
            if not ok[u]:
                continue
            v = hn[u]
            won = u == v
            if not won and vcmap[v] == MAX:
                vcmap[v] = MAX - 1
                won = True
            if won:
                cv = min(u, v)
                vcmap[u] = cv
                vcmap[v] = cv
            else:
                vcmap[u] = vcmap[v] if vcmap[v] < n else MAX
        perm = [u for u in perm if vcmap[u] >= n]
        swap ^= 1
    return dense(vcmap)


def dense(vcmap):
    n = len(vcmap)
    ids, nc = {}, 0
    for u in range(n):
        if vcmap[u] == u:
            ids[u] = nc
            nc += 1
    return [ids[vcmap[u]] for u in range(n)], nc


def pick(g, hn, vcmap, u, initial, uniform, r):
    rp, col, ew = g["rowptr"], g["col"], g["ew"]
    if not initial and (vcmap[u] != MAX or hn[u] == MAX or vcmap[hn[u]] == MAX):
        return
    h, max_ewt, tb = MAX, 0, 0
    for j in range(rp[u], rp[u + 1]):
        v = int(col[j])
        if initial or vcmap[v] == MAX:
            w = int(ew[j])
            if not uniform and max_ewt < w:
                max_ewt, h, tb = w, v, oracle.xorshift32((v + r) & 0xFFFFFFFF)
            elif uniform or max_ewt == w:
                s = oracle.xorshift32((v + r) & 0xFFFFFFFF)
                if s >= tb:
                    h, tb = v, s
    hn[u] = h


def match_hash(cands, hashes, nullkey, vcmap):
    """heuristics.hpp:283-318 under: non-null digests first, then null, each in candidate order."""
    m = len(cands)
    htable = [nullkey] * m
    twins = [-1] * m
    order = [x for x in range(m) if hashes[x] != nullkey] + [x for x in range(m) if hashes[x] == nullkey]
    for x in order:
        i, h = cands[x], hashes[x]
        key = h % m
        while True:
            if htable[key] == nullkey:
                htable[key] = h
            if htable[key] == h:
                break
            key = (key + 1) % m
        twin = twins[key]
        if twin == -1:
            twins[key] = i
        else:
            twins[key] = -1
            cv = min(twin, i)
            vcmap[twin] = cv
            vcmap[i] = cv


def coarsen_match(g, uniform, seed, level, two_hop):
    """heuristics.hpp:427-671; commit part 2 runs in increasing u."""
    rp, col, ew = g["rowptr"], g["col"], g["ew"]
    n = len(rp) - 1
    hn = [MAX] * n
    vcmap = [MAX] * n
    if uniform:
        for i in range(n):
            d = int(rp[i + 1] - rp[i])
            if d:
                hn[i] = int(col[rp[i] + oracle.rng(seed, level, 2, 0, i) % d])
    else:
        for u in range(n):
            pick(g, hn, vcmap, u, True, False, oracle.rng(seed, level, 3, 0, u) & 0xFFFFFFFF)
    wl = list(range(n))
    pas = 0
    while wl:
        for r in range(4):
            for u in wl:
                v = hn[u]
                if v == MAX or vcmap[u] != MAX:
                    continue
                cond = (oracle.pod_hash_i32(u + r) < oracle.pod_hash_i32(v + r)) if r > 0 else (u < v)
                if not cond:
                    vcmap[u] = MAX - 1
            for u in wl:
                v = hn[u]
                if v == MAX or vcmap[u] != MAX:
                    continue
                cv = min(u, v)
                if vcmap[v] == MAX - 1:
                    vcmap[v] = cv
                    vcmap[u] = cv
            for u in wl:
                if vcmap[u] == MAX - 1:
                    vcmap[u] = MAX
        pas += 1
        for u in wl:
            pick(g, hn, vcmap, u, False, uniform, oracle.rng(seed, level, 3, pas, u) & 0xFFFFFFFF)
        wl = [u for u in wl if vcmap[u] == MAX and hn[u] != MAX]
    if two_hop:
        def ratio():
            return sum(1 for x in vcmap if x == MAX) / n
        if ratio() > 0.25:
            c = [i for i in range(n) if vcmap[i] == MAX and rp[i + 1] - rp[i] == 1]
            match_hash(c, [int(col[rp[i]]) for i in c], MAX, vcmap)
        if ratio() > 0.25:
            c = [i for i in range(n) if vcmap[i] == MAX]
            hs = []
            for i in c:
                h = 0
                for j in range(rp[i], rp[i + 1]):
                    y = oracle.pod_hash_i32(int(col[j]))
                    h = (h + y * y + y) & (2**64 - 1)
                hs.append(h)
            match_hash(c, hs, 0, vcmap)
        if ratio() > 0.25:
            c = [i for i in range(n) if vcmap[i] == MAX]
            hs = []
            for u in c:
                h, max_wgt, min_deg = MAX, 0, MAX
                for j in range(rp[u], rp[u + 1]):
                    v = int(col[j])
                    vd = int(rp[v + 1] - rp[v])
                    if min_deg > vd:
                        min_deg, max_wgt, h = vd, int(ew[j]), v
                    elif min_deg == vd and max_wgt < int(ew[j]):
                        h, max_wgt = v, int(ew[j])
                hs.append(h)
            match_hash(c, hs, MAX, vcmap)
    for i in range(n):
        if vcmap[i] == MAX:
            vcmap[i] = i
    return dense(vcmap)
