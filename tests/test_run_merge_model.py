"""
The MSM accumulation's run-merging scheme, restated on the CPU (integers stand in for points).

The accumulation splits the sorted entries evenly over its threads; each thread writes the runs
strictly inside its range as finished buckets and leaves a summary of its first and last run
(one run if its range holds a single key). A boundary pass (k_fixup in csrc/zk_msm.hip) then adds
up every bucket cut into two fragments and writes it, and passes on only the fragments of longer
chains (hot buckets through whole ranges) as reduced summaries; trees of summaries (k_combine,
merge_runs) merge those level by level: each tree packs its non-empty summaries to the front first,
and below the last level writes the runs of its result that do not continue into the summary just
outside its range; the last summary's open runs are written at the end.

This checks the scheme's invariant on random layouts (uniform keys, hot keys, ranges shorter and
longer than buckets, empty ranges, trees of several widths, the boundary pass before the first
level, with and without a second one before the second level, as the library runs it): every
bucket is written exactly once, with the sum of its entries. The GPU tests check the kernels
themselves (tests/test_msm_gpu.py).
"""

import random

V, S = 1, 2  # summary flags: any entries; a single run covering the whole range


def empty():
    return dict(flags=0, fk=0, lk=0, fsum=0, lsum=0)


def thread_summaries(keys, vals, k, threads, buckets):
    out = []
    for t in range(threads):
        s = t * k
        if s >= len(keys):
            out.append(empty())
            continue
        runs = []
        for x in range(s, min(s + k, len(keys))):
            if runs and runs[-1][0] == keys[x]:
                runs[-1][1] += vals[x]
            else:
                runs.append([keys[x], vals[x]])
        for key, v in runs[1:-1]:
            write(buckets, key, v)
        if len(runs) == 1:
            out.append(dict(flags=V | S, fk=runs[0][0], lk=runs[0][0], fsum=runs[0][1], lsum=0))
        else:
            out.append(dict(flags=V, fk=runs[0][0], lk=runs[-1][0], fsum=runs[0][1], lsum=runs[-1][1]))
    return out


def write(buckets, key, v):
    assert key not in buckets, ("bucket written twice", key)
    buckets[key] = v


def boundary_pass(summ, buckets):
    """k_fixup: one thread per summary t, deciding from the keys of t - 2 .. t + 2."""
    n = len(summ)
    red = []
    for t in range(n):
        fk, ke, valid, single = [], [], [], []
        for k in range(5):
            i = t - 2 + k
            m = summ[i] if 0 <= i < n else empty()
            fk.append(m["fk"])
            valid.append(bool(m["flags"] & V))
            single.append(bool(m["flags"] & S))
            ke.append(m["fk"] if m["flags"] & S else m["lk"])
        jl = [k > 0 and valid[k - 1] and valid[k] and ke[k - 1] == fk[k] for k in range(5)]
        jr = [k < 4 and valid[k] and valid[k + 1] and ke[k] == fk[k + 1] for k in range(5)]
        r = empty()
        if not valid[2]:
            red.append(r)
            continue
        me = summ[t]
        add = False
        if single[2]:
            lo, hi = jl[2], jr[2]
            if not lo and not hi:
                write(buckets, fk[2], me["fsum"])
            elif (lo and hi) or (lo and single[1] and jl[1]) or (hi and single[3] and jr[3]):
                r = dict(flags=V | S, fk=fk[2], lk=fk[2], fsum=me["fsum"], lsum=0)
            else:
                add = lo
        else:
            first = last = False
            if not jl[2]:
                write(buckets, fk[2], me["fsum"])
            elif single[1] and jl[1]:
                first = True
            else:
                add = True
            if not jr[2]:
                write(buckets, me["lk"], me["lsum"])
            elif single[3] and jr[3]:
                last = True
            if first and last:
                r = dict(me)
            elif first or last:
                key = me["fk"] if first else me["lk"]
                r = dict(flags=V | S, fk=key, lk=key, fsum=me["fsum"] if first else me["lsum"], lsum=0)
        if add:
            left = summ[t - 1]["fsum"] if single[1] else summ[t - 1]["lsum"]
            write(buckets, fk[2], left + me["fsum"])
        red.append(r)
    return red


def merge(l, r, buckets):
    """merge_runs: at most one addition, closed runs written."""
    if not r["flags"] & V:
        return l
    if not l["flags"] & V:
        return dict(r)
    ls, rs = bool(l["flags"] & S), bool(r["flags"] & S)
    lk = l["fk"] if ls else l["lk"]
    l = dict(l)
    if lk == r["fk"]:
        joined = (l["fsum"] if ls else l["lsum"]) + r["fsum"]
        if ls and rs:
            l["fsum"] = joined
        elif ls:
            l.update(fsum=joined, lk=r["lk"], lsum=r["lsum"], flags=V)
        elif rs:
            l["lsum"] = joined
        else:
            write(buckets, lk, joined)
            l.update(lk=r["lk"], lsum=r["lsum"])
        return l
    if not ls:
        write(buckets, l["lk"], l["lsum"])
    if not rs:
        write(buckets, r["fk"], r["fsum"])
    l.update(lk=r["fk"] if rs else r["lk"], lsum=r["fsum"] if rs else r["lsum"], flags=V)
    return l


def close_complete(res, summ, b, width, buckets):
    """A tree's result below the last level: the runs that do not continue into the summary right
    outside the tree's range are written, the rest goes on."""
    if not res["flags"] & V:
        return res
    single = bool(res["flags"] & S)
    last = res["fk"] if single else res["lk"]
    cl = cr = False
    if b > 0:
        m = summ[b - 1]
        cl = bool(m["flags"] & V) and (m["fk"] if m["flags"] & S else m["lk"]) == res["fk"]
    if b + width < len(summ):
        m = summ[b + width]
        cr = bool(m["flags"] & V) and m["fk"] == last
    res = dict(res)
    if single:
        if not cl and not cr:
            write(buckets, res["fk"], res["fsum"])
            res = empty()
        return res
    if not cl:
        write(buckets, res["fk"], res["fsum"])
    if not cr:
        write(buckets, res["lk"], res["lsum"])
    if not cl and not cr:
        return empty()
    if not cl:
        return dict(flags=V | S, fk=res["lk"], lk=res["lk"], fsum=res["lsum"], lsum=0)
    if not cr:
        return dict(flags=V | S, fk=res["fk"], lk=res["fk"], fsum=res["fsum"], lsum=0)
    return res


def tree_level(summ, buckets, width, last_level):
    out = []
    for b in range(0, len(summ), width):
        sh = [x for x in summ[b : b + width] if x["flags"] & V]  # packed, in order
        w = 1
        while w < len(sh):
            w <<= 1
        sh = sh + [empty()] * (w - len(sh))
        d = 1
        while d < w:
            for i in range(w // (2 * d)):
                sh[2 * d * i] = merge(sh[2 * d * i], sh[2 * d * i + d], buckets)
            d <<= 1
        out.append(sh[0] if last_level else close_complete(sh[0], summ, b, width, buckets))
    return out


def run_case(seed):
    rng = random.Random(seed)
    nkeys = rng.choice([3, 10, 100, 1000])
    m = rng.randint(1, 4000)
    mode = rng.random()
    if mode < 0.3:
        keys = sorted(rng.randrange(nkeys) for _ in range(m))
    elif mode < 0.6:
        hot = rng.randrange(nkeys)
        keys = sorted(hot if rng.random() < 0.5 else rng.randrange(nkeys) for _ in range(m))
    else:
        keys = sorted(rng.choice([rng.randrange(nkeys), rng.randrange(3)]) for _ in range(m))
    vals = [rng.randrange(1, 1 << 30) for _ in range(m)]
    threads = rng.randint(1, 600)
    width = rng.choice([2, 4, 8, 256])
    buckets = {}
    summ = thread_summaries(keys, vals, (m + threads - 1) // threads, threads, buckets)
    second_pass = rng.random() < 0.5  # the library runs it when the threads outnumber the buckets
    level = 1
    while True:
        if level == 1 or (level == 2 and second_pass):
            summ = boundary_pass(summ, buckets)
        summ = tree_level(summ, buckets, width, len(summ) <= width)
        level += 1
        if len(summ) == 1:
            break
    last = summ[0]
    if last["flags"] & V:
        write(buckets, last["fk"], last["fsum"])
        if not last["flags"] & S:
            write(buckets, last["lk"], last["lsum"])
    want = {}
    for key, v in zip(keys, vals):
        want[key] = want.get(key, 0) + v
    assert buckets == want, seed


def test_every_bucket_written_once_with_its_sum():
    for seed in range(700):
        run_case(seed)
