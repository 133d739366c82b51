"""CPU model of the binned insert (goetia_amd/csrc/gt_kernels.hip, bin_* kernels).

Restates, step by step and with the kernels' own index arithmetic, what one
binned call does: the plan (buckets of 8192 bins, first / second digits), the
first pass (per 2048-start tile: events table-major, stable per-wave ranking
by first digit, region writes and the 257-entry digit-offset row), the
second-pass count and scan, the second-pass scatter (work items of 1024 tiles,
chunks of 4096 events, per-digit runs written from the chunk's sorted order),
the per-bucket replay with ranks inside 64-event chunks, and the per-read sum
of new flags or the insert_and_query minimum.  Every buffer index is asserted
against the buffer's size, so a plan or offset mistake fails here on the CPU.

Test infrastructure: the hashes come from the oracle (oracle/pyoracle.py).
"""
import numpy as np

from oracle import pyoracle as po

BIN_BITS, BUCKET_BINS = 13, 8192
DIGIT_BITS1, DIGITS1, DIGITS2 = 8, 256, 512
BIN_TILE, BIN_GROUP, BIN_CHUNK, TILE_MAX = 2048, 1024, 4096, 8192
MAX_TABLES, MAX_BUCKETS = 4, 1 << 17
NO_BIN = 0xFFFFFFFF


def _rank_chunk(keys):
    """Per lane of a 64-lane chunk: (lanes on the same key, earlier lanes on it)."""
    cnt, rank, seen = [], [], {}
    tot = {}
    for k in keys:
        tot[k] = tot.get(k, 0) + 1
    for k in keys:
        rank.append(seen.get(k, 0))
        seen[k] = seen.get(k, 0) + 1
        cnt.append(tot[k])
    return cnt, rank


def _windows(seq, off, K, mode):
    """Packed canonical hashes and validity of every window (oracle hashing)."""
    n_reads = len(off) - 1
    wins = np.maximum(np.diff(off) - K + 1, 0)
    kmer_off = np.concatenate([[0], np.cumsum(wins)]).astype(np.int64)
    nk = int(kmer_off[-1])
    hv = np.zeros(nk, np.uint64)
    valid = np.zeros(nk, bool)
    acgt = np.frombuffer(b'ACGT', np.uint8)
    for r in range(n_reads):
        s = bytes(seq[off[r]:off[r + 1]])
        if len(s) < K:
            continue
        fw, rc = po.hash_read(s, K)
        n = len(fw)
        ok = np.ones(n, bool)
        if mode == po.MODE_SKIP_N:
            bad = ~np.isin(np.frombuffer(s, np.uint8), acgt)
            c = np.concatenate([[0], np.cumsum(bad)])
            ok = (c[K:K + n] - c[:n]) == 0
        hv[kmer_off[r]:kmer_off[r] + n] = np.minimum(fw, rc)
        valid[kmer_off[r]:kmer_off[r] + n] = ok
    return wins, kmer_off, hv, valid


def binned_insert(kind, sizes, seq, off, K, mode, sub, iq=False):
    """Tables (4-byte padded), per-read n_new, n_unique, n_occupied, and the
    insert_and_query counts when iq."""
    off = np.asarray(off, np.int64)
    n_reads, n_bases, nt = len(off) - 1, int(off[-1]), len(sizes)
    assert 1 <= nt <= MAX_TABLES
    wins, kmer_off, hv, valid = _windows(seq, off, K, mode)
    nk = int(kmer_off[-1])
    # plan (binned_insert_impl)
    bfirst, gbase, nb = [], [], 0
    for p in sizes:
        bfirst.append(nb)
        gbase.append(nb << BIN_BITS)
        nb += (p + BUCKET_BINS - 1) // BUCKET_BINS
    bfirst.append(nb)
    assert nb <= MAX_BUCKETS
    bits = 0
    while (1 << bits) < nb:
        bits += 1
    shift2 = bits - DIGIT_BITS1 if bits > DIGIT_BITS1 else 0
    n_d1 = (nb + (1 << shift2) - 1) >> shift2
    mask2 = (1 << shift2) - 1
    sub = max((sub + TILE_MAX - 1) // TILE_MAX * TILE_MAX, TILE_MAX)
    sub = min(sub, (n_bases + TILE_MAX - 1) // TILE_MAX * TILE_MAX)
    tiles_max = sub // BIN_TILE
    groups_max = (tiles_max + BIN_GROUP - 1) // BIN_GROUP
    n_ent = tiles_max * nt * BIN_TILE
    cnt_max = (n_d1 << shift2) * groups_max
    if kind == po.BYTE:
        tabs = [np.zeros((p + 3) // 4 * 4, np.uint8) for p in sizes]
    else:
        tabs = [np.zeros((p // 8 + 1 + 3) // 4 * 4, np.uint8) for p in sizes]
    flags = np.zeros(n_bases, np.uint8)
    vq = np.zeros((max(n_bases, 1), MAX_TABLES), np.uint8)
    occ = 0
    for b0 in range(0, n_bases, sub):
        b1 = min(n_bases, b0 + sub)
        tile0 = b0 // BIN_TILE
        tiles1 = (b1 - b0 + BIN_TILE - 1) // BIN_TILE
        n_groups = (tiles1 + BIN_GROUP - 1) // BIN_GROUP
        r = int(np.searchsorted(off, b0, side='right') - 1)
        kbase = int(kmer_off[r] + min(max(b0 - off[r], 0), wins[r]))
        ent1 = np.full(n_ent, -1, np.int64)
        off1 = np.zeros((DIGITS1 + 1) * tiles_max, np.int64)
        # first pass
        for lt1 in range(tiles1):
            t0 = (tile0 + lt1) * BIN_TILE
            g = np.full(MAX_TABLES * BIN_TILE, NO_BIN, np.int64)
            ordv = np.zeros(BIN_TILE, np.int64)
            for w in range(BIN_TILE):
                s = t0 + w
                if s >= b1:
                    break
                rr = int(np.searchsorted(off, s, side='right') - 1)
                if s + K > off[rr + 1]:
                    continue
                k = int(kmer_off[rr] + (s - off[rr]))
                if not valid[k]:
                    continue
                ordv[w] = k - kbase
                assert 0 <= ordv[w] < 2 ** 32
                for t in range(nt):
                    g[t * BIN_TILE + w] = gbase[t] + int(hv[k]) % sizes[t]
            dig = [DIGITS1 if x == NO_BIN else (int(x) >> BIN_BITS) >> shift2 for x in g[:nt * BIN_TILE]]
            per_wave = nt * BIN_TILE // 4
            wpos = np.zeros((4, DIGITS1 + 1), np.int64)
            for wave in range(4):
                for d in dig[wave * per_wave:(wave + 1) * per_wave]:
                    if d != DIGITS1:
                        wpos[wave][d] += 1
            tot = wpos[:, :DIGITS1].sum(0)
            start = np.concatenate([[0], np.cumsum(tot)])[:DIGITS1]
            n_ev = int(tot.sum())
            for d in range(DIGITS1):
                assert start[d] <= 0xFFFF
                off1[d * tiles1 + lt1] = start[d]
                p = start[d]
                for wave in range(4):
                    c = wpos[wave][d]
                    wpos[wave][d] = p
                    p += c
            off1[DIGITS1 * tiles1 + lt1] = n_ev
            perm = np.full(nt * BIN_TILE, -1, np.int64)
            for wave in range(4):
                for c in range(0, per_wave, 64):
                    e0 = wave * per_wave + c
                    keys = dig[e0:e0 + 64]
                    cnt, rank = _rank_chunk(keys)
                    for i, d in enumerate(keys):
                        if d != DIGITS1:
                            pos = wpos[wave][d] + rank[i]
                            assert 0 <= pos < n_ev
                            perm[pos] = e0 + i
                    for i, d in enumerate(keys):
                        if d != DIGITS1 and rank[i] == 0:
                            wpos[wave][d] += cnt[i]
            base = lt1 * nt * BIN_TILE
            for i in range(n_ev):
                e = int(perm[i])
                assert base + i < n_ent
                ent1[base + i] = (int(g[e]) << 32) | int(ordv[e & (BIN_TILE - 1)])

        def runs(d1, grp):
            rs = []
            for j in range(BIN_GROUP):
                lt = grp * BIN_GROUP + j
                if lt < tiles1:
                    a, b = off1[d1 * tiles1 + lt], off1[(d1 + 1) * tiles1 + lt]
                    rs.append((lt * nt * BIN_TILE + a, b - a))
            return rs

        def item_events(d1, grp):
            out = []
            for st, ln in runs(d1, grp):
                for j in range(ln):
                    assert 0 <= st + j < n_ent and ent1[st + j] >= 0
                    out.append(int(ent1[st + j]))
            return out

        # second pass: count + scan
        cnt_len = (n_d1 << shift2) * n_groups
        assert cnt_len <= cnt_max
        cnt2 = np.zeros(cnt_len + 1, np.int64)
        for d1 in range(n_d1):
            for grp in range(n_groups):
                for e in item_events(d1, grp):
                    cnt2[(((d1 << shift2) | (((e >> 32) >> BIN_BITS) & mask2)) * n_groups) + grp] += 1
        cnt2 = np.concatenate([[0], np.cumsum(cnt2[:cnt_len])])
        # second pass: scatter, BIN_CHUNK events at a time
        ent2 = np.full(n_ent, -1, np.int64)
        for d1 in range(n_d1):
            for grp in range(n_groups):
                cur = [int(cnt2[((d1 << shift2) | d2) * n_groups + grp]) if d2 <= mask2 else 0
                       for d2 in range(DIGITS2)]
                ev_all = item_events(d1, grp)
                for c0 in range(0, len(ev_all), BIN_CHUNK):
                    ev = ev_all[c0:c0 + BIN_CHUNK]
                    m = len(ev)
                    d2s = [((e >> 32) >> BIN_BITS) & mask2 for e in ev]
                    tot = np.bincount(d2s, minlength=DIGITS2)
                    lstart = np.concatenate([[0], np.cumsum(tot)])
                    order = sorted(range(m), key=lambda j: (d2s[j], j))   # stable per digit
                    for i, j in enumerate(order):
                        d = d2s[j]
                        pos = cur[d] + (i - lstart[d])
                        assert cnt2[((d1 << shift2) | d) * n_groups + grp] <= pos < n_ent
                        ent2[pos] = ev[j]
                    for d in range(DIGITS2):
                        cur[d] += int(lstart[d + 1] - lstart[d])
        # replay
        for b in range(nb):
            lo_e, hi_e = cnt2[b * n_groups], cnt2[(b + 1) * n_groups]
            if lo_e == hi_e:
                continue
            t = max(i for i in range(nt) if b >= bfirst[i])
            bin0 = (b - bfirst[t]) << BIN_BITS
            nbins = min(BUCKET_BINS, sizes[t] - bin0)
            if kind == po.BYTE:
                nw = (nbins + 3) // 4
                assert nw <= BUCKET_BINS // 4 and bin0 + 4 * nw <= len(tabs[t])
                sl = tabs[t][bin0:bin0 + 4 * nw].astype(np.int64)
            else:
                nw = (nbins + 31) // 32
                assert nw <= BUCKET_BINS // 32 and bin0 // 8 + 4 * nw <= len(tabs[t])
                sl = np.unpackbits(tabs[t][bin0 // 8:bin0 // 8 + 4 * nw], bitorder='little').astype(np.int64)
            for c in range(lo_e, hi_e, 64):
                ev = [int(ent2[i]) for i in range(c, min(c + 64, hi_e))]
                keys = [(e >> 32) & (BUCKET_BINS - 1) for e in ev]
                cnt, rank = _rank_chunk(keys)
                pre = [int(sl[k]) for k in keys]
                for i, e in enumerate(ev):
                    assert 0 <= (e >> 32) - gbase[t] - bin0 < nbins
                    o = kbase + (e & 0xFFFFFFFF)
                    assert o < nk
                    if kind == po.BYTE:
                        if iq:
                            vq[o, t] = min(255, pre[i] + rank[i] + 1)
                        elif pre[i] + rank[i] == 0:
                            flags[o] = 1
                        if rank[i] == 0:
                            sl[keys[i]] = min(255, pre[i] + cnt[i])
                            occ += t == 0 and pre[i] == 0
                    elif not pre[i] and rank[i] == 0:
                        flags[o] = 1
                        sl[keys[i]] = 1
                        occ += t == 0
            if kind == po.BYTE:
                tabs[t][bin0:bin0 + 4 * nw] = sl.astype(np.uint8)
            else:
                tabs[t][bin0 // 8:bin0 // 8 + 4 * nw] = np.packbits(sl.astype(np.uint8), bitorder='little')
    if iq:
        w = vq[:nk, :nt]
        has = w.any(1)
        counts = np.where(has, w.min(1).astype(np.int16), -1).astype(np.int16)
        return tabs, None, int((counts == 1).sum()), occ, counts
    n_new = np.array([int(flags[kmer_off[r]:kmer_off[r + 1]].sum()) for r in range(n_reads)], np.uint64)
    return tabs, n_new, int(n_new.sum()), occ, None
