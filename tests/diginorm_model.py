"""CPU model of the batched diginorm (goetia_amd/csrc: gt_diginorm_dev and the
dn_* kernels), step for step: chunks of reads cut by first-pass tiles, the
chunk's (window, table) events by bucket in window order (what the binned
routing produces; tests/binned_model.py checks that routing), the replay in
64-event steps of whole reads (a read running past a step handled on its own),
the decision pass with its final / not-final bookkeeping, the guess cap that
cuts a chunk at its first undecided read, and the masked in-order apply.

Test infrastructure: the hashes come from the oracle (oracle/pyoracle.py).
"""
import numpy as np

from binned_model import BIN_BITS, BIN_TILE, BUCKET_BINS, _windows

KEPT, SETTLED = 1, 2


def _replay_bucket(ev, sl, state, cutoff, fails):
    """ev: list of (key, ord, rid) in window order; sl: the bucket's bins."""
    n = len(ev)
    pos = 0
    while pos < n:
        chunk = ev[pos:pos + 64]
        rn = ev[pos + 64][2] if pos + 64 < n else None
        np_ = len(chunk)
        first = [0] * np_
        for i in range(1, np_):
            first[i] = i if chunk[i][2] != chunk[i - 1][2] else first[i - 1]
        if np_ == 64 and chunk[63][2] == rn:
            s63 = first[63]
            if s63 == 0:
                pos = _long_read(ev, pos, sl, state, cutoff, fails)
                continue
            np_ = s63
        act = chunk[:np_]
        for i, (key, ordv, r) in enumerate(act):
            seen = int(sl[key]) + sum(1 for j in range(first[i]) if act[j][0] == key and state[act[j][2]] & KEPT)
            if not state[r] & SETTLED and min(seen, 255) < cutoff:
                fails[ordv] = 1
        add = {}
        for key, _, r in act:
            if state[r] & KEPT:
                add[key] = add.get(key, 0) + 1
        for key, c in add.items():
            sl[key] = min(255, int(sl[key]) + c)
        pos += np_


def _long_read(ev, pos, sl, state, cutoff, fails):
    rr = ev[pos][2]
    q = pos
    while q < len(ev) and ev[q][2] == rr:
        if not state[rr] & SETTLED and int(sl[ev[q][0]]) < cutoff:
            fails[ev[q][1]] = 1
        q += 1
    if state[rr] & KEPT:
        for key, _, _ in ev[pos:q]:
            sl[key] = min(255, int(sl[key]) + 1)
    return q


def diginorm(sizes, seq, off, K, mode, cutoff, tables=None, chunk_starts=1 << 25, cap=32):
    """(passed, tables, n_unique, n_occupied, guesses per chunk) starting from
    `tables` (count-min bytes per table, default empty) and zero counters."""
    off = np.asarray(off, np.int64)
    n_reads, nt = len(off) - 1, len(sizes)
    wins, kmer_off, hv, valid = _windows(seq, off, K, mode)
    tabs = [np.zeros(p, np.uint8) if tables is None else np.array(tables[i][:p], np.uint8)
            for i, p in enumerate(sizes)]
    bfirst, nb = [], 0
    for p in sizes:
        bfirst.append(nb)
        nb += (p + BUCKET_BINS - 1) // BUCKET_BINS
    wread = np.repeat(np.arange(n_reads), wins)

    def tiles(a, b):
        return (int(off[b]) + BIN_TILE - 1) // BIN_TILE - int(off[a]) // BIN_TILE

    tiles_max = max([max(chunk_starts, BIN_TILE) // BIN_TILE] + [tiles(r, r + 1) for r in range(n_reads)])
    state = np.zeros(n_reads, np.uint8)
    passed = np.zeros(n_reads, bool)
    n_unique = n_occ = 0
    guesses = []
    r0 = 0
    while r0 < n_reads:
        r1 = r0 + 1
        while r1 < n_reads and tiles(r0, r1 + 1) <= tiles_max:
            r1 += 1
        for r in range(r0, r1):
            state[r] = SETTLED if off[r + 1] - off[r] < K else 0
        r_end = r1
        k0, k1 = int(kmer_off[r0]), int(kmer_off[r1])
        if k1 > k0:
            buckets = {}
            for w in range(k0, k1):
                if not valid[w]:
                    continue
                for t, p in enumerate(sizes):
                    b = int(hv[w]) % p
                    buckets.setdefault(bfirst[t] + (b >> BIN_BITS), []).append((b & (BUCKET_BINS - 1), w - k0,
                                                                                int(wread[w])))
            hasev = np.zeros(k1 - k0, bool)
            for ev in buckets.values():
                for _, o, _ in ev:
                    hasev[o] = True

            def table_of(bk):
                return max(t for t in range(nt) if bfirst[t] <= bk)

            settled = False
            g = 0
            while g < cap and not settled:
                g += 1
                fails = np.zeros(k1 - k0, bool)
                for bk, ev in buckets.items():
                    t = table_of(bk)
                    lo = (bk - bfirst[t]) << BIN_BITS
                    sl = np.zeros(BUCKET_BINS, np.int64)
                    part = tabs[t][lo:lo + BUCKET_BINS]
                    sl[:len(part)] = part
                    _replay_bucket(ev, sl, state, cutoff, fails)
                changed, first_open = 0, None
                for r in range(r0, r1):
                    st = state[r]
                    if st & SETTLED:
                        continue
                    a, b = int(kmer_off[r]) - k0, int(kmer_off[r + 1]) - k0
                    c = int(np.sum(hasev[a:b] & ~fails[a:b]))
                    min_req = int(0.5 + float(np.float32(off[r + 1] - off[r] - K + 1) / np.float32(2)))
                    keep = c < min_req
                    fin = (not keep) if (g & 1) else keep
                    changed += keep != bool(st & KEPT)
                    state[r] = (KEPT if keep else 0) | (SETTLED if fin else 0)
                    if not fin and first_open is None:
                        first_open = r
                settled = changed == 0
            guesses.append(g)
            if not settled:
                assert first_open is not None and r0 < first_open <= r1
                r_end = first_open
            # masked apply in window order per bucket (bin_apply semantics)
            flags = np.zeros(k1 - k0, bool)
            for bk, ev in buckets.items():
                t = table_of(bk)
                lo = (bk - bfirst[t]) << BIN_BITS
                for key, o, r in ev:
                    if r >= r_end or not state[r] & KEPT:
                        continue
                    c = int(tabs[t][lo + key])
                    if c == 0:
                        flags[o] = True
                        n_occ += t == 0
                    tabs[t][lo + key] = min(255, c + 1)
            n_unique += int(flags[:int(kmer_off[r_end]) - k0].sum())
        for r in range(r0, r_end):
            passed[r] = bool(state[r] & KEPT)
        r0 = r_end
    return passed, tabs, n_unique, n_occ, guesses
