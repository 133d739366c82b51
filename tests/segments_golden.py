"""Reads tests/golden/segments_ref.npz (what goetia's compiled
StreamingCompactor::find_new_segments gave, see
tests/golden/make_segments_golden.py) into per-read records, and turns this
package's find_new_segments outputs into the same records for comparison."""
import os

import numpy as np

from conftest import GOLDEN_DIR

GOLD = np.load(os.path.join(GOLDEN_DIR, 'segments_ref.npz'), allow_pickle=False)
CASES = sorted(k[:-len('__params')] for k in GOLD.files if k.endswith('__params'))
SHORT = 'short'


def reads_of(seq, off):
    return [bytes(seq[off[i]:off[i + 1]]).decode() for i in range(len(off) - 1)]


def case(name):
    kind, K, max_table, n_tables, persist, insert_after = (int(v) for v in GOLD[f'{name}__params'])
    return dict(kind=kind, K=K, max_table=max_table, n_tables=n_tables, persist=bool(persist),
                insert_after=bool(insert_after),
                bg=reads_of(GOLD[f'{name}__bg_seq'], GOLD[f'{name}__bg_off']),
                reads=reads_of(GOLD[f'{name}__seq'], GOLD[f'{name}__off']),
                expected=parse(GOLD[f'{name}__out'], GOLD[f'{name}__kmers'], K, len(GOLD[f'{name}__off']) - 1,
                               bool(persist)))


def parse(out, kmers, K, n_reads, persist):
    """Per read: SHORT (the read raised SequenceLengthException) or
    (segments as 7-tuples, new-k-mer count, sorted new-k-mer values or None,
    sorted decision k-mer values, neighbour pairs as ([(value, kmer)...],
    [(value, kmer)...]))."""
    w = [int(x) for x in out]
    kb = bytes(kmers)
    i = kpos = 0
    res = []
    for r in range(n_reads):
        n_seg = w[i]
        i += 1
        if n_seg == (1 << 64) - 1:
            res.append(SHORT)
            continue
        segs = [tuple(w[i + 7 * j:i + 7 * j + 7]) for j in range(n_seg)]
        i += 7 * n_seg
        n_new = w[i]
        i += 1
        new = None
        if not persist or r + 1 == n_reads:
            new = sorted(w[i:i + n_new])
            i += n_new
        n_dec = w[i]
        i += 1
        dec = sorted(w[i:i + n_dec])
        i += n_dec
        n_pairs = w[i]
        i += 1
        pairs = []
        for _ in range(n_pairs):
            nl, nr = w[i], w[i + 1]
            i += 2
            sides = []
            for n in (nl, nr):
                side = []
                for _ in range(n):
                    side.append((w[i], kb[kpos:kpos + K].decode()))
                    i += 1
                    kpos += K
                sides.append(side)
            pairs.append(tuple(sides))
        res.append((segs, n_new, new, dec, pairs))
    assert i == len(w) and kpos == len(kb)
    return res


def record(segments, new_kmers, decision_kmers, pairs, with_new=True):
    """The same record from find_new_segments' outputs."""
    return ([s.as_tuple() for s in segments], len(new_kmers),
            sorted(h.value for h in new_kmers) if with_new else None,
            sorted(h.value for h in decision_kmers),
            [([(k.value, k.kmer) for k in left], [(k.value, k.kmer) for k in right]) for left, right in pairs])
