"""Writes what goetia's compiled ByteStorage does with bigcount on
(bytestorage.cc:60-139, storage.hh:129-135) to tests/golden/bigcount_ref.json,
through ref_set_use_bigcount / ref_insert_reads / ref_query_reads /
ref_insert_and_query_reads / ref_save in oracle/ref_driver.cc, and keeps one
saved count-min file holding bigcount pairs (tests/golden/data/bigcount/).

Run where the goetia checkout is (oracle/_ref built by `make -C oracle ref`):
    python tests/golden/make_bigcount_golden.py

Inputs: seeded reads in three batches, chosen to saturate bins: poly-A / poly-C
runs and short repeats (one k-mer many times), random reads on small tables
(bins saturated by many k-mers), a read holding N, and a 40,000-base poly-A
read whose k-mer's bigcount climbs past 32767 (goetia's int16 wraps to
-32768 and stays there).

Recorded per case and batch: per-read n_new, counters, the query counts of
every read so far (SHA-256 of the int16 array, and its first 1000
values), table SHA-256s and the bigcount map (from the saved file:
after the tables, n u64 then (k-mer u64, count i16) pairs; sorted by k-mer
and digested, with the first 200 pairs); and, on a second
storage, insert_and_query counts batch by batch (likewise digested).
"""
import ctypes as C
import hashlib
import json
import os
import struct
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_SO = os.path.join(ROOT, 'oracle', '_ref', 'libgoetia_ref.so')
OUT = os.path.join(HERE, 'bigcount_ref.json')
DATA = os.path.join(HERE, 'data', 'bigcount')


def batches(seed):
    rng = np.random.default_rng(seed)

    def dna(n):
        return ''.join(rng.choice(list('ACGT'), n))
    b1 = [dna(150) for _ in range(300)] + ['A' * 400, 'ACGT' * 150, dna(40) + 'NN' + dna(60), 'ACGTTGCA' * 40]
    b2 = ['A' * 40000] + [dna(int(rng.integers(30, 250))) for _ in range(200)]
    b3 = [dna(150) for _ in range(100)] + ['C' * 300, 'GATTACA' * 50, 'A' * 500]
    return [b1, b2, b3]


CASES = [
    # kind 2 only; canonical, K, max_table, n_tables, seed
    {'canonical': 1, 'K': 21, 'max_table': 1009, 'n_tables': 3, 'seed': 31},
    {'canonical': 0, 'K': 31, 'max_table': 50000, 'n_tables': 4, 'seed': 32},
    {'canonical': 1, 'K': 17, 'max_table': 4000, 'n_tables': 2, 'seed': 33},
    {'canonical': 1, 'K': 21, 'max_table': 101, 'n_tables': 3, 'seed': 34},   # every bin saturates
]


def digest(a):
    return hashlib.sha256(np.asarray(a, np.int16).tobytes()).hexdigest()


def pairs_digest(pairs):
    """SHA-256 of the key-sorted pairs as u64 keys then i16 values."""
    k = np.array([p[0] for p in pairs], np.uint64)
    v = np.array([p[1] for p in pairs], np.int16)
    return hashlib.sha256(k.tobytes() + v.tobytes()).hexdigest()


def pack(reads):
    seq = ''.join(reads).encode()
    off = np.zeros(len(reads) + 1, np.int64)
    off[1:] = np.cumsum([len(r) for r in reads])
    return seq, off


def parse_pairs(path):
    b = open(path, 'rb').read()
    assert b[:4] == b'OXLI'
    pos = 4 + 1 + 1 + 1 + 4
    n_tables = b[pos]
    pos += 1 + 8
    for _ in range(n_tables):
        (ts,) = struct.unpack_from('<Q', b, pos)
        pos += 8 + ts
    (n,) = struct.unpack_from('<Q', b, pos)
    pos += 8
    pairs = []
    for _ in range(n):
        k, v = struct.unpack_from('<Qh', b, pos)
        pos += 10
        pairs.append([k, v])
    assert pos == len(b)
    return sorted(pairs)


def main():
    if not os.path.exists(REF_SO):
        sys.exit(f'{REF_SO} is not built: make -C oracle ref REF=/root/reference')
    L = C.CDLL(REF_SO)
    L.ref_new.restype = C.c_void_p
    L.ref_new.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint16]
    L.ref_set_use_bigcount.argtypes = [C.c_void_p, C.c_int]
    L.ref_insert_reads.restype = C.c_uint64
    L.ref_insert_reads.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_char_p, C.c_void_p, C.c_uint64, C.c_void_p]
    L.ref_query_reads.restype = C.c_uint64
    L.ref_query_reads.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_char_p, C.c_void_p, C.c_uint64, C.c_void_p]
    L.ref_insert_and_query_reads.restype = C.c_uint64
    L.ref_insert_and_query_reads.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_void_p, C.c_uint64, C.c_void_p]
    L.ref_counters.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.ref_table.argtypes = [C.c_void_p, C.c_uint32, C.c_uint64, C.c_void_p]
    L.ref_primes.restype = C.c_uint32
    L.ref_primes.argtypes = [C.c_uint32, C.c_uint64, C.c_void_p]
    L.ref_save.argtypes = [C.c_void_p, C.c_char_p]
    os.makedirs(DATA, exist_ok=True)
    out = {'cases': []}
    tmp = tempfile.mkdtemp(prefix='bigcount_golden_')
    for ci, c in enumerate(CASES):
        K, can = c['K'], c['canonical']
        bs = batches(c['seed'])
        sizes = np.zeros(c['n_tables'], np.uint64)
        L.ref_primes(c['n_tables'], c['max_table'], sizes.ctypes.data)
        h = L.ref_new(2, can, K, c['max_table'], c['n_tables'])
        L.ref_set_use_bigcount(h, 1)
        rec = dict(c, sizes=sizes.tolist(), batches=[])
        seen = []
        for bi, reads in enumerate(bs):
            seq, off = pack(reads)
            n_new = np.zeros(len(reads), np.uint64)
            L.ref_insert_reads(h, K, 0, seq, off.ctypes.data, len(reads), n_new.ctypes.data)
            seen += reads
            qs, qo = pack(seen)
            nk = sum(max(0, len(r) - K + 1) for r in seen)
            counts = np.zeros(max(nk, 1), np.int16)
            L.ref_query_reads(h, K, 0, qs, qo.ctypes.data, len(seen), counts.ctypes.data)
            u, o = C.c_uint64(), C.c_uint64()
            L.ref_counters(h, C.byref(u), C.byref(o))
            shas = []
            for i, p in enumerate(sizes):
                t = np.zeros(int(p), np.uint8)
                L.ref_table(h, i, int(p), t.ctypes.data)
                shas.append(hashlib.sha256(t.tobytes()).hexdigest())
            path = os.path.join(tmp, f'c{ci}_b{bi}.oxli')
            L.ref_save(h, path.encode())
            pairs = parse_pairs(path)
            if ci == 0 and bi == len(bs) - 1:
                keep = os.path.join(DATA, 'cms_big.oxli')
                os.replace(path, keep)
            rec['batches'].append({'reads': reads, 'n_new': n_new.tolist(), 'counters': [u.value, o.value],
                                   'query_sha256': digest(counts[:nk]), 'query_head': counts[:1000].tolist(),
                                   'table_sha256': shas, 'bigcounts_n': len(pairs),
                                   'bigcounts_sha256': pairs_digest(pairs), 'bigcounts_head': pairs[:200]})
            print(f'case {ci} batch {bi}: {len(pairs)} bigcount pairs, min {min((v for _, v in pairs), default=None)}'
                  f', max {max((v for _, v in pairs), default=None)}')
        # insert_and_query on a fresh storage, batch by batch
        h2 = L.ref_new(2, can, K, c['max_table'], c['n_tables'])
        L.ref_set_use_bigcount(h2, 1)
        iq = []
        for reads in bs:
            seq, off = pack(reads)
            nk = sum(max(0, len(r) - K + 1) for r in reads)
            counts = np.zeros(max(nk, 1), np.int16)
            L.ref_insert_and_query_reads(h2, K, seq, off.ctypes.data, len(reads), counts.ctypes.data)
            iq.append({'sha256': digest(counts[:nk]), 'head': counts[:1000].tolist()})
        rec['insert_and_query'] = iq
        out['cases'].append(rec)
    with open(OUT, 'w') as f:
        json.dump(out, f)
        f.write('\n')
    print('wrote', OUT, os.path.getsize(OUT), 'bytes')


if __name__ == '__main__':
    main()
