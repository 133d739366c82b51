"""Writes paired-read fixtures and what goetia's compiled SplitPairedReader and
InserterProcessor make of them (tests/golden/pairs_ref.json), through
ref_paired_walk and ref_paired_insert in oracle/ref_driver.cc.

Run where the goetia checkout is (oracle/_ref built by `make -C oracle ref`):
    python tests/golden/make_pairs_golden.py

Inputs (tests/golden/data/):
  left.fq / right.fq      goetia's own paired test files (tests/test-data)
  pairs/edge_1.fq, _2.fq  eight steps: a good pair, an N on the left, a
                          10 bp right read, a right quality longer than its
                          sequence, unpaired names, lowercase bases, a good
                          pair, and one more left record than right
  pairs/many_1.fa, _2.fa  300 seeded pairs of 60-140 bp, names "p<i>/1" and
                          "p<i>/2" (with a few unpaired), for interval polling

Recorded per reader setting (strict, min_length, force_name_match): the
lengths SplitPairedReader::next returns step by step (-1 empty side, -2 a
rejected pair, -3 the reader raising); per insert case: every advance's
(n_sequences, time, remaining), whether the reader raised, n_skipped,
n_parsed, counters and table SHA-256s.
"""
import ctypes as C
import hashlib
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_SO = os.path.join(ROOT, 'oracle', '_ref', 'libgoetia_ref.so')
DATA = os.path.join(HERE, 'data')
PAIRS = os.path.join(DATA, 'pairs')
OUT = os.path.join(HERE, 'pairs_ref.json')
REFERENCE_DATA = '/root/reference/tests/test-data'


def fq(name, seq, qual=None):
    return f'@{name}\n{seq}\n+\n{qual if qual is not None else "I" * len(seq)}\n'


def write_edge(rng):
    def dna(n):
        return ''.join(rng.choice(list('ACGT'), n))
    left, right = [], []
    left.append(fq('r1/1', dna(60))); right.append(fq('r1/2', dna(55)))
    s = list(dna(50)); s[20] = 'N'
    left.append(fq('r2/1', ''.join(s))); right.append(fq('r2/2', dna(48)))
    left.append(fq('r3/1', dna(45))); right.append(fq('r3/2', dna(10)))
    r = dna(40)
    left.append(fq('r4/1', dna(52))); right.append(fq('r4/2', r, 'I' * (len(r) + 5)))
    left.append(fq('r5/1', dna(44))); right.append(fq('x5/2', dna(46)))
    left.append(fq('r6/1', dna(50).lower())); right.append(fq('r6/2', dna(51)))
    left.append(fq('r7/1', dna(70))); right.append(fq('r7/2', dna(65)))
    left.append(fq('r8/1', dna(58)))
    with open(os.path.join(PAIRS, 'edge_1.fq'), 'w') as f:
        f.write(''.join(left))
    with open(os.path.join(PAIRS, 'edge_2.fq'), 'w') as f:
        f.write(''.join(right))


def write_many(rng):
    lines = [[], []]
    for i in range(300):
        for k in range(2):
            n = int(rng.integers(60, 141))
            name = f'p{i}/{k + 1}' if i % 37 else f'p{i}/{k + 1}{"x" * k}'
            lines[k].append(f'>{name}\n' + ''.join(rng.choice(list('ACGT'), n)) + '\n')
    for k in range(2):
        with open(os.path.join(PAIRS, f'many_{k + 1}.fa'), 'w') as f:
            f.write(''.join(lines[k]))


FILES = {'goetia': ('left.fq', 'right.fq'), 'edge': ('pairs/edge_1.fq', 'pairs/edge_2.fq'),
         'many': ('pairs/many_1.fa', 'pairs/many_2.fa')}
READERS = [(0, 0, 0), (0, 20, 0), (1, 0, 0), (0, 0, 1), (1, 0, 1)]   # strict, min_length, force_name_match
INSERTS = [
    # files, kind, canonical, K, max_table, n_tables, strict, min_length, force, interval
    ('goetia', 1, 1, 21, 100_000, 4, 0, 0, 0, 10_000),
    ('goetia', 2, 1, 21, 50_000, 3, 0, 0, 0, 500),
    ('edge', 2, 1, 15, 20_000, 2, 0, 0, 0, 30),
    ('edge', 2, 1, 15, 20_000, 2, 0, 20, 0, 1),
    ('edge', 1, 0, 17, 20_000, 3, 1, 0, 0, 40),
    ('edge', 2, 1, 15, 20_000, 2, 0, 0, 1, 25),
    ('edge', 2, 1, 15, 20_000, 2, 1, 0, 1, 25),
    ('many', 2, 1, 21, 200_000, 4, 0, 0, 0, 1000),
    ('many', 1, 1, 25, 200_000, 4, 0, 0, 1, 777),
]


def main():
    if not os.path.exists(REF_SO):
        sys.exit(f'{REF_SO} is not built: make -C oracle ref REF=/root/reference')
    os.makedirs(PAIRS, exist_ok=True)
    for f in ('left.fq', 'right.fq'):
        shutil.copy(os.path.join(REFERENCE_DATA, f), os.path.join(DATA, f))
    rng = np.random.default_rng(515)
    write_edge(rng)
    write_many(rng)
    L = C.CDLL(REF_SO)
    L.ref_new.restype = C.c_void_p
    L.ref_new.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint16]
    L.ref_paired_walk.restype = C.c_uint64
    L.ref_paired_walk.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_uint32, C.c_int, C.c_void_p, C.c_uint64]
    L.ref_paired_insert.restype = C.c_uint64
    L.ref_paired_insert.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int, C.c_uint32, C.c_int, C.c_uint64,
                                    C.c_void_p, C.c_uint64, C.POINTER(C.c_int), C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_uint64)]
    L.ref_counters.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.ref_table.argtypes = [C.c_void_p, C.c_uint32, C.c_uint64, C.c_void_p]
    L.ref_primes.restype = C.c_uint32
    L.ref_primes.argtypes = [C.c_uint32, C.c_uint64, C.c_void_p]
    out = {'walks': {}, 'inserts': []}
    for key, (lf, rf) in FILES.items():
        lp, rp = os.path.join(DATA, lf).encode(), os.path.join(DATA, rf).encode()
        for strict, minlen, force in READERS:
            lens = np.zeros(2 * 4096, np.int64)
            n = L.ref_paired_walk(lp, rp, strict, minlen, force, lens.ctypes.data, 4096)
            out['walks'][f'{key}|{strict}|{minlen}|{force}'] = lens[:2 * n].tolist()
    for files, kind, can, K, mt, nt, strict, minlen, force, interval in INSERTS:
        lf, rf = FILES[files]
        h = L.ref_new(kind, can, K, mt, nt)
        steps = np.zeros(3 * 4096, np.uint64)
        err, skipped, parsed = C.c_int(), C.c_uint64(), C.c_uint64()
        n = L.ref_paired_insert(h, os.path.join(DATA, lf).encode(), os.path.join(DATA, rf).encode(), strict, minlen,
                                force, interval, steps.ctypes.data, 4096, C.byref(err), C.byref(skipped),
                                C.byref(parsed))
        u, o = C.c_uint64(), C.c_uint64()
        L.ref_counters(h, C.byref(u), C.byref(o))
        sizes = np.zeros(nt, np.uint64)
        L.ref_primes(nt, mt, sizes.ctypes.data)
        shas = []
        for i, p in enumerate(sizes):
            nb = int(p) // 8 + 1 if kind == 1 else int(p)
            t = np.zeros(nb, np.uint8)
            L.ref_table(h, i, nb, t.ctypes.data)
            shas.append(hashlib.sha256(t.tobytes()).hexdigest())
        out['inserts'].append({'files': files, 'kind': kind, 'canonical': can, 'K': K, 'max_table': mt,
                               'n_tables': nt, 'strict': strict, 'min_length': minlen, 'force_name_match': force,
                               'interval': interval, 'steps': steps[:3 * n].tolist(), 'error': err.value,
                               'n_skipped': skipped.value, 'n_parsed': parsed.value,
                               'counters': [u.value, o.value], 'table_sha256': shas})
        print(files, kind, strict, minlen, force, interval, '->', n, 'advances, error', err.value,
              'last', steps[3 * n - 3:3 * n].tolist() if n else None)
    with open(OUT, 'w') as f:
        json.dump(out, f, indent=0)
        f.write('\n')


if __name__ == '__main__':
    main()
