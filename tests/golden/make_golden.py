"""Writes the golden fixtures of tests/golden/ from the compiled goetia
reference (oracle/_ref/libgoetia_ref.so, built by `make -C oracle ref` from the
reference sources).  Run in a container that has the reference checkout:

    python tests/golden/make_golden.py

Inputs are the reference's own test data files (copied to tests/golden/data/)
and small seeded synthetic reads stored with the outputs.  Outputs: hash
vectors, primes, per-read n_new, tables, counters, query and
insert_and_query counts, neighbour hashes, degrees and decision k-mers.
"""
import ctypes as C
import hashlib
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_SO = os.path.join(ROOT, 'oracle', '_ref', 'libgoetia_ref.so')
DATA = os.path.join(HERE, 'data')


def ref():
    if not os.path.exists(REF_SO):
        subprocess.run(['make', '-C', os.path.join(ROOT, 'oracle'), '-j8', 'ref'], check=True)
    L = C.CDLL(REF_SO)
    u64p, i64p, i16p, u8p, u32p = (C.POINTER(t) for t in (C.c_uint64, C.c_int64, C.c_int16,
                                                          C.c_uint8, C.c_uint32))
    sig = {
        'ref_hash_read': (C.c_int, [C.c_int, C.c_char_p, C.c_size_t, u64p, u64p]),
        'ref_fwd_hash': (C.c_uint64, [C.c_int, C.c_char_p]),
        'ref_primes': (C.c_uint32, [C.c_uint32, C.c_uint64, u64p]),
        'ref_new': (C.c_void_p, [C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint16]),
        'ref_insert_reads': (C.c_uint64, [C.c_void_p, C.c_int, C.c_int, C.c_char_p, i64p,
                                          C.c_uint64, u64p]),
        'ref_insert_and_query_reads': (C.c_uint64, [C.c_void_p, C.c_int, C.c_char_p, i64p,
                                                    C.c_uint64, i16p]),
        'ref_query_reads': (C.c_uint64, [C.c_void_p, C.c_int, C.c_int, C.c_char_p, i64p,
                                         C.c_uint64, i16p]),
        'ref_query_hash': (C.c_int16, [C.c_void_p, C.c_uint64]),
        'ref_counters': (None, [C.c_void_p, u64p, u64p]),
        'ref_table': (None, [C.c_void_p, C.c_uint32, C.c_uint64, u8p]),
        'ref_kmer_extensions': (None, [C.c_void_p, C.c_char_p, C.c_int, u64p, u64p, u8p, u8p]),
        'ref_find_decision_kmers': (C.c_uint64, [C.c_void_p, C.c_char_p, C.c_size_t, u32p,
                                                 u64p, C.c_uint64]),
    }
    for n, (r, a) in sig.items():
        f = getattr(L, n)
        f.restype = r
        f.argtypes = a
    return L


def P(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def read_fasta(path):
    seqs, cur = [], []
    for line in open(path):
        line = line.strip()
        if line.startswith('>'):
            if cur:
                seqs.append(''.join(cur))
            cur = []
        elif line:
            cur.append(line)
    if cur:
        seqs.append(''.join(cur))
    return seqs


def read_fastq(path):
    lines = [l.rstrip('\n') for l in open(path)]
    return [lines[i + 1] for i in range(0, len(lines) - 3, 4)]


def pack(reads):
    enc = [r.encode() for r in reads]
    off = np.zeros(len(enc) + 1, np.int64)
    off[1:] = np.cumsum([len(e) for e in enc])
    return np.frombuffer(b''.join(enc), np.uint8).copy(), off


def synthetic(seed, genome_len, n_reads, lmin, lmax, err, n_rate=0.0):
    rng = np.random.default_rng(seed)
    genome = rng.choice(np.frombuffer(b'ACGT', np.uint8), genome_len)
    comp = np.zeros(256, np.uint8)
    for a, b in zip(b'ACGT', b'TGCA'):
        comp[a] = b
    reads = []
    for _ in range(n_reads):
        L = int(rng.integers(lmin, lmax + 1))
        s = int(rng.integers(0, genome_len - L + 1))
        r = genome[s:s + L].copy()
        if rng.random() < 0.5:
            r = comp[r[::-1]]
        m = rng.random(L) < err
        r[m] = np.frombuffer(b'ACGT', np.uint8)[(np.searchsorted(np.frombuffer(b'ACGT', np.uint8), r[m]) + rng.integers(1, 4, m.sum())) % 4]
        if n_rate:
            r[rng.random(L) < n_rate] = ord('N')
        reads.append(r.tobytes().decode())
    return reads


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run_storage(L, kind, canonical, K, mode, reads, max_table=0, n_tables=4, keep_tables=True):
    seq, off = pack(reads)
    n = len(reads)
    h = L.ref_new(kind, canonical, K, max_table, n_tables)
    nn = np.zeros(n, np.uint64)
    total = L.ref_insert_reads(h, K, mode, seq.ctypes.data_as(C.c_char_p), P(off, C.c_int64), n,
                               P(nn, C.c_uint64))
    nk = int(np.maximum(np.diff(off) - K + 1, 0).sum())
    q = np.zeros(max(nk, 1), np.int16)
    L.ref_query_reads(h, K, mode, seq.ctypes.data_as(C.c_char_p), P(off, C.c_int64), n, P(q, C.c_int16))
    u, o = C.c_uint64(), C.c_uint64()
    L.ref_counters(h, C.byref(u), C.byref(o))
    out = {'n_new': nn, 'n_kmers': np.array([total], np.uint64),
           'counters': np.array([u.value, o.value], np.uint64),
           'query_sha': np.frombuffer(bytes.fromhex(sha(q[:nk])), np.uint8)}
    if nk <= 20000:
        out['query'] = q[:nk]
    if kind in (1, 2):
        primes = np.zeros(n_tables, np.uint64)
        L.ref_primes(n_tables, max_table, P(primes, C.c_uint64))
        out['tablesizes'] = primes
        for i in range(n_tables):
            nb = int(primes[i]) // 8 + 1 if kind == 1 else int(primes[i])
            t = np.zeros(nb, np.uint8)
            L.ref_table(h, i, nb, P(t, C.c_uint8))
            if keep_tables:
                out[f'table{i}'] = t
            out[f'table{i}_sha'] = np.frombuffer(bytes.fromhex(sha(t)), np.uint8)
    return h, out


def main():
    L = ref()
    golden = {}
    # ---- hash vectors over every byte value (raw bytes, no NUL)
    rng = np.random.default_rng(7)
    for K in (1, 2, 21, 27, 31, 32, 33, 63, 64, 65, 101, 128, 200):
        s = rng.integers(1, 256, 400).astype(np.uint8).tobytes()
        n = len(s) - K + 1
        fw = np.zeros(n, np.uint64)
        rc = np.zeros(n, np.uint64)
        L.ref_hash_read(K, s, len(s), P(fw, C.c_uint64), P(rc, C.c_uint64))
        golden[f'hash_bytes_K{K}_seq'] = np.frombuffer(s, np.uint8)
        golden[f'hash_bytes_K{K}_fw'] = fw
        golden[f'hash_bytes_K{K}_rc'] = rc
    # ---- primes
    cases = [(4, 1_000_000), (4, 10_000_000), (4, 2 ** 28), (4, 2 ** 30), (3, 10), (5, 3), (2, 1),
             (4, 1000), (4, 5000), (4, 20000), (4, 50000), (1, 4)]
    for n, x in cases:
        out = np.zeros(max(n, 1), np.uint64)
        k = L.ref_primes(n, x, P(out, C.c_uint64))
        golden[f'primes_{n}_{x}'] = out[:k]
    # ---- storages over the reference's test data and synthetic reads
    rand20 = read_fasta(os.path.join(DATA, 'random-20-a.fa'))
    fastq = read_fastq(os.path.join(DATA, 'test-fastq-reads.fq'))
    synth = synthetic(11, 20000, 1500, 100, 150, 0.01)
    synth_n = synthetic(12, 5000, 800, 60, 150, 0.02, n_rate=0.01)
    golden['synth_seq'], golden['synth_off'] = pack(synth)
    golden['synthn_seq'], golden['synthn_off'] = pack(synth_n)
    runs = [
        ('rand20_bit_can21', 1, 1, 21, 0, rand20, 10_000, 4),
        ('rand20_byte_can21', 2, 1, 21, 0, rand20, 2_000, 4),
        ('rand20_set_can21', 3, 1, 21, 0, rand20, 0, 4),
        ('rand20_bit_fwd21', 1, 0, 21, 0, rand20, 10_000, 4),
        ('fastq_byte_can25_skipn', 2, 1, 25, 1, fastq, 1_000, 4),
        ('fastq_bit_can21_skipn', 1, 1, 21, 1, fastq, 5_000, 3),
        ('fastq_set_fwd21_skipn', 3, 0, 21, 1, fastq, 0, 4),
        ('synth_bit_can21', 1, 1, 21, 0, synth, 50_000, 4),
        ('synth_byte_can31', 2, 1, 31, 0, synth, 20_000, 4),
        ('synth_byte_fwd31', 2, 0, 31, 0, synth, 5_000, 2),
        ('synth_set_can25', 3, 1, 25, 0, synth, 0, 4),
        ('synthn_byte_can31_skipn', 2, 1, 31, 1, synth_n, 10_000, 4),
        ('synthn_set_can31_skipn', 3, 1, 31, 1, synth_n, 0, 4),
        ('synthn_bit_can31_raw', 1, 1, 31, 0, synth_n, 10_000, 4),
    ]
    meta = {}
    handles = {}
    for name, kind, can, K, mode, reads, mt, nt in runs:
        h, out = run_storage(L, kind, can, K, mode, reads, mt, nt, keep_tables=(kind == 2 and mt <= 5000))
        handles[name] = h
        for k, v in out.items():
            golden[f'{name}__{k}'] = v
        meta[name] = {'kind': kind, 'canonical': can, 'K': K, 'mode': mode, 'max_table': mt,
                      'n_tables': nt, 'input': name.split('_')[0]}
    # ---- count-min insert_and_query per read (sequential), raw mode
    for name, K, mt, reads in (('rand20_byteiq_can21', 21, 2000, rand20),
                               ('synth_byteiq_can31', 31, 5000, synth[:400])):
        seq, off = pack(reads)
        h = L.ref_new(2, 1, K, mt, 4)
        nk = int(np.maximum(np.diff(off) - K + 1, 0).sum())
        cnt = np.zeros(nk, np.int16)
        L.ref_insert_and_query_reads(h, K, seq.ctypes.data_as(C.c_char_p), P(off, C.c_int64),
                                     len(reads), P(cnt, C.c_int16))
        golden[f'{name}__counts'] = cnt
        u, o = C.c_uint64(), C.c_uint64()
        L.ref_counters(h, C.byref(u), C.byref(o))
        golden[f'{name}__counters'] = np.array([u.value, o.value], np.uint64)
        meta[name] = {'kind': 2, 'canonical': 1, 'K': K, 'max_table': mt, 'n_tables': 4,
                      'input': name.split('_')[0], 'n_reads': len(reads)}
    # ---- neighbours and decision k-mers on the synthetic graphs
    for name in ('synth_bit_can21', 'synth_byte_can31', 'synth_set_can25', 'synth_byte_fwd31'):
        h = handles[name]
        K = meta[name]['K']
        kmers = []
        for r in synth[:60]:
            for j in range(0, len(r) - K + 1, 7):
                kmers.append(r[j:j + K])
        left = np.zeros((len(kmers), 4), np.uint64)
        right = np.zeros((len(kmers), 4), np.uint64)
        ind = np.zeros(len(kmers), np.uint8)
        outd = np.zeros(len(kmers), np.uint8)
        for i, km in enumerate(kmers):
            l4 = np.zeros(4, np.uint64)
            r4 = np.zeros(4, np.uint64)
            a, b = C.c_uint8(), C.c_uint8()
            L.ref_kmer_extensions(h, km.encode(), K, P(l4, C.c_uint64), P(r4, C.c_uint64),
                                  C.byref(a), C.byref(b))
            left[i], right[i], ind[i], outd[i] = l4, r4, a.value, b.value
        golden[f'{name}__ext_kmers'] = np.frombuffer(''.join(kmers).encode(), np.uint8)
        golden[f'{name}__ext_left'] = left
        golden[f'{name}__ext_right'] = right
        golden[f'{name}__ext_in'] = ind
        golden[f'{name}__ext_out'] = outd
        pos_all, hs_all, read_idx = [], [], []
        for ri, r in enumerate(synth[:200]):
            cap = len(r)
            pos = np.zeros(cap, np.uint32)
            hs = np.zeros(cap, np.uint64)
            m = L.ref_find_decision_kmers(h, r.encode(), len(r), P(pos, C.c_uint32), P(hs, C.c_uint64), cap)
            pos_all.extend(pos[:m])
            hs_all.extend(hs[:m])
            read_idx.extend([ri] * m)
        golden[f'{name}__decision_read'] = np.array(read_idx, np.uint32)
        golden[f'{name}__decision_pos'] = np.array(pos_all, np.uint32)
        golden[f'{name}__decision_hash'] = np.array(hs_all, np.uint64)
    np.savez_compressed(os.path.join(HERE, 'golden_ref.npz'), **golden)
    with open(os.path.join(HERE, 'golden_ref.json'), 'w') as f:
        json.dump({'runs': meta,
                   'generator': 'tests/golden/make_golden.py with oracle/_ref/libgoetia_ref.so'},
                  f, indent=1, sort_keys=True)
    print('wrote', len(golden), 'arrays')


if __name__ == '__main__':
    sys.exit(main())
