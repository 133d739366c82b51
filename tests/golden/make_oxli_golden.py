"""Writes OXLI table files with goetia's compiled storages (dBG::save through
ref_save in oracle/ref_driver.cc) under tests/golden/data/oxli/, and
tests/golden/oxli_ref.json: for every file, what goetia's loader reads back
(ksize, table sizes, occupied bins, table SHA-256s; ref_load_file) and how the
tables were made, so the device storages can be filled the same way.

Run where the goetia checkout is (oracle/_ref built by `make -C oracle ref`):
    python tests/golden/make_oxli_golden.py

Files:
  bloom_k21.oxli      BitStorage(primes below 20,000, 4), CanLemireShifter,
                      K=21, goetia's random-20-a.fa reads
  bloom_k21.gz        the same storage saved under a .gz name (BitStorage::save
                      ignores the name and writes an uncompressed file)
  counts_k21.oxli     ByteStorage(primes below 5,000, 3), K=21, test-fastq-reads.fq
                      inserted 300 times (counts saturate at 255)
  counts_k21.oxli.gz  the same, gzip-compressed (ByteStorageGzFileWriter)
  counts_fwd_k17.gz   ByteStorage(primes below 3,001, 2), FwdLemireShifter, K=17,
                      random-20-a.fa once
"""
import ctypes as C
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, 'tests'))
sys.path.insert(0, ROOT)
REF_SO = os.path.join(ROOT, 'oracle', '_ref', 'libgoetia_ref.so')
DATA = os.path.join(HERE, 'data')
OUT_DIR = os.path.join(DATA, 'oxli')
OUT = os.path.join(HERE, 'oxli_ref.json')

CASES = [
    # file, kind, canonical, K, max_table, n_tables, input, repeats
    ('bloom_k21.oxli', 1, 1, 21, 20_000, 4, 'random-20-a.fa', 1),
    ('bloom_k21.gz', 1, 1, 21, 20_000, 4, 'random-20-a.fa', 1),
    ('counts_k21.oxli', 2, 1, 21, 5_000, 3, 'test-fastq-reads.fq', 300),
    ('counts_k21.oxli.gz', 2, 1, 21, 5_000, 3, 'test-fastq-reads.fq', 300),
    ('counts_fwd_k17.gz', 2, 0, 17, 3_001, 2, 'random-20-a.fa', 1),
]


def read_input(name):
    from conftest import read_fasta, read_fastq
    path = os.path.join(DATA, name)
    reads = read_fasta(path) if name.endswith('.fa') else read_fastq(path)
    enc = [r.encode() for r in reads]
    off = np.zeros(len(enc) + 1, np.int64)
    off[1:] = np.cumsum([len(e) for e in enc])
    return np.frombuffer(b''.join(enc), np.uint8).copy(), off


def main():
    if not os.path.exists(REF_SO):
        sys.exit(f'{REF_SO} is not built: make -C oracle ref REF=/root/reference')
    L = C.CDLL(REF_SO)
    L.ref_new.restype = C.c_void_p
    L.ref_new.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint16]
    L.ref_insert_reads.restype = C.c_uint64
    L.ref_insert_reads.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_char_p, C.c_void_p, C.c_uint64, C.c_void_p]
    L.ref_save.argtypes = [C.c_void_p, C.c_char_p]
    L.ref_load_file.restype = C.c_uint64
    L.ref_load_file.argtypes = [C.c_int, C.c_char_p, C.POINTER(C.c_uint16), C.POINTER(C.c_uint32), C.c_void_p,
                                C.POINTER(C.c_uint64), C.c_void_p, C.c_uint64]
    L.ref_free.argtypes = [C.c_void_p]
    os.makedirs(OUT_DIR, exist_ok=True)
    meta = {}
    for fname, kind, can, K, mt, nt, inp, reps in CASES:
        seq, off = read_input(inp)
        h = L.ref_new(kind, can, K, mt, nt)
        for _ in range(reps):
            L.ref_insert_reads(h, K, 0, seq.ctypes.data_as(C.c_char_p), off.ctypes.data, len(off) - 1, None)
        path = os.path.join(OUT_DIR, fname)
        L.ref_save(h, path.encode())
        L.ref_free(h)
        k, n, occ = C.c_uint16(), C.c_uint32(), C.c_uint64()
        sizes = np.zeros(16, np.uint64)
        buf = np.zeros(1 << 20, np.uint8)
        w = L.ref_load_file(kind, path.encode(), C.byref(k), C.byref(n), sizes.ctypes.data, C.byref(occ),
                            buf.ctypes.data, len(buf))
        assert w != 2**64 - 1 and w <= len(buf), fname
        tabs, pos = [], 0
        for i in range(n.value):
            nb = int(sizes[i]) // 8 + 1 if kind == 1 else int(sizes[i])
            tabs.append(hashlib.sha256(buf[pos:pos + nb].tobytes()).hexdigest())
            pos += nb
        raw = open(path, 'rb').read()
        meta[fname] = {'kind': kind, 'canonical': can, 'K': K, 'max_table': mt, 'n_tables': nt, 'input': inp,
                       'repeats': reps, 'ksize': k.value, 'tablesizes': [int(x) for x in sizes[:n.value]],
                       'occupied': occ.value, 'table_sha256': tabs,
                       'file_sha256': hashlib.sha256(raw).hexdigest(), 'gzip': raw[:2] == b'\x1f\x8b'}
        print(fname, len(raw), 'bytes', 'gzip' if meta[fname]['gzip'] else 'plain')
    with open(OUT, 'w') as f:
        json.dump(meta, f, indent=1)
        f.write('\n')


if __name__ == '__main__':
    main()
