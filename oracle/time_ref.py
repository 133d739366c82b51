"""Times the compiled goetia reference (oracle/_ref/libgoetia_ref.so, built by
`make -C oracle ref` from the goetia checkout) on samples of the benchmark
workloads.  Test and measurement infrastructure: it runs only where the goetia
checkout is, never on the GPU host, and nothing in the product imports it.

    python oracle/time_ref.py [cms_20m] [segments_5m]

cms_20m      insert + query_sequence of the first 200k reads of the count-min
             workload (skip-N: each ACGT run inserted with dBG::insert_sequence),
             1 thread, k-mers/s; the oracle port is timed beside it on the
             same reads and the query counts are compared.
segments_5m  StreamingCompactor::find_new_segments of the first 20k query
             reads against the graph of the segments workload, 1 thread,
             reads/s, with its segment and decision-k-mer counts (bench.py
             --workload segments_5m prints the GPU path's counts for the same
             reads as `first_reads`).
"""
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
REF_SO = os.path.join(HERE, '_ref', 'libgoetia_ref.so')


def load():
    if not os.path.exists(REF_SO):
        raise SystemExit(f'{REF_SO} is not built: make -C oracle ref REF=<goetia checkout>')
    L = C.CDLL(REF_SO)
    L.ref_new.restype = C.c_void_p
    L.ref_new.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint16]
    L.ref_insert_reads.restype = C.c_uint64
    L.ref_insert_reads.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_char_p, C.c_void_p, C.c_uint64, C.c_void_p]
    L.ref_query_reads.restype = C.c_uint64
    L.ref_query_reads.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_char_p, C.c_void_p, C.c_uint64, C.c_void_p]
    L.ref_find_new_segments.restype = C.c_uint64
    L.ref_find_new_segments.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_uint64, C.c_int, C.c_int,
                                        C.c_void_p, C.c_uint64, C.c_char_p, C.c_uint64, C.c_void_p]
    L.ref_free.argtypes = [C.c_void_p]
    return L


def cms_20m(L, n=200_000):
    from goetia_amd import workload
    from oracle import pyoracle as po
    cfg = dict(workload.CONFIGS['cms_20m'])
    _, batch = workload.make('cms_20m', n / 20_000_000)
    K = cfg['K']
    off = batch.offsets[:n + 1].astype(np.int64).copy()
    seq = batch.seq[:off[-1]].copy()
    nk = int(batch.kmer_counts(K)[:n].sum())
    h = L.ref_new(2, 1, K, cfg["max_table"], cfg["n_tables"])
    q_ref = np.zeros(nk, np.int16)
    t0 = time.perf_counter()
    hashed = L.ref_insert_reads(h, K, po.MODE_SKIP_N, seq.ctypes.data_as(C.c_char_p), off.ctypes.data, n, None)
    L.ref_query_reads(h, K, po.MODE_SKIP_N, seq.ctypes.data_as(C.c_char_p), off.ctypes.data, n, q_ref.ctypes.data)
    t_ref = time.perf_counter() - t0
    L.ref_free(h)
    st = po.Storage(po.BYTE, po.primes_near_x(cfg['n_tables'], cfg['max_table']))
    t0 = time.perf_counter()
    st.insert_reads(seq, off, K, True, po.MODE_SKIP_N)
    q_port = st.query_reads(seq, off, K, True, po.MODE_SKIP_N)
    t_port = time.perf_counter() - t0
    return {'workload': 'cms_20m', 'reads': n, 'kmers': int(hashed), 'cores': 1,
            'reference_kmers_per_s': round(hashed / t_ref, 1), 'reference_s': round(t_ref, 2),
            'port_kmers_per_s': round(hashed / t_port, 1), 'port_s': round(t_port, 2),
            'same_query_counts': bool(np.array_equal(q_ref[:len(q_port)], q_port))}


def segments_5m(L, n=20_000):
    from goetia_amd import workload
    cfg, bg, q = workload.segments_5m(1.0)
    K = cfg['K']
    h = L.ref_new(3, 0, K, 0, 0)
    L.ref_insert_reads(h, K, 0, bg.seq.ctypes.data_as(C.c_char_p), bg.offsets.ctypes.data, bg.n_reads, None)
    n = min(n, q.n_reads)
    cap, kcap = 1 << 24, 1 << 26
    words = np.zeros(cap, np.uint64)
    km = np.zeros(kcap, np.uint8)
    used = C.c_uint64()
    t0 = time.perf_counter()
    w = L.ref_find_new_segments(h, q.seq.ctypes.data_as(C.c_char_p), q.offsets.ctypes.data, n, 0, 0,
                                words.ctypes.data, cap, km.ctypes.data_as(C.c_char_p), kcap, C.byref(used))
    t_ref = time.perf_counter() - t0
    L.ref_free(h)
    # segment and decision-k-mer counts from the driver's record stream:
    # per read n_segments (all-ones for a skipped read), 7 words per segment,
    # the new and decision k-mer lists, then the decision neighbour pairs
    segs = decs = 0
    i = 0
    for _ in range(n):
        ns = int(words[i])
        i += 1
        if ns == (1 << 64) - 1:
            continue
        segs += ns
        i += 7 * ns
        i += 1 + int(words[i])
        i += 1 + int(words[i])
        npairs = int(words[i])
        decs += npairs
        i += 1
        for _ in range(npairs):
            i += 2 + int(words[i]) + int(words[i + 1])
    assert i == w
    return {'workload': 'segments_5m', 'reads': n, 'cores': 1, 'reference_reads_per_s': round(n / t_ref, 1),
            'reference_s': round(t_ref, 2), 'segments': segs, 'decision_kmers': decs}


def main():
    which = sys.argv[1:] or ['cms_20m', 'segments_5m']
    L = load()
    for w in which:
        print(json.dumps({'cms_20m': cms_20m, 'segments_5m': segments_5m}[w](L)), flush=True)


if __name__ == '__main__':
    main()
