"""Writes tests/golden/config_digests.json: SHA-256 digests of what the oracle
(oracle/oracle.c, the sequential restatement pinned to goetia's compiled
reference by tests/test_oracle.py) produces on two workloads at full size,
so the GPU path can be compared with it where the oracle would take minutes:

cms_20m       50 Mbp genome, 20M x 150 bp reads, 0.5% substitutions, 0.01% N
              (windows holding N skipped), K=31 canonical, ByteStorage(2^28, 4):
              insert every read in order (per-read n_new), the four tables,
              n_unique / n_occupied, query_sequence of every read, and a
              one-pass insert_and_query_sequence of every read on fresh tables
              (bytestorage.cc:60-150, dbg.hh:307-318,327-340,364-376).
neighbors_5m  20 Mbp genome with 1,000 copies of a 2 kb repeat and a 0.1% SNP
              haplotype, 5M x 150 bp reads at 1% error, K=31 canonical,
              ByteStorage(2^28, 4): insert every read (first-touch flag of
              every window), then in / out degrees at count >= 2 of every
              window that was new (unitig_walker.hh:322-330,1003-1045), and
              the decision nodes (in or out degree > 1) as a sorted set of
              canonical values.

The reads come from the library's seeded host generator (goetia_amd.workload),
whose digest is recorded too.  Array digests are over little-endian bytes:
n_new uint64, counts int16, flags and degrees uint8 (0xFF where a window is
skipped or not selected), decision nodes uint64.

Runs on the host only (about 30 minutes on 8 cores, 12 GB of memory):
    python tests/golden/make_config_digests.py [cms_20m] [neighbors_5m]
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
OUT = os.path.join(HERE, 'config_digests.json')
THREADS = min(8, os.cpu_count() or 1)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def input_digest(batch):
    h = hashlib.sha256(np.ascontiguousarray(batch.seq, np.uint8).tobytes())
    h.update(np.ascontiguousarray(batch.offsets, '<i8').tobytes())
    return h.hexdigest()


def log(*a):
    print(time.strftime('%H:%M:%S'), *a, flush=True)


def cms_20m():
    from goetia_amd import workload
    from oracle import pyoracle as po
    cfg, b = workload.make('cms_20m')
    K, mode = cfg['K'], po.MODE_SKIP_N
    sizes = po.primes_near_x(cfg['n_tables'], cfg['max_table'])
    out = {'reads': int(b.n_reads), 'bases': int(b.n_bases), 'K': K, 'mode': 'skip-N', 'tablesizes': sizes,
           'input_sha256': input_digest(b)}
    st = po.Storage(po.BYTE, sizes)
    log('cms_20m: insert')
    hashed, n_new = st.insert_reads(b.seq, b.offsets, K, True, mode)
    out.update(kmers=int(hashed), n_unique=st.n_unique, n_occupied=st.n_occupied,
               n_new_sum=int(n_new.sum()), n_new_sha256=sha(n_new.astype('<u8')),
               table_sha256=[sha(t) for t in st.tables()])
    del n_new
    log('cms_20m: query')
    q = st.query_reads_mt(b.seq, b.offsets, K, True, mode, threads=THREADS)
    out.update(windows=int(len(q)), query_sha256=sha(q.astype('<i2')),
               query_hist=np.bincount(q[q >= 0].astype(np.int64), minlength=256).tolist())
    del q, st
    log('cms_20m: insert_and_query')
    st = po.Storage(po.BYTE, sizes)
    iq = st.insert_and_query_reads(b.seq, b.offsets, K, True, mode)
    out.update(insert_and_query_sha256=sha(iq.astype('<i2')),
               insert_and_query_table_sha256=[sha(t) for t in st.tables()],
               insert_and_query_counters=[st.n_unique, st.n_occupied])
    log('cms_20m: done')
    return out


def neighbors_5m():
    from goetia_amd import workload
    from oracle import pyoracle as po
    cfg, b = workload.neighbors_5m()
    K, mc = cfg['K'], cfg['min_count']
    sizes = po.primes_near_x(cfg['n_tables'], cfg['max_table'])
    out = {'reads': int(b.n_reads), 'bases': int(b.n_bases), 'K': K, 'min_count': mc, 'tablesizes': sizes,
           'input_sha256': input_digest(b)}
    st = po.Storage(po.BYTE, sizes)
    log('neighbors_5m: insert')
    hashed, n_new, flags = st.insert_reads_flags(b.seq, b.offsets, K, True, po.MODE_RAW)
    out.update(kmers=int(hashed), n_unique=st.n_unique, n_occupied=st.n_occupied,
               n_new_sha256=sha(n_new.astype('<u8')), new_flags_sha256=sha(flags),
               new_windows=int(flags.sum()), table_sha256=[sha(t) for t in st.tables()])
    log('neighbors_5m: degrees')
    ind, outd = st.degrees_reads_mt(b.seq, b.offsets, K, True, po.MODE_RAW, mc, flags, threads=THREADS)
    out.update(in_deg_sha256=sha(ind), out_deg_sha256=sha(outd))
    dec = flags.astype(bool) & ((ind > 1) | (outd > 1))
    hv = po.hash_reads(b.seq, b.offsets, K, True, po.MODE_RAW)
    nodes = np.unique(hv[dec])
    out.update(decision_windows=int(dec.sum()), decision_nodes=int(len(nodes)),
               decision_nodes_sha256=sha(nodes.astype('<u8')),
               degree_pairs=np.bincount((ind[dec].astype(np.int64) * 5 + outd[dec]), minlength=25).tolist())
    log('neighbors_5m: done')
    return out


def main():
    which = sys.argv[1:] or ['cms_20m', 'neighbors_5m']
    res = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for w in which:
        res[w] = {'cms_20m': cms_20m, 'neighbors_5m': neighbors_5m}[w]()
        with open(OUT, 'w') as f:
            json.dump(res, f, indent=1)
            f.write('\n')
    print(f'wrote {OUT}')


if __name__ == '__main__':
    main()
