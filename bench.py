"""Benchmark of the streaming dBG hot path on one MI355X.

Workload (the reference's count-min configuration at full size): a seeded
50 Mbp uniform genome, 20,000,000 reads x 150 bp with 0.5% substitutions and
0.01% N bases (k-mers spanning N are skipped), canonical k = 31, ByteStorage
with 4 tables of the primes below 2^28 (1.07 GB of saturating 8-bit
counters).  One step = reset the tables, insert every read (per-read n_new,
exact first-touch semantics; the binned insert) and query_sequence every read
(per-k-mer counts).  Inputs are resident in HBM before the timed region.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line: k-mers/s = hashed k-mers / (insert + query time), the
roofline of the dominant kernel (algorithmic bytes per k-mer over its
HIP-event-timed duration, with the per-kernel breakdown), the one-pass
insert_and_query rate on the same reads, the plain chunk-at-a-time latency
and the CPU oracle timed on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
# Kernels of one step: the binned insert (hash + first partition pass, second
# pass count + scatter, per-bucket replay, per-read new counts) or, for small
# batches, the atomic insert (claim + apply), then the query.
KERNELS = ['bin_hash', 'bin_count', 'bin_route', 'bin_apply', 'bin_new', 'claim', 'apply', 'query']
# Algorithmic HBM bytes per k-mer.  Random table accesses count one 64-byte
# granule per table (twice for a read-modify-write).  The binned passes move
# 8-byte events, one per table per k-mer: write once when routed, read once
# by the next pass; the replay also reads and writes each table slice once per
# sub-batch (2 x 1.07 GB per 2^27 starts, 16 B per k-mer).
BYTES_PER_KMER = {'claim': 4 * 64, 'apply': 4 * 2 * 64, 'query': 4 * 64,
                  'bin_hash': 4 * 8, 'bin_count': 4 * 8, 'bin_route': 2 * 4 * 8,
                  'bin_apply': 4 * 8 + 16, 'bin_new': 1}
# HBM bytes per k-mer (read + write) measured with rocprofv3 FETCH_SIZE /
# WRITE_SIZE on this workload and scaled by a 1 GiB copy (the table and the
# commands are in profiles/cms_20m_hbm_traffic.md).
MEASURED_BYTES_PER_KMER = {'bin_hash': 36.3, 'bin_count': 37.1, 'bin_route': 93.6, 'bin_apply': 64.0, 'bin_new': 1.2,
                           'query': 515.5}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument('--gpus', type=int, default=1)
    p.add_argument('--steps', type=int, default=3)
    p.add_argument('--warmup', type=int, default=1)
    p.add_argument('--scale', type=float, default=1.0, help='fraction of the 20M reads (testing)')
    p.add_argument('--chunk-reads', type=int, default=1_000_000)
    p.add_argument('--cpu-sample-reads', type=int, default=500_000)
    p.add_argument('--no-cpu-baseline', action='store_true')
    p.add_argument('--no-chunks', action='store_true')
    p.add_argument('--no-iq', action='store_true', help='skip the one-pass insert_and_query timing')
    p.add_argument('--no-random-access', action='store_true',
                   help='skip the random-access roofline microbenchmark (4 GiB buffer)')
    p.add_argument('--random-access-sweep', action='store_true',
                   help='also time random reads over buffers of 16 MiB to 4 GiB')
    p.add_argument('--copy-calibration', action='store_true',
                   help='run one 1 GiB device copy first (reference kernel for HBM byte counters)')
    p.add_argument('--workload', default='cms_20m',
                   choices=['cms_20m', 'bloom_100k', 'exact_mixed', 'neighbors_5m', 'serving_64', 'fastx',
                            'diginorm', 'segments_5m'],
                   help='cms_20m is the headline; the others print their own JSON line')
    p.add_argument('--plain-chunk-kmers', type=int, default=500_000,
                   help='chunk size of the plain path in k-mers (goetia IntervalCounter default)')
    p.add_argument('--serving-chunks', type=int, default=400,
                   help='chunks of the served-vs-plain comparison (64 streams of BitStorage(2^30, 4)); 0 skips it')
    return p.parse_args()


def kernel_times(N, names):
    import ctypes as C
    out = {}
    for n in names:
        cnt, ms = C.c_uint64(), C.c_double()
        N.call('gt_profile_get', n.encode(), C.byref(cnt), C.byref(ms))
        out[n] = (cnt.value, ms.value)
    return out


SWEEP_MB = (16, 64, 256, 1024, 4096)


def random_access(N, table_bytes, n_ops=1 << 30, sweep=False):
    """The device's random-access roofline: uniform random 32-bit reads,
    atomicOr and atomicCAS (saturating byte increment) over a 4 GiB buffer
    from every CU (gt_membench_random), and random reads over a buffer of the
    tables' size; best of three runs each, in ops/s.  With `sweep`, random
    reads over buffers of SWEEP_MB MiB as well (the working set at
    which the rate stops depending on caches) and reads with the
    non-temporal cache policy over 4 GiB."""
    import ctypes as C

    import torch
    n_words = 1 << 30
    buf = torch.zeros(n_words, dtype=torch.int32, device='cuda')
    stream = torch.cuda.current_stream()

    def rate(words, kind):
        best = 0.0
        for rep in range(3):
            done = C.c_uint64()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            N.call('gt_membench_random', C.c_void_p(buf.data_ptr()), words, n_ops, kind, rep + 1,
                   C.c_void_p(stream.cuda_stream), C.byref(done))
            e1.record(stream)
            e1.synchronize()
            best = max(best, done.value / (e0.elapsed_time(e1) / 1e3))
        return round(best, 1)

    out = {'buffer_bytes': n_words * 4}
    for kind, name in ((0, 'read'), (1, 'atomic_or'), (2, 'atomic_cas')):
        out[f'{name}_ops_per_s'] = rate(n_words, kind)
    tw = max(1, min(n_words, table_bytes // 4))
    out['table_size_bytes'] = tw * 4
    out['read_ops_per_s_table_size'] = rate(tw, 0)
    # the same rates as bandwidth at a 64-byte access granule
    for k in ('read_ops_per_s', 'atomic_or_ops_per_s', 'atomic_cas_ops_per_s', 'read_ops_per_s_table_size'):
        out[k.replace('ops_per_s', 'gbps_64B')] = round(out[k] * 64 / 1e9, 1)
    if sweep:
        out['read_ops_per_s_by_mb'] = {mb: rate(min(n_words, mb * (1 << 20) // 4), 0) for mb in SWEEP_MB}
        out['nontemporal_read_ops_per_s'] = rate(n_words, 3)
    del buf
    torch.cuda.empty_cache()
    return out


def cpu_baseline(cfg, batch, sample_reads, mt_reads=2_000_000):
    """The oracle (sequential C restatement, 1 thread) on the first
    `sample_reads` reads: insert then query, k-mers/s; beside it the same on
    the host's cores (at most 16, the share of one GPU) with atomic table
    updates, as goetia's storages do, on the first `mt_reads` reads."""
    from oracle import pyoracle as po
    sizes = po.primes_near_x(cfg['n_tables'], cfg['max_table'])
    n = min(sample_reads, batch.n_reads)
    off = batch.offsets[:n + 1].copy()
    seq = batch.seq[:off[-1]].copy()
    st = po.Storage(po.BYTE, sizes)
    t0 = time.perf_counter()
    hashed, _ = st.insert_reads(seq, off, cfg['K'], True, po.MODE_SKIP_N)
    st.query_reads(seq, off, cfg['K'], True, po.MODE_SKIP_N)
    dt = time.perf_counter() - t0
    del st
    threads = max(1, min(16, os.cpu_count() or 1))
    n2 = min(mt_reads, batch.n_reads)
    off2 = batch.offsets[:n2 + 1].copy()
    seq2 = batch.seq[:off2[-1]].copy()
    st2 = po.Storage(po.BYTE, sizes)
    t0 = time.perf_counter()
    hashed2 = st2.insert_reads_mt(seq2, off2, cfg['K'], True, po.MODE_SKIP_N, threads)
    st2.query_reads_mt(seq2, off2, cfg['K'], True, po.MODE_SKIP_N, threads)
    dt2 = time.perf_counter() - t0
    return {'value': hashed / dt, 'unit': 'kmers/s', 'cores': 1, 'kind': 'port',
            'sample': f'first {n} reads of the workload ({hashed} k-mers), insert + query_sequence, '
                      f'oracle/oracle.c single-threaded, {dt:.1f} s',
            'all_cores': {'value': hashed2 / dt2, 'cores': threads,
                          'sample': f'first {n2} reads ({hashed2} k-mers), insert with atomic table updates '
                                    f'+ query_sequence on {threads} threads, {dt2:.1f} s'}}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({'error': 'this build runs on one GPU; sharding across GPUs (reads per GPU, '
                                   'tables merged by OR / saturating-add all-reduce) is not implemented',
                          'n_gpus': args.gpus}))
        sys.exit(2)
    if args.workload != 'cms_20m':
        import bench_configs
        return bench_configs.run(args)
    import torch
    from goetia_amd import _native as N
    from goetia_amd import libgoetia as lg
    from goetia_amd import workload

    cfg = dict(workload.CONFIGS['cms_20m'])
    t0 = time.perf_counter()
    _, batch = workload.make('cms_20m', args.scale)
    log(f'generated {batch.n_reads} reads / {batch.n_bases} bases in {time.perf_counter() - t0:.1f}s')
    K = cfg['K']
    dev = torch.device('cuda:0')
    reads = batch.to_device(dev)
    nk_slots = int(batch.kmer_counts(K).sum())
    n_new = torch.zeros(batch.n_reads, dtype=torch.int64, device=dev)
    counts = torch.empty(nk_slots, dtype=torch.int16, device=dev)
    store = lg.ByteStorage.build(cfg['max_table'], cfg['n_tables'])
    g = lg.dBG[lg.ByteStorage, lg.CanLemireShifter].build(store, K)
    g.mode = N.MODE_SKIP_N
    stream = torch.cuda.current_stream()
    if args.copy_calibration:
        src = torch.ones(1 << 30, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    def step():
        store.reset()
        g.insert_device(reads, n_new=n_new, stream=stream)
        g.query_device(reads, counts, stream=stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    N.call('gt_profile_enable', 1)
    N.call('gt_profile_reset')
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    N.call('gt_profile_enable', 0)
    n_kmers = int((counts >= 0).sum().item())
    total_new = int(n_new.sum().item())
    log(f'{n_kmers} k-mers per step, {total_new} new, n_unique={g.n_unique()}')
    kt = {k: v for k, v in kernel_times(N, KERNELS).items() if v[0]}
    log('kernel ms/step: ' + ', '.join(f'{k} {v[1] / args.steps:.1f} ({v[0]} launches)' for k, v in kt.items()))

    value = n_kmers * args.steps / elapsed
    # the insert pass and the query pass separately (N_kmers / T per pass, from
    # the kernels' HIP-event times; the step's wall clock adds the table reset)
    ins_ms = sum(v[1] for k, v in kt.items() if k != 'query') / args.steps
    q_ms = kt['query'][1] / args.steps if 'query' in kt else None
    passes = {'insert_ms': round(ins_ms, 2), 'query_ms': round(q_ms, 2) if q_ms else None,
              'insert_kmers_per_s': round(n_kmers / (ins_ms / 1e3), 1) if ins_ms else None,
              'query_kmers_per_s': round(n_kmers / (q_ms / 1e3), 1) if q_ms else None,
              'source': 'HIP events around each kernel, per step'}
    dom = max(kt, key=lambda k: kt[k][1])
    dom_ms = kt[dom][1] / args.steps
    achieved = n_kmers * BYTES_PER_KMER[dom] / (dom_ms / 1e3) / 1e9
    launches = max(kt[dom][0] // args.steps, 1)
    traffic = MEASURED_BYTES_PER_KMER.get(dom)
    roofline = {'bound': 'hbm', 'kernel': dom, 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBPS,
                'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBPS, 4),
                'traffic': round(traffic * n_kmers / launches) if traffic else None,
                'traffic_source': 'profile: rocprofv3 FETCH_SIZE + WRITE_SIZE bytes per k-mer of this kernel on '
                                  'this workload (profiles/cms_20m_hbm_traffic.md) x k-mers per launch',
                'algorithmic_bytes_per_launch': round(BYTES_PER_KMER[dom] * n_kmers / launches),
                'launches_per_step': launches,
                'bytes_per_kmer': BYTES_PER_KMER[dom],
                'kernel_ms_per_step': {k: round(v[1] / args.steps, 2) for k, v in kt.items()},
                'kernel_gbps': {k: round(n_kmers * BYTES_PER_KMER[k] / (v[1] / args.steps / 1e3) / 1e9, 1)
                                for k, v in kt.items()}}

    # One-pass insert_and_query (dbg.hh:327-340 semantics: each k-mer's count
    # just after its own insert), same reads, reset tables.
    iq_rate = None
    if not args.no_iq:
        iq_times = []
        for _ in range(2):
            store.reset()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            g.insert_and_query_device(reads, counts, stream=stream)
            torch.cuda.synchronize()
            iq_times.append(time.perf_counter() - t0)
        iq_rate = round(n_kmers / min(iq_times), 1)
        log(f'insert_and_query: {min(iq_times) * 1e3:.1f} ms, {iq_rate / 1e9:.2f} G k-mers/s')

    chunk = None
    if not args.no_chunks:
        # Plain path, one chunk of reads at a time: insert + query of a chunk,
        # synchronous (arrival = call, ready = results on the device).
        lat = []
        from goetia_amd.reads import DeviceReads
        store.reset()
        cr = args.chunk_reads
        for r0 in range(0, batch.n_reads, cr):
            r1 = min(batch.n_reads, r0 + cr)
            b0, b1 = int(batch.offsets[r0]), int(batch.offsets[r1])
            sub = DeviceReads(reads.seq[b0:b1], reads.offsets[r0:r1 + 1] - b0, b1 - b0)
            nk = int(batch.kmer_counts(K)[r0:r1].sum())
            torch.cuda.synchronize()
            c0 = time.perf_counter()
            g.insert_device(sub, n_new=n_new[r0:r1], stream=stream)
            g.query_device(sub, counts[:nk], stream=stream)
            torch.cuda.synchronize()
            lat.append((time.perf_counter() - c0) * 1e3)
        chunk = {'chunk_reads': cr, 'n_chunks': len(lat), 'p50_ms': round(float(np.percentile(lat, 50)), 2),
                 'p99_ms': round(float(np.percentile(lat, 99)), 2)}

    cpu = None if args.no_cpu_baseline else cpu_baseline(cfg, batch, args.cpu_sample_reads)

    # the query against the measured random-read roofline: one random table
    # byte per k-mer and table
    if roofline is not None and not args.no_random_access:
        ra = random_access(N, sum(store.table_bytes(i) for i in range(store.n_tables())),
                           sweep=args.random_access_sweep)
        q_ms = roofline['kernel_ms_per_step'].get('query')
        if q_ms:
            q_rate = n_kmers * store.n_tables() / (q_ms / 1e3)
            ra['query_random_reads_per_s'] = round(q_rate, 1)
            ra['query_frac_of_random_read'] = round(q_rate / ra['read_ops_per_s_table_size'], 4)
        roofline['random_access'] = ra

    tablesizes = store.get_tablesizes()
    # The served form against the plain one-chunk-at-a-time path on the same
    # chunks (the 64-stream serving configuration, fewer chunks): aggregate
    # rate ratio and chunk latency under Poisson arrivals.
    served = None
    if args.serving_chunks:
        del reads, counts, n_new
        torch.cuda.empty_cache()
        from goetia_amd import serving
        served = serving.serving_64(1.0, n_chunks=args.serving_chunks)
        log(f"served vs plain: {served['served_over_plain']}x, Poisson p99 "
            f"{served['poisson']['served']['p99_ms']} ms served, {served['poisson']['plain']['p99_ms']} ms plain")

    out = {
        'metric': 'canonical k-mers/s insert+query (1 GPU) vs plain; p50/p99 chunk latency',
        'value': round(value, 1),
        'unit': 'kmers/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(elapsed / args.steps * 1e3, 2),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'u64',
        'data': 'synthetic (seeded generator, inputs resident in HBM)',
        'config': {'workload': 'cms_20m: 50 Mbp genome, %d x 150 bp reads, 0.5%% subs, 0.01%% N (skip-N), '
                               'K=31 canonical, ByteStorage(2^28, 4); step = reset + insert (n_new) + '
                               'query_sequence' % batch.n_reads,
                   'reads': batch.n_reads, 'kmers_per_step': n_kmers, 'K': K,
                   'storage': 'ByteStorage', 'tablesizes': tablesizes},
        'roofline': roofline,
        'passes': passes,
        'insert_and_query_kmers_per_s': iq_rate,
        'chunk_latency': chunk,
        'vs_plain': served['served_over_plain'] if served else None,
        'served': None if served is None else {
            'workload': 'serving_64: 64 streams x BitStorage(2^30, 4), K=31, 16 GiB budget, '
                        f"{served['chunks']} lognormal chunks of 100-250 bp reads",
            'served_kmers_per_s': served['value'], 'plain_kmers_per_s': served['plain_kmers_per_s'],
            'served_over_plain': served['served_over_plain'],
            'p50_ms': served['poisson']['served']['p50_ms'], 'p99_ms': served['poisson']['served']['p99_ms'],
            'plain_p50_ms': served['poisson']['plain']['p50_ms'], 'plain_p99_ms': served['poisson']['plain']['p99_ms'],
            'arrivals_per_s': served['poisson']['rate_chunks_per_s'],
            'served_equals_plain': served['served_equals_plain']},
        'cpu_baseline': cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
