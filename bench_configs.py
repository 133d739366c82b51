"""The non-headline configurations of bench.py (--workload ...).

bloom_100k    1 Mbp genome, 100k x 150 bp, K=21, BitStorage(1e7, 4): insert
              with per-read n_new (tables are cache-resident).
exact_mixed   uneven reads (200k lognormal, median 8 kb, + 2M x 150 bp), K=25,
              exact set at load factor ~0.5: batched pipeline (hipGraph
              replays, autotuned tile shape) against the plain per-chunk path.
neighbors_5m  repeats + SNP bubbles, 5M x 150 bp, K=31, ByteStorage(2^28, 4):
              insert, then degrees (count >= 2) of every new k-mer and the
              decision nodes.
serving_64    64 sample streams, see goetia_amd/serving.py.
diginorm      the count-min workload's 20M x 150 bp reads (50 Mbp genome, about
              60x) through DiginormFilter, cutoff 20, K=31, ByteStorage(2^28, 4):
              reads/s with the reference's per-read decisions.
"""
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBPS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _sync():
    import torch
    torch.cuda.synchronize()


def bloom(args):
    import torch
    from goetia_amd import _native as N
    from goetia_amd import libgoetia as lg
    from goetia_amd import workload
    cfg, batch = workload.bloom_100k(args.scale)
    K = cfg['K']
    d = batch.to_device()
    store = lg.BitStorage.build(cfg['max_table'], cfg['n_tables'])
    g = lg.dBG[lg.BitStorage, lg.CanLemireShifter].build(store, K)
    n_new = torch.zeros(batch.n_reads, dtype=torch.int64, device='cuda')
    nk = int(batch.kmer_counts(K).sum())
    stream = torch.cuda.current_stream()

    def step():
        store.reset()
        g.insert_device(d, n_new=n_new, stream=stream)

    for _ in range(max(args.warmup, 1)):
        step()
    _sync()
    N.call('gt_profile_enable', 1)
    N.call('gt_profile_reset')
    steps = max(args.steps, 10)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    _sync()
    dt = (time.perf_counter() - t0) / steps
    N.call('gt_profile_enable', 0)
    # the same reads through the host-pointer entry point (PCIe included)
    g2 = lg.dBG[lg.BitStorage, lg.CanLemireShifter].build(lg.BitStorage.build(cfg['max_table'], 4), K)
    t0 = time.perf_counter()
    nn_host = g2.insert_sequences(batch)
    dt_host = time.perf_counter() - t0
    same = bool(np.array_equal(nn_host, n_new.cpu().numpy().astype(np.uint64)))
    out = {'metric': 'canonical k-mers/s insert with per-read n_new (1 GPU)', 'workload': 'bloom_100k',
           'value': round(nk / dt, 1), 'unit': 'kmers/s', 'ms_per_step': round(dt * 1e3, 3),
           'kmers': nk, 'n_unique': g.n_unique(), 'host_api_kmers_per_s': round(nk / dt_host, 1),
           'host_api_same_n_new': same, 'tables_bytes': sum(store.table_bytes(i) for i in range(4)),
           'note': 'tables (5 MB) stay in L2 / Infinity Cache; not an HBM-roofline case'}
    print(json.dumps(out), flush=True)


def exact(args):
    import torch
    from goetia_amd import libgoetia as lg
    from goetia_amd import workload
    from goetia_amd.pipeline import Pipeline, autotune
    from goetia_amd.reads import ReadBatch
    t0 = time.perf_counter()
    cfg, batch = workload.exact_mixed(args.scale)
    K = cfg['K']
    log(f'generated {batch.n_reads} reads, {batch.n_bases} bases in {time.perf_counter() - t0:.1f}s')
    nk = int(batch.kmer_counts(K).sum())
    hint = int(2_000_000_000 * args.scale) + 1
    lens = batch.lengths()

    def read_ranges(max_bases, max_reads):
        """Consecutive read ranges of at most max_bases bases / max_reads reads."""
        ranges, start, acc = [], 0, 0
        for r in range(batch.n_reads):
            L = int(lens[r])
            if r > start and (acc + L > max_bases or r - start >= max_reads):
                ranges.append((start, r))
                start, acc = r, 0
            acc += L
        ranges.append((start, batch.n_reads))
        return ranges

    def sub(a, b):
        return ReadBatch(batch.seq[batch.offsets[a]:batch.offsets[b]], batch.offsets[a:b + 1] - batch.offsets[a])

    # plain path: goetia's processor granularity, one synchronous host call per chunk
    chunk_bases = args.plain_chunk_kmers + 0
    plain_ranges = read_ranges(chunk_bases, 1 << 30)
    store = lg.SparseppSetStorage.build(capacity_hint=hint)
    g = lg.dBG[lg.SparseppSetStorage, lg.CanLemireShifter].build(store, K)
    plain_new = np.zeros(batch.n_reads, np.uint64)
    subs = [sub(a, b) for a, b in plain_ranges]
    _sync()
    t0 = time.perf_counter()
    for (a, b), sb in zip(plain_ranges, subs):
        plain_new[a:b] = g.insert_sequences(sb)
    _sync()
    t_plain = time.perf_counter() - t0
    plain_unique = g.n_unique()
    cap = g.n_occupied()
    del subs, g, store
    log(f'plain: {len(plain_ranges)} chunks, {t_plain:.2f}s, n_unique {plain_unique}')

    # batched path: autotuned tile shape, pipeline of large batches, graph replays
    store = lg.SparseppSetStorage.build(capacity_hint=hint)
    g = lg.dBG[lg.SparseppSetStorage, lg.CanLemireShifter].build(store, K)
    probe = sub(*read_ranges(64 << 20, 1 << 30)[0])
    shape = autotune(g, probe.to_device())
    log(f'autotuned tile shape (per, blocks/CU) = {shape}')
    max_bases = min(256 << 20, batch.n_bases)
    ranges = read_ranges(max_bases, 1 << 26)
    subs = [sub(a, b) for a, b in ranges]
    pipe = Pipeline(g, max(b - a for a, b in ranges), max(s.n_bases for s in subs))
    batched_new = np.zeros(batch.n_reads, np.uint64)
    outs = []
    _sync()
    t0 = time.perf_counter()
    for (a, b), sb in zip(ranges, subs):
        outs.append(((a, b), pipe.submit(sb)))
    pipe.flush()
    t_batched = time.perf_counter() - t0
    for (a, b), o in outs:
        batched_new[a:b] = o[:b - a]
    pipe.close()
    identical = bool(np.array_equal(plain_new, batched_new)) and g.n_unique() == plain_unique
    out = {'metric': 'canonical k-mers/s exact-set insert with per-read n_new, batched vs plain (1 GPU)',
           'workload': 'exact_mixed', 'value': round(nk / t_batched, 1), 'unit': 'kmers/s',
           'plain_kmers_per_s': round(nk / t_plain, 1), 'batched_over_plain': round(t_plain / t_batched, 3),
           'kmers': nk, 'reads': batch.n_reads, 'bases': batch.n_bases,
           'n_unique': plain_unique, 'slots': cap, 'load_factor': round(plain_unique / cap, 3),
           'sum_n_new_equals_n_unique': int(batched_new.sum()) == plain_unique,
           'plain_equals_batched': identical, 'tile_shape': {'per': shape[0], 'blocks_per_cu': shape[1]},
           'plain': f'{len(plain_ranges)} synchronous gt_dbg_insert calls of <= {chunk_bases} bases (host memory)',
           'batched': f'{len(ranges)} pipeline batches of <= {max_bases} bases (host memory, pinned double '
                      f'buffer, H2D overlapped, hipGraph replay)'}
    print(json.dumps(out), flush=True)


def neighbors(args):
    import torch
    from goetia_amd import _native as N
    from goetia_amd import libgoetia as lg
    from goetia_amd import workload
    t0 = time.perf_counter()
    cfg, batch = workload.neighbors_5m(args.scale)
    K = cfg['K']
    log(f'generated {batch.n_reads} reads in {time.perf_counter() - t0:.1f}s')
    d = batch.to_device()
    nk = int(batch.kmer_counts(K).sum())
    store = lg.ByteStorage.build(cfg['max_table'], cfg['n_tables'])
    g = lg.dBG[lg.ByteStorage, lg.CanLemireShifter].build(store, K)
    flags = torch.zeros(nk, dtype=torch.uint8, device='cuda')
    ind = torch.empty(nk, dtype=torch.uint8, device='cuda')
    outd = torch.empty(nk, dtype=torch.uint8, device='cuda')
    stream = torch.cuda.current_stream()
    N.call('gt_profile_enable', 1)
    N.call('gt_profile_reset')
    _sync()
    t0 = time.perf_counter()
    g.insert_device(d, new_flags=flags, stream=stream)
    _sync()
    t_ins = time.perf_counter() - t0
    t0 = time.perf_counter()
    g.degrees_device(d, ind, outd, min_count=cfg['min_count'], only=flags, stream=stream)
    _sync()
    t_deg = time.perf_counter() - t0
    N.call('gt_profile_enable', 0)
    import ctypes as C
    cnt, ms = C.c_uint64(), C.c_double()
    N.call('gt_profile_get', b'degrees', C.byref(cnt), C.byref(ms))
    n_new = int(flags.sum().item())
    sel = flags.bool()
    dec = sel & ((ind > 1) | (outd > 1)) & (ind != 255)
    n_dec_windows = int(dec.sum().item())
    # decision nodes as canonical hashes (unique)
    hashes = torch.empty(nk, dtype=torch.int64, device='cuda')
    counts = torch.empty(nk, dtype=torch.int16, device='cuda')
    g.query_device(d, counts, hashes=hashes, stream=stream)
    _sync()
    dec_nodes = int(torch.unique(hashes[dec]).numel())
    algo = n_new * 2048
    achieved = algo / (ms.value / 1e3) / 1e9
    out = {'metric': 'canonical k-mers/s insert + decision-node degrees of new k-mers (1 GPU)',
           'workload': 'neighbors_5m', 'value': round(nk / (t_ins + t_deg), 1), 'unit': 'kmers/s',
           'kmers': nk, 'new_kmers': n_new, 'n_unique': g.n_unique(),
           'decision_windows': n_dec_windows, 'decision_nodes': dec_nodes,
           'insert_ms': round(t_ins * 1e3, 2), 'degrees_ms': round(t_deg * 1e3, 2),
           'roofline': {'bound': 'hbm', 'kernel': 'degrees', 'achieved': round(achieved, 1),
                        'peak': HBM_PEAK_GBPS, 'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBPS, 4),
                        'bytes_per_new_kmer': 2048},
           'min_count': cfg['min_count']}
    print(json.dumps(out), flush=True)


def serving(args):
    from goetia_amd import serving as S
    out = S.bench(args)
    print(json.dumps(out), flush=True)


def fastx(args):
    """FASTQ file -> FastxParser -> InserterProcessor -> ByteStorage(2^28, 4):
    5M x 150 bp reads of the count-min workload's genome written as a plain
    FASTQ file, parsed alone, then inserted end to end with a large interval
    and with goetia's default one; tables checked against inserting the same
    reads from memory."""
    import ctypes as C
    import tempfile
    import torch
    from goetia_amd import _native as N
    from goetia_amd import libgoetia as lg
    from goetia_amd import workload
    n_reads = max(1000, int(5_000_000 * args.scale))
    g = workload.genome(2, 50_000_000)
    batch = workload.reads(2, g, n_reads, 150, 150, 0.005, 0.0)
    K = 31
    tmp = tempfile.mkdtemp(prefix='goetia_amd_fastx_')
    path = os.path.join(tmp, 'reads.fq')
    t0 = time.perf_counter()
    lens = batch.lengths()
    with open(path, 'wb') as f:
        chunk = []
        for i in range(batch.n_reads):
            s = batch.read(i)
            chunk.append(b'@r%d\n%s\n+\n%s\n' % (i, s, b'I' * int(lens[i])))
            if len(chunk) == 100_000:
                f.write(b''.join(chunk))
                chunk = []
        f.write(b''.join(chunk))
    size = os.path.getsize(path)
    log(f'wrote {size / 1e9:.2f} GB FASTQ in {time.perf_counter() - t0:.1f}s')
    nk = int(batch.kmer_counts(K).sum())
    # parse only (native batches into preallocated host buffers)
    lib = N.load()
    mr, mb = 1 << 16, 64 << 20
    seq = np.empty(mb, np.uint8)
    off = np.zeros(mr + 1, np.int64)
    h = N.vp()
    N.call('gt_fastx_open', path.encode(), 0, 0, C.byref(h))
    got = C.c_uint64()
    t0 = time.perf_counter()
    total = 0
    while True:
        N.check(lib.gt_fastx_next_batch(h, mr, mb, N.ptr(seq), N.ptr(off), None, 0, None, C.byref(got)))
        if got.value == 0:
            break
        total += got.value
    t_parse = time.perf_counter() - t0
    lib.gt_fastx_close(h)
    assert total == batch.n_reads
    out = {'metric': 'canonical k-mers/s, FASTQ file to count-min tables (1 GPU)', 'workload': 'fastx',
           'reads': batch.n_reads, 'kmers': nk, 'file_bytes': size,
           'parse_gbps': round(size / t_parse / 1e9, 2), 'parse_kmers_per_s': round(nk / t_parse, 1)}
    ref = lg.dBG[lg.ByteStorage, lg.CanLemireShifter].build(lg.ByteStorage.build(2 ** 28, 4), K)
    ref.insert_device(batch.to_device())
    torch.cuda.synchronize()
    ref_tables = ref.get_raw()
    del ref
    for label, interval in (('large_interval', 1 << 28), ('default_interval', 500_000)):
        gr = lg.dBG[lg.ByteStorage, lg.CanLemireShifter].build(lg.ByteStorage.build(2 ** 28, 4), K)
        proc = lg.InserterProcessor(gr, interval)
        t0 = time.perf_counter()
        n, t = proc.process(path)
        dt = time.perf_counter() - t0
        same = all(np.array_equal(a, b) for a, b in zip(gr.get_raw(), ref_tables))
        out[label] = {'interval': interval, 'kmers_per_s': round(nk / dt, 1), 'seconds': round(dt, 3),
                      'n_sequences': n, 'kmers': t, 'tables_equal_in_memory_insert': same}
        del gr, proc
    out['value'] = out['large_interval']['kmers_per_s']
    out['unit'] = 'kmers/s'
    out['note'] = ('plain FASTQ read by the block-parallel parser (host threads) into pinned double buffers of '
                   '128 MiB; the next batch is parsed while the GPU inserts the current one (PCIe included); '
                   'advance() returns at every interval with the reads it took held for the next batch (any call '
                   'on the storage inserts them first); the clock stops after the last insert has finished')
    os.remove(path)
    os.rmdir(tmp)
    print(json.dumps(out), flush=True)


def diginorm(args):
    """DiginormFilter over the count-min workload's reads (raw hashing, as
    the reference's k-mer iterator), cutoff 20: one call over all reads into
    fresh tables; the oracle (sequential C, 1 thread) on the first reads."""
    import ctypes as C

    import torch
    from goetia_amd import _native as N
    from goetia_amd import libgoetia as lg
    from goetia_amd import workload
    cfg, batch = workload.make('cms_20m', args.scale)
    K, cutoff = cfg['K'], 20
    d = batch.to_device()
    nk = int(batch.kmer_counts(K).sum())
    passed = torch.zeros(batch.n_reads, dtype=torch.uint8, device='cuda')
    store = lg.ByteStorage.build(cfg['max_table'], cfg['n_tables'])
    g = lg.dBG[lg.ByteStorage, lg.CanLemireShifter].build(store, K)
    f = lg.DiginormFilter.build(g, cutoff)
    stream = torch.cuda.current_stream()
    # warm-up on a small prefix (allocations), then a fresh run over all reads
    small = workload.make('cms_20m', min(args.scale, 0.005))[1].to_device()
    f.filter_device(small, torch.zeros(small.n_reads, dtype=torch.uint8, device='cuda'), stream=stream)
    store.reset()
    names = ['bin_hash', 'bin_count', 'bin_route', 'dn_replay', 'dn_decide', 'dn_apply']
    N.call('gt_profile_enable', 1)
    N.call('gt_profile_reset')
    _sync()
    t0 = time.perf_counter()
    f.filter_device(d, passed, stream=stream)
    _sync()
    dt = time.perf_counter() - t0
    N.call('gt_profile_enable', 0)
    phases = {}
    for n in names:
        cnt, ms = C.c_uint64(), C.c_double()
        N.call('gt_profile_get', n.encode(), C.byref(cnt), C.byref(ms))
        phases[n] = {'calls': cnt.value, 'ms': round(ms.value, 2)}
    n_passed = int(passed.sum().item())
    out = {'metric': 'reads/s through DiginormFilter (1 GPU)', 'workload': 'diginorm',
           'value': round(batch.n_reads / dt, 1), 'unit': 'reads/s', 'kmers_per_s': round(nk / dt, 1),
           'seconds': round(dt, 3), 'reads': batch.n_reads, 'kmers': nk, 'cutoff': cutoff,
           'passed': n_passed, 'n_unique': g.n_unique(), 'phases': phases,
           'chunks': phases['dn_apply']['calls'], 'guesses': phases['dn_replay']['calls']}
    if not args.no_cpu_baseline:
        from oracle import pyoracle as po
        n = min(args.cpu_sample_reads, batch.n_reads)
        st = po.Storage(po.BYTE, po.primes_near_x(cfg['n_tables'], cfg['max_table']))
        t0 = time.perf_counter()
        exp = st.diginorm_reads(batch.seq[:batch.offsets[n]], batch.offsets[:n + 1], K, cutoff)
        dtc = time.perf_counter() - t0
        out['cpu_baseline'] = {'value': round(n / dtc, 1), 'unit': 'reads/s', 'cores': 1, 'kind': 'port',
                               'sample': f'first {n} reads (below the cutoff: nearly every read is queried '
                                         f'and inserted), oracle/oracle.c single-threaded, {dtc:.1f} s'}
        out['cpu_sample_same_decisions'] = bool(np.array_equal(exp, passed[:n].cpu().numpy().astype(bool)))
    print(json.dumps(out), flush=True)


def segments(args):
    """StreamingCompactor::find_new_segments over a batch against a fixed
    graph (gt_find_new_segments: GPU view + the library's host passes).  The
    compiled goetia compactor is timed on the same reads on the host where the
    goetia checkout is (oracle/time_ref.py), never on the GPU host."""
    import ctypes as C
    from goetia_amd import _native as N
    from goetia_amd import libgoetia as lg
    from goetia_amd import workload
    t0 = time.perf_counter()
    cfg, bg, q = workload.segments_5m(args.scale)
    K = cfg['K']
    log(f'generated {bg.n_reads} + {q.n_reads} reads in {time.perf_counter() - t0:.1f}s')
    G = lg.dBG[lg.SparseppSetStorage, lg.FwdLemireShifter]
    g = G(lg.SparseppSetStorage.build(), K)
    g.insert_sequences(bg)
    nk = int(q.kmer_counts(K).sum())
    counts = np.zeros(nk, np.int16)
    hv = np.zeros(nk, np.uint64)
    seg_off = np.zeros(q.n_reads + 1, np.int64)
    dec_off = np.zeros(q.n_reads + 1, np.int64)
    seg_cap, dec_cap = 8 * q.n_reads + nk // 4, nk // 4 + 16
    segs = np.zeros(seg_cap * 7, np.uint64)
    decs = np.zeros(dec_cap * 11, np.uint64)
    n_segs, n_decs = C.c_uint64(), C.c_uint64()

    def call(n_reads):
        N.call('gt_find_new_segments', g.S.handle, K, 0, N.ptr(q.seq), N.ptr(q.offsets), n_reads, None, 0, 0, None,
               0, N.ptr(counts), N.ptr(hv), N.ptr(seg_off), N.ptr(segs), seg_cap, N.ptr(dec_off), N.ptr(decs),
               dec_cap, C.byref(n_segs), C.byref(n_decs))

    call(q.n_reads)                                        # warm-up
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        call(q.n_reads)
    dt = (time.perf_counter() - t0) / reps
    out = {'metric': 'reads/s find_new_segments against a fixed graph (1 GPU + host passes)',
           'workload': 'segments_5m', 'value': round(q.n_reads / dt, 1), 'unit': 'reads/s',
           'reads': q.n_reads, 'kmers': nk, 'graph_kmers': g.n_unique(), 'ms_per_batch': round(dt * 1e3, 2),
           'segments': int(n_segs.value), 'decision_kmers': int(n_decs.value),
           'new_kmers': int((counts == 0).sum())}
    # counts on the first 20k reads, the sample oracle/time_ref.py times the
    # compiled goetia compactor on
    n = min(q.n_reads, 20_000)
    call(n)
    out['first_reads'] = {'reads': n, 'segments': int(n_segs.value), 'decision_kmers': int(n_decs.value)}
    print(json.dumps(out), flush=True)


def run(args):
    return {'bloom_100k': bloom, 'exact_mixed': exact, 'neighbors_5m': neighbors,
            'serving_64': serving, 'fastx': fastx, 'diginorm': diginorm,
            'segments_5m': segments}[args.workload](args)
