"""The chunk driver on the GPU: InserterProcessor.advance returns exactly where
goetia's FileProcessor::advance does (after the read whose k-mers bring the
IntervalCounter to the interval) with the same running totals, and the
tables it builds equal inserting the parsed records directly and the
sequential oracle's."""
import gzip
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from goetia_amd import libgoetia as lg  # noqa: E402
from goetia_amd import workload  # noqa: E402
from goetia_amd.processors import IntervalCounter  # noqa: E402
from oracle import fastx as ofx  # noqa: E402

DATA = os.path.join(os.path.dirname(__file__), 'golden', 'data')


def expected_advances(sequences, K, interval):
    """FileProcessor::advance over the parsed records with an IntervalCounter."""
    timer = IntervalCounter(interval)
    out, n = [], 0
    for s in sequences:
        n += 1
        if timer.poll(max(0, len(s) - K + 1)):
            out.append((n, timer.total(), True))
    out.append((n, timer.total(), False))
    return out


def run(proc, parser):
    got = []
    while True:
        r = proc.advance(parser)
        got.append(r)
        if not r[2]:
            return got


def _write_fastq(path, batch):
    with gzip.open(path, 'wb') as f:
        for i in range(batch.n_reads):
            s = batch.read(i)
            f.write(b'@r%d\n%s\n+\n%s\n' % (i, s, b'I' * len(s)))


@pytest.mark.parametrize('name,interval', [('test-fastq-reads.fq', 777), ('random-20-a.fa', 1000),
                                           ('fastx/mixed.fq', 50), ('fastx/random_wrapped.fa', 5000)])
def test_advance_matches_interval_semantics(oracle, name, interval):
    path = os.path.join(DATA, name)
    K = 21
    seqs = ofx.parse(path)['sequences']
    sizes = oracle.primes_near_x(4, 100_000)
    g = lg.dBG[lg.ByteStorage, lg.CanLemireShifter].build(lg.ByteStorage(sizes), K)
    proc = lg.InserterProcessor.build(g, interval)
    got = run(proc, lg.FastxParser.build(path))
    assert got == expected_advances(seqs, K, interval)
    assert proc.n_sequences() == len(seqs) and proc.interval() == interval
    o = oracle.Storage(oracle.BYTE, sizes)
    for s in seqs:
        if len(s) >= K:
            o.insert_sequence(s, K, True)
    for a, b in zip(g.get_raw(), o.tables()):
        np.testing.assert_array_equal(a, b)
    assert (g.n_unique(), g.n_occupied()) == (o.n_unique, o.n_occupied)


def test_process_file_equals_direct_insert(tmp_path):
    """A 50k-read gzip FASTQ through process() with the default interval and
    small host batches builds the same Bloom tables as one batched insert."""
    genome = workload.genome(61, 200_000)
    batch = workload.reads(62, genome, 50_000, 80, 220, 0.01, 0.0)
    path = str(tmp_path / 'reads.fq.gz')
    _write_fastq(path, batch)
    K = 31
    g1 = lg.dBG[lg.BitStorage, lg.CanLemireShifter].build(lg.BitStorage.build(2_000_000, 4), K)
    proc = lg.InserterProcessor(g1, batch_bases=1 << 20)
    n, t = proc.process(path)
    assert n == batch.n_reads and t == int(batch.kmer_counts(K).sum())
    g2 = lg.dBG[lg.BitStorage, lg.CanLemireShifter].build(lg.BitStorage.build(2_000_000, 4), K)
    g2.insert_sequences(batch)
    for a, b in zip(g1.get_raw(), g2.get_raw()):
        np.testing.assert_array_equal(a, b)
    assert g1.n_unique() == g2.n_unique()


def test_chunked_process_yields_advances(oracle):
    """chunked_process (pythonize_processors.py:12-33) yields every advance
    that consumed reads, with the parser's skip count."""
    path = os.path.join(DATA, 'fastx', 'mixed.fq')
    K = 21
    seqs = ofx.parse(path)['sequences']
    g = lg.dBG[lg.ByteStorage, lg.CanLemireShifter].build(lg.ByteStorage(oracle.primes_near_x(4, 50_000)), K)
    proc = lg.InserterProcessor.build(g, 50)
    parser = lg.FastxParser.build(path)
    got = list(proc.chunked_process(parser))
    exp = [(n, t) for n, t, _ in expected_advances(seqs, K, 50)]
    dedup = [exp[0]] + [b for a, b in zip(exp, exp[1:]) if b[0] > a[0]]
    assert [(n, t) for n, t, _ in got] == dedup
    assert all(sk == parser.n_skipped() for _, _, sk in got[-1:])


def _solid_passes(o, s, K, min_prop, threshold):
    counts = o.insert_and_query_sequence(s, K, True)
    n_not = int(np.sum(counts.astype(np.int64) < threshold))
    ratio = np.float32(n_not) / np.float32(len(s) - K + 1)
    return not (float(ratio) >= 1.0 - float(np.float32(min_prop)))


@pytest.mark.parametrize('kind', ['diginorm', 'solid'])
@pytest.mark.parametrize('name,interval', [('test-fastq-reads.fq', 777), ('fastx/mixed.fq', 50),
                                           ('fastx/random_wrapped.fa', 5000)])
def test_filter_processor_vs_oracle(oracle, tmp_path, kind, name, interval):
    """FilterProcessor (processors.hh:355-440): the same advances as goetia's
    FileProcessor, the passed reads written as Record::write_fastx writes them,
    and the tables of filtering the reads one after another."""
    from goetia_amd.parsing import fastx_text
    path = os.path.join(DATA, name)
    K = 21
    parsed = ofx.parse(path, with_quality=True)
    sizes = oracle.primes_near_x(4, 50_000)
    G = lg.dBG[lg.ByteStorage, lg.CanLemireShifter]
    g = G.build(lg.ByteStorage(sizes), K)
    family = lg.DiginormFilter[G] if kind == 'diginorm' else lg.StreamingSolidFilter[G]
    f = family.Filter.build(g, 2) if kind == 'diginorm' else family.Filter.build(g, 0.75, 2)
    out = str(tmp_path / 'passed.fx')
    proc = family.Processor.build(f.__smartptr__(), out, interval)
    got = run(proc, lg.FastxParser.build(path))
    assert got == expected_advances(parsed['sequences'], K, interval)
    proc.close()
    o = oracle.Storage(oracle.BYTE, sizes)
    text = []
    for nm, s, q in zip(parsed['names'], parsed['sequences'], parsed['qualities']):
        if len(s) < K:
            continue
        ok = o.diginorm_filter(s, K, 2) == 1 if kind == 'diginorm' else _solid_passes(o, s, K, 0.75, 2)
        if ok:
            text.append(fastx_text(nm, s, q))
    assert open(out).read() == ''.join(text)
    assert proc.n_passed() == len(text) and proc.n_sequences() == len(parsed['sequences'])
    for a, b in zip(g.get_raw(), o.tables()):
        np.testing.assert_array_equal(a, b)
    assert len(text) > 0 or kind == 'solid'


def test_async_driver_over_samples(oracle):
    """AsyncSequenceProcessor (processors.py:133-351) driving an
    InserterProcessor over two samples: one Interval per advance with the
    running (sequence, t), then SampleFinished; the tables are the oracle's
    for both samples' reads in order."""
    from goetia_amd.messages import Interval, SampleFinished
    from goetia_amd.processors import AsyncSequenceProcessor
    K, interval = 21, 300
    paths = [os.path.join(DATA, 'test-fastq-reads.fq'), os.path.join(DATA, 'fastx', 'mixed.fq')]
    sizes = oracle.primes_near_x(4, 60_000)
    g = lg.dBG[lg.ByteStorage, lg.CanLemireShifter].build(lg.ByteStorage(sizes), K)
    proc = lg.InserterProcessor.build(g, interval)
    driver = AsyncSequenceProcessor(proc, iter([((p,), os.path.basename(p)) for p in paths]))
    got = []
    driver.add_listener('worker_q', 'test').on_messages([Interval, SampleFinished],
                                                        lambda m: got.append((type(m).__name__, m.sequence, m.t)))
    driver.run()
    seqs = [ofx.parse(p)['sequences'] for p in paths]
    timer = IntervalCounter(interval)
    exp, n = [], 0
    for sample in seqs:
        for s in sample:
            n += 1
            if timer.poll(max(0, len(s) - K + 1)):
                exp.append(('Interval', n, timer.total()))
        if not exp or exp[-1][1] != n:
            exp.append(('Interval', n, timer.total()))
        exp.append(('SampleFinished', n, timer.total()))
    assert got == exp
    o = oracle.Storage(oracle.BYTE, sizes)
    for sample in seqs:
        for s in sample:
            if len(s) >= K:
                o.insert_sequence(s, K, True)
    for a, b in zip(g.get_raw(), o.tables()):
        np.testing.assert_array_equal(a, b)


def _oracle_after(oracle, sizes, K, seqs):
    o = oracle.Storage(oracle.BYTE, sizes)
    for s in seqs:
        if len(s) >= K:
            o.insert_sequence(s, K, True)
    return o


def _assert_storage(g, o):
    for a, b in zip(g.get_raw(), o.tables()):
        np.testing.assert_array_equal(a, b)
    assert (g.n_unique(), g.n_occupied()) == (o.n_unique, o.n_occupied)


@pytest.mark.parametrize('batch_bases', [0, 4096])
def test_storage_sees_taken_reads_at_every_interval(oracle, batch_bases):
    """advance() returns with the reads it took held (queued with the next
    batch); every call on the storage queues them first, so counters, tables
    and queries right after each return equal inserting exactly the reads
    taken so far, as goetia's advance leaves them."""
    path = os.path.join(DATA, 'test-fastq-reads.fq')
    K = 21
    seqs = ofx.parse(path)['sequences']
    sizes = oracle.primes_near_x(3, 60_000)
    g = lg.dBG[lg.ByteStorage, lg.CanLemireShifter].build(lg.ByteStorage(sizes), K)
    proc = lg.InserterProcessor(g, 300, batch_bases=batch_bases)
    parser = lg.FastxParser.build(path)
    o = oracle.Storage(oracle.BYTE, sizes)
    done, steps = 0, 0
    while True:
        n, _, rem = proc.advance(parser)
        for s in seqs[done:n]:
            if len(s) >= K:
                o.insert_sequence(s, K, True)
        done = n
        if steps % 3 == 0:
            _assert_storage(g, o)
        else:
            probe = seqs[max(0, n - 1)]
            if len(probe) >= K:
                np.testing.assert_array_equal(np.array(g.query_sequence(probe)), o.query_sequence(probe, K, True)[0])
        steps += 1
        if not rem:
            break
    assert steps > 8
    _assert_storage(g, o)


def test_two_inserters_on_one_storage(oracle):
    """Two inserters advancing in turn over one storage: each advance's reads
    go in after those of every earlier advance, whichever inserter made it."""
    K = 21
    paths = [os.path.join(DATA, 'test-fastq-reads.fq'), os.path.join(DATA, 'fastx', 'random_wrapped.fa')]
    seqs = [ofx.parse(p)['sequences'] for p in paths]
    sizes = oracle.primes_near_x(3, 60_000)
    g = lg.dBG[lg.ByteStorage, lg.CanLemireShifter].build(lg.ByteStorage(sizes), K)
    procs = [lg.InserterProcessor(g, 700), lg.InserterProcessor(g, 1300, batch_bases=8192)]
    parsers = [lg.FastxParser.build(p) for p in paths]
    o = oracle.Storage(oracle.BYTE, sizes)
    done, live = [0, 0], [True, True]
    turn = 0
    while any(live):
        i = turn % 2
        turn += 1
        if not live[i]:
            continue
        n, _, rem = procs[i].advance(parsers[i])
        for s in seqs[i][done[i]:n]:
            if len(s) >= K:
                o.insert_sequence(s, K, True)
        done[i] = n
        live[i] = rem
        if turn % 4 == 0:
            assert (g.n_unique(), g.n_occupied()) == (o.n_unique, o.n_occupied)
    _assert_storage(g, o)


def test_freed_inserter_leaves_its_held_reads(oracle):
    """Reads an inserter holds when it is freed go into the storage."""
    path = os.path.join(DATA, 'test-fastq-reads.fq')
    K = 21
    seqs = ofx.parse(path)['sequences']
    sizes = oracle.primes_near_x(3, 60_000)
    g = lg.dBG[lg.ByteStorage, lg.CanLemireShifter].build(lg.ByteStorage(sizes), K)
    proc = lg.InserterProcessor(g, 2000)
    parser = lg.FastxParser.build(path)
    n, _, rem = proc.advance(parser)
    assert rem and 0 < n < len(seqs)
    del proc
    _assert_storage(g, _oracle_after(oracle, sizes, K, seqs[:n]))
