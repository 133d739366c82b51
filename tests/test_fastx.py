"""FASTX parsing (host code, no GPU): the oracle restatement and the native
parser against what goetia's compiled FastxParser<DNA_SIMPLE> returned on the
same files (tests/golden/fastx_ref.json), then the native parser against the
oracle on seeded random files, at batch sizes that split records across
calls."""
import gzip
import json
import os
import random

import numpy as np
import pytest

from goetia_amd import _native as N
from goetia_amd.parsing import FastxParser, NoMoreReadsAvailable
from oracle import fastx as ofx

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')
DATA = os.path.join(GOLDEN, 'data')
REF = json.load(open(os.path.join(GOLDEN, 'fastx_ref.json')))
CASES = sorted(REF)


def product_parse(path, strict, min_length, max_reads=1 << 16, max_bases=1 << 24):
    p = FastxParser(path, strict, min_length)
    out = {'sequences': [], 'names': [], 'strict_stopped': 0}
    while True:
        try:
            batch, names = p.next_batch(max_reads, max_bases, names=True)
        except N.InvalidCharacterException:
            out['strict_stopped'] = 1
            break
        if batch.n_reads == 0:
            break
        out['sequences'] += [batch.read(i).decode() for i in range(batch.n_reads)]
        out['names'] += names
    out['n_parsed'], out['n_skipped'], out['invalid_reads'], _ = p._stats()
    p.close()
    return out


@pytest.mark.parametrize('case', CASES)
def test_oracle_vs_reference(case):
    rel, strict, min_length = case.split('|')
    got = ofx.parse(os.path.join(DATA, rel), strict == '1', int(min_length))
    assert got == REF[case]


@pytest.mark.parametrize('case', CASES)
def test_parser_vs_reference(case):
    rel, strict, min_length = case.split('|')
    got = product_parse(os.path.join(DATA, rel), int(strict), int(min_length))
    assert got == REF[case]


def _random_file(path, seed, fastq, gz):
    rng = random.Random(seed)
    lines = []
    for i in range(300):
        L = rng.choice([0, 1, 3, 20, 31, 60, 150, 301, 2000])
        s = ''.join(rng.choice('ACGTacgt' if rng.random() < 0.3 else 'ACGT') for _ in range(L))
        if L and rng.random() < 0.1:
            j = rng.randrange(L)
            s = s[:j] + rng.choice('N n.-RYX ') + s[j + 1:]
        eol = '\r\n' if rng.random() < 0.1 else '\n'
        name = f'r{i}' + (f' c{i}' if rng.random() < 0.5 else '') + ('\tx' if rng.random() < 0.1 else '')
        if fastq:
            q = 'I' * L
            if rng.random() < 0.05:
                q = q[:max(0, L - rng.randint(1, 5))]
            w = rng.choice([L or 1, 40])
            lines.append('@' + name + eol)
            for k in range(0, max(L, 1), w):
                lines.append(s[k:k + w] + eol)
            lines.append('+' + (name if rng.random() < 0.3 else '') + eol)
            for k in range(0, max(len(q), 1), w):
                lines.append(q[k:k + w] + eol)
        else:
            w = rng.choice([60, 70, 1000])
            lines.append('>' + name + eol)
            for k in range(0, L, w):
                lines.append(s[k:k + w] + eol)
            if rng.random() < 0.05:
                lines.append(eol)
    data = ''.join(lines).encode()
    if gz:
        with gzip.GzipFile(path, 'wb', mtime=0) as f:
            f.write(data)
    else:
        with open(path, 'wb') as f:
            f.write(data)


@pytest.mark.parametrize('fastq', [False, True])
@pytest.mark.parametrize('gz', [False, True])
@pytest.mark.parametrize('seed', [1, 2])
def test_parser_vs_oracle_random(tmp_path, fastq, gz, seed):
    path = str(tmp_path / (('r.fq' if fastq else 'r.fa') + ('.gz' if gz else '')))
    _random_file(path, seed * 10 + fastq * 2 + gz, fastq, gz)
    for strict, min_length in [(0, 0), (0, 25), (1, 0)]:
        exp = ofx.parse(path, bool(strict), min_length)
        got = product_parse(path, strict, min_length, max_reads=7, max_bases=2500)  # records split across calls
        assert got == exp, (strict, min_length)


def test_record_iteration_and_end():
    p = FastxParser(os.path.join(DATA, 'random-20-a.fa'))
    recs = list(p)
    ref = REF['random-20-a.fa|0|0']
    assert [r.sequence for r in recs] == ref['sequences']
    assert [r.name for r in recs] == ref['names']
    assert p.is_complete() and p.n_parsed() == ref['n_parsed']
    with pytest.raises(NoMoreReadsAvailable):
        p.next()


def test_record_larger_than_buffers_is_kept(tmp_path):
    path = str(tmp_path / 'big.fa')
    with open(path, 'w') as f:
        f.write('>a\nACGT\n>b\n' + 'A' * 5000 + '\n>c\nGG\n')
    p = FastxParser(path)
    b = p.next_batch(16, 100)
    assert [b.read(i) for i in range(b.n_reads)] == [b'ACGT']
    with pytest.raises(ValueError):
        p.next_batch(16, 100)
    b = p.next_batch(16, 10000)
    assert [b.read(i) for i in range(b.n_reads)] == [b'A' * 5000, b'GG']
    assert p.next_batch(16, 10000).n_reads == 0


def test_missing_file():
    with pytest.raises(N.GoetiaFileException):
        FastxParser('/nonexistent/reads.fq')


def test_batch_layout_feeds_dbg_calls():
    """Batches use the ReadBatch layout of the batched dBG calls."""
    b = FastxParser(os.path.join(DATA, 'test-fastq-reads.fq')).next_batch()
    assert b.offsets[0] == 0 and np.all(np.diff(b.offsets) >= 0) and b.offsets[-1] == len(b.seq)
    assert set(np.unique(b.seq).tolist()) <= set(b'ACGT')


def _clean_file(path, seed, fastq, n=400, tail=b''):
    """Well-formed input for the block-parallel reader: 4-line FASTQ whose
    quality lines often start with '@' or '+', or multi-line FASTA with empty
    lines and empty records; lowercase and invalid bytes inside lines."""
    rng = random.Random(seed)
    out = []
    for i in range(n):
        L = rng.choice([1, 5, 21, 40, 150, 151, 600])
        s = ''.join(rng.choice('ACGTacgtACGT') for _ in range(L))
        if rng.random() < 0.08:
            j = rng.randrange(L)
            s = s[:j] + rng.choice('N.-* ') + s[j + 1:]
        name = f'c{i}' + (f' d{i}' if rng.random() < 0.5 else '')
        if fastq:
            q = ''.join(rng.choice('@+I#5') for _ in range(L))
            out.append(f'@{name}\n{s}\n+{name if rng.random() < 0.3 else ""}\n{q}\n')
        else:
            if L > 150 and rng.random() < 0.05:
                s = ''
            w = rng.choice([60, 61, 80])
            body = ''.join(s[k:k + w] + '\n' + ('\n' if rng.random() < 0.03 else '') for k in range(0, len(s), w))
            out.append(f'>{name}\n{body}')
    data = ''.join(out).encode() + tail
    if rng.random() < 0.5 and not tail:
        data = data.rstrip(b'\n')            # no final newline
    with open(path, 'wb') as f:
        f.write(data)


@pytest.mark.parametrize('fastq', [False, True])
@pytest.mark.parametrize('block,threads', [(64, 1), (200, 3), (1000, 8), (1 << 20, 16)])
def test_block_parallel_reader_vs_oracle(tmp_path, monkeypatch, fastq, block, threads):
    monkeypatch.setenv('GOETIA_AMD_FASTX_BLOCK', str(block))
    monkeypatch.setenv('GOETIA_AMD_FASTX_THREADS', str(threads))
    for k, tail in enumerate([b'', b'\n@odd\nACGT\n+\nII\n' if fastq else b'\n+odd line\nACGT\n']):
        path = str(tmp_path / f'c{k}.{"fq" if fastq else "fa"}')
        _clean_file(path, block + threads + k, fastq, tail=tail)
        for min_length in (0, 30):
            exp = ofx.parse(path, False, min_length)
            got = product_parse(path, 0, min_length, max_reads=13, max_bases=5000)
            assert got == exp, (k, min_length)


def _product_qualities(path, min_length, max_reads=13, max_bases=5000):
    p = FastxParser(path, False, min_length)
    seqs, quals = [], []
    while True:
        batch, _, qs = p.next_batch(max_reads, max_bases, quality=True)
        if batch.n_reads == 0:
            break
        seqs += [batch.read(i).decode() for i in range(batch.n_reads)]
        quals += qs
    p.close()
    return seqs, [q.decode() if q else None for q in quals]


@pytest.mark.parametrize('fastq', [False, True])
@pytest.mark.parametrize('block,threads', [(1 << 20, 1), (200, 3), (1 << 20, 16)])
def test_qualities_vs_oracle(tmp_path, monkeypatch, fastq, block, threads):
    """Record::quality through the batch reader (both the record-by-record
    and the block-parallel path) equals the restated kseq reader's."""
    monkeypatch.setenv('GOETIA_AMD_FASTX_BLOCK', str(block))
    monkeypatch.setenv('GOETIA_AMD_FASTX_THREADS', str(threads))
    for k, maker in enumerate([_clean_file, lambda p, s, fq: _random_file(p, s, fq, False)]):
        path = str(tmp_path / f'q{k}.{"fq" if fastq else "fa"}')
        maker(path, 40 + k + block % 7, fastq)
        for min_length in (0, 30):
            exp = ofx.parse(path, False, min_length, with_quality=True)
            seqs, quals = _product_qualities(path, min_length)
            assert seqs == exp['sequences'] and quals == exp['qualities'], (k, min_length)
            if fastq:
                assert any(q for q in quals)
            else:
                assert not any(quals)


def test_record_iteration_has_qualities():
    recs = list(FastxParser(os.path.join(DATA, 'test-fastq-reads.fq')))
    exp = ofx.parse(os.path.join(DATA, 'test-fastq-reads.fq'), with_quality=True)
    assert [r.quality for r in recs] == exp['qualities'] and all(len(r.quality) == len(r.sequence) for r in recs)


def _batches(path, min_length, max_reads, max_bases, names):
    """Each batch's reads and the parser's (n_parsed, n_skipped) after it."""
    p = FastxParser(path, False, min_length)
    out = []
    while True:
        r = p.next_batch(max_reads, max_bases, names=names)
        batch = r[0] if names else r
        if batch.n_reads == 0:
            break
        out.append(([batch.read(i) for i in range(batch.n_reads)], p._stats()[:2]))
    p.close()
    return out


@pytest.mark.parametrize('fastq', [False, True])
@pytest.mark.parametrize('block,threads', [(200, 3), (4096, 5), (1 << 20, 16)])
def test_bulk_batches_vs_record_path(tmp_path, monkeypatch, fastq, block, threads):
    """Batches without names or qualities (copied a run of records at a time
    from each parsed part) hold the same reads, cut at the same records, with
    the same parsed / skipped counts after every batch, as batches with names
    (copied record by record), and their concatenation is the oracle's."""
    monkeypatch.setenv('GOETIA_AMD_FASTX_BLOCK', str(block))
    monkeypatch.setenv('GOETIA_AMD_FASTX_THREADS', str(threads))
    for k, maker in enumerate([_clean_file, lambda p, s, fq: _random_file(p, s, fq, False)]):
        path = str(tmp_path / f'b{k}.{"fq" if fastq else "fa"}')
        maker(path, 70 + k + block % 11, fastq)
        for min_length in (0, 30):
            exp = [s.encode() for s in ofx.parse(path, False, min_length)['sequences']]
            for max_reads, max_bases in ((1, 1 << 20), (7, 2100), (13, 5000), (1 << 16, 1 << 24), (50, 2001)):
                bulk = _batches(path, min_length, max_reads, max_bases, False)
                rec = _batches(path, min_length, max_reads, max_bases, True)
                assert bulk == rec, (k, min_length, max_reads, max_bases)
                assert [s for b, _ in bulk for s in b] == exp


def test_bulk_batches_threaded_copy(tmp_path, monkeypatch):
    """A block large enough that its runs are copied on several threads."""
    monkeypatch.setenv('GOETIA_AMD_FASTX_BLOCK', str(1 << 23))
    monkeypatch.setenv('GOETIA_AMD_FASTX_THREADS', '8')
    rng = np.random.default_rng(91)
    path = str(tmp_path / 'big.fq')
    lens = rng.integers(20, 300, 30000)
    with open(path, 'w') as f:
        for i, n in enumerate(lens):
            s = ''.join(rng.choice(list('ACGTN' if i % 97 == 0 else 'ACGT'), n))
            f.write(f'@r{i}\n{s}\n+\n{"I" * n}\n')
    exp = [s.encode() for s in ofx.parse(path, False, 40)['sequences']]
    for max_reads, max_bases in ((1 << 16, 1 << 24), (5000, 1 << 24), (1 << 16, 1 << 20)):
        bulk = _batches(path, 40, max_reads, max_bases, False)
        assert bulk == _batches(path, 40, max_reads, max_bases, True)
        assert [s for b, _ in bulk for s in b] == exp
