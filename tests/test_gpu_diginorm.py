"""DiginormFilter on the GPU: pass flags, tables and counters equal what goetia's
compiled DiginormFilter left on the same reads (tests/golden/diginorm_ref.npz)
and what the sequential oracle gives on larger seeded inputs, for every chunk
size and guess cap (chunks of one tile and a cap of two guesses force the
chunk cuts and restarts), one batch or several, host or device reads."""
import hashlib
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from conftest import GOLDEN_DIR  # noqa: E402
from goetia_amd import _native as N  # noqa: E402
from goetia_amd import libgoetia as lg  # noqa: E402
from goetia_amd import workload  # noqa: E402
from goetia_amd.reads import ReadBatch  # noqa: E402

GOLD = np.load(os.path.join(GOLDEN_DIR, 'diginorm_ref.npz'), allow_pickle=False)
CASES = sorted(k[:-len('__params')] for k in GOLD.files if k.endswith('__params'))
ENVS = {'default': {}, 'tiny_chunks': {'GOETIA_AMD_DIGINORM_CHUNK': '2048'},
        'two_guesses': {'GOETIA_AMD_DIGINORM_GUESSES': '2'}}


def graph(sizes, K, canonical=True):
    sh = lg.CanLemireShifter if canonical else lg.FwdLemireShifter
    return lg.dBG[lg.ByteStorage, sh].build(lg.ByteStorage(list(sizes)), K)


@pytest.mark.parametrize('env', sorted(ENVS))
@pytest.mark.parametrize('name', CASES)
def test_vs_reference_golden(monkeypatch, name, env):
    for k, v in ENVS[env].items():
        monkeypatch.setenv(k, v)
    K, can, mt, nt, cutoff = (int(v) for v in GOLD[f'{name}__params'])
    inp = bytes(GOLD[f'{name}__input']).decode()
    batch = ReadBatch(GOLD[f'in_{inp}_seq'], GOLD[f'in_{inp}_off'])
    g = graph(GOLD[f'{name}__tablesizes'].tolist(), K, bool(can))
    f = lg.DiginormFilter.build(g, cutoff)
    got = f.filter_sequences(batch)
    np.testing.assert_array_equal(got, GOLD[f'{name}__passed'].astype(bool))
    assert [g.n_unique(), g.n_occupied()] == GOLD[f'{name}__counters'].tolist()
    for i, t in enumerate(g.get_raw()):
        assert hashlib.sha256(t.tobytes()).digest() == GOLD[f'{name}__table{i}_sha'].tobytes()


def _coverage_reads(seed, n, genome_len=100_000):
    genome = workload.genome(seed, genome_len)
    return workload.reads(seed + 1, genome, n, 60, 150, 0.01, 0.0)


@pytest.mark.parametrize('cutoff', [5, 20])
def test_vs_oracle_high_coverage(oracle, cutoff):
    """60k reads over 100 kbp (about 60x): most reads are dropped once the
    coverage passes the cutoff; two batches, the second starting from the
    tables the first left."""
    batch = _coverage_reads(71 + cutoff, 60_000)
    K = 21
    sizes = oracle.primes_near_x(4, 400_000)
    o = oracle.Storage(oracle.BYTE, sizes)
    exp = o.diginorm_reads(batch.seq, batch.offsets, K, cutoff)
    g = graph(sizes, K)
    f = lg.DiginormFilter.build(g, cutoff)
    h = batch.n_reads // 3
    first = ReadBatch(batch.seq[:batch.offsets[h]], batch.offsets[:h + 1])
    rest = ReadBatch(batch.seq[batch.offsets[h]:], batch.offsets[h:] - batch.offsets[h])
    got = np.concatenate([f.filter_sequences(first), f.filter_sequences(rest)])
    np.testing.assert_array_equal(got, exp)
    for a, b in zip(g.get_raw(), o.tables()):
        np.testing.assert_array_equal(a, b)
    assert (g.n_unique(), g.n_occupied()) == (o.n_unique, o.n_occupied)
    assert 0.05 < exp.mean() < 0.95


@pytest.mark.parametrize('K', [5, 64, 133, 255])
def test_every_k_vs_oracle(oracle, K):
    """Diginorm through the ordered binned passes at K from 5 to 255 (window
    starts stitched from 4-byte segments for every K) on high-coverage
    reads: pass flags, tables and counters equal the sequential oracle's."""
    genome = workload.genome(90 + K, 6_000)
    gen = workload.reads(91 + K, genome, 3000, K + 20, K + 400, 0.005, 0.0)
    batch = ReadBatch.from_strings([gen.read(r) for r in range(gen.n_reads)])
    sizes = oracle.primes_near_x(4, 100_003)
    o = oracle.Storage(oracle.BYTE, sizes)
    exp = o.diginorm_reads(batch.seq, batch.offsets, K, 5)
    g = graph(sizes, K)
    got = lg.DiginormFilter.build(g, 5).filter_sequences(batch)
    np.testing.assert_array_equal(got, exp)
    assert (g.n_unique(), g.n_occupied()) == (o.n_unique, o.n_occupied)
    for a, b in zip(g.get_raw(), o.tables()):
        np.testing.assert_array_equal(a, b)


def test_skip_n_mode_vs_oracle(oracle):
    """Skip-N input policy: windows holding a non-ACGT byte are neither
    counted nor inserted (min_req still counts them)."""
    genome = workload.genome(97, 30_000)
    batch = workload.reads(98, genome, 8_000, 30, 150, 0.01, 0.02)
    K = 21
    sizes = oracle.primes_near_x(4, 60_000)
    o = oracle.Storage(oracle.BYTE, sizes)
    exp = o.diginorm_reads(batch.seq, batch.offsets, K, 3, True, oracle.MODE_SKIP_N)
    g = lg.dBG[lg.ByteStorage, lg.CanLemireShifter](lg.ByteStorage(sizes), K, mode=N.MODE_SKIP_N)
    got = lg.DiginormFilter.build(g, 3).filter_sequences(batch)
    np.testing.assert_array_equal(got, exp)
    for a, b in zip(g.get_raw(), o.tables()):
        np.testing.assert_array_equal(a, b)
    assert (g.n_unique(), g.n_occupied()) == (o.n_unique, o.n_occupied)
    assert 0.05 < exp.mean() < 0.95


def test_device_reads_and_one_at_a_time(oracle):
    import torch
    batch = _coverage_reads(90, 5_000, 20_000)
    K = 25
    sizes = oracle.primes_near_x(3, 50_000)
    o = oracle.Storage(oracle.BYTE, sizes)
    exp = o.diginorm_reads(batch.seq, batch.offsets, K, 4)
    g = graph(sizes, K)
    d = batch.to_device()
    passed = torch.zeros(batch.n_reads, dtype=torch.uint8, device='cuda')
    lg.DiginormFilter.build(g, 4).filter_device(d, passed)
    np.testing.assert_array_equal(passed.cpu().numpy().astype(bool), exp)
    for a, b in zip(g.get_raw(), o.tables()):
        np.testing.assert_array_equal(a, b)
    # the per-read call gives the same decisions
    g2 = graph(sizes, K)
    f2 = lg.DiginormFilter.build(g2, 4)
    one = [f2.filter_sequence(batch.read(i))[0] for i in range(300)]
    np.testing.assert_array_equal(one, exp[:300])


def test_median_helpers(oracle):
    batch = _coverage_reads(95, 2_000, 10_000)
    K = 21
    sizes = oracle.primes_near_x(4, 30_000)
    g = graph(sizes, K)
    g.insert_sequences(batch)
    o = oracle.Storage(oracle.BYTE, sizes)
    o.insert_reads(batch.seq, batch.offsets, K)
    for i in range(0, 2000, 97):
        s = batch.read(i).decode()
        counts, hashes = o.query_sequence(s, K)
        min_req = int(0.5 + float(np.float32(len(counts)) / np.float32(2)))
        assert lg.DiginormFilter.median_count_at_least(s, 8, g) == (int(np.sum(counts >= 8)) >= min_req)
        med = int(np.sort(counts)[len(counts) // 2])
        ok, hs = lg.DiginormFilter.median_in_range(s, 3, 12, g)
        assert ok == (3 < med <= 12)
        assert [int(getattr(h, 'value', h)) for h in hs] == hashes.tolist()


def test_unsupported_storages():
    g = lg.dBG[lg.BitStorage, lg.CanLemireShifter].build(lg.BitStorage.build(100_000, 4), 21)
    with pytest.raises(N.UnsupportedError):
        lg.DiginormFilter.build(g, 5).filter_sequences(['ACGT' * 10])
    g5 = graph([100_003, 100_019, 100_043, 100_049, 100_057], 21)
    with pytest.raises(N.UnsupportedError):
        lg.DiginormFilter.build(g5, 5).filter_sequences(['ACGT' * 10])
