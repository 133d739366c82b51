"""GPU parity: the HIP path (through the C ABI) against the oracle and the
golden vectors of the compiled reference.  Integer work, so every
comparison is bit-exact."""
import hashlib

import numpy as np
import pytest

from conftest import golden_input, random_dna, revcomp

pytestmark = pytest.mark.gpu

from goetia_amd import libgoetia as lg  # noqa: E402
from goetia_amd import _native as N  # noqa: E402
from goetia_amd.hashing import hash_batch  # noqa: E402
from goetia_amd.reads import ReadBatch  # noqa: E402


def _storage(meta):
    if meta['kind'] == 3:
        return lg.SparseppSetStorage.build()
    cls = lg.BitStorage if meta['kind'] == 1 else lg.ByteStorage
    return cls.build(meta['max_table'], meta['n_tables'])


def _graph(meta):
    sh = lg.CanLemireShifter if meta['canonical'] else lg.FwdLemireShifter
    store = _storage(meta)
    return lg.dBG[type(store), sh].build(store, sh(meta['K'])), store


def _oracle_storage(oracle, meta):
    if meta['kind'] == 3:
        return oracle.Storage(oracle.EXACT)
    return oracle.Storage(meta['kind'], oracle.primes_near_x(meta['n_tables'], meta['max_table']))


def test_device_visible():
    import torch
    assert torch.cuda.is_available()
    N.call('gt_device_init', 0)


# ------------------------------------------------------------------ hashing

@pytest.mark.parametrize('K', [1, 2, 21, 27, 31, 32, 33, 63, 64, 65, 101, 128, 200])
def test_hash_all_bytes_vs_reference(golden, K):
    s = golden[f'hash_bytes_K{K}_seq'].tobytes()
    fw, rc, valid = hash_batch(ReadBatch.from_strings([s]), K)
    np.testing.assert_array_equal(fw, golden[f'hash_bytes_K{K}_fw'])
    np.testing.assert_array_equal(rc, golden[f'hash_bytes_K{K}_rc'])
    assert valid.all()


def test_known_vector_on_device():
    fw, rc, _ = hash_batch(ReadBatch.from_strings(['TCACCTGTGTTGTGCTACTTGCGGCGC']), 27)
    assert (int(fw[0]), int(rc[0])) == (13194817695400542713, 4324216031038051805)


@pytest.mark.parametrize('K', [5, 21, 31, 64, 65, 127])
@pytest.mark.parametrize('mode', [N.MODE_RAW, N.MODE_SKIP_N])
def test_hash_uneven_reads_vs_oracle(oracle, K, mode):
    """Empty, shorter-than-K, tile-spanning (10-40 kb) and N-holding reads."""
    rng = np.random.default_rng(K * 7 + mode)
    reads = []
    for i in range(300):
        L = int(rng.choice([0, 1, K - 1, K, K + 1, 150, 700, 5000, 12000, 40000],
                           p=[.05, .05, .05, .1, .1, .3, .15, .1, .07, .03]))
        r = list(random_dna(rng, max(L, 0)))
        for j in range(len(r)):
            u = rng.random()
            if u < 0.003:
                r[j] = 'N'
            elif u < 0.004:
                r[j] = 'a'
        reads.append(''.join(r))
    batch = ReadBatch.from_strings(reads)
    fw, rc, valid = hash_batch(batch, K, mode)
    exp_fw, exp_rc, exp_valid = [], [], []
    for r in reads:
        f, c = oracle.hash_read(r, K)
        exp_fw.append(f)
        exp_rc.append(c)
        n = max(0, len(r) - K + 1)
        v = np.ones(n, bool)
        if mode == N.MODE_SKIP_N and n:
            bad = ~np.isin(np.frombuffer(r.encode(), np.uint8), np.frombuffer(b'ACGT', np.uint8))
            c = np.concatenate([[0], np.cumsum(bad)])
            v = (c[K:K + n] - c[:n]) == 0
        exp_valid.append(v)
    np.testing.assert_array_equal(fw, np.concatenate(exp_fw))
    np.testing.assert_array_equal(rc, np.concatenate(exp_rc))
    np.testing.assert_array_equal(valid, np.concatenate(exp_valid))


def test_hash_many_tiny_reads():
    """More reads per tile than the LDS offset cache holds (global search path)."""
    rng = np.random.default_rng(2)
    reads = [random_dna(rng, int(rng.integers(0, 6))) for _ in range(20000)]
    batch = ReadBatch.from_strings(reads)
    fw, rc, _ = hash_batch(batch, 3)
    from oracle import pyoracle as po
    exp = np.concatenate([po.hash_read(r, 3)[0] for r in reads])
    np.testing.assert_array_equal(fw, exp)


# ------------------------------------------------------------------ storages

@pytest.mark.parametrize('binned', ['0', '1'])
def test_storages_vs_reference(golden, golden_meta, binned, monkeypatch):
    """Per-read n_new, byte-identical tables, counters and query counts for
    every golden run (Bloom, count-min, exact set; canonical and forward;
    raw and skip-N), through the atomic and the binned insert."""
    monkeypatch.setenv('GOETIA_AMD_BINNED', binned)
    for name, meta in golden_meta['runs'].items():
        if 'byteiq' in name:
            continue
        g, store = _graph(meta)
        g.mode = meta['mode']
        reads = golden_input(name, golden)
        n_new = g.insert_sequences(reads)
        np.testing.assert_array_equal(n_new, golden[f'{name}__n_new'], err_msg=name)
        assert g.n_unique() == int(golden[f'{name}__counters'][0]), name
        assert g.n_occupied() == int(golden[f'{name}__counters'][1]), name
        if meta['kind'] != 3:
            for i, t in enumerate(g.get_raw()):
                assert hashlib.sha256(t.tobytes()).digest() == golden[f'{name}__table{i}_sha'].tobytes(), (name, i)
                if f'{name}__table{i}' in golden:
                    np.testing.assert_array_equal(t, golden[f'{name}__table{i}'])
        q = g.query_sequences(reads)
        assert hashlib.sha256(q.tobytes()).digest() == golden[f'{name}__query_sha'].tobytes(), name
        if f'{name}__query' in golden:
            np.testing.assert_array_equal(q, golden[f'{name}__query'])


@pytest.mark.parametrize('binned', ['0', '1'])
@pytest.mark.parametrize('kind', [1, 2, 3])
@pytest.mark.parametrize('sub_batch', [0, 4096, 3 * 4096])
def test_new_flags_and_sub_batches_vs_oracle(oracle, kind, sub_batch, binned, monkeypatch):
    """Per-k-mer new flags match the sequential oracle, also when the batch
    is split into several first-touch sub-batches and the exact set grows
    (Bloom and count-min through both insert paths)."""
    if binned == '1' and kind == 3:
        pytest.skip('the exact set has one insert path')
    monkeypatch.setenv('GOETIA_AMD_BINNED', binned)
    if sub_batch:
        monkeypatch.setenv('GOETIA_AMD_SUB_BATCH', str(sub_batch))
    rng = np.random.default_rng(kind + sub_batch)
    genome = random_dna(rng, 3000)
    reads = []
    for _ in range(400):
        L = int(rng.integers(30, 120))
        s = int(rng.integers(0, len(genome) - L))
        r = genome[s:s + L]
        reads.append(revcomp(r) if rng.random() < 0.5 else r)
    meta = {'kind': kind, 'max_table': 4000, 'n_tables': 3, 'canonical': 1, 'K': 17}
    g, store = _graph(meta)
    n_new, flags = g.insert_sequences(reads, new_flags=True)
    o = _oracle_storage(oracle, meta)
    exp_nn, exp_flags = [], []
    for r in reads:
        _, nn, fl = o.insert_sequence(r, 17, True, flags=True)
        exp_nn.append(nn)
        exp_flags.append(fl)
    np.testing.assert_array_equal(n_new, exp_nn)
    np.testing.assert_array_equal(flags, np.concatenate(exp_flags).astype(bool))
    assert g.n_unique() == o.n_unique
    if kind == 3:
        np.testing.assert_array_equal(store.keys(), o.exact_keys())
        assert g.n_occupied() == o.n_occupied
    else:
        assert g.n_occupied() == o.n_occupied
        for a, b in zip(g.get_raw(), o.tables()):
            np.testing.assert_array_equal(a, b)


def test_byte_insert_and_query_vs_reference(golden, golden_meta):
    for name in ('rand20_byteiq_can21', 'synth_byteiq_can31'):
        meta = golden_meta['runs'][name]
        g, store = _graph(meta)
        reads = golden_input(name, golden)[:meta['n_reads']]
        got = g.insert_and_query_sequences(reads)
        np.testing.assert_array_equal(got, golden[f'{name}__counts'])
        assert [g.n_unique(), g.n_occupied()] == [int(v) for v in golden[f'{name}__counters']]


def test_count_saturates_at_255():
    g = lg.dBG[lg.ByteStorage, lg.CanLemireShifter].build(lg.ByteStorage.build(10007, 4), 21)
    read = 'ACGTTGCAAGGCTTACCGATAGCATTAC'
    g.insert_sequences([read] * 300)
    assert all(c == 255 for c in g.query_sequence(read))
    assert g.n_unique() == len(set(h.value for h in g.hashes(read)))
    for t in g.get_raw():
        assert t.max() == 255


@pytest.mark.parametrize('min_count', [1, 2])
def test_degrees_vs_oracle(oracle, golden, golden_meta, min_count):
    for name in ('synth_bit_can21', 'synth_byte_can31', 'synth_set_can25', 'synth_byte_fwd31'):
        meta = golden_meta['runs'][name]
        g, _ = _graph(meta)
        reads = golden_input(name, golden)
        g.insert_sequences(reads)
        o = _oracle_storage(oracle, meta)
        seq = np.frombuffer(''.join(reads).encode(), np.uint8)
        off = np.concatenate([[0], np.cumsum([len(r) for r in reads])]).astype(np.int64)
        o.insert_reads(seq, off, meta['K'], bool(meta['canonical']), 0)
        ind, outd = g.degrees(reads[:300], min_count=min_count)
        exp = [o.degrees_sequence(r, meta['K'], bool(meta['canonical']), min_count=min_count) for r in reads[:300]]
        np.testing.assert_array_equal(ind, np.concatenate([e[0] for e in exp]), err_msg=name)
        np.testing.assert_array_equal(outd, np.concatenate([e[1] for e in exp]), err_msg=name)


def test_decision_kmers_vs_reference(golden, golden_meta):
    for name in ('synth_bit_can21', 'synth_byte_can31', 'synth_set_can25', 'synth_byte_fwd31'):
        meta = golden_meta['runs'][name]
        g, _ = _graph(meta)
        reads = golden_input(name, golden)
        g.insert_sequences(reads)
        got_read, got_pos, got_hash = [], [], []
        for ri, r in enumerate(reads[:200]):
            pos, hs, _ = g.find_decision_kmers(r)
            got_read += [ri] * len(pos)
            got_pos += pos
            got_hash += [h.value for h in hs]
        np.testing.assert_array_equal(got_read, golden[f'{name}__decision_read'])
        np.testing.assert_array_equal(got_pos, golden[f'{name}__decision_pos'])
        np.testing.assert_array_equal(np.array(got_hash, np.uint64), golden[f'{name}__decision_hash'])


def test_extensions_vs_reference(golden, golden_meta):
    for name in ('synth_bit_can21', 'synth_set_can25', 'synth_byte_fwd31'):
        meta = golden_meta['runs'][name]
        g, _ = _graph(meta)
        g.insert_sequences(golden_input(name, golden))
        K = meta['K']
        kmers = golden[f'{name}__ext_kmers'].tobytes().decode()
        for i in range(0, len(kmers) // K, 5):
            km = kmers[i * K:(i + 1) * K]
            g.set_cursor(km)
            assert [e.value for e in g.left_extensions()] == [int(v) for v in golden[f'{name}__ext_left'][i]]
            assert [e.value for e in g.right_extensions()] == [int(v) for v in golden[f'{name}__ext_right'][i]]
            assert g.left_degree(km) == golden[f'{name}__ext_in'][i]
            assert g.right_degree(km) == golden[f'{name}__ext_out'][i]


# ------------------------------------------------------------------ per-hash API

@pytest.mark.parametrize('kind', [1, 2, 3])
def test_hash_api_vs_oracle(oracle, kind):
    rng = np.random.default_rng(kind)
    hs = rng.integers(0, 2 ** 63, 5000, dtype=np.uint64)
    hs = np.concatenate([hs, hs[:1000], np.array([0, 2 ** 64 - 1, 2 ** 64 - 1], np.uint64)])
    meta = {'kind': kind, 'max_table': 3000, 'n_tables': 4}
    store = _storage(meta)
    o = _oracle_storage(oracle, meta)
    got = store.insert_hashes(hs)
    exp = [o.insert(int(h)) for h in hs]
    np.testing.assert_array_equal(got, exp)
    probe = np.concatenate([hs[:2000], rng.integers(0, 2 ** 63, 2000, dtype=np.uint64)])
    np.testing.assert_array_equal(store.query_hashes(probe), [o.query(int(h)) for h in probe])
    assert store.n_unique_kmers() == o.n_unique
    iq = store.insert_and_query_hashes(probe[:500])
    np.testing.assert_array_equal(iq, [o.insert_and_query(int(h)) for h in probe[:500]])


def test_device_api_matches_host_api():
    import torch
    from goetia_amd import workload
    gnm = workload.genome(21, 100000)
    batch = workload.reads(22, gnm, 3000, 100, 150, 0.01, 0.001)
    K = 25
    g1 = lg.dBG[lg.ByteStorage, lg.CanLemireShifter].build(lg.ByteStorage.build(200000, 4), K)
    g2 = lg.dBG[lg.ByteStorage, lg.CanLemireShifter].build(lg.ByteStorage.build(200000, 4), K)
    g1.mode = g2.mode = N.MODE_SKIP_N
    nn_host = g1.insert_sequences(batch)
    d = batch.to_device()
    nn_dev = torch.zeros(batch.n_reads, dtype=torch.int64, device='cuda')
    g2.insert_device(d, n_new=nn_dev)
    nk = int(batch.kmer_counts(K).sum())
    counts = torch.zeros(nk, dtype=torch.int16, device='cuda')
    g2.query_device(d, counts)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(nn_dev.cpu().numpy().astype(np.uint64), nn_host)
    np.testing.assert_array_equal(counts.cpu().numpy(), g1.query_sequences(batch))
    for a, b in zip(g1.get_raw(), g2.get_raw()):
        np.testing.assert_array_equal(a, b)


def test_reset_clone_update_from_and_oxli(tmp_path):
    K = 21
    rng = np.random.default_rng(4)
    reads = [random_dna(rng, 100) for _ in range(200)]
    g = lg.dBG[lg.BitStorage, lg.CanLemireShifter].build(lg.BitStorage.build(20000, 4), K)
    g.insert_sequences(reads[:100])
    h = lg.dBG[lg.BitStorage, lg.CanLemireShifter].build(lg.BitStorage.build(20000, 4), K)
    h.insert_sequences(reads[100:])
    both = lg.dBG[lg.BitStorage, lg.CanLemireShifter].build(lg.BitStorage.build(20000, 4), K)
    both.insert_sequences(reads)
    occ_before = g.n_occupied()
    g.S.update_from(h.S)
    for a, b in zip(g.get_raw(), both.get_raw()):
        np.testing.assert_array_equal(a, b)
    assert g.n_occupied() == both.n_occupied() >= occ_before
    c = g.clone()
    assert c.n_unique() == 0 and all(t.max() == 0 for t in c.get_raw())
    path = str(tmp_path / 'g.oxli')
    g.save(path)
    c.load(path)
    for a, b in zip(g.get_raw(), c.get_raw()):
        np.testing.assert_array_equal(a, b)
    assert c.n_occupied() == g.n_occupied()
    g.reset()
    assert all(t.max() == 0 for t in g.get_raw())


@pytest.mark.parametrize('kind', [1, 2])
@pytest.mark.parametrize('sub_batch', [0, 8192 * 5])
def test_binned_insert_vs_oracle(oracle, kind, sub_batch, monkeypatch):
    """The binned insert on tables spanning more than 512 buckets (two-level
    routing), with hot repeated k-mers (ranks within one 64-event chunk and
    counts past 255), N bases and several sub-batches: tables, per-read n_new,
    flags and counters equal the sequential oracle's."""
    from goetia_amd import workload
    monkeypatch.setenv('GOETIA_AMD_BINNED', '1')
    if sub_batch:
        monkeypatch.setenv('GOETIA_AMD_SUB_BATCH', str(sub_batch))
    genome = workload.genome(31 + kind, 40_000)
    gen = workload.reads(32 + kind, genome, 3000, 40, 300, 0.01, 0.002)
    reads = [gen.read(r) for r in range(gen.n_reads)] + [b'ACGTTGCA' * 40] * 60 + [b'A' * 200] * 40
    batch = ReadBatch.from_strings(reads)
    sizes = oracle.primes_near_x(4, 3_000_000 if kind == 2 else 9_000_000)
    store = (lg.ByteStorage if kind == 2 else lg.BitStorage)(sizes)
    g = lg.dBG[type(store), lg.CanLemireShifter].build(store, 25)
    g.mode = N.MODE_SKIP_N
    n_new, flags = g.insert_sequences(batch, new_flags=True)
    o = oracle.Storage(kind, sizes)
    o_new, o_flags = [], []
    for r in reads:
        _, nn, fl = o.insert_sequence(r, 25, True, oracle.MODE_SKIP_N, flags=True)
        o_new.append(nn)
        o_flags.append(fl)
    np.testing.assert_array_equal(n_new, o_new)
    np.testing.assert_array_equal(flags, np.concatenate(o_flags).astype(bool))
    assert (g.n_unique(), g.n_occupied()) == (o.n_unique, o.n_occupied)
    for a, b in zip(g.get_raw(), o.tables()):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize('K', [1, 3, 4, 5, 8, 31, 33, 100, 131, 132, 133, 200, 255])
def test_binned_insert_every_k_vs_oracle(oracle, K, monkeypatch):
    """The binned plain insert for K from 1 to 255: the first pass stitches
    window starts from 4-byte segments up to K = 132 and rolls them beyond;
    tables, per-read n_new, flags and counters equal the sequential oracle's
    on reads of every length around K, with N bases."""
    monkeypatch.setenv('GOETIA_AMD_BINNED', '1')
    rng = np.random.default_rng(1000 + K)
    reads = []
    for _ in range(400):
        L = int(rng.choice([0, K - 1, K, K + 1, K + 7, 150, 600, 3000]))
        r = bytearray(random_dna(rng, max(L, 0)).encode())
        for j in np.nonzero(rng.random(len(r)) < 0.002)[0]:
            r[j] = ord('N')
        reads.append(bytes(r))
    reads += [b'ACGT' * 100] * 20
    for kind in (1, 2):
        sizes = oracle.primes_near_x(4, 200_000 if kind == 2 else 600_000)
        store = (lg.ByteStorage if kind == 2 else lg.BitStorage)(sizes)
        g = lg.dBG[type(store), lg.CanLemireShifter].build(store, K)
        g.mode = N.MODE_SKIP_N
        n_new, flags = g.insert_sequences(ReadBatch.from_strings(reads), new_flags=True)
        o = oracle.Storage(kind, sizes)
        o_new, o_flags = [], []
        for r in reads:
            _, nn, fl = o.insert_sequence(r, K, True, oracle.MODE_SKIP_N, flags=True)
            o_new.append(nn)
            o_flags.append(fl)
        np.testing.assert_array_equal(n_new, o_new)
        np.testing.assert_array_equal(flags, np.concatenate(o_flags).astype(bool))
        assert (g.n_unique(), g.n_occupied()) == (o.n_unique, o.n_occupied)
        for a, b in zip(g.get_raw(), o.tables()):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize('n_tables', [1, 2, 3])
@pytest.mark.parametrize('K', [5, 31])
def test_binned_insert_fewer_tables_vs_oracle(oracle, n_tables, K, monkeypatch):
    """The binned plain insert on 1 to 3 tables (a first-pass tile then holds
    2048 to 6144 events) over several sub-batches, with hot repeats: tables,
    per-read n_new, flags and counters equal the sequential oracle's."""
    monkeypatch.setenv('GOETIA_AMD_BINNED', '1')
    monkeypatch.setenv('GOETIA_AMD_SUB_BATCH', str(8192 * 3))
    rng = np.random.default_rng(50 * n_tables + K)
    reads = [random_dna(rng, int(rng.integers(0, 700))).encode() for _ in range(300)]
    reads += [b'ACGTTGCA' * 40] * 30 + [b'A' * 300] * 20
    for kind in (1, 2):
        sizes = oracle.primes_near_x(n_tables, 300_000 if kind == 2 else 900_000)
        store = (lg.ByteStorage if kind == 2 else lg.BitStorage)(sizes)
        g = lg.dBG[type(store), lg.CanLemireShifter].build(store, K)
        n_new, flags = g.insert_sequences(ReadBatch.from_strings(reads), new_flags=True)
        o = oracle.Storage(kind, sizes)
        o_new, o_flags = [], []
        for r in reads:
            _, nn, fl = o.insert_sequence(r, K, True, flags=True)
            o_new.append(nn)
            o_flags.append(fl)
        np.testing.assert_array_equal(n_new, o_new)
        np.testing.assert_array_equal(flags, np.concatenate(o_flags).astype(bool))
        assert (g.n_unique(), g.n_occupied()) == (o.n_unique, o.n_occupied)
        for a, b in zip(g.get_raw(), o.tables()):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize('sub_batch', [0, 8192 * 3])
@pytest.mark.parametrize('n_tables', [1, 3, 4, 5, 16])
def test_byte_insert_and_query_vs_oracle(oracle, sub_batch, n_tables, monkeypatch):
    """Count-min insert_and_query through the binned replay (up to four
    tables) and the sort-based path (five or more): every window's count just
    after its own insert (1 when new), with hot k-mers past 255, skip-N
    windows (-1) and several sub-batches, equals the sequential oracle;
    counters and tables too."""
    from goetia_amd import workload
    if sub_batch:
        monkeypatch.setenv('GOETIA_AMD_SUB_BATCH', str(sub_batch))
    genome = workload.genome(41 + n_tables, 20_000)
    gen = workload.reads(42 + n_tables, genome, 1500, 30, 250, 0.01, 0.003)
    reads = [gen.read(r) for r in range(gen.n_reads)] + [b'ACGTTGCA' * 30] * 50 + [b'C' * 150] * 30
    sizes = oracle.primes_near_x(n_tables, 200_000)
    g = lg.dBG[lg.ByteStorage, lg.CanLemireShifter].build(lg.ByteStorage(sizes), 21)
    g.mode = N.MODE_SKIP_N
    got = g.insert_and_query_sequences(ReadBatch.from_strings(reads))
    o = oracle.Storage(oracle.BYTE, sizes)
    exp = np.concatenate([o.insert_and_query_sequence(r, 21, True, oracle.MODE_SKIP_N) for r in reads])
    np.testing.assert_array_equal(got, exp)
    assert (g.n_unique(), g.n_occupied()) == (o.n_unique, o.n_occupied)
    for a, b in zip(g.get_raw(), o.tables()):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize('K', [1, 4, 5, 31, 64, 133, 255])
def test_byte_insert_and_query_every_k_vs_oracle(oracle, K):
    """Count-min insert_and_query through the ordered binned passes for K from
    1 to 255 (the ordered first pass stitches its window starts from 4-byte
    segments for every K), on reads of every length around K with N bases
    and hot repeats: every window's count, counters and tables equal the
    sequential oracle's."""
    rng = np.random.default_rng(2000 + K)
    reads = []
    for _ in range(300):
        L = int(rng.choice([0, K - 1, K, K + 1, K + 9, 150, 600, 2000]))
        r = bytearray(random_dna(rng, max(L, 0)).encode())
        for j in np.nonzero(rng.random(len(r)) < 0.003)[0]:
            r[j] = ord('N')
        reads.append(bytes(r))
    reads += [b'A' * (K + 300)] * 3 + [b'ACGT' * 80] * 20
    sizes = oracle.primes_near_x(4, 150_000)
    g = lg.dBG[lg.ByteStorage, lg.CanLemireShifter].build(lg.ByteStorage(sizes), K)
    g.mode = N.MODE_SKIP_N
    got = g.insert_and_query_sequences(ReadBatch.from_strings(reads))
    o = oracle.Storage(oracle.BYTE, sizes)
    exp = np.concatenate([o.insert_and_query_sequence(r, K, True, oracle.MODE_SKIP_N) for r in reads])
    np.testing.assert_array_equal(got, exp)
    assert (g.n_unique(), g.n_occupied()) == (o.n_unique, o.n_occupied)
    for a, b in zip(g.get_raw(), o.tables()):
        np.testing.assert_array_equal(a, b)


def test_byte_insert_and_query_large_table_vs_oracle(oracle):
    """A count-min storage with more than 2^30 bins (one table of 2^30 + 1
    bins beside a small one) takes the sort-based insert_and_query: results,
    counters and tables equal the sequential oracle's."""
    from goetia_amd import workload
    genome = workload.genome(61, 30_000)
    gen = workload.reads(62, genome, 2000, 30, 200, 0.01, 0.0)
    reads = [gen.read(r) for r in range(gen.n_reads)] + [b'ACGTTGCA' * 30] * 40
    sizes = [(1 << 30) + 1, 1_000_003]
    g = lg.dBG[lg.ByteStorage, lg.CanLemireShifter].build(lg.ByteStorage(sizes), 21)
    got = g.insert_and_query_sequences(ReadBatch.from_strings(reads))
    o = oracle.Storage(oracle.BYTE, sizes)
    exp = np.concatenate([o.insert_and_query_sequence(r, 21, True) for r in reads])
    np.testing.assert_array_equal(got, exp)
    assert (g.n_unique(), g.n_occupied()) == (o.n_unique, o.n_occupied)
    for a, b in zip(g.get_raw(), o.tables()):
        assert hashlib.sha256(a.tobytes()).digest() == hashlib.sha256(b.tobytes()).digest()


@pytest.mark.parametrize('n_tables', [2, 4, 7])
def test_byte_insert_and_query_hashes_vs_oracle(oracle, n_tables):
    """Storage::insert_and_query over an array of hashes in index order
    (sort-based): repeated hashes (runs past 255), colliding bins of small
    tables, all-ones and zero values; results, counters and tables equal the
    oracle's one-hash-at-a-time calls, over two calls."""
    rng = np.random.default_rng(70 + n_tables)
    pool = np.concatenate([rng.integers(0, 2**64 - 1, 3000, dtype=np.uint64),
                           np.array([0, 2**64 - 1], np.uint64)])
    hs = pool[rng.integers(0, len(pool), 20_000)]
    hs[5000:5400] = pool[7]
    sizes = oracle.primes_near_x(n_tables, 5_003)
    st = lg.ByteStorage(sizes)
    o = oracle.Storage(oracle.BYTE, sizes)
    for part in (hs[:12_000], hs[12_000:]):
        got = st.insert_and_query_hashes(part)
        np.testing.assert_array_equal(got, [o.insert_and_query(int(h)) for h in part])
    assert (st.n_unique_kmers(), st.n_occupied()) == (o.n_unique, o.n_occupied)
    for a, b in zip(st.get_raw_tables(), o.tables()):
        np.testing.assert_array_equal(a, b)
