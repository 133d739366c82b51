"""Pins the CPU oracle (oracle/oracle.c): against the golden vectors written
by the compiled goetia reference, the known-answer vector of
tests/test_hashing.py:15-18 in goetia, and what the mathematics fixes
(brute force on small inputs, closed forms, invariants)."""
import hashlib

import numpy as np
import pytest

from conftest import golden_input, random_dna, revcomp

KNOWN_KMER = 'TCACCTGTGTTGTGCTACTTGCGGCGC'
KNOWN_FW = 13194817695400542713
KNOWN_RC = 4324216031038051805


def test_known_vector(oracle):
    fw, rc = oracle.hash_kmer(KNOWN_KMER)
    assert (fw, rc) == (KNOWN_FW, KNOWN_RC)
    # hash of a longer sequence is the hash of its first K bytes (test_hashing.py:141-158)
    fw2, rc2 = oracle.hash_read(KNOWN_KMER + 'AA', 27)
    assert (int(fw2[0]), int(rc2[0])) == (KNOWN_FW, KNOWN_RC)


@pytest.mark.parametrize('K', [1, 2, 21, 27, 31, 32, 33, 63, 64, 65, 101, 128, 200])
def test_hash_all_bytes_vs_reference(oracle, golden, K):
    s = golden[f'hash_bytes_K{K}_seq'].tobytes()
    fw, rc = oracle.hash_read(s, K)
    np.testing.assert_array_equal(fw, golden[f'hash_bytes_K{K}_fw'])
    np.testing.assert_array_equal(rc, golden[f'hash_bytes_K{K}_rc'])


def _closed_form(kmer):
    """fw = XOR_i rotl(T[s_i], K-1-i); rc = XOR_i rotl(T[comp s_i], i)."""
    from oracle.pyoracle import table_value
    M = (1 << 64) - 1
    rotl = lambda x, r: ((x << (r % 64)) | (x >> ((64 - r % 64) % 64))) & M if r % 64 else x
    comp = {'A': 'T', 'C': 'G', 'G': 'C', 'T': 'A'}
    K = len(kmer)
    fw = rc = 0
    for i, c in enumerate(kmer):
        fw ^= rotl(table_value(ord(c)), K - 1 - i)
        rc ^= rotl(table_value(ord(comp.get(c, '\0'))), i)
    return fw, rc


@pytest.mark.parametrize('K', [5, 21, 31, 64, 101])
def test_rolling_equals_closed_form_and_direct(oracle, K):
    rng = np.random.default_rng(K)
    s = random_dna(rng, 300)
    fw, rc = oracle.hash_read(s, K)
    for i in range(len(s) - K + 1):
        d = oracle.hash_kmer(s[i:i + K])
        assert (int(fw[i]), int(rc[i])) == d
        assert d == _closed_form(s[i:i + K])


def test_canonical_is_min_of_strands(oracle):
    # test_hashing.py:161-172: canonical == min(fwd(kmer), fwd(revcomp(kmer)))
    rng = np.random.default_rng(3)
    s = random_dna(rng, 500)
    K = 21
    fw, rc = oracle.hash_read(s, K)
    for i in range(len(s) - K + 1):
        kmer = s[i:i + K]
        a = oracle.hash_kmer(kmer)[0]
        b = oracle.hash_kmer(revcomp(kmer))[0]
        assert min(a, b) == min(int(fw[i]), int(rc[i]))
        assert int(rc[i]) == b


def _brute_primes(n, x):
    out = []
    i = x - 1 if (x - 1) % 2 else x - 2
    while len(out) < n and i >= 1:
        if i > 1 and all(i % d for d in range(2, int(i ** 0.5) + 1)):
            out.append(i)
        if i == 1:
            break
        i -= 2
    return out


@pytest.mark.parametrize('n,x', [(4, 1000), (4, 5000), (3, 10), (4, 20000), (4, 50000), (1, 4)])
def test_primes_brute_force(oracle, n, x):
    assert oracle.primes_near_x(n, x) == _brute_primes(n, x)


def test_primes_vs_reference(oracle, golden):
    for key in golden:
        if key.startswith('primes_'):
            _, n, x = key.split('_')
            assert oracle.primes_near_x(int(n), int(x)) == [int(v) for v in golden[key]], key


def _storage_for(oracle, meta):
    if meta['kind'] == oracle.EXACT:
        return oracle.Storage(oracle.EXACT)
    return oracle.Storage(meta['kind'], oracle.primes_near_x(meta['n_tables'], meta['max_table']))


def _pack(reads):
    enc = [r.encode() for r in reads]
    off = np.zeros(len(enc) + 1, np.int64)
    off[1:] = np.cumsum([len(e) for e in enc])
    return np.frombuffer(b''.join(enc), np.uint8), off


def test_storages_vs_reference(oracle, golden, golden_meta):
    """Per-read n_new, tables, counters and query counts of every golden run."""
    checked = 0
    for name, meta in golden_meta['runs'].items():
        if 'byteiq' in name:
            continue
        reads = golden_input(name, golden)
        st = _storage_for(oracle, meta)
        seq, off = _pack(reads)
        total, nn = st.insert_reads(seq, off, meta['K'], bool(meta['canonical']), meta['mode'])
        np.testing.assert_array_equal(nn, golden[f'{name}__n_new'], err_msg=name)
        assert total == int(golden[f'{name}__n_kmers'][0]), name
        assert st.n_unique == int(golden[f'{name}__counters'][0]), name
        assert st.n_occupied == int(golden[f'{name}__counters'][1]), name
        if meta['kind'] != oracle.EXACT:
            for i, t in enumerate(st.tables()):
                assert hashlib.sha256(t.tobytes()).digest() == golden[f'{name}__table{i}_sha'].tobytes(), name
        q = st.query_reads(seq, off, meta['K'], bool(meta['canonical']), meta['mode'])
        assert hashlib.sha256(q.tobytes()).digest() == golden[f'{name}__query_sha'].tobytes(), name
        checked += 1
    assert checked >= 14


def test_byte_insert_and_query_vs_reference(oracle, golden, golden_meta):
    for name in ('rand20_byteiq_can21', 'synth_byteiq_can31'):
        meta = golden_meta['runs'][name]
        reads = golden_input(name, golden)[:meta['n_reads']]
        st = _storage_for(oracle, meta)
        got = np.concatenate([st.insert_and_query_sequence(r, meta['K']) for r in reads])
        np.testing.assert_array_equal(got, golden[f'{name}__counts'])
        assert [st.n_unique, st.n_occupied] == [int(v) for v in golden[f'{name}__counters']]


def test_bloom_table_is_or_of_events(oracle):
    """Bloom bit = OR over inserted hashes; query = AND over tables."""
    rng = np.random.default_rng(5)
    sizes = oracle.primes_near_x(3, 4000)
    st = oracle.Storage(oracle.BIT, sizes)
    hs = rng.integers(0, 2 ** 63, 3000, dtype=np.uint64) * 2 + rng.integers(0, 2, 3000, dtype=np.uint64)
    for h in hs:
        st.insert(int(h))
    for i, p in enumerate(sizes):
        bits = np.zeros(p // 8 + 1, np.uint8)
        for h in hs:
            b = int(h) % p
            bits[b // 8] |= 1 << (b % 8)
        np.testing.assert_array_equal(st.table(i), bits)
    assert st.n_occupied == int(sum(bin(x).count('1') for x in st.table(0)))


def test_cms_table_is_saturated_count(oracle):
    """Count-min byte = min(255, number of events in the bin)."""
    rng = np.random.default_rng(6)
    sizes = oracle.primes_near_x(4, 600)
    st = oracle.Storage(oracle.BYTE, sizes)
    hs = rng.integers(0, 2 ** 64 - 1, 40000, dtype=np.uint64)
    for h in hs:
        st.insert(int(h))
    for i, p in enumerate(sizes):
        cnt = np.bincount((hs % np.uint64(p)).astype(np.int64), minlength=p)
        np.testing.assert_array_equal(st.table(i), np.minimum(cnt, 255).astype(np.uint8))


def test_first_touch_brute_force(oracle):
    """n_new of a read = number of its k-mers that, inserted in order, found
    an empty bin in some table (simulated here with python sets)."""
    rng = np.random.default_rng(8)
    sizes = oracle.primes_near_x(2, 300)
    reads = [random_dna(rng, int(rng.integers(20, 60))) for _ in range(60)]
    K = 11
    st = oracle.Storage(oracle.BIT, sizes)
    seen = [set() for _ in sizes]
    for r in reads:
        _, nn = st.insert_sequence(r, K)
        fw, rc = oracle.hash_read(r, K)
        expect = 0
        for h in np.minimum(fw, rc):
            new = False
            for i, p in enumerate(sizes):
                b = int(h) % p
                if b not in seen[i]:
                    seen[i].add(b)
                    new = True
            expect += new
        assert nn == expect


def test_exact_n_unique_is_set_size(oracle):
    # test_dbg.py:276-284
    rng = np.random.default_rng(9)
    s = random_dna(rng, 5000)
    st = oracle.Storage(oracle.EXACT)
    st.insert_sequence(s, 21)
    fw, rc = oracle.hash_read(s, 21)
    assert st.n_unique == len(set(np.minimum(fw, rc).tolist()))
    np.testing.assert_array_equal(st.exact_keys(), np.unique(np.minimum(fw, rc)))


def test_extensions_vs_direct_hashing(oracle):
    """The 8 neighbour hashes equal direct hashes of the neighbour strings."""
    rng = np.random.default_rng(10)
    st = oracle.Storage(oracle.EXACT)
    for K in (21, 31, 33, 64, 65):
        for _ in range(10):
            kmer = random_dna(rng, K)
            _, _, left, right = st.kmer_degrees(kmer, K)
            for c, v in zip('ACGT', left):
                assert v == min(oracle.hash_kmer(c + kmer[:-1]))
            for c, v in zip('ACGT', right):
                assert v == min(oracle.hash_kmer(kmer[1:] + c))


def test_degrees_vs_reference(oracle, golden, golden_meta):
    for name in ('synth_bit_can21', 'synth_byte_can31', 'synth_set_can25', 'synth_byte_fwd31'):
        meta = golden_meta['runs'][name]
        K = meta['K']
        st = _storage_for(oracle, meta)
        seq, off = _pack(golden_input(name, golden))
        st.insert_reads(seq, off, K, bool(meta['canonical']), meta['mode'])
        kmers = golden[f'{name}__ext_kmers'].tobytes().decode()
        n = len(kmers) // K
        for i in range(n):
            ind, outd, left, right = st.kmer_degrees(kmers[i * K:(i + 1) * K], K, bool(meta['canonical']))
            assert left == [int(v) for v in golden[f'{name}__ext_left'][i]]
            assert right == [int(v) for v in golden[f'{name}__ext_right'][i]]
            assert (ind, outd) == (golden[f'{name}__ext_in'][i], golden[f'{name}__ext_out'][i])


def test_decision_kmers_vs_reference(oracle, golden, golden_meta):
    for name in ('synth_bit_can21', 'synth_byte_can31', 'synth_set_can25', 'synth_byte_fwd31'):
        meta = golden_meta['runs'][name]
        K = meta['K']
        st = _storage_for(oracle, meta)
        reads = golden_input(name, golden)
        seq, off = _pack(reads)
        st.insert_reads(seq, off, K, bool(meta['canonical']), meta['mode'])
        got_read, got_pos = [], []
        for ri, r in enumerate(reads[:200]):
            ind, outd = st.degrees_sequence(r, K, bool(meta['canonical']))
            for j in np.nonzero((ind > 1) | (outd > 1))[0]:
                got_read.append(ri)
                got_pos.append(int(j))
        np.testing.assert_array_equal(got_read, golden[f'{name}__decision_read'])
        np.testing.assert_array_equal(got_pos, golden[f'{name}__decision_pos'])


def test_skip_n_equals_run_splitting(oracle):
    """Skip-N on a read == inserting each ACGT run of length >= K in order."""
    rng = np.random.default_rng(12)
    sizes = oracle.primes_near_x(4, 5000)
    a = oracle.Storage(oracle.BYTE, sizes)
    b = oracle.Storage(oracle.BYTE, sizes)
    K = 15
    for _ in range(50):
        r = list(random_dna(rng, 120))
        for j in rng.integers(0, 120, 3):
            r[j] = 'N'
        r = ''.join(r)
        _, nn = a.insert_sequence(r, K, mode=oracle.MODE_SKIP_N)
        tot = 0
        for run in r.split('N'):
            if len(run) >= K:
                tot += b.insert_sequence(run, K)[1]
        assert nn == tot
    for ta, tb in zip(a.tables(), b.tables()):
        np.testing.assert_array_equal(ta, tb)


def test_planted_mistake_is_caught(oracle, golden):
    """A dropped term in the hash (here: the rc strand's complement) fails the
    reference vectors, so those vectors do pin the oracle."""
    s = golden['hash_bytes_K31_seq'].tobytes()
    fw, rc = oracle.hash_read(s, 31)
    wrong_rc = np.array([oracle.hash_kmer(s[i:i + 31][::-1])[0] for i in range(len(s) - 30)], np.uint64)
    assert not np.array_equal(wrong_rc, golden['hash_bytes_K31_rc'])
    np.testing.assert_array_equal(rc, golden['hash_bytes_K31_rc'])


@pytest.mark.parametrize('kind', [1, 2])
@pytest.mark.parametrize('threads', [1, 3, 8])
def test_multithreaded_tables_equal_sequential(oracle, kind, threads):
    """The all-cores CPU baseline (atomic table updates on several threads)
    leaves the sequential insert's tables and query counts."""
    rng = np.random.default_rng(kind * 10 + threads)
    genome = rng.choice(np.frombuffer(b'ACGT', np.uint8), 20000)
    reads = []
    for _ in range(3000):
        L = int(rng.integers(10, 200))
        s = int(rng.integers(0, len(genome) - L))
        r = genome[s:s + L].copy()
        if rng.random() < 0.05:
            r[int(rng.integers(0, L))] = ord('N')
        reads.append(r.tobytes())
    off = np.zeros(len(reads) + 1, np.int64)
    off[1:] = np.cumsum([len(r) for r in reads])
    seq = np.frombuffer(b''.join(reads), np.uint8).copy()
    sizes = oracle.primes_near_x(4, 5000)
    for mode in (oracle.MODE_RAW, oracle.MODE_SKIP_N):
        a, b = oracle.Storage(kind, sizes), oracle.Storage(kind, sizes)
        ha, _ = a.insert_reads(seq, off, 21, True, mode)
        hb = b.insert_reads_mt(seq, off, 21, True, mode, threads)
        assert ha == hb
        for x, y in zip(a.tables(), b.tables()):
            np.testing.assert_array_equal(x, y)
        np.testing.assert_array_equal(a.query_reads(seq, off, 21, True, mode),
                                      b.query_reads_mt(seq, off, 21, True, mode, threads))
