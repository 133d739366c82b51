"""ByteStorage bigcount (bytestorage.cc:60-139, storage.hh:129-135) against
what goetia's compiled storage did with it on (tests/golden/bigcount_ref.json
and the file it saved, tests/golden/data/bigcount/cms_big.oxli, both written
by tests/golden/make_bigcount_golden.py).

The inputs saturate bins: poly-A / poly-C runs and short repeats, random reads
on small tables, and one 40,000-base poly-A read whose k-mer's count passes
32767, where goetia's int16 wraps to -32768 and stays.

Host (CPU): the oracle (oracle/oracle.c) gives goetia's per-read n_new,
counters, tables, queries, bigcount map and insert_and_query results.
GPU: the device storage gives the same, loads goetia's saved file (tables,
map and the queries they answer), writes a file whose map goetia's format
holds, and the paths that do not keep the map refuse a bigcount storage.
"""
import hashlib
import json
import os
import struct

import numpy as np
import pytest

from conftest import GOLDEN_DIR

GOLD = json.load(open(os.path.join(GOLDEN_DIR, 'bigcount_ref.json')))
CASES = GOLD['cases']
BIG_FILE = os.path.join(GOLDEN_DIR, 'data', 'bigcount', 'cms_big.oxli')


def digest(a):
    return hashlib.sha256(np.asarray(a, np.int16).tobytes()).hexdigest()


def pairs_digest(keys, vals):
    return hashlib.sha256(np.asarray(keys, np.uint64).tobytes() + np.asarray(vals, np.int16).tobytes()).hexdigest()


def packed(reads):
    seq = np.frombuffer(''.join(reads).encode(), np.uint8)
    off = np.zeros(len(reads) + 1, np.int64)
    off[1:] = np.cumsum([len(r) for r in reads])
    return seq, off


def file_pairs(path):
    """The (k-mer, count) pairs after a count-min file's tables."""
    b = open(path, 'rb').read()
    pos = 4 + 1 + 1 + 1 + 4
    n_tables = b[pos]
    pos += 1 + 8
    for _ in range(n_tables):
        (ts,) = struct.unpack_from('<Q', b, pos)
        pos += 8 + ts
    (n,) = struct.unpack_from('<Q', b, pos)
    pos += 8
    pairs = [struct.unpack_from('<Qh', b, pos + 10 * j) for j in range(n)]
    assert pos + 10 * n == len(b)
    return pairs


@pytest.mark.parametrize('ci', range(len(CASES)))
def test_oracle_vs_reference(oracle, ci):
    c = CASES[ci]
    K, can = c['K'], bool(c['canonical'])
    st = oracle.Storage(oracle.BYTE, c['sizes'])
    st.set_use_bigcount(True)
    seen = []
    for b in c['batches']:
        seq, off = packed(b['reads'])
        _, n_new = st.insert_reads(seq, off, K, can)
        assert n_new.tolist() == b['n_new']
        seen += b['reads']
        q = st.query_reads(*packed(seen), K, can)
        assert q[:1000].tolist() == b['query_head']
        assert digest(q) == b['query_sha256']
        assert [st.n_unique, st.n_occupied] == b['counters']
        assert [hashlib.sha256(t.tobytes()).hexdigest() for t in st.tables()] == b['table_sha256']
        keys, vals = st.bigcount_pairs()
        assert len(keys) == b['bigcounts_n']
        assert [[int(k), int(v)] for k, v in zip(keys[:200], vals[:200])] == b['bigcounts_head']
        assert pairs_digest(keys, vals) == b['bigcounts_sha256']
    st2 = oracle.Storage(oracle.BYTE, c['sizes'])
    st2.set_use_bigcount(True)
    for b, want in zip(c['batches'], c['insert_and_query']):
        got = st2.insert_and_query_reads(*packed(b['reads']), K, can)
        assert got[:1000].tolist() == want['head']
        assert digest(got) == want['sha256']


def test_reference_file_holds_the_map():
    """goetia's saved file carries use_bigcount = 1 and case 0's final map."""
    b = open(BIG_FILE, 'rb').read()
    assert b[:4] == b'OXLI' and b[6] == 1
    pairs = sorted(file_pairs(BIG_FILE))
    last = CASES[0]['batches'][-1]
    assert len(pairs) == last['bigcounts_n']
    assert pairs_digest([p[0] for p in pairs], [p[1] for p in pairs]) == last['bigcounts_sha256']


def _graph(lg, c):
    store = lg.ByteStorage(c['sizes'])
    sh = lg.CanLemireShifter if c['canonical'] else lg.FwdLemireShifter
    g = lg.dBG[lg.ByteStorage, sh].build(store, c['K'])
    store.set_use_bigcount(True)
    return g, store


@pytest.mark.gpu
@pytest.mark.parametrize('ci', range(len(CASES)))
def test_device_vs_reference(ci):
    from goetia_amd import libgoetia as lg
    c = CASES[ci]
    g, store = _graph(lg, c)
    seen = []
    for b in c['batches']:
        n_new = g.insert_sequences(b['reads'])
        assert [int(x) for x in n_new] == b['n_new']
        seen += b['reads']
        q = g.query_sequences(seen)
        assert q[:1000].tolist() == b['query_head']
        assert digest(q) == b['query_sha256']
        assert [g.n_unique(), g.n_occupied()] == b['counters']
        assert [hashlib.sha256(t.tobytes()).hexdigest() for t in g.get_raw()] == b['table_sha256']
        keys, vals = store.bigcounts()
        assert len(keys) == b['bigcounts_n']
        assert pairs_digest(keys, vals) == b['bigcounts_sha256']
    g2, _ = _graph(lg, c)
    for b, want in zip(c['batches'], c['insert_and_query']):
        got = g2.insert_and_query_sequences(b['reads'])
        assert got[:1000].tolist() == want['head']
        assert digest(got) == want['sha256']


@pytest.mark.gpu
@pytest.mark.parametrize('ci', range(len(CASES)))
def test_device_insert_and_query_in_window_chunks(ci, monkeypatch):
    """insert_and_query with bigcount answers in window chunks, each chunk's
    full windows joining the map before the next is answered; chunks of
    8192 windows (GOETIA_AMD_SUB_BATCH) give goetia's results."""
    from goetia_amd import libgoetia as lg
    monkeypatch.setenv('GOETIA_AMD_SUB_BATCH', '8192')
    c = CASES[ci]
    g, _ = _graph(lg, c)
    for b, want in zip(c['batches'], c['insert_and_query']):
        got = g.insert_and_query_sequences(b['reads'])
        assert got[:1000].tolist() == want['head']
        assert digest(got) == want['sha256']


@pytest.mark.gpu
@pytest.mark.parametrize('sub_batch', [None, '8192'])
def test_device_per_hash_calls_match_oracle(oracle, sub_batch, monkeypatch):
    """Per-k-mer insert / query / insert_and_query on hash values (the
    storage API) with bigcount on equal the oracle's, across the wrap, also
    answered in chunks of 8192 hashes."""
    from goetia_amd import libgoetia as lg
    if sub_batch:
        monkeypatch.setenv('GOETIA_AMD_SUB_BATCH', sub_batch)
    sizes = oracle.primes_near_x(3, 211)
    rng = np.random.default_rng(7)
    hashes = np.concatenate([rng.integers(0, 2 ** 63, 3000, dtype=np.uint64),
                             np.full(33000, 12345, np.uint64), np.full(400, 2 ** 64 - 1, np.uint64)])
    rng.shuffle(hashes)
    o = oracle.Storage(oracle.BYTE, sizes)
    o.set_use_bigcount(True)
    want_iq = [o.insert_and_query(int(h)) for h in hashes[:20000]]
    for h in hashes[20000:]:
        o.insert(int(h))
    s = lg.ByteStorage(sizes)
    s.set_use_bigcount(True)
    got_iq = s.insert_and_query_hashes(hashes[:20000])
    assert got_iq.tolist() == want_iq
    s.insert_hashes(hashes[20000:])
    probe = np.unique(hashes)
    assert s.query_hashes(probe).tolist() == [o.query(int(h)) for h in probe]
    keys, vals = s.bigcounts()
    ok, ov = o.bigcount_pairs()
    assert keys.tolist() == ok.tolist() and vals.tolist() == ov.tolist()
    assert -32768 in vals.tolist()


@pytest.mark.gpu
def test_device_loads_and_saves_bigcount_files(tmp_path):
    """goetia's saved file loads with its tables and map, answering case 0's
    queries as goetia's storage did; a saved file holds the same map (pairs
    in k-mer order) and loads back."""
    from goetia_amd import libgoetia as lg
    c = CASES[0]
    last = c['batches'][-1]
    g = lg.dBG[lg.ByteStorage, lg.CanLemireShifter].build(lg.ByteStorage.build(1000, 1), c['K'])
    g.load(BIG_FILE)
    assert g.S.get_use_bigcount()
    assert [hashlib.sha256(t.tobytes()).hexdigest() for t in g.get_raw()] == last['table_sha256']
    keys, vals = g.S.bigcounts()
    assert pairs_digest(keys, vals) == last['bigcounts_sha256']
    seen = [r for b in c['batches'] for r in b['reads']]
    assert digest(g.query_sequences(seen)) == last['query_sha256']
    out = str(tmp_path / 'saved.oxli')
    g.save(out)
    b = open(out, 'rb').read()
    assert b[6] == 1
    pairs = file_pairs(out)
    assert pairs == sorted(pairs)
    assert [p[0] for p in pairs] == keys.tolist() and [p[1] for p in pairs] == vals.tolist()
    g2 = lg.dBG[lg.ByteStorage, lg.CanLemireShifter].build(lg.ByteStorage.build(1000, 1), c['K'])
    g2.load(out)
    k2, v2 = g2.S.bigcounts()
    assert k2.tolist() == keys.tolist() and v2.tolist() == vals.tolist()


@pytest.mark.gpu
def test_paths_without_the_map_refuse():
    from goetia_amd import _native as N
    from goetia_amd import libgoetia as lg
    g, store = _graph(lg, CASES[0])
    g.insert_sequences(['A' * 600])
    with pytest.raises(N.UnsupportedError):
        g.degrees(['A' * 100])
    store.set_use_bigcount(False)
    assert not store.get_use_bigcount()
    g.degrees(['A' * 100])
