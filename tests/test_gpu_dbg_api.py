"""goetia's own dBG test assertions (tests/test_dbg.py, test_hashing.py and
test_traversal.py of the reference), restated against the GPU-backed
goetia_amd surface."""
import numpy as np
import pytest

from conftest import random_dna

pytestmark = pytest.mark.gpu

from goetia_amd import libgoetia as lg  # noqa: E402
from goetia_amd import _native as N  # noqa: E402

PRESENCE = [lg.BitStorage, lg.SparseppSetStorage, lg.PHMapStorage, lg.BTreeStorage]
COUNTING = [lg.ByteStorage]
ALL = PRESENCE + COUNTING
SHIFTERS = [lg.FwdLemireShifter, lg.CanLemireShifter]


def graph(storage_type, shifter_type, K):
    store = storage_type.build()
    return lg.dBG[storage_type, shifter_type].build(store, shifter_type(K))


def kmers(s, K):
    for i in range(len(s) - K + 1):
        yield s[i:i + K]


@pytest.mark.parametrize('st', PRESENCE)
@pytest.mark.parametrize('sh', SHIFTERS)
@pytest.mark.parametrize('K', [21, 101])
def test_presence_insert(st, sh, K):
    g = graph(st, sh, K)
    s = random_dna(np.random.default_rng(K), 160)
    for kmer in list(kmers(s, K))[:25]:
        h = g.hash(kmer)
        assert g.query(kmer) == 0 and g.query(h) == 0
        assert g.insert(kmer) is True
        assert g.query(kmer) == 1 and g.query(h) == 1
        assert g.insert(kmer) is False
        assert g.query(kmer) == 1


@pytest.mark.parametrize('sh', SHIFTERS)
@pytest.mark.parametrize('K', [21, 101])
def test_counting_insert_and_query(sh, K):
    g = graph(lg.ByteStorage, sh, K)
    s = random_dna(np.random.default_rng(K + 1), 160)
    for kmer in list(kmers(s, K))[:25]:
        h = g.hash(kmer)
        assert g.query(kmer) == 0
        assert g.insert(kmer) is True
        assert g.query(h) == 1
        assert g.insert(kmer) is False
        assert g.query(h) == 2
    g2 = graph(lg.ByteStorage, sh, K)
    for kmer in list(kmers(s, K))[:25]:
        assert g2.insert_and_query(kmer) == 1
        assert g2.insert_and_query(g2.hash(kmer)) == 2


@pytest.mark.parametrize('st', PRESENCE)
def test_presence_insert_and_query_is_one(st):
    g = graph(st, lg.CanLemireShifter, 21)
    for kmer in list(kmers(random_dna(np.random.default_rng(3), 80), 21))[:10]:
        assert g.insert_and_query(kmer) == 1
        assert g.insert_and_query(g.hash(kmer)) == 1


@pytest.mark.parametrize('K', [21, 101])
def test_counting_count_add_sequence(K):
    g = graph(lg.ByteStorage, lg.CanLemireShifter, K)
    seq = random_dna(np.random.default_rng(K), 400)
    ks = list(kmers(seq, K))
    for it in range(10):
        counts = g.query_sequence(seq)
        assert min(counts) >= it
        assert g.query(ks[0]) >= it
        g.insert_sequence(seq)


def test_hash_bad_length_and_short_sequence():
    g = graph(lg.BitStorage, lg.CanLemireShifter, 21)
    assert g.hash('A' * 21) == g.hash('A' * 21 + 'TTTT')
    with pytest.raises(N.SequenceLengthException):
        g.insert_sequence('ATGCA')
    with pytest.raises(Exception):
        g.hash('AAAAA')


def test_get_kmer_counts():
    # test_dbg.py:230-249
    g = graph(lg.ByteStorage, lg.CanLemireShifter, 21)
    g.insert_sequence('A' * 21)
    assert g.query_sequence('A' * 21) == [1]
    g.insert_sequence('A' * 21)
    assert g.query_sequence('A' * 21)[0] >= 1
    g.insert_sequence('A' * 21 + 'T')
    c = g.query_sequence('A' * 21 + 'T')
    assert len(c) == 2 and c[0] >= 1 and c[1] == 1


@pytest.mark.parametrize('st', ALL)
def test_insert_sequence_overload_and_hashes(st):
    # test_dbg.py:199-209, 252-273
    K = 21
    x = random_dna(np.random.default_rng(5), 50)
    g = graph(st, lg.CanLemireShifter, K)
    hashes, report = [], []
    n = g.insert_sequence(x, hashes, report)
    assert n == len(x) - K + 1
    assert sum(report) == len(x) - K + 1
    for start in range(len(x) - K + 1):
        assert g.query(x[start:start + K]) == 1
    hs = list(g.hashes(x))
    assert hs == [g.hash(k) for k in kmers(x, K)]
    assert hs == hashes


@pytest.mark.parametrize('st', [lg.SparseppSetStorage, lg.PHMapStorage, lg.BTreeStorage])
def test_n_unique_exact(st):
    # test_dbg.py:276-284
    g = graph(st, lg.CanLemireShifter, 21)
    seq = random_dna(np.random.default_rng(6), 50000)
    kmer_set = set(g.hash(k).value for k in kmers(seq[:3000], 21))
    g.insert_sequence(seq[:3000])
    assert len(kmer_set) == g.n_unique()


@pytest.mark.parametrize('st', ALL)
def test_get_counts_all_positive(st):
    g = graph(st, lg.CanLemireShifter, 21)
    seq = random_dna(np.random.default_rng(7), 50000)
    g.insert_sequence(seq)
    assert all(c > 0 for c in g.query_sequence(seq))


@pytest.mark.parametrize('st', ALL)
@pytest.mark.parametrize('K', [21, 101])
def test_clone_is_empty(st, K):
    g = graph(st, lg.CanLemireShifter, K)
    g.insert_sequence(random_dna(np.random.default_rng(8), 300))
    g2 = g.clone()
    assert g.n_unique() > 0
    assert g2.n_unique() == 0


@pytest.mark.parametrize('st', ALL)
def test_linear_path_neighbors(st):
    # test_dbg.py:182-196: a k-mer inside a linear path has one left and one
    # right neighbour
    g = graph(st, lg.CanLemireShifter, 21)
    s = random_dna(np.random.default_rng(9), 23)
    g.insert_sequence(s)
    assert len(g.out_neighbors(s[1:22])) == 1
    left, right = g.neighbors(s[1:22])
    assert len(left) == 1 and len(right) == 1


@pytest.mark.parametrize('st', ALL)
def test_fork_degree(st):
    # test_traversal.py:123-151: a right fork has out-degree 2 at the fork
    K = 21
    rng = np.random.default_rng(10)
    core = random_dna(rng, 40)
    branch_a = core + random_dna(rng, 30)
    alt = 'A' if branch_a[40] != 'A' else 'C'
    branch_b = core + alt + random_dna(rng, 29)
    g = graph(st, lg.CanLemireShifter, K)
    g.insert_sequence(branch_a)
    g.insert_sequence(branch_b)
    fork = core[-K:]
    assert g.right_degree(fork) == 2
    assert g.left_degree(fork) == 1
    assert g.is_decision_kmer(fork)
    pos, hs, deg = g.find_decision_kmers(branch_a)
    assert len(core) - K in pos
    # get_decision_neighbors (unitig_walker.hh:1085-1117): the two right
    # neighbours are suffix(fork) + the two branch bases, as k-mers with
    # their hashes; a k-mer inside the core is no decision k-mer
    is_dec, (left, right) = g.get_decision_neighbors(fork)
    assert is_dec
    assert sorted(k.kmer for k in right) == sorted([fork[1:] + branch_a[40], fork[1:] + alt])
    assert [k.kmer for k in left] == [core[-K - 1:-1]]
    for k in left + right:
        assert k.value == g.hash(k.kmer).value
    assert g.get_decision_neighbors(core[5:5 + K]) == (False, ([], []))
    # the cursor walks the core (hashextender.hh:327-353)
    g.set_cursor(core[:K])
    for i in range(1, len(core) - K + 1):
        assert g.shift_right(core[i + K - 1]).value == g.hash(core[i:i + K]).value
        assert g.get_cursor() == core[i:i + K]
    assert g.shift_left(core[len(core) - K - 1]).value == g.hash(core[-K - 1:-1]).value
    pos2, hs2, deg2, nbs = g.find_decision_kmers(branch_a, neighbors=True)
    assert pos2 == pos and len(nbs) == len(pos)
    i = pos.index(len(core) - K)
    assert sorted(k.kmer for k in nbs[i][1]) == sorted(k.kmer for k in right)
    # union nodes (unitig_walker.hh:336-348, 457-468, 1047-1082): a third
    # branch not in the graph counts once it is among the extras
    alt2 = [c for c in 'ACGT' if c not in (branch_a[40], alt)][0]
    third = g.hash(fork[1:] + alt2)
    g.set_cursor(fork)
    assert g.out_degree() == 2 and g.out_degree({third}) == 3 and g.in_degree({third}) == 1
    assert g.degree({third.value}) == 4
    is_dec, (left3, right3) = g.get_decision_neighbors(fork, union_nodes={third})
    assert is_dec and sorted(k.kmer for k in right3) == sorted([fork[1:] + branch_a[40], fork[1:] + alt,
                                                                fork[1:] + alt2])
    assert len(g.filter_nodes(g.right_extensions(), [third])) == 3
    inner = core[5:5 + K]
    nxt = g.hash(inner[1:] + ('A' if core[5 + K] != 'A' else 'C'))
    assert g.get_decision_neighbors(inner, union_nodes={nxt})[0]


def test_kmer_iterator():
    K = 27
    s = random_dna(np.random.default_rng(11), 300)
    for sh in SHIFTERS:
        it = lg.KmerIterator[sh](s, K)
        got = []
        while not it.done():
            got.append(it.next().value)
        hasher = sh(K)
        assert got == [hasher.hash(k).value for k in kmers(s, K)]
    with pytest.raises(N.SequenceLengthException):
        lg.KmerIterator[lg.CanLemireShifter]('ACGT', K)


def test_storage_build_params_and_describe():
    b = lg.BitStorage.build()
    assert b.get_tablesizes() == [999983, 999979, 999961, 999959]
    c = lg.ByteStorage.build(10 ** 7, 4)
    assert c.get_tablesizes() == [9999991, 9999973, 9999971, 9999943]
    assert 'Counting:      True' in c.describe()
    assert not c.get_use_bigcount()
    c.set_use_bigcount(True)
    assert c.get_use_bigcount()
    with pytest.raises(lg.GoetiaException):
        b.set_use_bigcount(True)
