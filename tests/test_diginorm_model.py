"""The CPU model of the batched diginorm (tests/diginorm_model.py) against the
sequential oracle: same pass flags, tables and counters for every chunk size
and guess cap, on reads with duplicates, tandem repeats (a read with hundreds
of events on one bin), short reads and reads spanning chunks.  Runs the
kernels' step logic on the CPU before any of it runs on a GPU."""
import numpy as np
import pytest

from diginorm_model import diginorm


def _reads(seed):
    rng = np.random.default_rng(seed)
    genome = ''.join(rng.choice(list('ACGT'), 1500))
    reads = []
    for i in range(500):
        kind = i % 9
        if kind == 0:
            reads.append('A' * int(rng.integers(15, 400)))
        elif kind == 1:
            unit = ''.join(rng.choice(list('ACGT'), int(rng.integers(2, 4))))
            reads.append((unit * 200)[:int(rng.integers(30, 300))])
        elif kind == 2:
            reads.append(genome[100:220])
        else:
            L = int(rng.integers(10, 160))
            s = int(rng.integers(0, len(genome) - L))
            reads.append(genome[s:s + L])
    enc = [r.encode() for r in reads]
    off = np.zeros(len(enc) + 1, np.int64)
    off[1:] = np.cumsum([len(e) for e in enc])
    return np.frombuffer(b''.join(enc), np.uint8).copy(), off


@pytest.mark.parametrize('chunk,cap', [(1 << 25, 32), (2048, 32), (4096, 2), (1 << 25, 3)])
@pytest.mark.parametrize('cutoff,max_table,n_tables', [(3, 3000, 2), (1, 20000, 4), (6, 9000, 3)])
def test_model_vs_oracle(oracle, chunk, cap, cutoff, max_table, n_tables):
    seq, off = _reads(cutoff * 7 + n_tables)
    K = 21
    sizes = oracle.primes_near_x(n_tables, max_table)
    o = oracle.Storage(oracle.BYTE, sizes)
    exp = o.diginorm_reads(seq, off, K, cutoff)
    # a first half already in the tables: the chunk replays start from them
    h = len(off) // 2
    o2 = oracle.Storage(oracle.BYTE, sizes)
    o2.diginorm_reads(seq[:off[h]], off[:h + 1], K, cutoff)
    got, tabs, nu, no, guesses = diginorm(sizes, seq[off[h]:], off[h:] - off[h], K, 0, cutoff, tables=o2.tables(),
                                          chunk_starts=chunk, cap=cap)
    np.testing.assert_array_equal(got, exp[h:])
    for a, b in zip(tabs, o.tables()):
        np.testing.assert_array_equal(a, b[:len(a)])
    assert nu == o.n_unique - o2.n_unique and no == o.n_occupied - o2.n_occupied
    assert 0 < exp.sum() < len(exp)
    if cap == 2:
        assert max(guesses) == 2
