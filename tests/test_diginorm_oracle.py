"""The oracle's DiginormFilter (oracle.c or_diginorm_*) against what goetia's
compiled DiginormFilter left on the same reads (tests/golden/diginorm_ref.npz:
pass flags, counters, tables), and against properties the rule fixes: n
copies of a read pass exactly `cutoff` times, and the early return of
median_count_at_least (diginorm.hh:47-73) decides as the full count of k-mers
at the cutoff would."""
import ctypes as C
import hashlib
import os

import numpy as np
import pytest

from conftest import GOLDEN_DIR

GOLD = np.load(os.path.join(GOLDEN_DIR, 'diginorm_ref.npz'), allow_pickle=False)
CASES = sorted(k[:-len('__params')] for k in GOLD.files if k.endswith('__params'))
REF_SO = os.path.join(os.path.dirname(GOLDEN_DIR), '..', 'oracle', '_ref', 'libgoetia_ref.so')


def case_inputs(name):
    K, can, mt, nt, cutoff = (int(v) for v in GOLD[f'{name}__params'])
    inp = bytes(GOLD[f'{name}__input']).decode()
    return GOLD[f'in_{inp}_seq'], GOLD[f'in_{inp}_off'], K, can, mt, nt, cutoff


@pytest.mark.parametrize('name', CASES)
def test_oracle_vs_reference(oracle, name):
    seq, off, K, can, mt, nt, cutoff = case_inputs(name)
    o = oracle.Storage.build(oracle.BYTE, mt, nt)
    passed = o.diginorm_reads(seq, off, K, cutoff, bool(can))
    np.testing.assert_array_equal(passed, GOLD[f'{name}__passed'].astype(bool))
    assert [o.n_unique, o.n_occupied] == GOLD[f'{name}__counters'].tolist()
    for i, t in enumerate(o.tables()):
        assert hashlib.sha256(t.tobytes()).digest() == GOLD[f'{name}__table{i}_sha'].tobytes()
        if f'{name}__table{i}' in GOLD.files:
            np.testing.assert_array_equal(t, GOLD[f'{name}__table{i}'])


def _random_read(rng, n):
    return ''.join(rng.choice(list('ACGT'), n))


@pytest.mark.parametrize('cutoff', [0, 1, 3, 7, 20])
def test_copies_pass_cutoff_times(oracle, cutoff):
    rng = np.random.default_rng(cutoff)
    r = _random_read(rng, 150)
    o = oracle.Storage.build(oracle.BYTE, 1_000_000, 4)
    got = [o.diginorm_filter(r, 21, cutoff) for _ in range(30)]
    assert got == [1] * min(cutoff, 30) + [0] * (30 - min(cutoff, 30))


@pytest.mark.parametrize('cutoff', [1, 2, 4])
def test_early_return_equals_full_count(oracle, cutoff):
    """A second oracle storage decides each read from its full query_sequence
    counts (count >= cutoff for at least min_req k-mers) and inserts the kept
    ones: same decisions and tables as the filter's early-return walk."""
    rng = np.random.default_rng(100 + cutoff)
    genome = _random_read(rng, 3000)
    reads = []
    for _ in range(1500):
        L = int(rng.integers(10, 200))
        s = int(rng.integers(0, len(genome) - L))
        reads.append(genome[s:s + L])
    K = 21
    a = oracle.Storage.build(oracle.BYTE, 4000, 3)
    b = oracle.Storage.build(oracle.BYTE, 4000, 3)
    for r in reads:
        got = a.diginorm_filter(r, K, cutoff)
        if len(r) < K:
            assert got == -1
            continue
        counts, _ = b.query_sequence(r, K)
        min_req = int(0.5 + float(np.float32(len(r) - K + 1) / np.float32(2)))
        keep = int(np.sum(counts >= cutoff)) < min_req
        if keep:
            b.insert_sequence(r, K)
        assert got == int(keep)
    for x, y in zip(a.tables(), b.tables()):
        np.testing.assert_array_equal(x, y)


@pytest.mark.skipif(not os.path.exists(REF_SO), reason='compiled reference (oracle/_ref) not built here')
@pytest.mark.parametrize('seed', [1, 2])
def test_oracle_vs_compiled_reference_random(oracle, seed):
    L = C.CDLL(REF_SO)
    L.ref_new.restype = C.c_void_p
    L.ref_new.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint16]
    L.ref_diginorm_reads.restype = C.c_uint64
    L.ref_diginorm_reads.argtypes = [C.c_void_p, C.c_uint, C.c_char_p, C.POINTER(C.c_int64), C.c_uint64,
                                     C.POINTER(C.c_uint8)]
    L.ref_free.argtypes = [C.c_void_p]
    rng = np.random.default_rng(seed)
    genome = _random_read(rng, 2000)
    reads = []
    for _ in range(800):
        n = int(rng.integers(15, 300))
        s = int(rng.integers(0, len(genome) - n))
        r = list(genome[s:s + n])
        if rng.random() < 0.1:
            r[int(rng.integers(0, n))] = 'N'
        reads.append(''.join(r))
    enc = [r.encode() for r in reads]
    off = np.zeros(len(enc) + 1, np.int64)
    off[1:] = np.cumsum([len(e) for e in enc])
    seq = np.frombuffer(b''.join(enc), np.uint8).copy()
    h = L.ref_new(2, 1, 21, 3000, 3)
    exp = np.zeros(len(reads), np.uint8)
    L.ref_diginorm_reads(h, 3, seq.ctypes.data_as(C.c_char_p), off.ctypes.data_as(C.POINTER(C.c_int64)),
                         len(reads), exp.ctypes.data_as(C.POINTER(C.c_uint8)))
    L.ref_free(h)
    o = oracle.Storage.build(oracle.BYTE, 3000, 3)
    np.testing.assert_array_equal(o.diginorm_reads(seq, off, 21, 3), exp.astype(bool))
