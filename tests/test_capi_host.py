"""CPU checks of the C ABI library: it loads, exports every entry point of
include/goetia_amd.h, its host-only functions agree with the oracle, and its
device entry points fail loudly (no fallback) when no GPU is visible."""
import numpy as np
import pytest

from goetia_amd import _native as N

try:
    import torch
    HAS_GPU = torch.cuda.is_available()
except Exception:  # pragma: no cover
    HAS_GPU = False


def test_library_exports_every_header_symbol():
    lib = N.load()
    names = N.header_symbols()
    assert len(names) >= 40
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(N.SIGNATURES) | {'gt_last_error', 'gt_server_last_error'}
    assert lib.gt_abi_version() == 1


@pytest.mark.parametrize('n,x', [(4, 10 ** 6), (4, 10 ** 7), (4, 2 ** 28), (4, 2 ** 30), (3, 10), (2, 1)])
def test_primes_match_oracle(oracle, n, x):
    from goetia_amd.storage import get_n_primes_near_x
    assert get_n_primes_near_x(n, x) == oracle.primes_near_x(n, x)


def test_shifter_host_state_matches_oracle(oracle):
    from goetia_amd.hashing import CanLemireShifter, FwdLemireShifter
    rng = np.random.default_rng(1)
    for K in (1, 21, 27, 31, 63, 64, 65, 101):
        s = ''.join(rng.choice(list('ACGT'), 200))
        fw, rc = oracle.hash_read(s, K)
        sh = CanLemireShifter(K)
        got = [sh.hash_base(s[:K])]
        for i in range(1, len(s) - K + 1):
            got.append(sh.shift_right(s[i - 1], s[i + K - 1]))
        assert [g.fw_hash for g in got] == [int(v) for v in fw]
        assert [g.rc_hash for g in got] == [int(v) for v in rc]
        # right-left-right round trip (test_hashing.py:303-329)
        back = [sh.hash_base(s[-K:])]
        for i in range(len(s) - K - 1, -1, -1):
            back.append(sh.shift_left(s[i], s[i + K]))
        assert [b.value for b in back[::-1]] == [g.value for g in got]
        f = FwdLemireShifter(K)
        assert f.hash(s[:K]).value == int(fw[0])


def test_shifter_errors():
    from goetia_amd.hashing import CanLemireShifter
    sh = CanLemireShifter(21)
    with pytest.raises(N.UninitializedShifterException):
        sh.shift_right('A', 'C')
    with pytest.raises(N.InvalidSequenceException):
        sh.hash_base('ACGT')
    with pytest.raises(N.InvalidSequenceException):
        CanLemireShifter.hash('AAAAA', 7)
    assert CanLemireShifter.hash('TCACCTGTGTTGTGCTACTTGCGGCGC', 27).value == 4324216031038051805


def test_generator_is_deterministic_and_shaped():
    from goetia_amd import workload
    g1 = workload.genome(5, 10000)
    g2 = workload.genome(5, 10000)
    assert g1.tobytes() == g2.tobytes()
    assert set(g1.tobytes()) <= set(b'ACGT')
    b1 = workload.reads(7, g1, 500, 100, 150, 0.01, 0.01)
    b2 = workload.reads(7, g1, 500, 100, 150, 0.01, 0.01)
    assert b1.seq.tobytes() == b2.seq.tobytes()
    np.testing.assert_array_equal(b1.offsets, b2.offsets)
    lens = b1.lengths()
    assert lens.min() >= 100 and lens.max() <= 150
    frac_n = (b1.seq == ord('N')).mean()
    assert 0.002 < frac_n < 0.03


def test_generator_error_rate_and_strands():
    from goetia_amd import workload
    g = workload.genome(3, 200000)
    b = workload.reads(4, g, 2000, 150, 150, 0.05, 0.0)
    gs = g.tobytes().decode()
    rc = gs[::-1].translate(str.maketrans('ACGT', 'TGCA'))
    mism, total, rev = 0, 0, 0
    for r in range(200):
        s = b.read(r).decode()
        # locate by the best of both strands over a 20-mer seed search
        best = None
        for strand, ref in ((0, gs), (1, rc)):
            for j in range(0, 150 - 20, 10):
                p = ref.find(s[j:j + 20])
                if p >= 0 and p - j >= 0:
                    d = sum(a != c for a, c in zip(s, ref[p - j:p - j + 150]))
                    if best is None or d < best[0]:
                        best = (d, strand)
                    break
        if best is not None:
            mism += best[0]
            total += 150
            rev += best[1]
    assert 0.03 < mism / total < 0.07
    assert 60 < rev < 140


@pytest.mark.skipif(HAS_GPU, reason='checks the no-GPU behaviour')
def test_device_calls_fail_loudly_without_gpu():
    from goetia_amd import libgoetia
    with pytest.raises(N.DeviceError):
        libgoetia.BitStorage.build()
    from goetia_amd.hashing import hash_batch
    from goetia_amd.reads import ReadBatch
    with pytest.raises(N.DeviceError):
        hash_batch(ReadBatch.from_strings(['ACGTACGTACGTACGTACGTACGT']), 21)


def test_read_batch_validation():
    from goetia_amd.reads import ReadBatch
    b = ReadBatch.from_strings(['ACGT', '', 'ACGTACGTAC'])
    assert b.n_reads == 3 and b.n_bases == 14
    np.testing.assert_array_equal(b.kmer_offsets(4), [0, 1, 1, 8])
    with pytest.raises(ValueError):
        ReadBatch(np.zeros(4, np.uint8), np.array([1, 4], np.int64))
    with pytest.raises(ValueError):
        ReadBatch(np.zeros(4, np.uint8), np.array([0, 3, 2], np.int64))
