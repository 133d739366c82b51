"""The CPU model of the binned insert (tests/binned_model.py) against the
sequential oracle: tables, per-read n_new, counters and count-min
insert_and_query, with every buffer index of the kernels' plan asserted."""
import numpy as np
import pytest

from binned_model import binned_insert
from goetia_amd import workload
from oracle import pyoracle as po


def _batch(seed, n_reads):
    g = workload.genome(seed, 20000)
    return workload.reads(seed + 1, g, n_reads, 60, 200, 0.02, 0.01)


@pytest.mark.parametrize('kind,sizes,n_reads,sub,mode', [
    (po.BYTE, po.primes_near_x(4, 40000), 90, 8192, po.MODE_SKIP_N),          # 3 sub-batches
    (po.BIT, po.primes_near_x(3, 30000), 70, 16384, po.MODE_SKIP_N),
    (po.BYTE, po.primes_near_x(4, 2_200_000), 40, 8192, po.MODE_SKIP_N),      # two-level routing
    (po.BYTE, [997, 1009], 60, 8192, po.MODE_RAW),                            # saturated tiny tables
])
def test_binned_model_insert_vs_oracle(kind, sizes, n_reads, sub, mode):
    batch = _batch(5, n_reads)
    tabs, n_new, nu, occ, _ = binned_insert(kind, sizes, batch.seq, batch.offsets, 21, mode, sub)
    o = po.Storage(kind, sizes)
    _, o_new = o.insert_reads(batch.seq, batch.offsets, 21, True, mode)
    for a, b in zip(tabs, o.tables()):
        np.testing.assert_array_equal(a[:len(b)], b)
    np.testing.assert_array_equal(n_new, o_new)
    assert (nu, occ) == (o.n_unique, o.n_occupied)


def test_binned_model_insert_and_query_vs_oracle():
    batch = _batch(9, 60)
    sizes = po.primes_near_x(3, 5000)
    tabs, _, nu, occ, counts = binned_insert(po.BYTE, sizes, batch.seq, batch.offsets, 21, po.MODE_SKIP_N, 8192,
                                             iq=True)
    o = po.Storage(po.BYTE, sizes)
    exp = np.concatenate([o.insert_and_query_sequence(batch.read(r), 21, True, po.MODE_SKIP_N)
                          for r in range(batch.n_reads)])
    np.testing.assert_array_equal(counts, exp)
    assert (nu, occ) == (o.n_unique, o.n_occupied)
    for a, b in zip(tabs, o.tables()):
        np.testing.assert_array_equal(a[:len(b)], b)
