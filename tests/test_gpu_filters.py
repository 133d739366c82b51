"""StreamingSolidFilter on the GPU against goetia's compiled filter
(tests/golden/solid_ref.npz, written by tests/golden/make_solid_golden.py):
every read's pass decision (solidifier.hh:58-73: insert_and_query in read
order, then goetia's float ratio of non-solid k-mers against
1.0 - min_prop_solid; reads shorter than K never pass), and the tables and
counters goetia's graph was left with -- batched in two calls, and one read
at a time."""
import hashlib
import os

import numpy as np
import pytest

from conftest import GOLDEN_DIR

GOLD = np.load(os.path.join(GOLDEN_DIR, 'solid_ref.npz'), allow_pickle=False)
CASES = sorted(k[:-len('__params')] for k in GOLD.files if k.endswith('__params'))


def golden_inputs():
    from goetia_amd import workload
    genome = workload.genome(81, 30_000)
    b = workload.reads(82, genome, 2500, 10, 200, 0.02, 0.0)
    h = hashlib.sha256(np.ascontiguousarray(b.seq, np.uint8).tobytes())
    h.update(np.ascontiguousarray(b.offsets, '<i8').tobytes())
    return b, h.digest()


def test_golden_inputs_regenerate():
    _, d = golden_inputs()
    assert d == GOLD['input_sha256'].tobytes()


@pytest.mark.gpu
@pytest.mark.parametrize('name', CASES)
def test_solid_filter_vs_reference(name):
    from goetia_amd import libgoetia as lg
    kind, can, K, mt, nt, thr = (int(v) for v in GOLD[f'{name}__params'])
    prop = float(GOLD[f'{name}__min_prop'][0])
    exp = GOLD[f'{name}__passed'].astype(bool)
    batch, _ = golden_inputs()
    reads = [batch.read(i) for i in range(batch.n_reads)]
    cls = lg.ByteStorage if kind == 2 else lg.BitStorage
    sh = lg.CanLemireShifter if can else lg.FwdLemireShifter
    g = lg.dBG[cls, sh].build(cls.build(mt, nt), K)
    f = lg.StreamingSolidFilter.build(g, prop, thr)
    got = np.concatenate([f.filter_sequences(reads[:1200]), f.filter_sequences(reads[1200:])])
    np.testing.assert_array_equal(got, exp)
    assert [g.n_unique(), g.n_occupied()] == [int(v) for v in GOLD[f'{name}__counters']]
    for i, t in enumerate(g.get_raw()):
        assert hashlib.sha256(t.tobytes()).digest() == GOLD[f'{name}__table_sha'][i].tobytes(), (name, i)
    # one read at a time gives the same decisions
    g2 = lg.dBG[cls, sh].build(cls.build(mt, nt), K)
    f2 = lg.StreamingSolidFilter.build(g2, prop, thr)
    one = [f2.filter_sequence(r)[0] if len(r) >= K else False for r in reads[:300]]
    np.testing.assert_array_equal(one, exp[:300])
