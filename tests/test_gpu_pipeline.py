"""The batched pipeline (captured hipGraph replays over double-buffered
batches) gives exactly the oracle's sequential results (per-read n_new of
every batch, tables or set contents, counters), the plain path's too, and so
does every tile shape; the uneven-read configuration at 1/100 scale through
the pipeline equals the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from goetia_amd import libgoetia as lg  # noqa: E402
from goetia_amd import workload  # noqa: E402
from goetia_amd.pipeline import Pipeline, autotune, tune_get, tune_set  # noqa: E402
from goetia_amd.reads import ReadBatch  # noqa: E402


def uneven_batch(seed, n_long, n_short, genome_len=200_000, sub=0.05):
    g = workload.genome(seed, genome_len)
    lens = workload.lognormal_lengths(seed, n_long, 3000, 1.0, 150, 30_000)
    a = workload.reads(seed + 1, g, n_long, 0, 0, sub, 0.0, lengths=lens)
    b = workload.reads(seed + 2, g, n_short, 150, 150, sub, 0.0)
    return a, b


def split(batch, n_parts):
    """Consecutive read ranges of a ReadBatch."""
    edges = np.linspace(0, batch.n_reads, n_parts + 1).astype(int)
    out = []
    for a, b in zip(edges[:-1], edges[1:]):
        o = batch.offsets[a:b + 1] - batch.offsets[a]
        out.append(ReadBatch(batch.seq[batch.offsets[a]:batch.offsets[b]], o))
    return out


def oracle_storage(oracle, kind):
    if kind == 'bit':
        return oracle.Storage(oracle.BIT, oracle.primes_near_x(4, 2_000_003))
    if kind == 'byte':
        return oracle.Storage(oracle.BYTE, oracle.primes_near_x(4, 1_000_003))
    return oracle.Storage(oracle.EXACT)


def check_against_oracle(o, graph, kind):
    assert graph.n_unique() == o.n_unique
    assert graph.n_occupied() == o.n_occupied
    if kind == 'exact':
        np.testing.assert_array_equal(graph.S.keys(), o.exact_keys())
    else:
        for x, y in zip(graph.get_raw(), o.tables()):
            np.testing.assert_array_equal(x, y)


def make_graph(kind, K=25, cap=0):
    if kind == 'bit':
        st = lg.BitStorage.build(2_000_003, 4)
    elif kind == 'byte':
        st = lg.ByteStorage.build(1_000_003, 4)
    else:
        st = lg.SparseppSetStorage.build(capacity_hint=cap)
    return lg.dBG[type(st), lg.CanLemireShifter].build(st, K)


@pytest.mark.parametrize('kind', ['bit', 'byte', 'exact'])
def test_pipeline_matches_plain_path(oracle, kind):
    a, b = uneven_batch(31, 60, 3000)
    parts = split(a, 3) + split(b, 4)
    plain = make_graph(kind)
    piped = make_graph(kind)
    exp = [plain.insert_sequences(p) for p in parts]
    max_reads = max(p.n_reads for p in parts)
    max_bases = max(p.n_bases for p in parts)
    with Pipeline(piped, max_reads, max_bases) as pipe:
        outs = [pipe.submit(p) for p in parts]
    o = oracle_storage(oracle, kind)
    for e, out, p in zip(exp, outs, parts):
        _, o_nn = o.insert_reads(p.seq, p.offsets, 25, True, oracle.MODE_RAW)
        np.testing.assert_array_equal(out[:p.n_reads], o_nn)
        np.testing.assert_array_equal(out[:p.n_reads], e)
    check_against_oracle(o, piped, kind)
    assert piped.n_unique() == plain.n_unique()
    if kind == 'exact':
        np.testing.assert_array_equal(piped.S.keys(), plain.S.keys())
    else:
        for x, y in zip(piped.get_raw(), plain.get_raw()):
            np.testing.assert_array_equal(x, y)
        assert piped.n_occupied() == plain.n_occupied()


def test_pipeline_uneven_reads_vs_oracle(oracle):
    """The uneven-read configuration (lognormal long reads, median 8 kb,
    clipped to 150 bp-100 kb, among 150 bp reads, 5% substitutions, K=25,
    exact set) at 1/100 scale, in 8 batches through the pipeline: every
    batch's n_new, the set contents and counters equal the oracle's
    sequential insert."""
    cfg, batch = workload.exact_mixed(0.01)
    K = cfg['K']
    parts = split(batch, 8)
    g = make_graph('exact', K=K)
    with Pipeline(g, max(p.n_reads for p in parts), max(p.n_bases for p in parts)) as pipe:
        outs = [pipe.submit(p) for p in parts]
    o = oracle.Storage(oracle.EXACT)
    for out, p in zip(outs, parts):
        _, o_nn = o.insert_reads(p.seq, p.offsets, K, True, oracle.MODE_RAW)
        np.testing.assert_array_equal(out[:p.n_reads], o_nn)
    check_against_oracle(o, g, 'exact')


def test_pipeline_exact_set_grows_between_batches():
    a, b = uneven_batch(41, 30, 2000)
    parts = split(a, 4) + split(b, 3)
    plain = make_graph('exact')
    piped = make_graph('exact', cap=0)   # starts at 1024 slots, must grow
    exp = [plain.insert_sequences(p) for p in parts]
    pipe = Pipeline(piped, max(p.n_reads for p in parts), max(p.n_bases for p in parts))
    outs = [pipe.submit(p) for p in parts]
    pipe.flush()
    pipe.close()
    for e, o, p in zip(exp, outs, parts):
        np.testing.assert_array_equal(o[:p.n_reads], e)
    np.testing.assert_array_equal(piped.S.keys(), plain.S.keys())


@pytest.mark.parametrize('per', [8, 16, 32])
@pytest.mark.parametrize('bpc', [4, 16])
def test_tile_shapes_give_identical_results(per, bpc):
    a, b = uneven_batch(51, 20, 1500)
    saved = {k: tune_get(k) for k in (1, 2, 3)}
    try:
        results = {}
        for kind, code in (('bit', 1), ('byte', 2), ('exact', 3)):
            tune_set(code, 16, 8)
            ref = make_graph(kind)
            r1 = ref.insert_sequences(a), ref.insert_sequences(b), ref.query_sequences(b)
            tune_set(code, per, bpc)
            g = make_graph(kind)
            r2 = g.insert_sequences(a), g.insert_sequences(b), g.query_sequences(b)
            for x, y in zip(r1, r2):
                np.testing.assert_array_equal(x, y)
            results[kind] = True
    finally:
        for k, (p, c) in saved.items():
            tune_set(k, p, c)
    assert len(results) == 3


def test_autotune_keeps_results():
    a, _ = uneven_batch(61, 40, 10)
    saved = tune_get(2)
    try:
        g = make_graph('byte')
        g.insert_sequences(a)
        before = g.get_raw()
        per, bpc = autotune(g, a.to_device())
        assert per in (8, 16, 32) and bpc in (4, 8, 16)
        assert tune_get(2) == (per, bpc)
        for x, y in zip(before, g.get_raw()):
            np.testing.assert_array_equal(x, y)
        h = make_graph('byte')
        h.insert_sequences(a)
        for x, y in zip(before, h.get_raw()):
            np.testing.assert_array_equal(x, y)
    finally:
        tune_set(2, *saved)
